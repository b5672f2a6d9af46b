// npy.cpp -- see npy.h.
#include "npy.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>

static std::string field(const std::string& hdr, const std::string& key)
{
    size_t k = hdr.find("'" + key + "'");
    if (k == std::string::npos) throw std::runtime_error("npy: header lacks " + key);
    size_t c = hdr.find(':', k);
    size_t s = hdr.find_first_not_of(" ", c + 1);
    if (hdr[s] == '\'') {
        size_t e = hdr.find('\'', s + 1);
        return hdr.substr(s + 1, e - s - 1);
    }
    if (hdr[s] == '(') {
        size_t e = hdr.find(')', s);
        return hdr.substr(s + 1, e - s - 1);
    }
    size_t e = hdr.find_first_of(",}", s);
    return hdr.substr(s, e - s);
}

NpyArray npy_load(const std::string& path)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("npy: cannot open " + path);
    unsigned char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
        std::fclose(f);
        throw std::runtime_error("npy: not a .npy file: " + path);
    }
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        if (std::fread(b, 1, 2, f) != 2) { std::fclose(f); throw std::runtime_error("npy: short header"); }
        hlen = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        if (std::fread(b, 1, 4, f) != 4) { std::fclose(f); throw std::runtime_error("npy: short header"); }
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string hdr(hlen, '\0');
    if (std::fread(&hdr[0], 1, hlen, f) != hlen) { std::fclose(f); throw std::runtime_error("npy: short header"); }
    NpyArray a;
    a.dtype = field(hdr, "descr");
    if (field(hdr, "fortran_order").find("True") != std::string::npos) {
        std::fclose(f);
        throw std::runtime_error("npy: Fortran order not supported: " + path);
    }
    std::string sh = field(hdr, "shape");
    size_t p = 0;
    while (p < sh.size()) {
        size_t q = sh.find(',', p);
        std::string tok = sh.substr(p, q == std::string::npos ? std::string::npos : q - p);
        if (tok.find_first_of("0123456789") != std::string::npos) a.shape.push_back(std::stoll(tok));
        if (q == std::string::npos) break;
        p = q + 1;
    }
    size_t esz = 0;
    if (a.dtype == "<f8" || a.dtype == "<i8") esz = 8;
    else if (a.dtype == "<f4") esz = 4;
    else { std::fclose(f); throw std::runtime_error("npy: unsupported dtype " + a.dtype); }
    size_t bytes = (size_t)a.count() * esz;
    a.data.resize(bytes);
    if (bytes && std::fread(a.data.data(), 1, bytes, f) != bytes) {
        std::fclose(f);
        throw std::runtime_error("npy: truncated data in " + path);
    }
    std::fclose(f);
    return a;
}

void npy_save(const std::string& path, const std::string& dtype, const std::vector<int64_t>& shape,
              const void* data)
{
    std::string sh = "(";
    for (size_t i = 0; i < shape.size(); ++i) sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    sh += ")";
    std::string hdr = "{'descr': '" + dtype + "', 'fortran_order': False, 'shape': " + sh + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ') + "\n";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("npy: cannot write " + path);
    unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, 0, 0};
    pre[8] = (unsigned char)(hdr.size() & 0xff);
    pre[9] = (unsigned char)(hdr.size() >> 8);
    std::fwrite(pre, 1, 10, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    int64_t c = 1;
    for (int64_t s : shape) c *= s;
    size_t esz = dtype == "<f4" ? 4 : 8;
    if (c) std::fwrite(data, esz, (size_t)c, f);
    std::fclose(f);
}

std::vector<double> npy_as_f64(const NpyArray& a)
{
    std::vector<double> v((size_t)a.count());
    if (a.dtype == "<f8") std::memcpy(v.data(), a.data.data(), v.size() * 8);
    else if (a.dtype == "<f4") {
        const float* s = reinterpret_cast<const float*>(a.data.data());
        for (size_t i = 0; i < v.size(); ++i) v[i] = s[i];
    } else {
        const int64_t* s = reinterpret_cast<const int64_t*>(a.data.data());
        for (size_t i = 0; i < v.size(); ++i) v[i] = (double)s[i];
    }
    return v;
}

std::vector<int64_t> npy_as_i64(const NpyArray& a)
{
    std::vector<int64_t> v((size_t)a.count());
    if (a.dtype == "<i8") std::memcpy(v.data(), a.data.data(), v.size() * 8);
    else {
        std::vector<double> d = npy_as_f64(a);
        for (size_t i = 0; i < v.size(); ++i) v[i] = (int64_t)d[i];
    }
    return v;
}
