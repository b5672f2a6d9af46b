// npy.h -- minimal .npy (v1/v2/v3) reader and v1 writer for the drivers:
// little-endian f8 / f4 / i8 arrays in C order, 1-D or 2-D.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct NpyArray {
    std::string dtype;            // "<f8", "<f4" or "<i8"
    std::vector<int64_t> shape;
    std::vector<char> data;
    int64_t count() const
    {
        int64_t c = 1;
        for (int64_t s : shape) c *= s;
        return c;
    }
    int64_t rows() const { return shape.empty() ? 0 : shape[0]; }
    int64_t cols() const { return shape.size() > 1 ? shape[1] : 1; }
};

// Throws std::runtime_error on malformed input.
NpyArray npy_load(const std::string& path);
void npy_save(const std::string& path, const std::string& dtype, const std::vector<int64_t>& shape,
              const void* data);

// Convenience: load a 2-D float array as FP64 or keep FP32.
std::vector<double> npy_as_f64(const NpyArray& a);
std::vector<int64_t> npy_as_i64(const NpyArray& a);
