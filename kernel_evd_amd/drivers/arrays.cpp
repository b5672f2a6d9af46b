// arrays.cpp -- .npy / HDF5 dispatch and the reference-layout point loader
// of the drivers (see arrays.h).
#include "arrays.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <sys/stat.h>

#include "kevd.h"

#ifdef KEVD_HAVE_HDF5
#include <hdf5.h>
#endif

namespace {

// "file.h5:/name" -> ("file.h5", "/name"); no suffix -> (spec, dflt)
void split_spec(const std::string& spec, const std::string& dflt, std::string* path, std::string* dset)
{
    for (const char* ext : {".h5:", ".hdf5:"}) {
        size_t p = spec.find(ext);
        if (p != std::string::npos) {
            size_t c = p + std::strlen(ext) - 1;
            *path = spec.substr(0, c);
            *dset = spec.substr(c + 1);
            if (dset->empty() || (*dset)[0] != '/') *dset = "/" + *dset;
            return;
        }
    }
    *path = spec;
    *dset = dflt;
}

bool ends_with(const std::string& s, const std::string& e)
{
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

#ifdef KEVD_HAVE_HDF5
struct H5Quiet {
    H5Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

void h5_check(bool ok, const std::string& what)
{
    if (!ok) throw std::runtime_error("HDF5: " + what);
}

NpyArray h5_load(const std::string& path, const std::string& dset)
{
    static H5Quiet quiet;
    hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    h5_check(f >= 0, "cannot open " + path);
    hid_t d = H5Dopen2(f, dset.c_str(), H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        throw std::runtime_error("HDF5: no dataset " + dset + " in " + path);
    }
    hid_t sp = H5Dget_space(d);
    int nd = H5Sget_simple_extent_ndims(sp);
    std::vector<hsize_t> dims(nd > 0 ? nd : 1, 1);
    if (nd > 0) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
    hid_t ty = H5Dget_type(d);
    H5T_class_t cls = H5Tget_class(ty);
    size_t sz = H5Tget_size(ty);
    NpyArray a;
    for (int i = 0; i < nd; ++i) a.shape.push_back((int64_t)dims[i]);
    if (nd == 0) a.shape.push_back(1);
    hid_t mem;
    if (cls == H5T_FLOAT) {
        a.dtype = sz == 4 ? "<f4" : "<f8";
        mem = sz == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    } else if (cls == H5T_INTEGER) {
        a.dtype = "<i8";
        mem = H5T_NATIVE_INT64;
    } else {
        H5Tclose(ty);
        H5Sclose(sp);
        H5Dclose(d);
        H5Fclose(f);
        throw std::runtime_error("HDF5: " + dset + " in " + path + " is neither float nor integer");
    }
    a.data.resize((size_t)a.count() * (a.dtype == "<f4" ? 4 : 8));
    herr_t st = H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.data.data());
    H5Tclose(ty);
    H5Sclose(sp);
    H5Dclose(d);
    H5Fclose(f);
    h5_check(st >= 0, "cannot read " + dset + " from " + path);
    return a;
}

void h5_save(const std::string& path, const std::string& dset, const std::string& dtype, int64_t count,
             const void* data, bool append)
{
    static H5Quiet quiet;
    struct stat sb;
    hid_t f = append && stat(path.c_str(), &sb) == 0 ? H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT)
                                                      : H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT,
                                                                  H5P_DEFAULT);
    h5_check(f >= 0, "cannot create or open " + path);
    if (H5Lexists(f, dset.c_str(), H5P_DEFAULT) > 0) H5Ldelete(f, dset.c_str(), H5P_DEFAULT);
    hsize_t dims[1] = {(hsize_t)count};
    hid_t sp = H5Screate_simple(1, dims, nullptr);
    hid_t ty = dtype == "<f4" ? H5T_NATIVE_FLOAT : dtype == "<f8" ? H5T_NATIVE_DOUBLE : H5T_NATIVE_UINT64;
    hid_t d = H5Dcreate2(f, dset.c_str(), ty, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t st = d >= 0 ? H5Dwrite(d, ty, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : -1;
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    H5Fclose(f);
    h5_check(st >= 0, "cannot write " + dset + " to " + path);
}
#endif

}  // namespace

bool is_h5(const std::string& spec)
{
    std::string path, dset;
    split_spec(spec, "", &path, &dset);
    return ends_with(path, ".h5") || ends_with(path, ".hdf5");
}

bool have_hdf5()
{
#ifdef KEVD_HAVE_HDF5
    return true;
#else
    return false;
#endif
}

NpyArray load_array(const std::string& spec, const std::string& dset)
{
    if (!is_h5(spec)) return npy_load(spec);
#ifdef KEVD_HAVE_HDF5
    std::string path, name;
    split_spec(spec, dset, &path, &name);
    return h5_load(path, name);
#else
    (void)dset;
    throw std::runtime_error("this driver was built without HDF5 support: " + spec);
#endif
}

void save_array(const std::string& spec, const std::string& dtype, const std::vector<int64_t>& shape,
                const void* data, const std::string& dset, bool append)
{
    if (!is_h5(spec)) {
        npy_save(spec, dtype == "<u8" ? "<i8" : dtype, shape, data);
        return;
    }
#ifdef KEVD_HAVE_HDF5
    std::string path, name;
    split_spec(spec, dset, &path, &name);
    int64_t count = 1;
    for (int64_t s : shape) count *= s;
    h5_save(path, name, dtype, count, data, append);
#else
    (void)shape, (void)data, (void)dset, (void)append, (void)dtype;
    throw std::runtime_error("this driver was built without HDF5 support: " + spec);
#endif
}

std::vector<double> load_points(const std::string& spec, int64_t delay, int64_t grid_points, int64_t* n,
                                int32_t* dim, bool* f32)
{
    if (!is_h5(spec)) {
        NpyArray X = load_array(spec);
        if (X.shape.size() != 2) throw std::runtime_error("--points must be a 2-D array: " + spec);
        *f32 = X.dtype == "<f4";
        std::vector<double> x = npy_as_f64(X);
        if (delay > 0) {
            // a 2-D array of snapshots (time x grid) delay embedded here
            int64_t nt = X.rows(), M = X.cols();
            if (nt < delay) throw std::runtime_error("--delay exceeds the number of snapshots");
            std::vector<double> out((size_t)(nt - delay + 1) * M * delay);
            if (kevd_delay_embed(x.data(), nt, (int32_t)M, (int32_t)delay, out.data()) != KEVD_OK)
                throw std::runtime_error("kevd_delay_embed failed");
            *n = (nt - delay + 1) * M;
            *dim = (int32_t)delay;
            return out;
        }
        *n = X.rows();
        *dim = (int32_t)X.cols();
        return x;
    }
    std::string path, name;
    split_spec(spec, "/u", &path, &name);
    NpyArray U = load_array(path + ":" + name);
    *f32 = U.dtype == "<f4";
    std::vector<double> u = npy_as_f64(U);
    int64_t total = U.count(), M = grid_points;
    if (M <= 0) {
        NpyArray T = load_array(path + ":/t");
        if (T.count() < 1 || total % T.count() != 0)
            throw std::runtime_error("cannot infer the grid size from /t; give --grid-points");
        M = total / T.count();
    }
    if (M < 1 || total % M != 0) throw std::runtime_error("/u length is not a multiple of the grid size");
    int64_t nt = total / M;
    // series[t][g] = u[g * nt + t] (time fastest in the file)
    std::vector<double> series((size_t)total);
    for (int64_t g = 0; g < M; ++g)
        for (int64_t t = 0; t < nt; ++t) series[(size_t)(t * M + g)] = u[(size_t)(g * nt + t)];
    if (delay > 0) {
        if (nt < delay) throw std::runtime_error("--delay exceeds the number of snapshots");
        std::vector<double> out((size_t)(nt - delay + 1) * M * delay);
        if (kevd_delay_embed(series.data(), nt, (int32_t)M, (int32_t)delay, out.data()) != KEVD_OK)
            throw std::runtime_error("kevd_delay_embed failed");
        *n = (nt - delay + 1) * M;
        *dim = (int32_t)delay;
        return out;
    }
    *n = nt;
    *dim = (int32_t)M;
    return series;
}

int set_points_from(kevd_ctx* ctx, const std::string& spec, int64_t delay, int64_t grid_points, int64_t* n,
                    int32_t* dim)
{
    std::vector<double> x;
    bool f32 = false;
    try {
        x = load_points(spec, delay, grid_points, n, dim, &f32);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return KEVD_EINVAL;
    }
    if (f32) {
        std::vector<float> f(x.begin(), x.end());
        return kevd_set_points_f32(ctx, f.data(), *n, *dim);
    }
    return kevd_set_points(ctx, x.data(), *n, *dim);
}
