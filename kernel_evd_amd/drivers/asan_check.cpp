// asan_check.cpp -- exercises the drivers' host code (KS generation, delay
// embedding, .npy and HDF5 I/O, point loading) in a build with the address
// and undefined-behaviour sanitizers.  No GPU call is made.
//   asan_check <scratch directory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "arrays.h"
#include "kevd.h"
#include "npy.h"

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                    \
        }                                                               \
    } while (0)

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: asan_check <scratch directory>\n");
        return 2;
    }
    const std::string dir = argv[1];
    // KS snapshots and the delay windows
    kevd_ks_params kp{22.0, 32, 0.25, 200, 4, 7, 0.1};
    const int64_t n = 300;
    std::vector<double> X((size_t)n * kp.M);
    CHECK(kevd_ks_generate(&kp, n, X.data()) == KEVD_OK);
    bool finite = true;
    for (double v : X) finite = finite && std::isfinite(v);
    CHECK(finite);
    const int nq = 8;
    std::vector<double> D((size_t)(n - nq + 1) * kp.M * nq);
    CHECK(kevd_delay_embed(X.data(), n, kp.M, nq, D.data()) == KEVD_OK);
    CHECK(D[(size_t)(5 * (n - nq + 1) + 17) * nq + 3] == X[(size_t)(17 + 3) * kp.M + 5]);
    CHECK(kevd_ks_generate(&kp, -1, X.data()) != KEVD_OK);
    // .npy round trips in the three dtypes
    npy_save(dir + "/x.npy", "<f8", {n, (int64_t)kp.M}, X.data());
    NpyArray a = npy_load(dir + "/x.npy");
    CHECK(a.dtype == "<f8" && a.rows() == n && a.cols() == kp.M);
    CHECK(std::memcmp(a.data.data(), X.data(), X.size() * sizeof(double)) == 0);
    std::vector<float> xf(X.begin(), X.end());
    npy_save(dir + "/xf.npy", "<f4", {n, (int64_t)kp.M}, xf.data());
    std::vector<double> back = npy_as_f64(npy_load(dir + "/xf.npy"));
    CHECK(back.size() == xf.size() && back[123] == (double)xf[123]);
    std::vector<int64_t> piv = {3, 1, 4, 1, 5, 9, 2, 6};
    npy_save(dir + "/p.npy", "<i8", {(int64_t)piv.size()}, piv.data());
    CHECK(npy_as_i64(npy_load(dir + "/p.npy")) == piv);
    // points through the drivers' loader (delay windows of the snapshots)
    int64_t np_ = 0;
    int32_t dim = 0;
    bool f32 = false;
    std::vector<double> P = load_points(dir + "/x.npy", nq, 0, &np_, &dim, &f32);
    CHECK(dim == nq && !f32 && (int64_t)P.size() == np_ * dim);
    // a truncated file is rejected with an exception, not a crash
    {
        FILE* f = std::fopen((dir + "/bad.npy").c_str(), "wb");
        std::fwrite("\x93NUMPY\x01\x00\x10\x00{'descr'", 1, 18, f);
        std::fclose(f);
        bool threw = false;
        try {
            npy_load(dir + "/bad.npy");
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
    }
    if (have_hdf5()) {
        save_array(dir + "/f.h5", "<f8", {n, (int64_t)kp.M}, X.data(), "/u");
        save_array(dir + "/f.h5", "<u8", {(int64_t)piv.size()}, piv.data(), "/spiv", true);
        NpyArray h = load_array(dir + "/f.h5", "/u");
        CHECK(h.count() == n * kp.M);
        CHECK(npy_as_i64(load_array(dir + "/f.h5:/spiv", "/u")) == piv);
    }
    std::printf("asan_check: %s (%d failed checks)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
