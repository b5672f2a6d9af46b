// kevd-bandwidth -- the ks_bandwidth.cu stage: kernel sums over a grid of
// bandwidths from one fused launch, then the reference's post-processing
// (ks_bandwidth.cu:609-658) into a bw0.out-style text file:
//   "%.5e %.5e %.5e" (bandwidth, log of the mean row sum, slope) for
//   i = 1..n-2, a blank line, eopt, mdim.
//
//   kevd-bandwidth --points X.npy|u.h5 [--delay Nq] [--grid-points M]
//                  [--rows 16384 [--row-seed 0] | --all-pairs]
//                  [--grid log [--lo 1e-3] [--hi 1e3] [--count 64] | --grid ref36]
//                  [--scale median|<value>] [--precision mixed|fp64]
//                  [--out bw0.out] [--sums-out S.npy] [--device 0]
// --grid log: eps_j log-spaced over [lo, hi] x scale (scale = the median
// squared distance by default).  --grid ref36: the reference's 36 values
// bw = {1..9} x 10^{-3..0} with eps = dim * bw (its kernel is
// exp(-sum dx^2 / (Nq bw))); the file then lists bw.  --all-pairs sums
// every row: the reference's strict lower triangle is (S - N) / 2.
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "args.h"
#include "kevd.h"
#include "arrays.h"

static uint64_t philox_u64(uint64_t seed, uint32_t stream, uint32_t round, uint32_t idx)
{
    uint32_t c0 = idx, c1 = round, c2 = stream, c3 = 0, k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ((uint64_t)c1 << 32) | c0;
}

static int fail(kevd_ctx* c, kevd_status s, const char* what)
{
    std::fprintf(stderr, "%s: %s (%s)\n", what, kevd_status_string(s), c ? kevd_last_error(c) : "");
    if (c) kevd_destroy(c);
    return exit_code(s);
}

int main(int argc, char** argv)
{
    Args a(argc, argv);
    if (a.has("help")) {
        std::printf("kevd-bandwidth --points X.npy|u.h5 [--delay Nq] [--grid-points M]\n"
                    "               [--rows 16384 [--row-seed 0] | --all-pairs]\n"
                    "               [--grid log [--lo 1e-3] [--hi 1e3] [--count 64] | --grid ref36]\n"
                    "               [--scale median|<value>] [--precision mixed|fp64] [--out bw0.out]\n");
        return 0;
    }
    kevd_ctx* c = nullptr;
    kevd_status s = kevd_create((int)a.integer("device", 0), &c);
    if (s != KEVD_OK) return fail(nullptr, s, "kevd_create");
    int64_t n = 0;
    int32_t dim = 0;
    s = (kevd_status)set_points_from(c, a.need("points"), a.integer("delay", 0), a.integer("grid-points", 0), &n,
                                     &dim);
    if (s != KEVD_OK) return fail(c, s, "kevd_set_points");
    // rows: all, or a Philox-keyed partial Fisher-Yates sample without replacement
    std::vector<int64_t> rows;
    if (a.has("all-pairs")) {
        rows.resize((size_t)n);
        std::iota(rows.begin(), rows.end(), 0);
    } else {
        int64_t nr = a.integer("rows", 16384);
        if (nr > n) nr = n;
        std::vector<int64_t> idx((size_t)n);
        std::iota(idx.begin(), idx.end(), 0);
        uint64_t seed = (uint64_t)a.integer("row-seed", 0);
        for (int64_t t = 0; t < nr; ++t) {
            uint64_t u = philox_u64(seed, 4u, 0u, (uint32_t)t);
            int64_t j = t + (int64_t)(((unsigned __int128)u * (uint64_t)(n - t)) >> 64);
            std::swap(idx[t], idx[j]);
        }
        rows.assign(idx.begin(), idx.begin() + nr);
    }
    std::vector<double> eps, label;
    std::string grid = a.str("grid", "log");
    if (grid == "ref36") {
        for (int e = -3; e <= 0; ++e)
            for (int m = 1; m <= 9; ++m) {
                double bw = m * std::pow(10.0, e);
                label.push_back(bw);
                eps.push_back(dim * bw);
            }
    } else {
        double scale = 0.0;
        std::string sc = a.str("scale", "median");
        if (sc == "median") {
            s = kevd_median_sqdist(c, 0, n <= 8192 ? 0 : ((int64_t)1 << 22), &scale);
            if (s != KEVD_OK) return fail(c, s, "kevd_median_sqdist");
        } else {
            scale = std::atof(sc.c_str());
        }
        int cnt = (int)a.integer("count", 64);
        double lo = std::log10(a.num("lo", 1e-3)), hi = std::log10(a.num("hi", 1e3));
        for (int j = 0; j < cnt; ++j) {
            double e = scale * std::pow(10.0, lo + (hi - lo) * j / (cnt > 1 ? cnt - 1 : 1));
            eps.push_back(e);
            label.push_back(e);
        }
        std::printf("scale (median squared distance) %.6g\n", scale);
    }
    kevd_precision prec = a.str("precision", "mixed") == "fp64" ? KEVD_FP64 : KEVD_MIXED;
    std::vector<double> S(eps.size());
    s = kevd_bandwidth_sweep(c, rows.data(), (int64_t)rows.size(), eps.data(), (int32_t)eps.size(), prec, S.data());
    if (s != KEVD_OK) return fail(c, s, "kevd_bandwidth_sweep");
    kevd_destroy(c);
    int ne = (int)eps.size();
    std::vector<double> logS((size_t)ne), slope((size_t)(ne > 2 ? ne - 2 : 1));
    double eopt = 0, mdim = 0;
    if (kevd_bandwidth_post(eps.data(), S.data(), ne, logS.data(), slope.data(), &eopt, &mdim) != KEVD_OK) {
        std::fprintf(stderr, "post-processing needs >= 3 bandwidths with positive sums\n");
        return 1;
    }
    double eopt_label = label[0];
    for (int i = 0; i < ne; ++i)
        if (eps[i] == eopt) eopt_label = label[i];
    std::string out = a.str("out", "bw0.out");
    FILE* f = std::fopen(out.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        return 1;
    }
    // Column 2 is the log of the mean row sum, log(S / rows): with every row
    // (--all-pairs) that is the reference's log(2 S_lower / NM + 1)
    // (ks_bandwidth.cu:615-618, whose summands carry a 1/NM), the full sum
    // being 2 S_lower + NM.  The constant does not change the slope.
    const double log_rows = std::log((double)rows.size());
    for (int i = 1; i < ne - 1; ++i)
        std::fprintf(f, "%.5e %.5e %.5e\n", label[i], logS[i] - log_rows, slope[i - 1]);
    std::fprintf(f, "\n%.5e\n%.5e\n", eopt_label, mdim);
    std::fclose(f);
    if (a.has("sums-out")) npy_save(a.str("sums-out"), "<f8", {(int64_t)ne}, S.data());
    std::printf("rows %zu  bandwidths %d  eopt %.5e  mdim %.5f -> %s\n", rows.size(), ne, eopt_label, mdim,
                out.c_str());
    return 0;
}
