// kevd-arpc -- the ks_arpclm.cu stage with its constants as flags: points
// file in; eps, rank, block size and seed as parameters; factor, pivot and
// L^-1 files out; the reference's per-round table on stdout
// (ks_arpclm.cu:1129-1136).
//
//   kevd-arpc --points X.npy|u.h5 [--delay Nq] [--grid-points M]
//             (--eps E | --eps-median [--median-pairs P])
//             --rank R [--block 128] [--seed 0] [--tol 1e-10]
//             [--precision mixed|fp64] [--max-rounds K] [--inject-pivots S.npy|spiv.h5]
//             [--regenerate] [--factor-out F.npy] [--pivots-out S.npy|spiv.h5]
//             [--linv-out Linv.npy|linv.h5] [--device 0]
// --regenerate runs the reference's own LM form (S and L^-1 kept, K(R,S)
// regenerated every round; memory O(n + r^2)): no factor file then.
// File arguments are .npy, or the reference's HDF5 files (ks_arpclm.cu:
// 628-644): the snapshot file's /u (delay embedded with --delay), and /spiv
// (unsigned 64-bit) and /linv (FP64, r x r column-major, flattened) out.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "args.h"
#include "kevd.h"
#include "arrays.h"

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
        std::printf("kevd-arpc --points X.npy|u.h5 [--delay Nq] [--grid-points M] (--eps E | --eps-median)\n"
                    "          --rank R [--block 128] [--seed 0] [--tol 1e-10] [--precision mixed|fp64]\n"
                    "          [--max-rounds K] [--inject-pivots S.npy|spiv.h5] [--regenerate] [--factor-out F.npy]\n"
                    "          [--pivots-out S.npy|spiv.h5] [--linv-out Linv.npy|linv.h5]\n"
                    "          [--median-pairs 4194304] [--device 0]\n");
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    kevd_ctx* c = nullptr;
    kevd_status s = kevd_create((int)a.integer("device", 0), &c);
    if (s != KEVD_OK) return fail(nullptr, s, "kevd_create");
    int64_t npts = 0;
    int32_t dim = 0;
    s = (kevd_status)set_points_from(c, a.need("points"), a.integer("delay", 0), a.integer("grid-points", 0), &npts,
                                     &dim);
    if (s != KEVD_OK) return fail(c, s, "kevd_set_points");
    double eps = a.num("eps", 0.0);
    if (a.has("eps-median") || !(eps > 0.0)) {
        int64_t pairs = a.integer("median-pairs", (int64_t)1 << 22);
        s = kevd_median_sqdist(c, (uint64_t)a.integer("seed", 0), pairs, &eps);
        if (s != KEVD_OK) return fail(c, s, "kevd_median_sqdist");
    }
    kevd_arpc_params p{};
    p.eps = eps;
    p.rank = a.integer("rank", 0);
    p.block = (int32_t)a.integer("block", 128);
    p.seed = (uint64_t)a.integer("seed", 0);
    p.tol = a.num("tol", 1e-10);
    p.prec = a.str("precision", "mixed") == "fp64" ? KEVD_FP64 : KEVD_MIXED;
    p.max_rounds = (int32_t)a.integer("max-rounds", 0);
    p.regenerate = a.has("regenerate") ? 1 : 0;
    if (p.regenerate && a.has("factor-out")) {
        std::fprintf(stderr, "--regenerate keeps no factor: drop --factor-out\n");
        kevd_destroy(c);
        return 1;
    }
    std::vector<int64_t> inj;
    if (a.has("inject-pivots")) {
        try {
            inj = npy_as_i64(load_array(a.str("inject-pivots"), "/spiv"));
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            kevd_destroy(c);
            return 1;
        }
        p.inject_pivots = inj.data();
        p.n_inject = (int64_t)inj.size();
    }
    int32_t mr = p.max_rounds > 0 ? p.max_rounds : (int32_t)(4 * ((p.rank + p.block - 1) / (p.block > 0 ? p.block : 1)) + 64);
    std::vector<int64_t> piv((size_t)(p.rank > 0 ? p.rank : 1));
    std::vector<int32_t> acc((size_t)mr);
    std::vector<double> err((size_t)mr);
    kevd_arpc_result r{};
    r.pivots = piv.data();
    r.accepted = acc.data();
    r.trace_err = err.data();
    r.rounds_cap = mr;
    std::printf("Points: %lld x %lld  eps %.6g  rank %lld  block %d  seed %llu  precision %s\n\n",
                (long long)npts, (long long)dim, eps, (long long)p.rank, p.block,
                (unsigned long long)p.seed, p.prec == KEVD_FP64 ? "fp64" : "mixed");
    s = kevd_arpc(c, &p, &r);
    if (s != KEVD_OK) return fail(c, s, "kevd_arpc");
    std::printf("Accepted rank: %lld in %d rounds\n", (long long)r.rank, r.rounds);
    std::printf("# round accepted cumrank relerr\n");
    long long cum = 0;
    for (int i = 0; i < r.rounds; ++i) {
        cum += acc[i];
        std::printf("%d %d %lld %.5e\n", i, acc[i], cum, err[i]);
    }
    try {
        if (a.has("factor-out")) {
            bool f32 = p.prec == KEVD_MIXED;
            std::vector<char> F((size_t)npts * r.rank * (f32 ? 4 : 8));
            s = kevd_get_factor(c, F.data(), r.rank, p.prec);
            if (s != KEVD_OK) return fail(c, s, "kevd_get_factor");
            save_array(a.str("factor-out"), f32 ? "<f4" : "<f8", {npts, r.rank}, F.data(), "/F");
        }
        if (a.has("pivots-out")) save_array(a.str("pivots-out"), "<u8", {r.rank}, piv.data(), "/spiv");
        if (a.has("linv-out")) {
            std::vector<double> L((size_t)r.rank * r.rank);
            s = kevd_get_linv(c, L.data());
            if (s != KEVD_OK) return fail(c, s, "kevd_get_linv");
            // column-major r x r, flattened (the reference's /linv layout)
            save_array(a.str("linv-out"), "<f8", {r.rank * r.rank}, L.data(), "/linv");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        kevd_destroy(c);
        return 1;
    }
    kevd_timings tm;
    kevd_get_timings(c, &tm);
    kevd_destroy(c);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("ARPC time: %.4f s  Total time: %.2f s\n", tm.arpc_ms / 1e3, secs);
    return 0;
}
