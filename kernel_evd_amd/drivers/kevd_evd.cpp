// kevd-evd -- the ks_kevd.cu stage: factor (or pivot list) in, eigenvalues
// (all r, ascending, FP64; the reference's /lam) and the leading k_out
// eigenvectors (N x k_out FP32 row-major, columns ascending; its /U) out.
// Exit codes follow ks_kevd.cu's aborts: 2 non-positive degree, 3 potrf,
// 4 non-positive q, 5 eigensolver.
//
//   kevd-evd --points X.npy (--factor F.npy | --pivots S.npy (--eps E | --eps-median))
//            [--kout 100] [--norm ref|sinkhorn] [--sinkhorn-tol 1e-12]
//            [--precision mixed|fp64] [--all-eigenvalues] [--lam-out lam.npy]
//            [--U-out U.npy] [--device 0]
// Without --all-eigenvalues only the leading eigenvalues are computed and
// the rest of lam is NaN.
// With --pivots, F = K(:,S) L^-T is rebuilt by the ARPC kernels with the
// pivot list injected (K(S,S) = L L^T), which is what the reference's kEVD
// regenerates panel by panel from /spiv and /linv.
#include <chrono>
#include <cstdio>
#include <vector>

#include "args.h"
#include "kevd.h"
#include "npy.h"

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
        std::printf("kevd-evd --points X.npy (--factor F.npy | --pivots S.npy (--eps E | --eps-median))\n"
                    "         [--kout 100] [--norm ref|sinkhorn] [--sinkhorn-tol 1e-12]\n"
                    "         [--precision mixed|fp64] [--all-eigenvalues] [--lam-out lam.npy] [--U-out U.npy]\n");
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    NpyArray X;
    try {
        X = npy_load(a.need("points"));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    kevd_precision prec = a.str("precision", "mixed") == "fp64" ? KEVD_FP64 : KEVD_MIXED;
    kevd_ctx* c = nullptr;
    kevd_status s = kevd_create((int)a.integer("device", 0), &c);
    if (s != KEVD_OK) return fail(nullptr, s, "kevd_create");
    if (X.dtype == "<f4")
        s = kevd_set_points_f32(c, reinterpret_cast<const float*>(X.data.data()), X.rows(), (int32_t)X.cols());
    else {
        std::vector<double> x = npy_as_f64(X);
        s = kevd_set_points(c, x.data(), X.rows(), (int32_t)X.cols());
    }
    if (s != KEVD_OK) return fail(c, s, "kevd_set_points");
    int64_t r = 0;
    if (a.has("factor")) {
        NpyArray F = npy_load(a.str("factor"));
        if (F.shape.size() != 2 || F.rows() != X.rows()) {
            std::fprintf(stderr, "--factor must be N x r with N = points\n");
            kevd_destroy(c);
            return 1;
        }
        r = F.cols();
        if (prec == KEVD_MIXED) {
            std::vector<float> f((size_t)F.count());
            if (F.dtype == "<f4") f.assign(reinterpret_cast<const float*>(F.data.data()),
                                           reinterpret_cast<const float*>(F.data.data()) + f.size());
            else {
                std::vector<double> d = npy_as_f64(F);
                for (size_t i = 0; i < f.size(); ++i) f[i] = (float)d[i];
            }
            s = kevd_set_factor(c, f.data(), F.rows(), r, r, KEVD_MIXED);
        } else {
            std::vector<double> d = npy_as_f64(F);
            s = kevd_set_factor(c, d.data(), F.rows(), r, r, KEVD_FP64);
        }
        if (s != KEVD_OK) return fail(c, s, "kevd_set_factor");
    } else {
        std::vector<int64_t> S = npy_as_i64(npy_load(a.need("pivots")));
        double eps = a.num("eps", 0.0);
        if (a.has("eps-median") || !(eps > 0.0)) {
            s = kevd_median_sqdist(c, (uint64_t)a.integer("seed", 0), a.integer("median-pairs", (int64_t)1 << 22), &eps);
            if (s != KEVD_OK) return fail(c, s, "kevd_median_sqdist");
        }
        kevd_arpc_params p{};
        p.eps = eps;
        p.rank = (int64_t)S.size();
        p.block = 128;
        p.tol = a.num("tol", 1e-10);
        p.prec = prec;
        p.inject_pivots = S.data();
        p.n_inject = (int64_t)S.size();
        int32_t mr = (int32_t)(4 * ((p.rank + 127) / 128) + 64);
        std::vector<int64_t> piv(S.size());
        std::vector<int32_t> acc((size_t)mr);
        std::vector<double> err((size_t)mr);
        kevd_arpc_result res{};
        res.pivots = piv.data();
        res.accepted = acc.data();
        res.trace_err = err.data();
        res.rounds_cap = mr;
        s = kevd_arpc(c, &p, &res);
        if (s != KEVD_OK) return fail(c, s, "kevd_arpc (pivot injection)");
        r = res.rank;
    }
    kevd_evd_params ep{};
    ep.k_out = (int32_t)a.integer("kout", 100);
    if (ep.k_out > r) ep.k_out = (int32_t)r;
    ep.norm = a.str("norm", "ref") == "sinkhorn" ? KEVD_NORM_SINKHORN : KEVD_NORM_REF;
    ep.sinkhorn_tol = a.num("sinkhorn-tol", 1e-12);
    ep.sinkhorn_max_iter = (int32_t)a.integer("sinkhorn-max-iter", 1000);
    ep.prec = prec;
    ep.all_eigenvalues = a.has("all-eigenvalues") ? 1 : 0;
    std::vector<double> lam((size_t)r);
    std::vector<float> U;
    kevd_evd_result er{};
    er.lam = lam.data();
    if (a.has("U-out")) {
        U.resize((size_t)X.rows() * ep.k_out);
        er.U = U.data();
    }
    s = kevd_evd(c, &ep, &er);
    if (s != KEVD_OK) return fail(c, s, "kevd_evd");
    std::printf("rank %lld  k_out %d  norm %s  shift %.6e  min degree %.6e  row-sum deviation %.3e\n",
                (long long)r, ep.k_out, ep.norm ? "sinkhorn" : "ref", er.shift, er.min_degree, er.rowsum_dev);
    std::printf("leading eigenvalues:");
    for (int i = 0; i < 8 && i < r; ++i) std::printf(" %.10f", lam[r - 1 - i]);
    std::printf("\n");
    if (a.has("lam-out")) npy_save(a.str("lam-out"), "<f8", {r}, lam.data());
    if (a.has("U-out")) npy_save(a.str("U-out"), "<f4", {X.rows(), (int64_t)ep.k_out}, U.data());
    kevd_timings tm;
    kevd_get_timings(c, &tm);
    kevd_destroy(c);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("EVD time: %.4f s  Total time: %.2f s\n", tm.evd_ms / 1e3, secs);
    return 0;
}
