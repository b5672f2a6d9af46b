// kevd-evd -- the ks_kevd.cu stage: factor (or pivot list) in, eigenvalues
// (all r, ascending, FP64; the reference's /lam) and the leading k_out
// eigenvectors (N x k_out FP32 row-major, columns ascending; its /U) out.
// Exit codes follow ks_kevd.cu's aborts: 2 non-positive degree, 3 potrf,
// 4 non-positive q, 5 eigensolver.
//
//   kevd-evd --points X.npy|u.h5 [--delay Nq] [--grid-points M]
//            (--factor F.npy | --pivots S.npy|spiv.h5 [--linv Linv.npy|linv.h5]
//             (--eps E | --eps-median) [--regenerate])
//            [--kout 100] [--norm ref|sinkhorn] [--sinkhorn-tol 1e-12]
//            [--precision mixed|fp64] [--all-eigenvalues] [--lam-out lam.npy]
//            [--U-out U.npy] [--out evd.h5 [--leading-only]]
//            [--dump kevd_dump.h5|DIR] [--device 0]
// Without --all-eigenvalues only the leading eigenvalues are computed and
// the rest of lam is NaN.  --out writes the reference's EVD file
// (ks_kevd.cu:488-501, 1930-1938, 2300-2306): /lam (all r eigenvalues,
// ascending, FP64) and /U (N x k_out FP32, row-major, flattened); it
// implies --all-eigenvalues unless --leading-only.  --linv reads the
// reference's /linv and reports how far the rebuilt L^-1 is from it.
// With --pivots, F = K(:,S) L^-T is rebuilt by the ARPC kernels with the
// pivot list injected (K(S,S) = L L^T).  With --pivots --regenerate no
// factor is stored at all: the LM form of ARPC keeps S and L^-1 and the
// kEVD regenerates K(R,S) panel by panel in every pass, as the reference's
// kEVD does from /spiv and /linv (FP64 only).
// --dump writes the reference's kevd_dump file (its dump_h5 switch,
// ks_kevd.cu:484, 531-542 and the later h5add_double sites): c1, p1, dtil,
// cq, pq, Atil, A, R, G, qtil, Btil, W, V, lam, Z (FP64) and U (FP32), each
// one flat 1-D array, matrices column-major (read back with order='F'), in
// the reference's kernel basis K(:,S).  A name ending in .h5 is one HDF5
// file of those datasets; anything else is a directory of NAME.npy files.
// It needs --pivots (a factor file carries no kernel basis) and the ref
// normalisation, and implies --all-eigenvalues unless --leading-only (V is
// then the r x k_out block).
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "args.h"
#include "arrays.h"
#include "kevd.h"

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
        std::printf("kevd-evd --points X.npy|u.h5 [--delay Nq] [--grid-points M]\n"
                    "         (--factor F.npy | --pivots S.npy|spiv.h5 [--linv linv.h5] (--eps E | --eps-median)\n"
                    "          [--regenerate])\n"
                    "         [--kout 100] [--norm ref|sinkhorn] [--sinkhorn-tol 1e-12]\n"
                    "         [--precision mixed|fp64] [--all-eigenvalues] [--lam-out lam.npy] [--U-out U.npy]\n"
                    "         [--out evd.h5 [--leading-only]] [--dump kevd_dump.h5|DIR] [--device 0]\n");
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    kevd_precision prec = a.str("precision", "mixed") == "fp64" ? KEVD_FP64 : KEVD_MIXED;
    kevd_ctx* c = nullptr;
    kevd_status s = kevd_create((int)a.integer("device", 0), &c);
    if (s != KEVD_OK) return fail(nullptr, s, "kevd_create");
    int64_t npts = 0;
    int32_t dim = 0;
    s = (kevd_status)set_points_from(c, a.need("points"), a.integer("delay", 0), a.integer("grid-points", 0), &npts,
                                     &dim);
    if (s != KEVD_OK) return fail(c, s, "kevd_set_points");
    try {
        int64_t r = 0;
        if (a.has("factor")) {
            NpyArray F = load_array(a.str("factor"), "/F");
            if (F.shape.size() != 2 || F.rows() != npts) {
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
            std::vector<int64_t> S = npy_as_i64(load_array(a.need("pivots"), "/spiv"));
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
            // --regenerate: keep only S and L^-1 (the LM form) and let the
            // kEVD regenerate K(R,S) panels instead of storing F (FP64)
            p.regenerate = a.has("regenerate") ? 1 : 0;
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
            if (a.has("linv")) {
                std::vector<double> Lf = npy_as_f64(load_array(a.str("linv"), "/linv"));
                if ((int64_t)Lf.size() != r * r) {
                    std::fprintf(stderr, "--linv holds %zu values, the pivot list needs %lld\n", Lf.size(),
                                 (long long)(r * r));
                    kevd_destroy(c);
                    return 1;
                }
                std::vector<double> L((size_t)(r * r));
                s = kevd_get_linv(c, L.data());
                if (s != KEVD_OK) return fail(c, s, "kevd_get_linv");
                double num = 0.0, den = 0.0;
                for (size_t i = 0; i < L.size(); ++i) {
                    num += (L[i] - Lf[i]) * (L[i] - Lf[i]);
                    den += Lf[i] * Lf[i];
                }
                std::printf("rebuilt L^-1 vs --linv: relative Frobenius difference %.3e\n", std::sqrt(num / den));
            }
        }
        kevd_evd_params ep{};
        ep.k_out = (int32_t)a.integer("kout", 100);
        if (ep.k_out > r) ep.k_out = (int32_t)r;
        ep.norm = a.str("norm", "ref") == "sinkhorn" ? KEVD_NORM_SINKHORN : KEVD_NORM_REF;
        ep.sinkhorn_tol = a.num("sinkhorn-tol", 1e-12);
        ep.sinkhorn_max_iter = (int32_t)a.integer("sinkhorn-max-iter", 1000);
        ep.prec = prec;
        const bool h5out = a.has("out");
        const bool dump = a.has("dump");
        ep.all_eigenvalues = a.has("all-eigenvalues") || ((h5out || dump) && !a.has("leading-only")) ? 1 : 0;
        if (dump) {
            if (ep.norm != KEVD_NORM_REF) {
                std::fprintf(stderr, "--dump records the ref normalisation's intermediates\n");
                kevd_destroy(c);
                return 1;
            }
            s = kevd_set_dump(c, 1);
            if (s != KEVD_OK) return fail(c, s, "kevd_set_dump");
        }
        std::vector<double> lam((size_t)r);
        std::vector<float> U;
        kevd_evd_result er{};
        er.lam = lam.data();
        if (a.has("U-out") || h5out || dump) {
            U.resize((size_t)npts * ep.k_out);
            er.U = U.data();
        }
        s = kevd_evd(c, &ep, &er);
        if (s != KEVD_OK) return fail(c, s, "kevd_evd");
        std::printf("rank %lld  k_out %d  norm %s  shift %.6e  min degree %.6e  row-sum deviation %.3e\n",
                    (long long)r, ep.k_out, ep.norm ? "sinkhorn" : "ref", er.shift, er.min_degree, er.rowsum_dev);
        std::printf("leading eigenvalues:");
        for (int i = 0; i < 8 && i < r; ++i) std::printf(" %.10f", lam[r - 1 - i]);
        std::printf("\n");
        if (a.has("lam-out")) save_array(a.str("lam-out"), "<f8", {r}, lam.data(), "/lam");
        if (a.has("U-out")) save_array(a.str("U-out"), "<f4", {npts, (int64_t)ep.k_out}, U.data(), "/U");
        if (h5out) {
            if (!is_h5(a.str("out"))) {
                std::fprintf(stderr, "--out names an HDF5 file (.h5)\n");
                kevd_destroy(c);
                return 1;
            }
            save_array(a.str("out"), "<f8", {r}, lam.data(), "/lam", false);
            save_array(a.str("out"), "<f4", {npts, (int64_t)ep.k_out}, U.data(), "/U", true);
        }
        if (dump) {
            const std::string path = a.str("dump");
            const bool h5 = is_h5(path);
            if (!h5 && mkdir(path.c_str(), 0755) != 0 && errno != EEXIST) {
                std::fprintf(stderr, "cannot create the directory %s\n", path.c_str());
                kevd_destroy(c);
                return 1;
            }
            bool first = true;
            std::vector<double> v;
            for (const char* name : {"c1", "p1", "dtil", "cq", "pq", "Atil", "A", "R", "G", "qtil", "Btil", "W",
                                     "V", "lam", "Z"}) {
                int64_t cnt = 0;
                s = kevd_get_intermediate(c, name, nullptr, &cnt);
                if (s != KEVD_OK) return fail(c, s, "kevd_get_intermediate");
                v.resize((size_t)cnt);
                s = kevd_get_intermediate(c, name, v.data(), &cnt);
                if (s != KEVD_OK) return fail(c, s, "kevd_get_intermediate");
                const std::string ds = std::string("/") + name;
                save_array(h5 ? path : path + ds + ".npy", "<f8", {cnt}, v.data(), ds, h5 && !first);
                first = false;
            }
            save_array(h5 ? path : path + "/U.npy", "<f4", {npts * (int64_t)ep.k_out}, U.data(), "/U", h5);
            std::printf("intermediates written to %s\n", path.c_str());
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
    std::printf("EVD time: %.4f s  Total time: %.2f s\n", tm.evd_ms / 1e3, secs);
    return 0;
}
