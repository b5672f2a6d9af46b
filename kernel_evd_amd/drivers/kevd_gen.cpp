// kevd-gen -- Kuramoto-Sivashinsky snapshot file (the ks_datagen.py stage,
// ks_datagen.py:5-107), optionally delay embedded the way ks_arpclm.cu reads
// its input (point g*N + t is u[g][t .. t+Nq), ks_arpclm.cu:263-268).
//
//   kevd-gen --n N [--L 22] [--M 64] [--dt 0.25] [--discard 10000]
//            [--stride 4] [--seed 0] [--ic-scale 0.1] [--delay Nq] [--f32]
//            --out X.npy|u.h5
// Without --delay: N x M snapshots.  With --delay Nq: N + Nq - 1 snapshots
// are generated and N*M points of dimension Nq written.
// An .h5 output is the reference's snapshot file (ks_datagen.py:91-103): /u
// holds the N snapshots flattened with the time index fastest (FP64, or
// FP32 with --f32 as in its _u32 file) and /t their times (FP64); delay
// embedding then happens where the file is read (--delay of the other
// drivers), so --delay is not taken with an .h5 output.
#include <chrono>
#include <cstdio>
#include <vector>

#include "args.h"
#include "kevd.h"
#include "arrays.h"

int main(int argc, char** argv)
{
    Args a(argc, argv);
    if (a.has("help")) {
        std::printf("kevd-gen --n N [--L 22] [--M 64] [--dt 0.25] [--discard 10000] [--stride 4] "
                    "[--seed 0] [--ic-scale 0.1] [--delay Nq] [--f32] --out X.npy\n");
        return 0;
    }
    kevd_ks_params p;
    p.L = a.num("L", 22.0);
    p.M = (int32_t)a.integer("M", 64);
    p.dt = a.num("dt", 0.25);
    p.n_discard = a.integer("discard", 10000);
    p.stride = (int32_t)a.integer("stride", 4);
    p.seed = (uint64_t)a.integer("seed", 0);
    p.ic_scale = a.num("ic-scale", 0.1);
    int64_t n = a.integer("n", 0);
    int64_t nq = a.integer("delay", 0);
    std::string out = a.need("out");
    if (n < 1) {
        std::fprintf(stderr, "--n must be >= 1\n");
        return 1;
    }
    const bool h5 = is_h5(out);
    if (h5 && nq > 0) {
        std::fprintf(stderr, "an .h5 output holds the snapshot series; give --delay to the driver that reads it\n");
        return 1;
    }
    int64_t nsnap = nq > 0 ? n + nq - 1 : n;
    std::vector<double> U((size_t)nsnap * p.M);
    auto t0 = std::chrono::steady_clock::now();
    kevd_status s = kevd_ks_generate(&p, nsnap, U.data());
    if (s != KEVD_OK) {
        std::fprintf(stderr, "kevd_ks_generate: %s\n", kevd_status_string(s));
        return exit_code(s);
    }
    if (h5) {
        // u[g * n + t] = snapshot t at grid point g; t_i = (discard + i stride) dt
        std::vector<double> u((size_t)n * p.M), t((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            t[(size_t)i] = (double)(p.n_discard + i * p.stride) * p.dt;
            for (int32_t g = 0; g < p.M; ++g) u[(size_t)(g * n + i)] = U[(size_t)(i * p.M + g)];
        }
        try {
            if (a.has("f32")) {
                std::vector<float> f(u.begin(), u.end());
                save_array(out, "<f4", {n * p.M}, f.data(), "/u", false);
            } else {
                save_array(out, "<f8", {n * p.M}, u.data(), "/u", false);
            }
            save_array(out, "<f8", {n}, t.data(), "/t", true);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
        std::printf("wrote %s: /u %lld x %d (time fastest), /t %lld\n", out.c_str(), (long long)n, p.M,
                    (long long)n);
        return 0;
    }
    std::vector<double> X;
    int64_t rows = nsnap, cols = p.M;
    if (nq > 0) {
        rows = n * p.M;
        cols = nq;
        X.resize((size_t)rows * cols);
        s = kevd_delay_embed(U.data(), nsnap, p.M, (int32_t)nq, X.data());
        if (s != KEVD_OK) return exit_code(s);
    } else {
        X.swap(U);
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (a.has("f32")) {
        std::vector<float> f(X.begin(), X.end());
        save_array(out, "<f4", {rows, cols}, f.data(), "/X");
    } else {
        save_array(out, "<f8", {rows, cols}, X.data(), "/X");
    }
    std::printf("wrote %s: %lld x %lld (%.2f s)\n", out.c_str(), (long long)rows, (long long)cols, secs);
    return 0;
}
