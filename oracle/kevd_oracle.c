/*
 * kevd_oracle.c -- CPU FP64 restatement of the reference's ARPC, kernel and
 * bandwidth arithmetic.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check
 * and bench.py's cpu_baseline leg load it as the checker.  The product
 * library never links or calls it.
 *
 * Written in plain C (OpenMP for the row loops).  Every function names the
 * reference lines it restates.  Where the build departs from the reference
 * by design (Philox instead of cuRAND, fixed-point CDF instead of a float
 * scan, F-form instead of L^-1 form) the departure is the one libkevd makes,
 * so pivots are comparable bit for bit.
 *
 * Pinning (tests/test_oracle_*.py): Philox against the Random123
 * known-answer vectors; the kernel block against brute-force NumPy; the
 * sampler against known answers (one-hot residual, exact integer targets);
 * the rejection Cholesky against numpy.linalg.cholesky of H[acc,acc] and
 * the b = 1 reduction to plain RPCholesky; the factor against the
 * interpolation identities F F[S]^T = K(:,S), and with r = n against
 * FF^T = K; bandwidth sums against brute force and the eps -> inf limit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ RNG */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), replacing the cuRAND
 * XORWOW states of ks_arpclm.cu:175-180. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum { ORC_STREAM_PROPOSE = 0, ORC_STREAM_REJECT = 1, ORC_STREAM_GAUSS = 2,
       ORC_STREAM_MEDIAN = 3, ORC_STREAM_ROWS = 4 };

static void draw(uint64_t seed, uint32_t stream, uint32_t round, uint32_t idx,
                 uint32_t out[4])
{
    uint32_t ctr[4] = { idx, round, stream, 0u };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox4x32_10(ctr, key, out);
}

static uint64_t draw_u64(uint64_t seed, uint32_t stream, uint32_t round, uint32_t idx)
{
    uint32_t o[4];
    draw(seed, stream, round, idx, o);
    return ((uint64_t)o[1] << 32) | o[0];
}

/* uniform double in (0,1]: the range of curand_uniform_double used by
 * rejection_sample (ks_arpclm.cu:329). */
static double u64_to_unit(uint64_t u) { return (double)((u >> 11) + 1) * 0x1.0p-53; }

uint64_t orc_draw_u64(uint64_t seed, uint32_t stream, uint32_t round, uint32_t idx)
{
    return draw_u64(seed, stream, round, idx);
}

double orc_uniform01(uint64_t seed, uint32_t stream, uint32_t round, uint32_t idx)
{
    return u64_to_unit(draw_u64(seed, stream, round, idx));
}

static uint64_t mulhi64(uint64_t a, uint64_t b)
{
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

/* ---------------------------------------------------------- kernel value */

/* K(i,j) = exp(-sum_k (x_ik - x_jk)^2 / eps), direct differences in FP64:
 * compute_kblock / kexp_t (ks_arpclm.cu:226-280) and compute_kernel_mat
 * (ks_reference.py:51-58) with eps = Nq * bw. */
static double wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

static double kval(const double* X, int32_t dim, int64_t i, int64_t j, double eps)
{
    if (i == j) return 1.0;
    const double* a = X + i * (int64_t)dim;
    const double* b = X + j * (int64_t)dim;
    double s = 0.0;
    for (int32_t k = 0; k < dim; ++k) {
        double t = a[k] - b[k];
        s += t * t;
    }
    return exp(-s / eps);
}

static double sqdist(const double* X, int32_t dim, int64_t i, int64_t j)
{
    const double* a = X + i * (int64_t)dim;
    const double* b = X + j * (int64_t)dim;
    double s = 0.0;
    for (int32_t k = 0; k < dim; ++k) {
        double t = a[k] - b[k];
        s += t * t;
    }
    return s;
}

void orc_kernel_block(const double* X, int64_t n, int32_t dim,
                      const int64_t* rows, int64_t nr, const int64_t* cols,
                      int64_t nc, double eps, double* out)
{
    (void)n;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nr; ++r)
        for (int64_t c = 0; c < nc; ++c)
            out[r * nc + c] = kval(X, dim, rows[r], cols[c], eps);
}

/* ---------------------------------------------------------------- median */

static int cmp_double(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* Median squared pairwise distance.  The reference has no median (it fixes
 * bw, ks_arpclm.cu:602); this is the build's estimator: exact over all
 * pairs i<j, or over nsample Philox pairs (i != j). */
int orc_median_sqdist(const double* X, int64_t n, int32_t dim, uint64_t seed,
                      int64_t nsample, double* out)
{
    int64_t npairs = n * (n - 1) / 2;
    int exact = (nsample <= 0 || nsample >= npairs);
    int64_t m = exact ? npairs : nsample;
    if (m <= 0) return 1;
    double* v = (double*)malloc((size_t)m * sizeof(double));
    if (!v) return 2;
    if (exact) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 1; i < n; ++i) {
            int64_t base = i * (i - 1) / 2;
            for (int64_t j = 0; j < i; ++j) v[base + j] = sqdist(X, dim, i, j);
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < m; ++t) {
            uint32_t o[4];
            draw(seed, ORC_STREAM_MEDIAN, 0u, (uint32_t)t, o);
            int64_t i = (int64_t)(((uint64_t)o[0] * (uint64_t)n) >> 32);
            int64_t j = (int64_t)(((uint64_t)o[1] * (uint64_t)(n - 1)) >> 32);
            if (j >= i) ++j;
            v[t] = sqdist(X, dim, i, j);
        }
    }
    qsort(v, (size_t)m, sizeof(double), cmp_double);
    *out = (m & 1) ? v[m / 2] : 0.5 * (v[m / 2 - 1] + v[m / 2]);
    free(v);
    return 0;
}

/* -------------------------------------------------------------- sampling */

static int ceil_log2_i64(int64_t n)
{
    int c = 0;
    while (((int64_t)1 << c) < n) ++c;
    return c;
}

/* Fixed-point weights of the residual diagonal: w_i = floor(d_i 2^(K-e)),
 * K = 62 - ceil(log2 n), d_max = m 2^e with m in [0.5,1).  Returns the total
 * W (< 2^62) and writes the inclusive prefix.  Replaces the float CDF scan
 * of propose_block (ks_arpclm.cu:193-224), whose sum was order dependent. */
uint64_t orc_fixed_weights(const double* d, int64_t n, uint64_t* prefix)
{
    double dmax = 0.0;
    for (int64_t i = 0; i < n; ++i) if (d[i] > dmax) dmax = d[i];
    if (!(dmax > 0.0)) {
        for (int64_t i = 0; i < n; ++i) prefix[i] = 0;
        return 0;
    }
    int e;
    frexp(dmax, &e);
    int shift = 62 - ceil_log2_i64(n) - e;
    uint64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        double di = d[i] > 0.0 ? d[i] : 0.0;
        acc += (uint64_t)ldexp(di, shift);
        prefix[i] = acc;
    }
    return acc;
}

/* Step 1 of a round (ks_arpclm.cu:193-224; README_arpclm.md:23-24): b iid
 * draws i ~ d_i / sum d by inverse CDF.  Target t = floor(U W / 2^64), the
 * draw is the smallest i with prefix[i] > t. */
int orc_propose(const double* d, int64_t n, uint64_t seed, uint32_t round,
                int32_t b, int64_t* out, uint64_t* W_out)
{
    uint64_t* pre = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    if (!pre) return 2;
    uint64_t W = orc_fixed_weights(d, n, pre);
    if (W_out) *W_out = W;
    if (W == 0) { free(pre); return 3; }
    for (int32_t j = 0; j < b; ++j) {
        uint64_t t = mulhi64(draw_u64(seed, ORC_STREAM_PROPOSE, round, (uint32_t)j), W);
        int64_t lo = 0, hi = n - 1;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (pre[mid] > t) hi = mid; else lo = mid + 1;
        }
        out[j] = lo;
    }
    free(pre);
    return 0;
}

/* ------------------------------------------------------- rejection block */

/* Algorithm 2.1 thinning on the b x b residual block H (row-major, consumed)
 * followed by the compacted factor's inverse: rejection_sample,
 * ks_arpclm.cu:300-408.  Acceptance of proposal i: u_i * rnd < H_ii and
 * H_ii > tol, with u = diag(H) before any update and rnd in (0,1] from the
 * REJECT stream; with `inject` set, only H_ii > tol is tested.  Outputs the
 * accepted positions, the compacted lower factor Lc (k x k, row-major, ld b)
 * and its inverse Lcinv (same layout). */
int32_t orc_rejchol(double* H, int32_t b, uint64_t seed, uint32_t round,
                    double tol, int inject, int32_t* accpos, double* Lc,
                    double* Lcinv)
{
    double* u = (double*)malloc((size_t)b * sizeof(double));
    double* L = (double*)calloc((size_t)b * b, sizeof(double));
    for (int32_t i = 0; i < b; ++i) u[i] = H[i * b + i];
    int32_t k = 0;
    for (int32_t i = 0; i < b; ++i) {
        double rnd = orc_uniform01(seed, ORC_STREAM_REJECT, round, (uint32_t)i);
        double hii = H[i * b + i];
        int acc = inject ? (hii > tol) : ((u[i] * rnd < hii) && (hii > tol));
        if (!acc) continue;
        accpos[k++] = i;
        double f = 1.0 / sqrt(hii);
        for (int32_t j = i; j < b; ++j) L[j * b + i] = H[j * b + i] * f;
        for (int32_t j = i + 1; j < b; ++j)
            for (int32_t m = i + 1; m <= j; ++m)
                H[j * b + m] -= L[j * b + i] * L[m * b + i];
    }
    for (int32_t a = 0; a < k; ++a)
        for (int32_t c = 0; c < b; ++c) {
            Lc[a * b + c] = (c <= a && c < k) ? L[accpos[a] * b + accpos[c]] : 0.0;
            Lcinv[a * b + c] = 0.0;
        }
    /* Column t of Lc^-1 by forward substitution (ks_arpclm.cu:378-393). */
    for (int32_t t = 0; t < k; ++t) {
        Lcinv[t * b + t] = 1.0 / Lc[t * b + t];
        for (int32_t a = t + 1; a < k; ++a) {
            double s = 0.0;
            for (int32_t e = t; e < a; ++e) s += Lc[a * b + e] * Lcinv[e * b + t];
            Lcinv[a * b + t] = -s / Lc[a * b + a];
        }
    }
    free(u);
    free(L);
    return k;
}

/* ------------------------------------------------------------------ ARPC */

/* Accelerated RPCholesky in F-form.  One round (README_arpclm.md:23-39,
 * ks_arpclm.cu:793-1046):
 *   S'  = propose(d)                                      (193-224)
 *   H   = K(S',S') - F(S',:) F(S',:)^T                     (807-874)
 *   k, S_i, Lc = rejchol(H)                                (300-408)
 *   G   = (K(:,S_i) - F F(S_i,:)^T) Lc^-T                  (929-986)
 *   F(:, Nr:Nr+k) = G;  d = max(d - rowsq(G), 0)           (455-478)
 *   err = sum(d) / n                                       (536-564)
 * The reference keeps L^-1 and regenerates K(R,S) (its "LM" form); with
 * F = K(:,S) L^-T both give the same G.  Rounds stop at the target rank
 * (the last block truncated), zero residual mass, a maximum residual
 * entry <= tol (nothing could be accepted any more), or max_rounds.
 * F is n x rank row-major (caller-zeroed).  phase_s (optional, 3 values)
 * receives the wall seconds of the sampling, of the pivot block (H and the
 * rejection Cholesky) and of the downdate (the n-row update).  Returns 0,
 * or 2 on allocation failure. */
int orc_arpc(const double* X, int64_t n, int32_t dim, double eps, int64_t rank,
             int32_t b, uint64_t seed, double tol, int32_t max_rounds,
             const int64_t* inject, int64_t n_inject, double* F,
             int64_t* pivots, int32_t* accepted, double* err,
             int32_t* rounds_out, int64_t* rank_out, double* d_out, double* phase_s)
{
    double* d = (double*)malloc((size_t)n * sizeof(double));
    int64_t* Sp = (int64_t*)malloc((size_t)b * sizeof(int64_t));
    double* H = (double*)malloc((size_t)b * b * sizeof(double));
    double* Lc = (double*)malloc((size_t)b * b * sizeof(double));
    double* Lcinv = (double*)malloc((size_t)b * b * sizeof(double));
    int32_t* accpos = (int32_t*)malloc((size_t)b * sizeof(int32_t));
    double* FS = (double*)malloc((size_t)b * (rank > 0 ? rank : 1) * sizeof(double));
    if (b < 1 || b > 1024) return 1;
    if (!d || !Sp || !H || !Lc || !Lcinv || !accpos || !FS) return 2;
    for (int64_t i = 0; i < n; ++i) d[i] = 1.0; /* K_ii = 1 (ks_arpclm.cu:182-191) */

    int64_t Nr = 0, inj_pos = 0;
    int32_t round = 0;
    double t_sample = 0.0, t_block = 0.0, t_down = 0.0, t0;
    for (; round < max_rounds && Nr < rank; ++round) {
        t0 = wtime();
        /* no residual entry above tol: Algorithm 2.1 accepts only
         * H_ii > tol and H_ii <= d_i, so no later round could accept */
        double dmax = 0.0;
        for (int64_t i = 0; i < n; ++i) dmax = d[i] > dmax ? d[i] : dmax;
        if (!(dmax > tol)) break;
        int32_t bb = b;
        if (inject) {
            if (inj_pos >= n_inject) break;
            if (n_inject - inj_pos < bb) bb = (int32_t)(n_inject - inj_pos);
            for (int32_t j = 0; j < bb; ++j) Sp[j] = inject[inj_pos + j];
            inj_pos += bb;
        } else {
            uint64_t W = 0;
            if (orc_propose(d, n, seed, (uint32_t)round, b, Sp, &W) != 0) break;
        }
        t_sample += wtime() - t0;
        t0 = wtime();
        /* H = K(S',S') - F(S',:Nr) F(S',:Nr)^T */
#pragma omp parallel for schedule(static)
        for (int32_t p = 0; p < bb; ++p)
            for (int32_t q = 0; q < bb; ++q) {
                double s = kval(X, dim, Sp[p], Sp[q], eps);
                const double* fp = F + Sp[p] * rank;
                const double* fq = F + Sp[q] * rank;
                double t = 0.0;
                for (int64_t m = 0; m < Nr; ++m) t += fp[m] * fq[m];
                H[p * bb + q] = s - t;
            }
        int32_t k = orc_rejchol(H, bb, seed, (uint32_t)round, tol, inject != NULL,
                                accpos, Lc, Lcinv);
        if (k > rank - Nr) k = (int32_t)(rank - Nr);
        accepted[round] = k;
        t_block += wtime() - t0;
        t0 = wtime();
        if (k > 0) {
            int64_t Si[1024];
            for (int32_t a = 0; a < k; ++a) Si[a] = Sp[accpos[a]];
            for (int32_t a = 0; a < k; ++a)
                memcpy(FS + (int64_t)a * rank, F + Si[a] * rank, (size_t)Nr * sizeof(double));
            /* rows in blocks of RB: each pivot row of F(S_i,:) is loaded once
             * per block and used for RB rows (register blocking) */
            enum { RB = 4 };
#pragma omp parallel
            {
                double T[RB][1024], G[1024];
#pragma omp for schedule(static)
                for (int64_t i0 = 0; i0 < n; i0 += RB) {
                    const int nrb = n - i0 < RB ? (int)(n - i0) : RB;
                    if (nrb == RB) {
                        const double* f0 = F + i0 * rank;
                        const double* f1 = f0 + rank;
                        const double* f2 = f1 + rank;
                        const double* f3 = f2 + rank;
                        for (int32_t c = 0; c < k; ++c) {
                            const double* fs = FS + (int64_t)c * rank;
                            double t0r = 0.0, t1r = 0.0, t2r = 0.0, t3r = 0.0;
#pragma omp simd reduction(+ : t0r, t1r, t2r, t3r)
                            for (int64_t m = 0; m < Nr; ++m) {
                                const double v = fs[m];
                                t0r += f0[m] * v;
                                t1r += f1[m] * v;
                                t2r += f2[m] * v;
                                t3r += f3[m] * v;
                            }
                            T[0][c] = kval(X, dim, i0, Si[c], eps) - t0r;
                            T[1][c] = kval(X, dim, i0 + 1, Si[c], eps) - t1r;
                            T[2][c] = kval(X, dim, i0 + 2, Si[c], eps) - t2r;
                            T[3][c] = kval(X, dim, i0 + 3, Si[c], eps) - t3r;
                        }
                    } else {
                        for (int r = 0; r < nrb; ++r) {
                            const double* fi = F + (i0 + r) * rank;
                            for (int32_t c = 0; c < k; ++c) {
                                const double* fs = FS + (int64_t)c * rank;
                                double t = 0.0;
#pragma omp simd reduction(+ : t)
                                for (int64_t m = 0; m < Nr; ++m) t += fi[m] * fs[m];
                                T[r][c] = kval(X, dim, i0 + r, Si[c], eps) - t;
                            }
                        }
                    }
                    for (int r = 0; r < nrb; ++r) {
                        const int64_t i = i0 + r;
                        double sq = 0.0;
                        for (int32_t a = 0; a < k; ++a) {
                            double g = 0.0;
                            for (int32_t c = 0; c <= a; ++c) g += Lcinv[a * bb + c] * T[r][c];
                            G[a] = g;
                            sq += g * g;
                        }
                        double* fo = F + i * rank + Nr;
                        for (int32_t a = 0; a < k; ++a) fo[a] = G[a];
                        double dd = d[i] - sq;
                        d[i] = dd >= 0.0 ? dd : 0.0;
                    }
                }
            }
            for (int32_t a = 0; a < k; ++a) pivots[Nr + a] = Si[a];
            Nr += k;
        }
        t_down += wtime() - t0;
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += d[i];
        err[round] = s / (double)n;
    }
    *rounds_out = round;
    *rank_out = Nr;
    if (d_out) memcpy(d_out, d, (size_t)n * sizeof(double));
    if (phase_s) {
        phase_s[0] = t_sample;
        phase_s[1] = t_block;
        phase_s[2] = t_down;
    }
    free(d); free(Sp); free(H); free(Lc); free(Lcinv); free(accpos); free(FS);
    return 0;
}

/* ------------------------------------------------------------- bandwidth */

/* S(eps_j) = sum_{i in rows} sum_{j'} exp(-||x_i - x_j'||^2 / eps_j) in
 * FP64: the quantity dmat_comp + kernelsum_comp accumulate
 * (ks_bandwidth.cu:181-325), over a row subset times all columns instead of
 * the strictly lower triangle. */
void orc_bandwidth(const double* X, int64_t n, int32_t dim, const int64_t* rows,
                   int64_t nrows, const double* eps, int32_t neps, double* S)
{
    for (int32_t e = 0; e < neps; ++e) S[e] = 0.0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double* part = (double*)calloc((size_t)nt * neps, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* my = part + (size_t)tid * neps;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < nrows; ++r)
            for (int64_t j = 0; j < n; ++j) {
                double dd = sqdist(X, dim, rows[r], j);
                for (int32_t e = 0; e < neps; ++e) my[e] += exp(-dd / eps[e]);
            }
    }
    for (int t = 0; t < nt; ++t)
        for (int32_t e = 0; e < neps; ++e) S[e] += part[(size_t)t * neps + e];
    free(part);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
