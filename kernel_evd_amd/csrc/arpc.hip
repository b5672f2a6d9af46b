// arpc.hip -- accelerated randomly pivoted Cholesky on one GPU, F-form.
//
// One round (README_arpclm.md:23-39; ks_arpclm.cu:793-1046):
//   k_reduce_d + k_weights_scan + k_draw   S' ~ d / sum d      (propose_block 193-224)
//   k_form_H                               H = K(S',S') - F(S',:) F(S',:)^T
//                                                                (extract_H + Dsyrk 820-874)
//   k_rejchol                              accept / Cholesky / L_c^-1  (rejection_sample 300-408)
//   k_gather_pivots                        X(S_i,:), F(S_i,:), L_c^-1 in the GEMM precision
//   k_downdate                             G = (K(:,S_i) - F F(S_i,:)^T) L_c^-T,
//                                          F(:,Nr:Nr+k) = G, d = max(d - rowsq(G), 0)
//                                                                (929-995, 455-478)
//   k_commit                               S = S u S_i, Nr += k   (1017-1019)
// The per-round state (rank, accepted count, stop flag) lives on the device,
// so a round issues no host synchronisation; the host polls a pinned copy of
// the state two rounds behind and stops launching once the device says so.
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gemm.h"
#include "internal.h"

namespace {

constexpr int SCAN_CHUNK = 4096;       // residual entries per scan block
constexpr int MAX_SCAN_BLOCKS = 8192;  // => n <= 33,554,432
constexpr int BMAX = 128;              // largest block size b

struct ArpcState {
    int32_t done;
    int32_t stop_reason;  // 1 rank reached, 2 zero residual mass, 4 pivot list exhausted
    int32_t k;            // accepted this round, after truncation to the target rank
    int32_t nprop;        // proposals this round
    int64_t Nr;           // rank so far
    int64_t rank;         // target
    int64_t inj_pos;      // next injected pivot
    int32_t round;        // rounds completed
    int32_t pad;
    uint64_t W;           // fixed-point residual mass of the last draw
};

// ---------------------------------------------------------------- sampling

// Per-block max and FP64 sum of the residual diagonal (fixed reduction shape,
// so the sums are bit-reproducible).
__global__ __launch_bounds__(256) void k_reduce_d(const double* __restrict__ d, int64_t n,
                                                  double* pmax, double* psum,
                                                  const ArpcState* st, int force)
{
    if (!force && st->done) return;
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
    double mx = 0.0, sm = 0.0;
#pragma unroll 4
    for (int j = 0; j < SCAN_CHUNK / 256; ++j) {
        int64_t i = base + threadIdx.x + 256 * j;
        if (i < n) {
            double v = d[i];
            mx = fmax(mx, v);
            sm += v;
        }
    }
    __shared__ double smx[4], ssm[4];
    mx = wave_max(mx);
    sm = wave_sum(sm);
    if ((threadIdx.x & 63) == 0) {
        smx[threadIdx.x >> 6] = mx;
        ssm[threadIdx.x >> 6] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = fmax(fmax(smx[0], smx[1]), fmax(smx[2], smx[3]));
        psum[blockIdx.x] = (ssm[0] + ssm[1]) + (ssm[2] + ssm[3]);
    }
}

// Fixed-point weights w_i = floor(d_i 2^(shift_base - e)), d_max = m 2^e,
// then a block-local inclusive scan (integer, hence order independent).
__global__ __launch_bounds__(256) void k_weights_scan(const double* __restrict__ d, int64_t n,
                                                      const double* __restrict__ pmax, int nb,
                                                      int shift_base, uint64_t* prefix,
                                                      uint64_t* btot, const ArpcState* st)
{
    if (st->done) return;
    __shared__ uint64_t w[SCAN_CHUNK];
    __shared__ uint64_t wtot[4];
    __shared__ double gmx_s[4];
    double mx = 0.0;
    for (int j = threadIdx.x; j < nb; j += 256) mx = fmax(mx, pmax[j]);
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) gmx_s[threadIdx.x >> 6] = mx;
    __syncthreads();
    double gmax = fmax(fmax(gmx_s[0], gmx_s[1]), fmax(gmx_s[2], gmx_s[3]));
    int e = 0;
    frexp(gmax, &e);
    int shift = shift_base - e;
    bool live = gmax > 0.0;
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
    for (int j = 0; j < SCAN_CHUNK / 256; ++j) {
        int idx = threadIdx.x + 256 * j;
        int64_t i = base + idx;
        uint64_t wi = 0;
        if (live && i < n) {
            double v = d[i];
            v = v > 0.0 ? v : 0.0;
            wi = (uint64_t)ldexp(v, shift);
        }
        w[idx] = wi;
    }
    __syncthreads();
    // each thread scans 16 consecutive weights
    constexpr int PER = SCAN_CHUNK / 256;
    uint64_t loc[PER];
    uint64_t run = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        run += w[threadIdx.x * PER + q];
        loc[q] = run;
    }
    // exclusive scan of the 256 thread totals
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wtot[wv] = incl;
    __syncthreads();
    uint64_t off = incl - run;
    for (int q = 0; q < wv; ++q) off += wtot[q];
#pragma unroll
    for (int q = 0; q < PER; ++q) w[threadIdx.x * PER + q] = loc[q] + off;
    __syncthreads();
    for (int j = 0; j < SCAN_CHUNK / 256; ++j) {
        int idx = threadIdx.x + 256 * j;
        int64_t i = base + idx;
        if (i < n) prefix[i] = w[idx];
    }
    if (threadIdx.x == 0) btot[blockIdx.x] = wtot[0] + wtot[1] + wtot[2] + wtot[3];
}

// Record the previous round's trace error, stop on zero mass, then draw the
// b proposals by integer inverse CDF (or take them from the injected list).
__global__ __launch_bounds__(256) void k_draw(const double* __restrict__ psum,
                                              const uint64_t* __restrict__ btot, int nb,
                                              const uint64_t* __restrict__ prefix, int64_t n,
                                              double trK, uint64_t seed, int round, int b,
                                              int64_t* Sprime, ArpcState* st, double* err,
                                              const int64_t* inject, int64_t n_inject)
{
    if (st->done) return;
    __shared__ uint64_t bpre[MAX_SCAN_BLOCKS];
    __shared__ int nprop_s;
    if (threadIdx.x == 0) {
        double s = 0.0;
        uint64_t run = 0;
        for (int j = 0; j < nb; ++j) {
            s += psum[j];
            run += btot[j];
            bpre[j] = run;
        }
        if (round > 0) err[round - 1] = s / trK;
        st->W = run;
        int np = b;
        if (inject) {
            int64_t left = n_inject - st->inj_pos;
            np = (int)(left < (int64_t)b ? left : (int64_t)b);
        }
        if (run == 0) {
            st->done = 1;
            st->stop_reason = 2;
            np = 0;
        } else if (np <= 0) {
            st->done = 1;
            st->stop_reason = 4;
            np = 0;
        }
        st->nprop = np;
        nprop_s = np;
    }
    __syncthreads();
    int np = nprop_s;
    if (np == 0) return;
    int t = threadIdx.x;
    if (inject) {
        if (t < np) Sprime[t] = inject[st->inj_pos + t];
        __syncthreads();
        if (t == 0) st->inj_pos += np;
        return;
    }
    if (t >= np) return;
    uint64_t W = bpre[nb - 1];
    uint64_t T = kevd_mulhi64(kevd_draw_u64(seed, KEVD_STREAM_PROPOSE, (uint32_t)round, (uint32_t)t), W);
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (bpre[mid] > T) hi = mid; else lo = mid + 1;
    }
    uint64_t Tl = T - (lo > 0 ? bpre[lo - 1] : 0);
    int64_t b0 = (int64_t)lo * SCAN_CHUNK;
    int64_t len = n - b0 < SCAN_CHUNK ? n - b0 : SCAN_CHUNK;
    int64_t l2 = 0, h2 = len - 1;
    while (l2 < h2) {
        int64_t mid = (l2 + h2) >> 1;
        if (prefix[b0 + mid] > Tl) h2 = mid; else l2 = mid + 1;
    }
    Sprime[t] = b0 + l2;
}

// ------------------------------------------------------------ H = residual

// H(p,q) = K(S'_p, S'_q) - sum_{m<Nr} F(S'_p, m) F(S'_q, m) in FP64, one
// 16 x 16 tile per block.  K from the FP64 centred points, direct
// differences.  Rows/columns past the proposal count are zero.
template <typename T>
__global__ __launch_bounds__(256) void k_form_H(const double* __restrict__ X64, int32_t dp,
                                                const T* __restrict__ F, int64_t ldf,
                                                const int64_t* __restrict__ Sprime,
                                                const ArpcState* st, double neg_inv_eps,
                                                double* H)
{
    if (st->done) return;
    int np = st->nprop;
    int64_t Nr = st->Nr;
    int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    int p = blockIdx.y * 16 + ty, q = blockIdx.x * 16 + tx;
    if (blockIdx.y * 16 >= np || blockIdx.x * 16 >= np) {
        if (p < BMAX && q < BMAX) H[p * BMAX + q] = 0.0;
        return;
    }
    __shared__ double Fp[16][65], Fq[16][65];
    int64_t sp = Sprime[p < np ? p : 0], sq = Sprime[q < np ? q : 0];
    double acc = 0.0;
    for (int64_t m0 = 0; m0 < Nr; m0 += 64) {
        // stage 16 rows x 64 columns of F(S'_p) and F(S'_q)
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            int rr = e >> 6, cc = e & 63;
            int pp = blockIdx.y * 16 + rr, qq = blockIdx.x * 16 + rr;
            int64_t m = m0 + cc;
            double vp = 0.0, vq = 0.0;
            if (m < Nr) {
                if (pp < np) vp = (double)F[Sprime[pp] * ldf + m];
                if (qq < np) vq = (double)F[Sprime[qq] * ldf + m];
            }
            Fp[rr][cc] = vp;
            Fq[rr][cc] = vq;
        }
        __syncthreads();
#pragma unroll 8
        for (int c = 0; c < 64; ++c) acc += Fp[ty][c] * Fq[tx][c];
        __syncthreads();
    }
    double kv;
    if (sp == sq) {
        kv = 1.0;
    } else {
        const double* a = X64 + sp * dp;
        const double* bq = X64 + sq * dp;
        double s = 0.0;
        for (int j = 0; j < dp; ++j) {
            double dd = a[j] - bq[j];
            s += dd * dd;
        }
        kv = exp(s * neg_inv_eps);
    }
    if (p < np && q < np) H[p * BMAX + q] = kv - acc;
    else H[p * BMAX + q] = 0.0;
}

// ------------------------------------------------------- rejection block

// Algorithm 2.1 on H (rejection_sample, ks_arpclm.cu:300-408) in one
// workgroup with H resident in LDS.  Every thread evaluates the acceptance
// test itself (same Philox counter, same LDS values), so a rejected proposal
// costs no barrier.  Column i of L is kept transposed in the unused upper
// triangle of H, its diagonal in Ld.  Work inside a step is spread over a
// 16 x 16 thread grid (no integer division).  After the loop the compacted
// inverse X = L_c^-1 is formed in place in the lower triangle by a
// right-looking triangular inversion (k sequential steps, each a parallel
// rank-1 update) instead of one thread per column.
template <typename T>
__global__ __launch_bounds__(256) void k_rejchol(const double* __restrict__ H, ArpcState* st,
                                                 uint64_t seed, int round, double tol, int inject,
                                                 const int64_t* __restrict__ Sprime,
                                                 int64_t* Si, T* LinvT, double* Linv64)
{
    if (st->done) return;
    __shared__ double Hs[BMAX][BMAX + 1];
    __shared__ double u[BMAX], Lcol[BMAX], Ld[BMAX];
    __shared__ int accpos[BMAX];
    int np = st->nprop;
    int tid = threadIdx.x;
    int tx = tid & 15, ty = tid >> 4;
    for (int p = ty; p < np; p += 16)
        for (int q = tx; q < np; q += 16) Hs[p][q] = H[p * BMAX + q];
    __syncthreads();
    if (tid < np) u[tid] = Hs[tid][tid];
    __syncthreads();
    int k = 0;
    for (int i = 0; i < np; ++i) {
        double rnd = kevd_u64_to_unit(kevd_draw_u64(seed, KEVD_STREAM_REJECT, (uint32_t)round, (uint32_t)i));
        double hii = Hs[i][i];
        bool acc = inject ? (hii > tol) : ((u[i] * rnd < hii) && (hii > tol));
        if (!acc) continue;
        double f = 1.0 / sqrt(hii);
        if (tid >= i && tid < np) Lcol[tid] = Hs[tid][i] * f;
        __syncthreads();
        for (int jj = i + 1 + ty; jj < np; jj += 16) {
            double lj = Lcol[jj];
            for (int mm = i + 1 + tx; mm <= jj; mm += 16) Hs[jj][mm] -= lj * Lcol[mm];
        }
        if (tid > i && tid < np) Hs[i][tid] = Lcol[tid];
        if (tid == i) Ld[i] = Lcol[i];
        if (tid == 0) accpos[k] = i;
        ++k;
        __syncthreads();
    }
    int64_t room = st->rank - st->Nr;
    if ((int64_t)k > room) k = (int)room;
    // L_c(a, c) = L(accpos[a], accpos[c]) = Hs[accpos[c]][accpos[a]] for a > c,
    // L_c(a, a) = Ld[accpos[a]].  X = L_c^-1 lives at Hs[a][t], a >= t
    // (compacted indices, strictly below the upper-triangle storage).
    for (int a = ty; a < k; a += 16)
        for (int t = tx; t <= a; t += 16) Hs[a][t] = (a == t) ? 1.0 : 0.0;
    __syncthreads();
    for (int a = 0; a < k; ++a) {
        double inv = 1.0 / Ld[accpos[a]];
        if (tid <= a) Hs[a][tid] *= inv;  // finalise row a (k <= 128 < 256 threads)
        __syncthreads();
        int pa = accpos[a];
        for (int j = a + 1 + ty; j < k; j += 16) {
            double l = Hs[pa][accpos[j]];
            for (int t = tx; t <= a; t += 16) Hs[j][t] -= l * Hs[a][t];
        }
        __syncthreads();
    }
    for (int e = tid; e < BMAX * BMAX; e += 256) {
        int a = e / BMAX, c = e % BMAX;
        double v = (a < k && c <= a) ? Hs[a][c] : 0.0;
        LinvT[e] = (T)v;
        Linv64[e] = v;
    }
    if (tid < BMAX) Si[tid] = tid < k ? Sprime[accpos[tid]] : -1;
    if (tid == 0) st->k = k;
}

// Pivot rows in the GEMM precision: X(S_i,:), |x|^2, F(S_i, 0:Nr).  Rows
// past k are zero.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_pivots(const T* __restrict__ Xsrc,
                                                       const double* __restrict__ nrm, int32_t dp,
                                                       const T* __restrict__ F, int64_t ldf,
                                                       const int64_t* __restrict__ Si,
                                                       const ArpcState* st, T* XS, double* nrmS,
                                                       T* FS)
{
    if (st->done) return;
    int c = blockIdx.x;
    int k = st->k;
    int64_t Nr = st->Nr;
    int64_t s = c < k ? Si[c] : -1;
    for (int j = threadIdx.x; j < dp; j += 256) XS[c * dp + j] = s >= 0 ? Xsrc[s * dp + j] : (T)0;
    if (threadIdx.x == 0) nrmS[c] = s >= 0 ? nrm[s] : 0.0;
    for (int64_t m = threadIdx.x; m < Nr; m += 256) FS[c * ldf + m] = s >= 0 ? F[s * ldf + m] : (T)0;
}

// ----------------------------------------------------------------- downdate

// The fused Step-3 update for BM rows (ks_arpclm.cu:929-995 and
// update_diag_block 455-478):
//   acc  = X_R X_{S_i}^T                    (MFMA, K = dp)
//   acc  = -K(R, S_i)                       (epilogue: norms, exp)
//   acc += F_R F_{S_i}^T                    (MFMA, K = Nr)
//   T    = -acc -> LDS
//   acc  = T L_c^-T                         (MFMA, K = k)
//   F(R, Nr:Nr+k) = acc;  d(R) = max(d(R) - rowsq(acc), 0)
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_downdate(const T* __restrict__ X, int32_t dp,
                                                  const double* __restrict__ nrm,
                                                  T* __restrict__ F, int64_t ldf,
                                                  double* __restrict__ d, int64_t n,
                                                  const T* __restrict__ XS,
                                                  const double* __restrict__ nrmS,
                                                  const int64_t* __restrict__ Si,
                                                  const T* __restrict__ FS,
                                                  const T* __restrict__ LinvT,
                                                  const ArpcState* st, GaussScale gs)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    constexpr int TP = BN + 16 / (int)sizeof(T);  // T tile pitch
    constexpr int LDS = G::LDS_ELEMS > BM * TP ? G::LDS_ELEMS : BM * TP;
    constexpr int WN = BN / (G::NTN * Mfma<T>::MT);
    __shared__ __attribute__((aligned(16))) T lds[LDS];
    __shared__ double rs[WN][BM];
    if (st->done) return;
    int k = st->k;
    if (k == 0) return;
    int64_t Nr = st->Nr;
    int64_t r0 = (int64_t)blockIdx.x * BM;
    int64_t rows = n - r0;
    G g;
    g.init();
    g.set_live_cols(k);
    // distance product
    g.mainloop(lds, X + r0 * dp, dp, rows, XS, dp, BN, dp);
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < Mfma<T>::ACC; ++v) {
                int64_t row = r0 + g.row_of(a, v);
                int col = g.col_of(b);
                double kv = 0.0;
                if (row < n && col < k) {
                    double D = nrm[row] + nrmS[col] - 2.0 * (double)g.acc[a][b][v];
                    D = D > 0.0 ? D : 0.0;
                    kv = (row == Si[col]) ? 1.0 : (double)gauss_val(D, gs, (T*)nullptr);
                }
                g.acc[a][b][v] = (T)(-kv);
            }
    // overlap with the current factor
    if (Nr > 0) g.mainloop_live(lds, F + r0 * ldf, ldf, rows, FS, ldf, BN, Nr);
    // T = K(R,S_i) - F_R F_{S_i}^T into LDS
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < Mfma<T>::ACC; ++v)
                lds[g.row_of(a, v) * TP + g.col_of(b)] = -g.acc[a][b][v];
    __syncthreads();
    g.zero();
    g.lds_global_tri_live(lds, TP, LinvT, BMAX, (k + 7) & ~7);  // L_c^-T is triangular
    // write the new columns and the squared row norms
    double part[G::NTM][Mfma<T>::ACC];
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < Mfma<T>::ACC; ++v) part[a][v] = 0.0;
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < Mfma<T>::ACC; ++v) {
                int64_t row = r0 + g.row_of(a, v);
                int col = g.col_of(b);
                T gv = g.acc[a][b][v];
                if (col < k) {
                    if (row < n) F[row * ldf + Nr + col] = gv;
                    part[a][v] += (double)gv * (double)gv;
                }
            }
    // reduce over the MT lanes that share a row (lane % MT = column)
    constexpr int MT = Mfma<T>::MT;
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < Mfma<T>::ACC; ++v) {
            double s = part[a][v];
#pragma unroll
            for (int o = 1; o < MT; o <<= 1) s += __shfl_xor(s, o, 64);
            if (g.li == 0) rs[g.wn][g.row_of(a, v)] = s;
        }
    __syncthreads();
    for (int r = threadIdx.x; r < BM; r += 256) {
        int64_t row = r0 + r;
        if (row < n) {
            double sq = 0.0;
#pragma unroll
            for (int q = 0; q < WN; ++q) sq += rs[q][r];
            double dd = d[row] - sq;
            d[row] = dd >= 0.0 ? dd : 0.0;
        }
    }
}

__global__ void k_commit(ArpcState* st, const int64_t* __restrict__ Si, int64_t* pivots,
                         int32_t* accepted, int round)
{
    if (st->done) return;
    int k = st->k;
    int64_t Nr = st->Nr;
    if ((int)threadIdx.x < k) pivots[Nr + threadIdx.x] = Si[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        accepted[round] = k;
        st->Nr = Nr + k;
        st->round = round + 1;
        if (st->Nr >= st->rank) {
            st->done = 1;
            st->stop_reason = 1;
        }
    }
}

__global__ void k_final_err(const double* psum, int nb, double trK, const ArpcState* st,
                            double* err)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int j = 0; j < nb; ++j) s += psum[j];
    if (st->round > 0) err[st->round - 1] = s / trK;
}

__global__ void k_fill(double* p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int ceil_log2(int64_t n)
{
    int c = 0;
    while (((int64_t)1 << c) < n) ++c;
    return c;
}

template <typename T>
kevd_status run_rounds(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds,
                       const int64_t* inject_dev, int64_t n_inject, ArpcState* st,
                       int64_t* pivots_dev, int32_t* acc_dev, double* err_dev, double* d)
{
    int64_t n = ctx->n;
    int nb = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    double *pmax, *psum, *H, *nrmS, *Linv64;
    uint64_t *prefix, *btot;
    int64_t *Sprime, *Si;
    T *XS, *FS, *LinvT;
    KEVD_TRY(ctx_buf_t(ctx, "a_pmax", nb, &pmax));
    KEVD_TRY(ctx_buf_t(ctx, "a_psum", nb, &psum));
    KEVD_TRY(ctx_buf_t(ctx, "a_prefix", n, &prefix));
    KEVD_TRY(ctx_buf_t(ctx, "a_btot", nb, &btot));
    KEVD_TRY(ctx_buf_t(ctx, "a_Sprime", BMAX, &Sprime));
    KEVD_TRY(ctx_buf_t(ctx, "a_Si", BMAX, &Si));
    KEVD_TRY(ctx_buf_t(ctx, "a_H", BMAX * BMAX, &H));
    KEVD_TRY(ctx_buf_t(ctx, "a_Linv64", BMAX * BMAX, &Linv64));
    KEVD_TRY(ctx_buf_t(ctx, "a_nrmS", BMAX, &nrmS));
    KEVD_TRY(ctx_buf_t(ctx, "a_XS", (size_t)BMAX * ctx->dp, &XS));
    KEVD_TRY(ctx_buf_t(ctx, "a_FS", (size_t)BMAX * ctx->ldf, &FS));
    KEVD_TRY(ctx_buf_t(ctx, "a_LinvT", BMAX * BMAX, &LinvT));

    const T* Xsrc = std::is_same<T, float>::value ? (const T*)ctx->X32 : (const T*)ctx->X64;
    const double* nrm = std::is_same<T, float>::value ? ctx->nrm32 : ctx->nrm64;
    T* F = (T*)ctx->F;
    GaussScale gs{-1.0 / p->eps, -1.4426950408889634074 / p->eps};
    int shift_base = 62 - ceil_log2(n);
    double trK = (double)n;  // K_ii = 1
    hipStream_t s = ctx->stream;

    ArpcState* host_st = nullptr;
    KEVD_TRY(ctx_pinned(ctx, 4 * sizeof(ArpcState), (void**)&host_st));
    std::memset(host_st, 0, 4 * sizeof(ArpcState));
    hipEvent_t ev[4];
    for (int i = 0; i < 4; ++i) KEVD_HIP(ctx, hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    kevd_status status = KEVD_OK;
    unsigned gdown = (unsigned)((n + Cfg<T>::BM - 1) / Cfg<T>::BM);
    for (int rd = 0; rd < max_rounds; ++rd) {
        {
            ProfScope ps(ctx, TC_SAMPLE);
            k_reduce_d<<<nb, 256, 0, s>>>(d, n, pmax, psum, st, 0);
            k_weights_scan<<<nb, 256, 0, s>>>(d, n, pmax, nb, shift_base, prefix, btot, st);
            k_draw<<<1, 256, 0, s>>>(psum, btot, nb, prefix, n, trK, p->seed, rd, p->block, Sprime,
                                     st, err_dev, inject_dev, n_inject);
        }
        {
            ProfScope ps(ctx, TC_REJECT);
            k_form_H<T><<<dim3(BMAX / 16, BMAX / 16), 256, 0, s>>>(ctx->X64, ctx->dp, F, ctx->ldf,
                                                                  Sprime, st, gs.neg_inv_eps, H);
            k_rejchol<T><<<1, 256, 0, s>>>(H, st, p->seed, rd, p->tol, inject_dev != nullptr,
                                           Sprime, Si, LinvT, Linv64);
            k_gather_pivots<T><<<BMAX, 256, 0, s>>>(Xsrc, nrm, ctx->dp, F, ctx->ldf, Si, st, XS,
                                                    nrmS, FS);
        }
        {
            ProfScope ps(ctx, TC_DOWNDATE);
            k_downdate<T><<<gdown, 256, 0, s>>>(Xsrc, ctx->dp, nrm, F, ctx->ldf, d, n, XS, nrmS, Si,
                                                FS, LinvT, st, gs);
        }
        k_commit<<<1, BMAX, 0, s>>>(st, Si, pivots_dev, acc_dev, rd);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            status = kevd_hip_fail(ctx, le, "ARPC round launch", __FILE__, __LINE__);
            break;
        }
        int slot = rd & 3;
        hipError_t ce = hipMemcpyAsync(&host_st[slot], st, sizeof(ArpcState), hipMemcpyDeviceToHost, s);
        if (ce == hipSuccess) ce = hipEventRecord(ev[slot], s);
        if (ce != hipSuccess) {
            status = kevd_hip_fail(ctx, ce, "ARPC status copy", __FILE__, __LINE__);
            break;
        }
        if (rd >= 2) {
            int old = (rd - 2) & 3;
            hipError_t we = hipEventSynchronize(ev[old]);
            if (we != hipSuccess) {
                status = kevd_hip_fail(ctx, we, "ARPC status wait", __FILE__, __LINE__);
                break;
            }
            if (host_st[old].done) break;
        }
    }
    if (status == KEVD_OK) {
        k_reduce_d<<<nb, 256, 0, s>>>(d, n, pmax, psum, st, 1);
        k_final_err<<<1, 64, 0, s>>>(psum, nb, trK, st, err_dev);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) status = kevd_hip_fail(ctx, le, "ARPC final error", __FILE__, __LINE__);
    }
    hipError_t se = hipStreamSynchronize(s);
    if (status == KEVD_OK && se != hipSuccess) status = kevd_hip_fail(ctx, se, "ARPC sync", __FILE__, __LINE__);
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
    return status;
}

}  // namespace

extern "C" kevd_status kevd_arpc(kevd_ctx* ctx, const kevd_arpc_params* p, kevd_arpc_result* res)
{
    if (!ctx || !p || !res) return kevd_fail(ctx, KEVD_EINVAL, "arpc: null argument");
    if (ctx->n < 2) return kevd_fail(ctx, KEVD_EINVAL, "arpc: no points");
    if (!(p->eps > 0.0) || !std::isfinite(p->eps))
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: eps must be positive and finite");
    if (p->rank < 1 || p->rank > ctx->n)
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: need 1 <= rank <= n");
    if (p->block < 1 || p->block > BMAX)
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: block must be in [1, 128]");
    if (!(p->tol >= 0.0)) return kevd_fail(ctx, KEVD_EINVAL, "arpc: tol must be >= 0");
    if (p->prec != KEVD_FP64 && p->prec != KEVD_MIXED)
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: unknown precision");
    if (ctx->n > (int64_t)SCAN_CHUNK * MAX_SCAN_BLOCKS)
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: n exceeds 33,554,432");
    if (p->inject_pivots) {
        if (p->n_inject < 1) return kevd_fail(ctx, KEVD_EINVAL, "arpc: empty pivot list");
        for (int64_t i = 0; i < p->n_inject; ++i)
            if (p->inject_pivots[i] < 0 || p->inject_pivots[i] >= ctx->n)
                return kevd_fail(ctx, KEVD_EINVAL, "arpc: injected pivot out of range");
    }
    int max_rounds = p->max_rounds > 0 ? p->max_rounds
                                       : (int)(4 * ((p->rank + p->block - 1) / p->block) + 64);
    if (!res->pivots || !res->accepted || !res->trace_err || res->rounds_cap < max_rounds)
        return kevd_fail(ctx, KEVD_EINVAL, "arpc: result arrays missing or rounds_cap < max_rounds");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->tm = kevd_timings{};
    prof_reset(ctx);

    // factor storage
    int64_t ldf = KEVD_ROUND_UP(p->rank, 64);
    size_t esz = p->prec == KEVD_MIXED ? sizeof(float) : sizeof(double);
    size_t fbytes = (size_t)ctx->n * ldf * esz;
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->F && ctx->F_bytes < fbytes) {
        KEVD_HIP(ctx, hipFree(ctx->F));
        ctx->F = nullptr;
        ctx->F_bytes = 0;
    }
    if (!ctx->F) {
        KEVD_HIP(ctx, hipMalloc(&ctx->F, fbytes));
        ctx->F_bytes = fbytes;
    }
    ctx->have_factor = false;
    ctx->ldf = ldf;
    ctx->fprec = p->prec;
    ctx->rank = 0;
    ctx->pivots.clear();

    double* d = nullptr;
    ArpcState* st = nullptr;
    int64_t* piv = nullptr;
    int32_t* accd = nullptr;
    double* errd = nullptr;
    int64_t* inj = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "a_d", ctx->n, &d));
    KEVD_TRY(ctx_buf_t(ctx, "a_state", 1, &st));
    KEVD_TRY(ctx_buf_t(ctx, "a_piv", p->rank, &piv));
    KEVD_TRY(ctx_buf_t(ctx, "a_acc", max_rounds, &accd));
    KEVD_TRY(ctx_buf_t(ctx, "a_err", max_rounds, &errd));
    if (p->inject_pivots) {
        KEVD_TRY(ctx_buf_t(ctx, "a_inj", p->n_inject, &inj));
        KEVD_HIP(ctx, hipMemcpyAsync(inj, p->inject_pivots, p->n_inject * sizeof(int64_t),
                                     hipMemcpyHostToDevice, ctx->stream));
    }
    ArpcState init{};
    init.rank = p->rank;
    KEVD_HIP(ctx, hipMemcpyAsync(st, &init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipMemsetAsync(accd, 0, max_rounds * sizeof(int32_t), ctx->stream));
    KEVD_HIP(ctx, hipMemsetAsync(errd, 0, max_rounds * sizeof(double), ctx->stream));
    k_fill<<<(unsigned)((ctx->n + 255) / 256), 256, 0, ctx->stream>>>(d, ctx->n, 1.0);
    KEVD_HIP(ctx, hipGetLastError());

    kevd_status s = p->prec == KEVD_MIXED
                        ? run_rounds<float>(ctx, p, max_rounds, inj, p->n_inject, st, piv, accd, errd, d)
                        : run_rounds<double>(ctx, p, max_rounds, inj, p->n_inject, st, piv, accd, errd, d);
    if (s != KEVD_OK) return s;
    ArpcState fin{};
    KEVD_HIP(ctx, hipMemcpy(&fin, st, sizeof(fin), hipMemcpyDeviceToHost));
    res->rank = fin.Nr;
    res->rounds = fin.round;
    KEVD_HIP(ctx, hipMemcpy(res->pivots, piv, fin.Nr * sizeof(int64_t), hipMemcpyDeviceToHost));
    KEVD_HIP(ctx, hipMemcpy(res->accepted, accd, max_rounds * sizeof(int32_t), hipMemcpyDeviceToHost));
    KEVD_HIP(ctx, hipMemcpy(res->trace_err, errd, max_rounds * sizeof(double), hipMemcpyDeviceToHost));
    ctx->rank = fin.Nr;
    ctx->pivots.assign(res->pivots, res->pivots + fin.Nr);
    ctx->have_factor = fin.Nr > 0;
    ctx->factor_eps = p->eps;
    auto t1 = std::chrono::steady_clock::now();
    ctx->tm.arpc_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    // Algorithmic flop of the downdate: per round 2 n k Nr (overlap) +
    // 2 n k dim (distances) + n k^2 (triangular solve) + 2 n k (norms).
    double fl = 0.0, nr = 0.0;
    for (int r = 0; r < fin.round; ++r) {
        double k = res->accepted[r];
        fl += (double)ctx->n * k * (2.0 * nr + 2.0 * ctx->dim + k + 2.0);
        nr += k;
    }
    ctx->tm.downdate_flop = fl;
    prof_resolve(ctx);
    return KEVD_OK;
}
