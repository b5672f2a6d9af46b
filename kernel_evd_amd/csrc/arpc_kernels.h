// arpc_kernels.h -- device bodies of one accelerated-RPCholesky round,
// shared by the single-problem kernels (arpc.hip) and the batched ensemble
// kernels (ensemble.hip).  Each body works on one problem given by plain
// pointers and sizes; the block index it would take from blockIdx is an
// argument, so a batched kernel can map its blocks onto (problem, block)
// pairs.  The round index comes from the device state, which makes a round
// free of host-side per-round arguments (capturable as a graph).
//
// One round (README_arpclm.md:23-39; ks_arpclm.cu:793-1046):
//   reduce_d + weights_scan + draw   S' ~ d / sum d      (propose_block 193-224)
//   form_H                           H = K(S',S') - F(S',:) F(S',:)^T
//                                                         (extract_H + Dsyrk 820-874)
//   rejchol                          accept / Cholesky / L_c^-1  (rejection_sample 300-408)
//   gather_pivots                    X(S_i,:), F(S_i,:), L_c^-1 in the GEMM precision
//   downdate                         G = (K(:,S_i) - F F(S_i,:)^T) L_c^-T,
//                                    F(:,Nr:Nr+k) = G, d = max(d - rowsq(G), 0)
//                                                         (929-995, 455-478)
//   commit                           S = S u S_i, Nr += k   (1017-1019)
#pragma once

#include <type_traits>

#include "bf16x6.h"
#include "gemm.h"
#include "internal.h"

namespace arpck {

constexpr int SCAN_CHUNK = 4096;       // residual entries per scan block
constexpr int MAX_SCAN_BLOCKS = 8192;  // => n <= 33,554,432
constexpr int BMAX = 128;              // largest block size b

struct ArpcState {
    int32_t done;
    int32_t stop_reason;  // 1 rank reached, 2 zero residual mass, 3 max residual <= tol,
                          // 4 pivot list exhausted, 8 round budget used (ensemble)
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
__device__ inline void reduce_d_body(const double* __restrict__ d, int64_t n, double* pmax, double* psum,
                                     const ArpcState* st, int force, int blk)
{
    if (!force && st->done) return;
    int64_t base = (int64_t)blk * SCAN_CHUNK;
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
        pmax[blk] = fmax(fmax(smx[0], smx[1]), fmax(smx[2], smx[3]));
        psum[blk] = (ssm[0] + ssm[1]) + (ssm[2] + ssm[3]);
    }
}

// Fixed-point weights w_i = floor(d_i 2^(shift_base - e)), d_max = m 2^e,
// then a block-local inclusive scan (integer, hence order independent).
__device__ inline void weights_scan_body(const double* __restrict__ d, int64_t n,
                                         const double* __restrict__ pmax, int nb, int shift_base,
                                         uint64_t* prefix, uint64_t* btot, const ArpcState* st, int blk)
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
    int64_t base = (int64_t)blk * SCAN_CHUNK;
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
    if (threadIdx.x == 0) btot[blk] = wtot[0] + wtot[1] + wtot[2] + wtot[3];
}

// Record the previous round's trace error, stop on zero mass or when no
// residual entry exceeds tol (Algorithm 2.1 accepts only H_ii > tol, and
// H_ii is at most the residual diagonal, so no later round could accept),
// then draw the b proposals by integer inverse CDF (or take them from the
// injected list).  The round index is the device state's, so a captured
// round replays as is.
__device__ inline void draw_body(const double* __restrict__ psum, const double* __restrict__ pmax,
                                 const uint64_t* __restrict__ btot, int nb,
                                 const uint64_t* __restrict__ prefix, int64_t n, double trK, uint64_t seed,
                                 int b, double tol, int64_t* Sprime, ArpcState* st, double* err,
                                 const int64_t* inject, int64_t n_inject)
{
    if (st->done) return;
    const int round = st->round;
    __shared__ uint64_t bpre[MAX_SCAN_BLOCKS];
    __shared__ double spsum[MAX_SCAN_BLOCKS];
    __shared__ double wmax[4];
    __shared__ int nprop_s;
    // the block partials are loaded by every thread, then summed in block
    // order by one (the same order, so the sums are unchanged)
    double mx = 0.0;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        bpre[j] = btot[j];
        spsum[j] = psum[j];
        mx = fmax(mx, pmax[j]);
    }
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double dmax = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        double s = 0.0;
        uint64_t run = 0;
        for (int j = 0; j < nb; ++j) {
            s += spsum[j];
            run += bpre[j];
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
        } else if (!(dmax > tol)) {
            st->done = 1;
            st->stop_reason = 3;
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
// 16 x 16 tile per block of FORMH_THREADS.  K from the FP64 centred points,
// direct differences.  Rows/columns past the proposal count are zero.  With
// frow_by_position the rows of F are the proposals' own rows (row p).
// Each of the 8 waves takes an eighth of m (64-column chunks, staged in its
// own LDS slice, so no block barrier per chunk; the next chunk's loads are
// in flight while the current one is multiplied) and each lane a 2 x 2
// block of the tile; the partial sums are added in wave order at the end.
constexpr int FORMH_THREADS = 512;
template <typename T>
__device__ inline void form_H_body(const double* __restrict__ X64, int32_t dp, const T* __restrict__ F,
                                   int64_t ldf, const int64_t* __restrict__ Sprime, const ArpcState* st,
                                   double neg_inv_eps, double* H, int bx, int by, bool frow_by_position = false)
{
    constexpr int NW = FORMH_THREADS / 64;
    if (st->done) return;
    const int np = st->nprop;
    const int64_t Nr = st->Nr;
    const int tid = threadIdx.x;
    // threads < 256 own one entry (ty, tx) of the tile each
    const bool owner = tid < 256;
    const int ty = (tid >> 4) & 15, tx = tid & 15;
    const int p = by * 16 + ty, q = bx * 16 + tx;
    if (by * 16 >= np || bx * 16 >= np) {
        if (owner && p < BMAX && q < BMAX) H[p * BMAX + q] = 0.0;
        return;
    }
    __shared__ double Fp[NW][16][65], Fq[NW][16][65];
    __shared__ double part[NW][16][17];
    const int wv = tid >> 6, lane = tid & 63;
    // this wave's m range, in whole 64-column chunks
    const int64_t nch = (Nr + 63) / 64, per = (nch + NW - 1) / NW;
    const int64_t c_beg = wv * per, c_end = c_beg + per < nch ? c_beg + per : nch;
    // lane -> rows 2 rp, 2 rp + 1 of the tile, columns 2 cp, 2 cp + 1
    const int rp = lane >> 3, cp = lane & 7;
    // staging: lane holds column `lane` of the chunk for all 16 rows
    const T* rowp[16];
    const T* rowq[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int pp = by * 16 + r, qq = bx * 16 + r;
        rowp[r] = pp < np ? F + (frow_by_position ? (int64_t)pp : Sprime[pp]) * ldf : nullptr;
        rowq[r] = qq < np ? F + (frow_by_position ? (int64_t)qq : Sprime[qq]) * ldf : nullptr;
    }
    T vp[16], vq[16];
    auto gload = [&](int64_t c) {
        const int64_t m = c * 64 + lane;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            vp[r] = (m < Nr && rowp[r]) ? rowp[r][m] : (T)0;
            vq[r] = (m < Nr && rowq[r]) ? rowq[r][m] : (T)0;
        }
    };
    if (c_beg < c_end) gload(c_beg);
    double acc00 = 0.0, acc01 = 0.0, acc10 = 0.0, acc11 = 0.0;
    for (int64_t c = c_beg; c < c_end; ++c) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            Fp[wv][r][lane] = (double)vp[r];
            Fq[wv][r][lane] = (double)vq[r];
        }
        if (c + 1 < c_end) gload(c + 1);
        // the slice is this wave's own: its writes are visible to its reads
        // in order, with no block barrier
        __builtin_amdgcn_wave_barrier();
#pragma unroll 8
        for (int cc = 0; cc < 64; ++cc) {
            const double a0 = Fp[wv][2 * rp][cc], a1 = Fp[wv][2 * rp + 1][cc];
            const double b0 = Fq[wv][2 * cp][cc], b1 = Fq[wv][2 * cp + 1][cc];
            acc00 += a0 * b0;
            acc01 += a0 * b1;
            acc10 += a1 * b0;
            acc11 += a1 * b1;
        }
        __builtin_amdgcn_wave_barrier();
    }
    part[wv][2 * rp][2 * cp] = acc00;
    part[wv][2 * rp][2 * cp + 1] = acc01;
    part[wv][2 * rp + 1][2 * cp] = acc10;
    part[wv][2 * rp + 1][2 * cp + 1] = acc11;
    __syncthreads();
    if (!owner) return;
    double acc = part[0][ty][tx];
#pragma unroll
    for (int w = 1; w < NW; ++w) acc += part[w][ty][tx];
    const int64_t sp = Sprime[p < np ? p : 0], sq = Sprime[q < np ? q : 0];
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
// workgroup of 256 threads.  The Schur complement lives in registers: on a
// 16 x 16 thread grid, thread (tx, ty) owns H(ty + 16a, tx + 16b) for
// a, b < 8, so a step's rank-1 update is 64 register FMAs after 16 LDS
// reads.  Every wave keeps its own copy of the current diagonal in LDS and
// updates it with the same FMA the owners apply to their registers, and the
// proposals' uniforms sit in LDS, so every thread evaluates the acceptance
// test itself: a rejected proposal costs no barrier and an accepted one
// costs one (its column of L published).  Column c of L (the c-th
// acceptance) is kept by proposal row in Lst[c][.].  The compacted inverse X = L_c^-1 is then
// formed in registers on the same grid by a right-looking triangular
// inversion: step c finalises row c (its owners divide by L_c(c, c) and
// publish it) and every later row subtracts L_c(a, c) times it, one barrier
// per step (the published row is double buffered).
template <typename T>
__device__ __attribute__((always_inline)) inline void rejchol_body(const double* __restrict__ H, ArpcState* st, uint64_t seed, double tol,
                                                                   int inject, const int64_t* __restrict__ Sprime,
                                                                   int64_t* Si, T* LinvT, double* Linv64,
                                                                   int32_t* Spos = nullptr)
{
    static_assert(BMAX == 128, "the 16 x 16 grid of 8 x 8 register blocks covers 128 x 128");
    if (st->done) return;
    const int round = st->round;
    __shared__ double Lst[BMAX][BMAX];
    __shared__ double dgw[4][BMAX], u0[BMAX], rndv[BMAX], xrow[2][BMAX], dinv[BMAX];
    __shared__ int accpos[BMAX];
    const int np = st->nprop;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    double h[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) h[a][b] = H[(ty + 16 * a) * BMAX + tx + 16 * b];
    const int wv = tid >> 6, lane = tid & 63;
    // this wave's diagonal copy (H is zero past the proposals)
    dgw[wv][lane] = H[lane * BMAX + lane];
    dgw[wv][lane + 64] = H[(lane + 64) * BMAX + lane + 64];
    if (tx == ty) {
#pragma unroll
        for (int a = 0; a < 8; ++a) u0[ty + 16 * a] = h[a][a];
    }
    if (tid < BMAX)
        rndv[tid] = kevd_u64_to_unit(kevd_draw_u64(seed, KEVD_STREAM_REJECT, (uint32_t)round, (uint32_t)tid));
    __syncthreads();
    int k = 0;
    // Proposal i lives in register column bi = i / 16 of its owners; the
    // outer loop over bi is unrolled, so the owners read h[a][bi] from a
    // fixed register (a runtime column index costs a select chain over all
    // eight).  Column k of L is written whole: its rows below i are never
    // read (the update, the diagonal copies and the inverse below read rows
    // past the pivot only).
#pragma unroll
    for (int bi = 0; bi < 8; ++bi) {
        const int i_end = np < 16 * (bi + 1) ? np : 16 * (bi + 1);
        for (int i = 16 * bi; i < i_end; ++i) {
            const double hii = dgw[wv][i];
            const bool acc = inject ? (hii > tol) : ((u0[i] * rndv[i] < hii) && (hii > tol));
            if (!acc) continue;
            const double f = 1.0 / sqrt(hii);
            if (tx == (i & 15)) {
#pragma unroll
                for (int a = 0; a < 8; ++a) Lst[k][ty + 16 * a] = h[a][bi] * f;
            }
            if (tid == 0) accpos[k] = i;
            __syncthreads();
            // every row of column k was written: plain reads, then selects
            double lr[8], lc[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                int row = ty + 16 * a, col = tx + 16 * a;
                const double vr = Lst[k][row], vc = Lst[k][col];
                lr[a] = (row > i && row < np) ? vr : 0.0;
                lc[a] = (col > i && col < np) ? vc : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) h[a][b] -= lr[a] * lc[b];
            // the wave's diagonal copy, as the owners' h(j, j) -= L(j) L(j)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int j = lane + 64 * q;
                if (j > i && j < np) {
                    const double l = Lst[k][j];
                    dgw[wv][j] -= l * l;
                }
            }
            ++k;
        }
    }
    int64_t room = st->rank - st->Nr;
    if ((int64_t)k > room) k = (int)room;
    // X = L_c^-1 with L_c(a, c) = Lst[c][accpos[a]] (a >= c); thread (tx, ty)
    // owns X(ty + 16a, tx + 16b).
    int apos[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) apos[a] = ty + 16 * a < k ? accpos[ty + 16 * a] : 0;
    // the diagonal's reciprocals, all at once (one division per thread)
    if (tid < k) dinv[tid] = 1.0 / Lst[tid][accpos[tid]];
    __syncthreads();
    double x[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) x[a][b] = (ty + 16 * a == tx + 16 * b && ty + 16 * a < k) ? 1.0 : 0.0;
    // row c lives in register row ac = c / 16 of its owners: unrolled as in
    // the acceptance loop
#pragma unroll
    for (int ac = 0; ac < 8; ++ac) {
        const int c_end = k < 16 * (ac + 1) ? k : 16 * (ac + 1);
        for (int c = 16 * ac; c < c_end; ++c) {
            const int buf = c & 1;
            if (ty == (c & 15)) {
                const double inv = dinv[c];
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    x[ac][b] *= inv;
                    xrow[buf][tx + 16 * b] = x[ac][b];
                }
            }
            __syncthreads();
            double l[8], xr[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                int row = ty + 16 * a;
                const double v = Lst[c][apos[a]];
                l[a] = (row > c && row < k) ? v : 0.0;
                xr[a] = xrow[buf][tx + 16 * a];
            }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) x[a][b] -= l[a] * xr[b];
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            int row = ty + 16 * a, col = tx + 16 * b;
            double v = (row < k && col <= row) ? x[a][b] : 0.0;
            LinvT[row * BMAX + col] = (T)v;
            Linv64[row * BMAX + col] = v;
        }
    if (tid < BMAX) Si[tid] = tid < k ? Sprime[accpos[tid]] : -1;
    // positions in S' of the accepted proposals, when asked for
    if (Spos && tid < BMAX) Spos[tid] = tid < k ? accpos[tid] : 0;
    if (tid == 0) st->k = k;
}

// The bf16x6 parts of the pivot rows follow FS (BMAX x ldf FP32) in the
// same buffer, chunk-major for the downdate's loads: the parts of columns
// [32 c, 32 c + 32) of all BMAX rows are one contiguous 24 KB block in MFMA
// fragment order (X6_CHUNK elements, x6_idx), so that every load of the
// downdate reads whole cache lines.
__host__ __device__ inline size_t x6_fs_bytes(int64_t ldf) { return (size_t)BMAX * ldf * (4 + 3 * 2); }
__device__ inline __bf16* x6_fsx(float* FS, int64_t ldf) { return reinterpret_cast<__bf16*>(FS + (int64_t)BMAX * ldf); }
__device__ inline const __bf16* x6_fsx(const float* FS, int64_t ldf)
{
    return reinterpret_cast<const __bf16*>(FS + (int64_t)BMAX * ldf);
}

// Likewise the parts of the pivot points X(S_i,:) follow XS (BMAX x dp FP32),
// chunk-major: the downdate's distance product runs as bf16x6 too.
__host__ __device__ inline size_t x6_xs_bytes(int32_t dp) { return (size_t)BMAX * dp * (4 + 3 * 2); }
__device__ inline __bf16* x6_xsx(float* XS, int32_t dp) { return reinterpret_cast<__bf16*>(XS + (int64_t)BMAX * dp); }
__device__ inline const __bf16* x6_xsx(const float* XS, int32_t dp)
{
    return reinterpret_cast<const __bf16*>(XS + (int64_t)BMAX * dp);
}

// And the parts of L_c^-1 follow LinvT (BMAX x BMAX FP32, row-major, lower
// triangular), chunk-major in the same fragment order: the downdate's
// product with L_c^-T runs as bf16x6.
__host__ __device__ inline size_t x6_linv_bytes() { return (size_t)BMAX * BMAX * (4 + 3 * 2); }
__device__ inline __bf16* x6_lx(float* LinvT) { return reinterpret_cast<__bf16*>(LinvT + BMAX * BMAX); }
__device__ inline const __bf16* x6_lx(const float* LinvT)
{
    return reinterpret_cast<const __bf16*>(LinvT + BMAX * BMAX);
}

// Pivot rows in the GEMM precision: X(S_i,:), |x|^2, F(S_i, 0:Nr).  Rows
// past k are zero.  X6: also the bf16x6 parts of X(S_i,:) and of F(S_i,:),
// the latter zero from Nr up to a multiple of 32, and of row c of L_c^-1.
template <typename T, bool X6 = false>
__device__ inline void gather_pivots_body(const T* __restrict__ Xsrc, const double* __restrict__ nrm,
                                          int32_t dp, const T* __restrict__ F, int64_t ldf,
                                          const int64_t* __restrict__ Si, const ArpcState* st, T* XS,
                                          double* nrmS, T* FS, T* LinvT, int c)
{
    if (st->done) return;
    int k = st->k;
    int64_t Nr = st->Nr;
    int64_t s = c < k ? Si[c] : -1;
    for (int j = threadIdx.x; j < dp; j += 256) XS[c * dp + j] = s >= 0 ? Xsrc[s * dp + j] : (T)0;
    if (threadIdx.x == 0) nrmS[c] = s >= 0 ? nrm[s] : 0.0;
    for (int64_t m = threadIdx.x; m < Nr; m += 256) FS[c * ldf + m] = s >= 0 ? F[s * ldf + m] : (T)0;
    if constexpr (X6 && std::is_same<T, float>::value) {
        __bf16* xx = x6_xsx(XS, dp);  // dp is a multiple of 32
        for (int j = threadIdx.x; j < dp; j += 256) {
            __bf16 h, md, l;
            bf16_split3(s >= 0 ? Xsrc[s * dp + j] : 0.0f, h, md, l);
            __bf16* e = xx + (j / 32) * X6_CHUNK + x6_idx(c, j % 32);
            e[0] = h;
            e[X6_CHUNK / 3] = md;
            e[2 * X6_CHUNK / 3] = l;
        }
        __bf16* fx = x6_fsx(FS, ldf);
        const int64_t kp = (Nr + 31) / 32 * 32;
        for (int64_t m = threadIdx.x; m < kp; m += 256) {
            float v = s >= 0 && m < Nr ? F[s * ldf + m] : 0.0f;
            __bf16 h, md, l;
            bf16_split3(v, h, md, l);
            __bf16* e = fx + (m / 32) * X6_CHUNK + x6_idx(c, (int)(m % 32));
            e[0] = h;
            e[X6_CHUNK / 3] = md;
            e[2 * X6_CHUNK / 3] = l;
        }
        // row c of L_c^-1 (zero past the diagonal and for c >= k)
        __bf16* lx = x6_lx(LinvT);
        for (int m = threadIdx.x; m < BMAX; m += 256) {
            __bf16 h, md, l;
            bf16_split3(LinvT[c * BMAX + m], h, md, l);
            __bf16* e = lx + (m / 32) * X6_CHUNK + x6_idx(c, m % 32);
            e[0] = h;
            e[X6_CHUNK / 3] = md;
            e[2 * X6_CHUNK / 3] = l;
        }
    }
}

// ----------------------------------------------------------------- downdate

// LDS of one downdate workgroup: the GEMM stages, later the T tile, and the
// per-wave row-norm partials.
template <typename T> struct DowndateLds {
    typedef typename Cfg<T>::G G;
    static constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    static constexpr int TP = BN + 16 / (int)sizeof(T);  // T tile pitch
    static constexpr int ELEMS = G::LDS_ELEMS > BM * TP ? G::LDS_ELEMS : BM * TP;
    static constexpr int WN = BN / (G::NTN * G::M::MT);
};

// Epilogue constants of one downdate row block, staged in LDS before the
// distance product so their global loads overlap it: the block's row norms
// and residual diagonal, and the pivots' norms and indices.
template <typename T> struct DowndateEp {
    double nrm[Cfg<T>::BM];
    double d[Cfg<T>::BM];
    double nrmS[Cfg<T>::BN];
    int64_t Si[Cfg<T>::BN];
};

// First stage of the Step-3 update of BM rows: the epilogue constants go to
// LDS and acc = -K(R, S_i) (distance product on MFMA, K = dp, then norms
// and the compensated exp; exactly 1 on a row's own pivot).
// XSx (FP32 only, else null): the pivot points' bf16x6 parts, chunk-major;
// the distance product then runs as bf16x6 on the bf16 matrix cores.
template <typename T>
__device__ inline void dd_prologue(typename Cfg<T>::G& g, const T* __restrict__ X, int32_t dp,
                                   const double* __restrict__ nrm, const double* __restrict__ d, int64_t n,
                                   const T* __restrict__ XS, const double* __restrict__ nrmS,
                                   const int64_t* __restrict__ Si, int k, GaussScale gs, int64_t r0, T* lds,
                                   DowndateEp<T>* ep, const __bf16* XSx = nullptr)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    int64_t rows = n - r0;
    for (int i = threadIdx.x; i < BM; i += 256) {
        ep->nrm[i] = r0 + i < n ? nrm[r0 + i] : 0.0;
        ep->d[i] = r0 + i < n ? d[r0 + i] : 0.0;
    }
    for (int i = threadIdx.x; i < BN; i += 256) {
        ep->nrmS[i] = nrmS[i];
        ep->Si[i] = Si[i];
    }
    g.init();
    g.set_live_cols(k);
    // distance product (its barriers also publish ep)
    if constexpr (std::is_same<T, float>::value) {
        if (XSx)
            x6_nt_mainloop_live<G>(g, reinterpret_cast<unsigned char*>(lds), X + r0 * dp, dp, rows, XSx, dp);
        else
            g.mainloop(lds, X + r0 * dp, dp, rows, XS, dp, BN, dp);
    } else {
        g.mainloop(lds, X + r0 * dp, dp, rows, XS, dp, BN, dp);
    }
    // The lane's row and column constants are read once; the element loop
    // is then branch-free (out-of-range rows and columns are masked to 0).
    double nr[G::NTM][G::M::ACC], ns[G::NTN];
    int rl[G::NTM][G::M::ACC], sp[G::NTN];
    bool rv[G::NTM][G::M::ACC], cv[G::NTN];
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < G::M::ACC; ++v) {
            rl[a][v] = g.row_of(a, v);
            rv[a][v] = rl[a][v] < rows;
            nr[a][v] = ep->nrm[rl[a][v]];
        }
#pragma unroll
    for (int b = 0; b < G::NTN; ++b) {
        int col = g.col_of(b);
        cv[b] = col < k;
        ns[b] = ep->nrmS[col];
        int64_t s = ep->Si[col] - r0;  // the pivot's row within this block, if it lies here
        sp[b] = s >= 0 && s < BM ? (int)s : -1;
    }
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                double D = nr[a][v] + ns[b] - 2.0 * (double)g.acc[a][b][v];
                D = D > 0.0 ? D : 0.0;
                T kv = gauss_val(D, gs, (T*)nullptr);
                kv = rl[a][v] == sp[b] ? (T)1 : kv;
                kv = cv[b] && rv[a][v] ? kv : (T)0;
                g.acc[a][b][v] = -kv;
            }
}

// Last stage: with acc = -(K(R,S_i) - F_R F_{S_i}^T) (the overlap added by
// the caller, which leaves the LDS free behind a barrier), T goes to LDS,
// G = T L_c^-T (L_c^-T is triangular), G is written to F(R, Nr:Nr+k) when F
// is given, and d(R) = max(d(R) - rowsq(G), 0) (the old d(R) from ep).
// csrow (optional): this block's row of the F^T 1 partials; its entries
// Nr .. Nr+k-1 get the column sums of G over the block's rows (FP64 sums of
// the FP32 or FP64 values written to F), so the kEVD's degree pass F^T 1
// need not read F again.
template <typename T>
__device__ inline void dd_finish(typename Cfg<T>::G& g, T* __restrict__ F, int64_t ldf, double* __restrict__ d,
                                 int64_t n, const T* __restrict__ LinvT, int k, int64_t Nr, int64_t r0, T* lds,
                                 double (*rs)[Cfg<T>::BM], const DowndateEp<T>* ep, double* __restrict__ csrow = nullptr,
                                 const __bf16* Lx = nullptr)
{
    typedef typename Cfg<T>::G G;
    typedef DowndateLds<T> DL;
    constexpr int BM = DL::BM, TP = DL::TP, WN = DL::WN;
    // T = K(R,S_i) - F_R F_{S_i}^T into LDS
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v)
                lds[g.row_of(a, v) * TP + g.col_of(b)] = -g.acc[a][b][v];
    __syncthreads();
    g.zero();
    // L_c^-T is triangular; Lx (FP32 only): its bf16x6 parts, the product
    // then runs on the bf16 matrix cores
    if constexpr (std::is_same<T, float>::value) {
        if (Lx)
            x6_tri_live<G>(g, lds, TP, Lx);
        else
            g.lds_global_tri_live(lds, TP, LinvT, BMAX, (k + G::KSUB - 1) / G::KSUB * G::KSUB);
    } else {
        g.lds_global_tri_live(lds, TP, LinvT, BMAX, (k + G::KSUB - 1) / G::KSUB * G::KSUB);
    }
    // write the new columns and the squared row norms
    double part[G::NTM][G::M::ACC];
    int ro[G::NTM][G::M::ACC];  // element offset of the lane's rows from F(r0, Nr)
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < G::M::ACC; ++v) {
            part[a][v] = 0.0;
            ro[a][v] = g.row_of(a, v) * (int)ldf;
        }
    if (F) {
        // buffer stores over F(r0 : r0 + valid rows, Nr : ldf): a row past n
        // falls outside the range and is dropped by the hardware, and the
        // offsets are 32-bit (BM ldf elements per block)
        int64_t vrows = n - r0 < BM ? n - r0 : BM;
        __amdgpu_buffer_rsrc_t rsF = __builtin_amdgcn_make_buffer_rsrc(
            F + r0 * ldf + Nr, (short)0, (int)((vrows * ldf - Nr) * (int64_t)sizeof(T)), 0x00020000);
#pragma unroll
        for (int a = 0; a < G::NTM; ++a)
#pragma unroll
            for (int b = 0; b < G::NTN; ++b)
#pragma unroll
                for (int v = 0; v < G::M::ACC; ++v) {
                    int col = g.col_of(b);
                    if (col < k) buffer_store(g.acc[a][b][v], rsF, (uint32_t)((ro[a][v] + col) * (int)sizeof(T)));
                }
    }
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                double gv = (double)g.acc[a][b][v];
                part[a][v] += g.col_of(b) < k ? gv * gv : 0.0;
            }
    // reduce over the MT lanes that share a row (lane % MT = column)
    constexpr int MT = G::M::MT;
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < G::M::ACC; ++v) {
            double s = part[a][v];
#pragma unroll
            for (int o = 1; o < MT; o <<= 1) s += __shfl_xor(s, o, 64);
            if (g.li == 0) rs[g.wn][g.row_of(a, v)] = s;
        }
    // column sums of G: a lane's rows, then the MT-lane groups that share a
    // column (lane % MT), in a fixed order
    double cs[G::NTN];
    if (csrow) {
#pragma unroll
        for (int b = 0; b < G::NTN; ++b) {
            double c = 0.0;
#pragma unroll
            for (int a = 0; a < G::NTM; ++a)
#pragma unroll
                for (int v = 0; v < G::M::ACC; ++v) c += (double)g.acc[a][b][v];
#pragma unroll
            for (int o = MT; o < 64; o <<= 1) c += __shfl_xor(c, o, 64);
            cs[b] = c;
        }
    }
    __syncthreads();  // every wave is past its product: the LDS tile is free
    if (csrow) {
        // the waves' column sums through LDS, added in wave order
        constexpr int BN = Cfg<T>::BN, NW = 4 / WN;
        double* cw = reinterpret_cast<double*>(lds);
        if (g.h == 0)
#pragma unroll
            for (int b = 0; b < G::NTN; ++b) cw[g.wm * BN + g.col_of(b)] = cs[b];
        __syncthreads();
        for (int c = threadIdx.x; c < k; c += 256) {
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) sum += cw[w * BN + c];
            csrow[Nr + c] = sum;
        }
    }
    for (int r = threadIdx.x; r < BM; r += 256) {
        int64_t row = r0 + r;
        if (row < n) {
            double sq = 0.0;
#pragma unroll
            for (int q = 0; q < WN; ++q) sq += rs[q][r];
            double dd = ep->d[r] - sq;
            d[row] = dd >= 0.0 ? dd : 0.0;
        }
    }
}

// The fused Step-3 update for BM rows (ks_arpclm.cu:929-995 and
// update_diag_block 455-478):
//   acc  = X_R X_{S_i}^T                    (MFMA, K = dp; bf16x6 with X6)
//   acc  = -K(R, S_i)                       (epilogue: norms, exp)
//   acc += F_R F_{S_i}^T                    (MFMA, K = Nr)
//   T    = -acc -> LDS
//   acc  = T L_c^-T                         (MFMA, K = k; bf16x6 with X6)
//   F(R, Nr:Nr+k) = acc;  d(R) = max(d(R) - rowsq(acc), 0)
// This is row block blk (rows blk * BM ...) of one problem; it returns
// early (before any barrier) when the problem is done or accepted nothing
// this round.  cspart (optional): the F^T 1 partials, row blk at
// cspart + blk * ldf (dd_finish).
// X6: the distance product and the overlap of an FP32 factor on the bf16
// matrix cores (bf16x6 products, bf16x6.h) instead of FP32 MFMA.
template <typename T, bool X6 = false>
__device__ inline void downdate_body(const T* __restrict__ X, int32_t dp, const double* __restrict__ nrm,
                                     T* __restrict__ F, int64_t ldf, double* __restrict__ d, int64_t n,
                                     const T* __restrict__ XS, const double* __restrict__ nrmS,
                                     const int64_t* __restrict__ Si, const T* __restrict__ FS,
                                     const T* __restrict__ LinvT, const ArpcState* st, GaussScale gs,
                                     int64_t blk, T* lds, double (*rs)[Cfg<T>::BM], DowndateEp<T>* ep,
                                     double* __restrict__ cspart)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    if (st->done) return;
    int k = st->k;
    if (k == 0) return;
    int64_t Nr = st->Nr;
    int64_t r0 = blk * BM;
    G g;
    const __bf16* XSx = nullptr;
    if constexpr (X6 && std::is_same<T, float>::value) XSx = x6_xsx(XS, dp);
    dd_prologue<T>(g, X, dp, nrm, d, n, XS, nrmS, Si, k, gs, r0, lds, ep, XSx);
    // overlap with the current factor (ends with a barrier)
    if (Nr > 0) {
        if constexpr (X6 && std::is_same<T, float>::value) {
            static_assert(X6Stage<G>::BYTES <= DowndateLds<T>::ELEMS * (int)sizeof(T), "the x6 stage fits");
            x6_nt_mainloop_live<G>(g, reinterpret_cast<unsigned char*>(lds), F + r0 * ldf, ldf, n - r0,
                                   x6_fsx(FS, ldf), Nr);
        } else {
            g.mainloop_live(lds, F + r0 * ldf, ldf, n - r0, FS, ldf, BN, Nr);
        }
    }
    const __bf16* Lx = nullptr;
    if constexpr (X6 && std::is_same<T, float>::value) Lx = x6_lx(LinvT);
    dd_finish<T>(g, F, ldf, d, n, LinvT, k, Nr, r0, lds, rs, ep, cspart ? cspart + blk * ldf : nullptr, Lx);
}

// ------------------------------------------- regenerating (LM) variant
// The form ks_arpclm.cu itself runs: no factor is stored.  With the pivots S
// (Nr so far) and L^-1 (K(S,S) = L L^T) kept, F = K(:,S) L^-T is never
// formed; its rows and products are regenerated from the points:
//   F(S',:)         = K(S',S) L^-T                     (FP64, proposals)
//   F_R F_{S_i}^T   = K(R,S) M,  M = L^-T F_{S_i}^T      (Sgemm T, :965-972)
//   L^-1 <- [[L^-1, 0], [-L_c^-1 F_{S_i} L^-1, L_c^-1]]  (append_Linv_*, :480-534)
// Memory is O(n + r^2) instead of the factor's n r.

// Pivot chunk of the regenerated overlap: LM_CH pivots per distance tile.
constexpr int LM_CH = 32;
template <typename T> struct LmCfg;
template <> struct LmCfg<float> {
    typedef NTGemm<float, 128, LM_CH, 32, 4, 1, 2, 2, Mfma16f> GD;
};
template <> struct LmCfg<double> {
    typedef NTGemm<double, 64, LM_CH, 16, 4, 1, 1, 2> GD;
};
template <typename T> struct DowndateLmLds {
    typedef typename LmCfg<T>::GD GD;
    static constexpr int TPC = LM_CH + 16 / (int)sizeof(T);  // K-chunk tile pitch
    static constexpr int ELEMS = GD::LDS_ELEMS + Cfg<T>::BM * TPC;
    static_assert(ELEMS <= DowndateLds<T>::ELEMS, "the chunk stage fits the downdate LDS");
};

// P(p, m) = K(S'_p, S_m) in FP64 (direct differences of the FP64 points),
// row-major BMAX x ldp; zero for p >= nprop.  Thread per (p, m), m < Nr.
__device__ inline void lm_kprop_body(const double* __restrict__ X64, int32_t dp, const int64_t* __restrict__ Sprime,
                                     const int64_t* __restrict__ piv, const ArpcState* st, double neg_inv_eps,
                                     double* P, int64_t ldp, int p, int64_t m)
{
    if (st->done || m >= st->Nr) return;
    double kv = 0.0;
    if (p < st->nprop) {
        int64_t sp = Sprime[p], sm = piv[m];
        if (sp == sm) {
            kv = 1.0;
        } else {
            const double* a = X64 + sp * dp;
            const double* b = X64 + sm * dp;
            double q = 0.0;
            for (int j = 0; j < dp; ++j) {
                double t = a[j] - b[j];
                q += t * t;
            }
            kv = exp(q * neg_inv_eps);
        }
    }
    P[p * ldp + m] = kv;
}

// Rows of the accepted proposals: FSd(c, :) = F(S'_{pos_c}, :) (FP64, rows
// past k zero), and the accepted points appended to the pivot points XP and
// their norms (at Nr + c).  Block per c < BMAX.
template <typename T>
__device__ inline void lm_gather_body(const double* __restrict__ P2, int64_t ldp, const int32_t* __restrict__ Spos,
                                      const T* __restrict__ Xsrc, const double* __restrict__ nrm, int32_t dp,
                                      const int64_t* __restrict__ Si, const ArpcState* st, double* FSd, T* XP,
                                      double* nrmP, int c)
{
    if (st->done) return;
    int k = st->k;
    int64_t Nr = st->Nr;
    const double* src = c < k ? P2 + (int64_t)Spos[c] * ldp : nullptr;
    for (int64_t m = threadIdx.x; m < Nr; m += 256) FSd[c * ldp + m] = src ? src[m] : 0.0;
    if (c < k) {
        int64_t s = Si[c];
        for (int j = threadIdx.x; j < dp; j += 256) XP[(Nr + c) * dp + j] = Xsrc[s * dp + j];
        if (threadIdx.x == 0) nrmP[Nr + c] = nrm[s];
    }
}

// L^-1 (row-major, lower, FP64, ld ldl) gains rows Nr .. Nr+k-1:
// -L_c^-1 MTd in columns < Nr (MTd = F_{S_i} L^-1, k x Nr) and L_c^-1 on the
// diagonal block.  Thread per (c, m), m < Nr + k.
__device__ inline void lm_linv_append_body(const double* __restrict__ Lc, const double* __restrict__ MTd,
                                           int64_t ldp, const ArpcState* st, double* Linv, int64_t ldl, int c,
                                           int64_t m)
{
    if (st->done) return;
    int k = st->k;
    int64_t Nr = st->Nr;
    if (c >= k || m >= Nr + k) return;
    double v;
    if (m < Nr) {
        v = 0.0;
        for (int q = 0; q <= c; ++q) v -= Lc[c * BMAX + q] * MTd[q * ldp + m];
    } else {
        int q = (int)(m - Nr);
        v = q <= c ? Lc[c * BMAX + q] : 0.0;
    }
    Linv[(Nr + c) * ldl + m] = v;
}

// The Step-3 update of the LM form: as downdate_body, with the overlap
// F_R F_{S_i}^T = K(R,S) M regenerated LM_CH pivots at a time (distance tile
// of X_R against the pivot points XP, norms and exp into LDS, then MFMA with
// the rows of M^T = MT, read from L2), and no factor written.
template <typename T>
__device__ inline void downdate_lm_body(const T* __restrict__ X, int32_t dp, const double* __restrict__ nrm,
                                        double* __restrict__ d, int64_t n, const T* __restrict__ XS,
                                        const double* __restrict__ nrmS, const int64_t* __restrict__ Si,
                                        const T* __restrict__ XP, const double* __restrict__ nrmP,
                                        const int64_t* __restrict__ piv, const T* __restrict__ MT, int64_t ldm,
                                        const T* __restrict__ LinvT, const ArpcState* st, GaussScale gs,
                                        int64_t blk, T* lds, double (*rs)[Cfg<T>::BM], DowndateEp<T>* ep)
{
    typedef typename Cfg<T>::G G;
    typedef typename LmCfg<T>::GD GD;
    constexpr int BM = Cfg<T>::BM, TPC = DowndateLmLds<T>::TPC;
    if (st->done) return;
    int k = st->k;
    if (k == 0) return;
    int64_t Nr = st->Nr;
    int64_t r0 = blk * BM;
    G g;
    dd_prologue<T>(g, X, dp, nrm, d, n, XS, nrmS, Si, k, gs, r0, lds, ep);
    T* ldsD = lds;
    T* Tc = lds + GD::LDS_ELEMS;
    for (int64_t c0 = 0; c0 < Nr; c0 += LM_CH) {
        const int cw = Nr - c0 < LM_CH ? (int)(Nr - c0) : LM_CH;
        GD gd;
        gd.init();
        // the lane's pivot norms and pivot rows (within this block, else -1)
        // are loaded ahead of the distance product, which covers their latency
        double ns[GD::NTN];
        int sp[GD::NTN];
#pragma unroll
        for (int b = 0; b < GD::NTN; ++b) {
            int col = gd.col_of(b);
            bool live = col < cw;
            ns[b] = live ? nrmP[c0 + col] : 0.0;
            int64_t s = live ? piv[c0 + col] - r0 : -1;
            sp[b] = s >= 0 && s < BM ? (int)s : -1;
        }
        // the previous chunk's reads of ldsD ended behind its barrier; those
        // of Tc end before any wave passes this mainloop's barriers
        gd.mainloop(ldsD, X + r0 * dp, dp, n - r0, XP + c0 * dp, dp, cw, dp);
        double nr[GD::NTM][GD::M::ACC];
#pragma unroll
        for (int a = 0; a < GD::NTM; ++a)
#pragma unroll
            for (int v = 0; v < GD::M::ACC; ++v) nr[a][v] = ep->nrm[gd.row_of(a, v)];
        // branch-free element loop: out-of-range rows and columns masked to 0
#pragma unroll
        for (int a = 0; a < GD::NTM; ++a)
#pragma unroll
            for (int b = 0; b < GD::NTN; ++b)
#pragma unroll
                for (int v = 0; v < GD::M::ACC; ++v) {
                    int rl = gd.row_of(a, v);
                    int col = gd.col_of(b);
                    double D = nr[a][v] + ns[b] - 2.0 * (double)gd.acc[a][b][v];
                    D = D > 0.0 ? D : 0.0;
                    T kv = gauss_val(D, gs, (T*)nullptr);
                    kv = rl == sp[b] ? (T)1 : kv;
                    kv = r0 + rl < n && col < cw ? kv : (T)0;
                    Tc[rl * TPC + col] = kv;
                }
        __syncthreads();
        g.lds_global(Tc, TPC, MT + c0, ldm, LM_CH);
    }
    __syncthreads();  // every wave is done with Tc before T overwrites the LDS
    dd_finish<T>(g, nullptr, 0, d, n, LinvT, k, Nr, r0, lds, rs, ep);
}

// Block of BMAX threads.
__device__ inline void commit_body(ArpcState* st, const int64_t* __restrict__ Si, int64_t* pivots,
                                   int32_t* accepted)
{
    if (st->done) return;
    const int round = st->round;
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

__device__ inline void final_err_body(const double* psum, int nb, double trK, const ArpcState* st, double* err)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int j = 0; j < nb; ++j) s += psum[j];
    if (st->round > 0) err[st->round - 1] = s / trK;
}


inline int ceil_log2(int64_t n)
{
    int c = 0;
    while (((int64_t)1 << c) < n) ++c;
    return c;
}

// ------------------------------------------------------------ host side

// Device buffers of one problem's ARPC run (named context scratch).
struct ArpcBufs {
    double* d = nullptr;          // residual diagonal, n
    ArpcState* st = nullptr;
    int64_t* piv = nullptr;       // rank
    int32_t* acc = nullptr;       // max_rounds
    double* err = nullptr;        // max_rounds
    int64_t* inj = nullptr;       // injected pivots or null
    int nb = 0;                   // scan chunks
    double *pmax = nullptr, *psum = nullptr, *H = nullptr, *nrmS = nullptr, *Linv64 = nullptr;
    uint64_t *prefix = nullptr, *btot = nullptr;
    int64_t *Sprime = nullptr, *Si = nullptr;
    void *XS = nullptr, *FS = nullptr, *LinvT = nullptr;  // GEMM precision
    double* cspart = nullptr;     // F^T 1 partials, ceil(n / BM) x ldf (F form)
};

// Checks the parameters and result arrays as kevd_arpc does; returns the
// round cap.
kevd_status arpc_validate(kevd_ctx* ctx, const kevd_arpc_params* p, const kevd_arpc_result* res,
                          int* max_rounds);
// Sizes the factor, allocates the run and round buffers and queues the
// initial state (d = 1, counters zero) on the context stream.
kevd_status arpc_prepare(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, ArpcBufs* b);
// After the rounds (stream synchronised): copies rank, rounds, pivots,
// accepted counts and trace errors out and records the factor in ctx.
kevd_status arpc_finish(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, const ArpcBufs& b,
                        kevd_arpc_result* res);

}  // namespace arpck
