// evd.hip -- bistochastic kernel EVD on the resident factor F (n x r).
//
// The reference streams K(R,S) panels through four passes and two dense
// phases in the L^-1 form (README_kevd.md:22-69; ks_kevd.cu:755-2308).  With
// F resident, the same quantities are (K^ = F F^T):
//   t  = F^T 1,        d~ = F t                    pass 1 + pass 2 matvec
//   cq = F^T d~^-1,    q~ = F cq                   pass 2 / pass 3 matvecs
//   A  = F^T D~^-2 F,  B = F^T Q~^-1 F             Atil / Btil (one pass, two Grams)
//   s  = DBL_EPSILON tr(A);  R = chol(A + sI)      dense 1 (ks_kevd.cu:1465-1520)
//   W  = R B R^T;  W = V diag(lam) V^T             dense 2 (1764-1913)
//   U  = D~^-1 F R^-1 V_k                          Z step + pass 4 (1973-2275)
// plus P 1 - 1 (row-sum deviation) from one extra output column of the U GEMM.
// Sinkhorn mode: s .* (K^ s) = 1 by a Chebyshev-accelerated, scale-free
// Sinkhorn-Knopp iteration (one pass over F each), one Gram F^T S^2 F,
// U = S F V_k Lambda_k^-1/2.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <map>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <vector>

#include "bf16x6.h"
#include "gemm.h"
#include "internal.h"

rocblas_handle blas_handle(kevd_ctx* ctx);

namespace {

// ------------------------------------------------------------- matvecs

constexpr int COLSUM_ROWS = 512;  // rows of F per column-sum block

// part[b][j] = sum_{rows of block b} w_i F(i, j), FP64 (w = 1 if null).
// Each thread owns 16-byte column groups; a block streams `rows`
// contiguous rows (COLSUM_ROWS, fewer for a small n so that the grid still
// fills the device), so every wave load is one contiguous 4 KiB segment.
template <typename T>
__global__ __launch_bounds__(256) void k_colsum(const T* __restrict__ F, int64_t ldf, int64_t n,
                                                int64_t r, const double* __restrict__ w,
                                                double* part, int64_t rpad, int rows)
{
    constexpr int V = 16 / (int)sizeof(T);
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type vec_t;
    int64_t r0 = (int64_t)blockIdx.x * rows;
    int64_t r1 = r0 + rows < n ? r0 + rows : n;
    for (int64_t cg = (int64_t)threadIdx.x * V; cg < r; cg += 256 * V) {
        double acc[V];
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = 0.0;
        for (int64_t i = r0; i < r1; ++i) {
            double wi = w ? w[i] : 1.0;
            vec_t v = *reinterpret_cast<const vec_t*>(F + i * ldf + cg);
            const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
            for (int q = 0; q < V; ++q) acc[q] += wi * (double)e[q];
        }
#pragma unroll
        for (int q = 0; q < V; ++q)
            if (cg + q < r) part[(int64_t)blockIdx.x * rpad + cg + q] = acc[q];
    }
}

// c[j] = sum_b part[b][j], deterministic: a block owns 64 columns; its 16
// row groups sum the blocks b = g (mod 16) in order, then group 0 adds the
// 16 group sums in order.
__device__ inline void colsum_reduce_body(const double* part, int nb, int64_t r, int64_t rpad, double* c, int bx)
{
    __shared__ double red[16][64];
    int col = threadIdx.x & 63, g = threadIdx.x >> 6;
    int64_t j = (int64_t)bx * 64 + col;
    double s = 0.0;
    if (j < r) {
        int b = g;
        for (; b + 48 < nb; b += 64) {
            double v0 = part[(int64_t)b * rpad + j], v1 = part[(int64_t)(b + 16) * rpad + j];
            double v2 = part[(int64_t)(b + 32) * rpad + j], v3 = part[(int64_t)(b + 48) * rpad + j];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; b < nb; b += 16) s += part[(int64_t)b * rpad + j];
    }
    red[g][col] = s;
    __syncthreads();
    if (g == 0 && j < r) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) t += red[q][col];
        c[j] = t;
    }
}
__global__ __launch_bounds__(1024) void k_colsum_reduce(const double* part, int nb, int64_t r,
                                                        int64_t rpad, double* c)
{
    colsum_reduce_body(part, nb, r, rpad, c, blockIdx.x);
}

constexpr int ROWDOT_LDS_MAX = 8192;  // c staged in LDS up to this many columns

// y_i = F_i . c (FP64).  A block walks groups of 16 rows (grid-stride);
// each wave takes 4 rows of a group at once, so a lane has 4 independent
// 16-byte loads in flight per column step, with c read from LDS (STAGED)
// or from global memory.  mode 0: plain; mode 1 (Sinkhorn step): with s
// the current scaling, also dev_i = |s_i y_i - 1| and s_new_i =
// sqrt(s_i / y_i).  Per-block min(y) and max(dev) go to bmin/bdev.
template <typename T, bool STAGED>
__global__ __launch_bounds__(256) void k_rowdot(const T* __restrict__ F, int64_t ldf, int64_t n,
                                                int64_t r, const double* __restrict__ c,
                                                double* y, double* bmin, int mode,
                                                const double* s, double* s_new, double* bdev)
{
    constexpr int V = 16 / (int)sizeof(T), RW = 4;
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type vec_t;
    extern __shared__ double csh[];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (STAGED) {
        for (int64_t j = threadIdx.x; j < r; j += 256) csh[j] = c[j];
        __syncthreads();
    }
    double mn = INFINITY, dv = 0.0;
    int64_t ngroups = (n + 15) / 16;
    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        int64_t row0 = grp * 16 + wv * RW;
        double acc[RW];
#pragma unroll
        for (int i = 0; i < RW; ++i) acc[i] = 0.0;
        for (int64_t cg = (int64_t)lane * V; cg < r; cg += 64 * V) {
            double cv[V];
#pragma unroll
            for (int q = 0; q < V; ++q) cv[q] = cg + q < r ? (STAGED ? csh[cg + q] : c[cg + q]) : 0.0;
            vec_t v[RW];
#pragma unroll
            for (int i = 0; i < RW; ++i)
                if (row0 + i < n) v[i] = *reinterpret_cast<const vec_t*>(F + (row0 + i) * ldf + cg);
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                if (row0 + i >= n) continue;
                const T* e = reinterpret_cast<const T*>(&v[i]);
#pragma unroll
                for (int q = 0; q < V; ++q)
                    if (cg + q < r) acc[i] += (double)e[q] * cv[q];
            }
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            int64_t row = row0 + i;
            double a = wave_sum(acc[i]);
            if (row >= n) continue;
            mn = fmin(mn, a);
            if (lane == 0) {
                y[row] = a;
                if (mode == 1) {
                    double si = s[row];
                    dv = fmax(dv, fabs(si * a - 1.0));
                    s_new[row] = a > 0.0 ? sqrt(si / a) : si;
                }
            }
        }
    }
    __shared__ double smn[4], sdv[4];
    if (lane == 0) {
        smn[wv] = mn;
        sdv[wv] = dv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bmin[blockIdx.x] = fmin(fmin(smn[0], smn[1]), fmin(smn[2], smn[3]));
        if (mode == 1) bdev[blockIdx.x] = fmax(fmax(sdv[0], sdv[1]), fmax(sdv[2], sdv[3]));
    }
}

// One fused Sinkhorn pass over F (symmetric scaling of K^ = F F^T).
// With the iterate s and u = F^T s, each row i gives y_i = F_i . u =
// (K^ s)_i, the scaled row sum rs_i = s_i y_i and the vector the next
// iterate is built from,
//   q_i = 1 / y_i            (Sinkhorn-Knopp / Chebyshev steps)
//   q_i = sqrt(s_i / y_i)    (damped),
// and the same loaded row adds q_i F_i to this block's partial of F^T q
// -- so one iteration streams F once instead of twice.  A block owns
// SK_ROWS rows and walks them SKR at a time; each thread keeps VPT 16-byte
// column vectors of u and of its F^T q partial in registers.
// Per block: bstat = [min y, max |rs - 1|, sum rs, min s, sum (rs - 1)^2].
constexpr int SK_NSTAT = 5;

constexpr int SK_ROWS = 512;

template <typename T> struct SkRows;
template <> struct SkRows<float> { static constexpr int R = 4; };
template <> struct SkRows<double> { static constexpr int R = 2; };

template <typename T, int VPT>
__global__ __launch_bounds__(256) void k_sinkhorn_pass(const T* __restrict__ F, int64_t ldf, int64_t n,
                                                       int64_t r, const double* __restrict__ u,
                                                       const double* __restrict__ s, int damped, double* q_out,
                                                       double* part, int64_t rpad, double* bstat)
{
    constexpr int V = 16 / (int)sizeof(T), R = SkRows<T>::R;
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type vec_t;
    __shared__ double red[2][4][R];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t r0 = (int64_t)blockIdx.x * SK_ROWS;
    int64_t r1 = r0 + SK_ROWS < n ? r0 + SK_ROWS : n;
    double uv[VPT][V], acc[VPT][V];
#pragma unroll
    for (int q = 0; q < VPT; ++q)
#pragma unroll
        for (int e = 0; e < V; ++e) {
            int64_t c = ((int64_t)threadIdx.x + 256 * q) * V + e;
            uv[q][e] = c < r ? u[c] : 0.0;
            acc[q][e] = 0.0;
        }
    // two register sets, each indexed only by compile-time constants
    vec_t f0[R][VPT], f1[R][VPT];
    auto load = [&](vec_t (&f)[R][VPT], int64_t i0) {
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                int64_t c = ((int64_t)threadIdx.x + 256 * q) * V;
                if (i0 + rr < r1 && c < r)
                    f[rr][q] = *reinterpret_cast<const vec_t*>(F + (i0 + rr) * ldf + c);
                else
                    f[rr][q] = vec_t{};
            }
    };
    double mn = INFINITY, dv = 0.0, srs = 0.0, smin = INFINITY, se2 = 0.0;
    auto group = [&](const vec_t (&f)[R][VPT], int64_t i0, double (&rd)[4][R]) {
        double fv[R][VPT][V];
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                const T* e = reinterpret_cast<const T*>(&f[rr][q]);
                int64_t c = ((int64_t)threadIdx.x + 256 * q) * V;
#pragma unroll
                for (int k = 0; k < V; ++k) fv[rr][q][k] = c + k < r ? (double)e[k] : 0.0;
            }
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < VPT; ++q)
#pragma unroll
                for (int k = 0; k < V; ++k) d += fv[rr][q][k] * uv[q][k];
            d = wave_sum(d);
            if (lane == 0) rd[wv][rr] = d;
        }
        // the iterate's entries for these rows, read before the barrier
        double sv[R];
#pragma unroll
        for (int rr = 0; rr < R; ++rr) sv[rr] = i0 + rr < r1 ? s[i0 + rr] : 1.0;
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            int64_t i = i0 + rr;
            if (i >= r1) break;
            double y = rd[0][rr] + rd[1][rr] + rd[2][rr] + rd[3][rr];
            double si = sv[rr];
            double rs = si * y;
            double bn = damped ? sqrt(si / y) : 1.0 / y;
            if (threadIdx.x == 0) {
                q_out[i] = bn;
                mn = fmin(mn, y);
                dv = fmax(dv, fabs(rs - 1.0));
                srs += rs;
                smin = fmin(smin, si);
                se2 += (rs - 1.0) * (rs - 1.0);
            }
#pragma unroll
            for (int q = 0; q < VPT; ++q)
#pragma unroll
                for (int k = 0; k < V; ++k) acc[q][k] += bn * fv[rr][q][k];
        }
    };
    // red[0] / red[1] alternate between groups: a group's partials are
    // written before its barrier and read after it, and the next write to
    // the same slot comes after the following barrier.
    int ng = (int)((r1 - r0 + R - 1) / R);
    if (ng > 0) load(f0, r0);
    for (int g = 0; g < ng; g += 2) {
        int64_t i0 = r0 + (int64_t)g * R;
        if (g + 1 < ng) load(f1, i0 + R);
        group(f0, i0, red[0]);
        if (g + 1 >= ng) break;
        if (g + 2 < ng) load(f0, i0 + 2 * R);
        group(f1, i0 + R, red[1]);
    }
#pragma unroll
    for (int q = 0; q < VPT; ++q)
#pragma unroll
        for (int e = 0; e < V; ++e) {
            int64_t c = ((int64_t)threadIdx.x + 256 * q) * V + e;
            if (c < r) part[(int64_t)blockIdx.x * rpad + c] = acc[q][e];
        }
    if (threadIdx.x == 0) {
        bstat[SK_NSTAT * blockIdx.x + 0] = mn;
        bstat[SK_NSTAT * blockIdx.x + 1] = dv;
        bstat[SK_NSTAT * blockIdx.x + 2] = srs;
        bstat[SK_NSTAT * blockIdx.x + 3] = smin;
        bstat[SK_NSTAT * blockIdx.x + 4] = se2;
    }
}

// The same Sinkhorn pass with every wave on whole rows of its own: lane l
// holds the 16-byte column vectors l, l + 64, ... (VPQ of them) of u and of
// its F^T s' partial in registers, a row's dot product is a wave reduction,
// and the next row (with its iterate entry) loads while the current one is
// processed, so no barrier is taken per row.  The four waves' partials and
// statistics are combined once per block.  Rows: block b owns SK_ROWS
// consecutive rows, wave w every fourth of them.  At VPQ = 8 (r <= 2048) it
// needs ~310 VGPRs, one wave per SIMD; capping it at two waves spills (and
// moving u to LDS does not lower the count), so it runs at one.  FP32 only:
// the FP64 factor uses k_sinkhorn_pass.
// (body: row block blk; the batched ensemble pass runs it too)
template <typename T, int VPQ>
__device__ __attribute__((always_inline)) inline void sinkhorn_rows_body(
    const T* __restrict__ F, int64_t ldf, int64_t n, int64_t r, const double* __restrict__ u,
    const double* __restrict__ s, int damped, double* q_out, double* part, int64_t rpad, double* bstat, int blk)
{
    constexpr int V = 16 / (int)sizeof(T);
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type vec_t;
    __shared__ double red[4][SK_NSTAT];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blk * SK_ROWS;
    const int64_t r1 = r0 + SK_ROWS < n ? r0 + SK_ROWS : n;
    double uv[VPQ][V], acc[VPQ][V];
#pragma unroll
    for (int q = 0; q < VPQ; ++q)
#pragma unroll
        for (int e = 0; e < V; ++e) {
            int64_t c = ((int64_t)lane + 64 * q) * V + e;
            uv[q][e] = c < r ? u[c] : 0.0;
            acc[q][e] = 0.0;
        }
    auto load = [&](vec_t (&f)[VPQ], double& sb, int64_t i) {
#pragma unroll
        for (int q = 0; q < VPQ; ++q) {
            int64_t c = ((int64_t)lane + 64 * q) * V;
            f[q] = (i < r1 && c < r) ? *reinterpret_cast<const vec_t*>(F + i * ldf + c) : vec_t{};
        }
        sb = i < r1 ? s[i] : 1.0;
    };
    double mn = INFINITY, dv = 0.0, srs = 0.0, smin = INFINITY, se2 = 0.0;
    // F values are widened twice (dot, then accumulation) rather than held
    // as doubles, which would cost VPQ * V * 2 more registers
    auto fval = [&](const vec_t& f, int q, int k) {
        int64_t c = ((int64_t)lane + 64 * q) * V + k;
        return c < r ? (double)reinterpret_cast<const T*>(&f)[k] : 0.0;
    };
    auto row = [&](const vec_t (&f)[VPQ], double sb, int64_t i) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < VPQ; ++q)
#pragma unroll
            for (int k = 0; k < V; ++k) d += fval(f[q], q, k) * uv[q][k];
        double y = wave_sum(d);
        double si = sb;
        double rs = si * y;
        double bn = damped ? sqrt(si / y) : 1.0 / y;
        if (lane == 0) {
            q_out[i] = bn;
            mn = fmin(mn, y);
            dv = fmax(dv, fabs(rs - 1.0));
            srs += rs;
            smin = fmin(smin, si);
            se2 += (rs - 1.0) * (rs - 1.0);
        }
#pragma unroll
        for (int q = 0; q < VPQ; ++q)
#pragma unroll
            for (int k = 0; k < V; ++k) acc[q][k] += bn * fval(f[q], q, k);
    };
    vec_t f0[VPQ], f1[VPQ];
    double s0 = 1.0, s1 = 1.0;
    int64_t i = r0 + wv;
    if (i < r1) load(f0, s0, i);
    for (; i < r1; i += 8) {
        if (i + 4 < r1) load(f1, s1, i + 4);
        row(f0, s0, i);
        if (i + 4 >= r1) break;
        if (i + 8 < r1) load(f0, s0, i + 8);
        row(f1, s1, i + 4);
    }
    // combine the four waves: partials through LDS one wave at a time, in
    // wave order (deterministic)
    __shared__ double cpart[64 * VPQ * V];
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int q = 0; q < VPQ; ++q)
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    int idx = (q * 64 + lane) * V + k;
                    cpart[idx] = w == 0 ? acc[q][k] : cpart[idx] + acc[q][k];
                }
            if (lane == 0) {
                red[w][0] = mn;
                red[w][1] = dv;
                red[w][2] = srs;
                red[w][3] = smin;
                red[w][4] = se2;
            }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < 64 * VPQ * V; idx += 256) {
        int q = idx / (64 * V), rem = idx % (64 * V), ln = rem / V, k = rem % V;
        int64_t c = ((int64_t)ln + 64 * q) * V + k;
        if (c < r) part[(int64_t)blk * rpad + c] = cpart[idx];
    }
    if (threadIdx.x == 0) {
        bstat[SK_NSTAT * blk + 0] = fmin(fmin(red[0][0], red[1][0]), fmin(red[2][0], red[3][0]));
        bstat[SK_NSTAT * blk + 1] = fmax(fmax(red[0][1], red[1][1]), fmax(red[2][1], red[3][1]));
        bstat[SK_NSTAT * blk + 2] = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
        bstat[SK_NSTAT * blk + 3] = fmin(fmin(red[0][3], red[1][3]), fmin(red[2][3], red[3][3]));
        bstat[SK_NSTAT * blk + 4] = (red[0][4] + red[1][4]) + (red[2][4] + red[3][4]);
    }
}
template <typename T, int VPQ>
__global__ __launch_bounds__(256) void k_sinkhorn_rows(const T* __restrict__ F, int64_t ldf, int64_t n,
                                                       int64_t r, const double* __restrict__ u,
                                                       const double* __restrict__ s, int damped, double* q_out,
                                                       double* part, int64_t rpad, double* bstat)
{
    sinkhorn_rows_body<T, VPQ>(F, ldf, n, r, u, s, damped, q_out, part, rpad, bstat, blockIdx.x);
}

// [min, max, sum, min, sum] over the nb per-block Sinkhorn statistics, in
// block order.
__device__ inline void stat_reduce_body(const double* bstat, int nb, double* out)
{
    __shared__ double smn[256], sdv[256], ssm[256], ssn[256], se2[256];
    double mn = INFINITY, dv = 0.0, sm = 0.0, sn = INFINITY, e2 = 0.0;
    for (int j = threadIdx.x; j < nb; j += 256) {
        mn = fmin(mn, bstat[SK_NSTAT * j]);
        dv = fmax(dv, bstat[SK_NSTAT * j + 1]);
        sm += bstat[SK_NSTAT * j + 2];
        sn = fmin(sn, bstat[SK_NSTAT * j + 3]);
        e2 += bstat[SK_NSTAT * j + 4];
    }
    smn[threadIdx.x] = mn;
    sdv[threadIdx.x] = dv;
    ssm[threadIdx.x] = sm;
    ssn[threadIdx.x] = sn;
    se2[threadIdx.x] = e2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + o]);
            sdv[threadIdx.x] = fmax(sdv[threadIdx.x], sdv[threadIdx.x + o]);
            ssm[threadIdx.x] += ssm[threadIdx.x + o];
            ssn[threadIdx.x] = fmin(ssn[threadIdx.x], ssn[threadIdx.x + o]);
            se2[threadIdx.x] += se2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smn[0];
        out[1] = sdv[0];
        out[2] = ssm[0];
        out[3] = ssn[0];
        out[4] = se2[0];
    }
}
__global__ __launch_bounds__(256) void k_stat_reduce(const double* bstat, int nb, double* out)
{
    stat_reduce_body(bstat, nb, out);
}

// out = a x + b y + c z over r entries (a Sinkhorn iterate over the n
// points, or its F^T image over the rank); a term with a zero coefficient
// is not read (its vector may be unset).  out may alias x, y or z.
__global__ void k_lincomb3(int64_t r, double a, const double* x, double b, const double* y, double c,
                           const double* z, double* out)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= r) return;
    double v = a != 0.0 ? a * x[j] : 0.0;
    if (b != 0.0) v += b * y[j];
    if (c != 0.0) v += c * z[j];
    out[j] = v;
}

// Per-block minimum of x (grid-stride), for a second-level k_block_minmax.
__global__ __launch_bounds__(256) void k_min_partial(const double* x, int64_t n, double* part)
{
    __shared__ double sm[256];
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmin(m, x[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_block_minmax(const double* bmin, const double* bdev, int nb, double* out)
{
    __shared__ double smn[256], sdv[256];
    double mn = INFINITY, dv = 0.0;
    for (int j = threadIdx.x; j < nb; j += 256) {
        mn = fmin(mn, bmin[j]);
        if (bdev) dv = fmax(dv, bdev[j]);
    }
    smn[threadIdx.x] = mn;
    sdv[threadIdx.x] = dv;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + o]);
            sdv[threadIdx.x] = fmax(sdv[threadIdx.x], sdv[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smn[0];
        out[1] = sdv[0];
    }
}

// Elementwise weights: mode 0: w1 = 1/d^2 (A), 1: w = 1/q (B), 2: w = 1/d,
// 3: w = s^2, 4: copy.
__global__ void k_weights(const double* x, int64_t n, int mode, double* w)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = x[i];
    w[i] = mode == 0 ? 1.0 / (v * v) : mode == 1 ? 1.0 / v : mode == 2 ? 1.0 / v
           : mode == 3 ? v * v : v;
}

// --------------------------------------------------------- weighted Gram

// G_g = F^T diag(w_g) F for NG weight vectors, lower-triangle tiles.
// Block tile BI x BJ of the r x r output, K = rows of F.  Both panels are
// staged into LDS unweighted, row-major [k][col] exactly as F lies in HBM,
// with the chunk's row weights beside them; a lane scales its A operand by
// w_g(k) in registers, so one staged copy serves all NG Grams.  Each lane
// feeds one K value per MFMA; it reads IL = 2 adjacent columns with one
// 8/16-byte LDS load, which serve two MFMA tiles whose rows (columns)
// interleave: tile q of a pair covers panel columns IL*i + q.  Every
// PROMOTE rows the MFMA accumulators are added into this block's FP64
// partial tile (a private slice of `part`), which bounds the length of any
// single FP32 accumulation.
// NT threads per workgroup in WGM x WGN waves, WG_PER_CU workgroups resident
// per CU.
template <typename T, int NG> struct GramCfg;
#ifndef KEVD_GRAM_PROMOTE
#define KEVD_GRAM_PROMOTE 2048
#endif
// The bf16x6 Gram promotes every 4096 rows: half the partial tile's
// read-modify-write traffic of 2048 (about 9.5 of its 47.7 GB at
// N = 1M, r = 2000; Gram 24.6 -> 23.6 ms) with the top-32 eigenvalue error
// against FP64 unchanged (1.16e-6 -> 1.14e-6; two-step normalisation
// 2.20e-6 -> 2.13e-6).
#ifndef KEVD_GRAM_X6_PROMOTE
#define KEVD_GRAM_X6_PROMOTE 4096
#endif
#ifndef KEVD_GRAM_X6_PINGPONG
#define KEVD_GRAM_X6_PINGPONG 1
#endif
#ifndef KEVD_GRAM1_CFG
#define KEVD_GRAM1_CFG 0
#endif
// rows of F per split (0: chosen from the tile count and the CU count)
#ifndef KEVD_GRAM_SPLIT_ROWS
#define KEVD_GRAM_SPLIT_ROWS 0
#endif
// workgroups per tile in the split reduction
constexpr int GRAM_REDUCE_SLICES = 8;
// slices for ntiles tiles: at least 8, and enough workgroups for 4 per CU
inline int gram_reduce_slices(kevd_ctx* ctx, int ntiles)
{
    int s = (4 * ctx->num_cu + ntiles - 1) / ntiles;
    return std::max(GRAM_REDUCE_SLICES, std::min(s, 64));
}
template <int NG> struct GramCfg<float, NG> {
    static constexpr int BI = 128, BJ = 128, KC = 32, TM = 2, TN = 2, PROMOTE = KEVD_GRAM_PROMOTE;
    static constexpr int NT = 256, WGM = 2, WGN = 2, WG_PER_CU = 2, WPE = 2, PB = 16;
};
#if KEVD_GRAM1_CFG == 1
template <> struct GramCfg<float, 1> {
    static constexpr int BI = 256, BJ = 128, KC = 32, TM = 2, TN = 2, PROMOTE = KEVD_GRAM_PROMOTE;
    static constexpr int NT = 512, WGM = 4, WGN = 2, WG_PER_CU = 1, WPE = 2, PB = 16;
};
#elif KEVD_GRAM1_CFG == 2
template <> struct GramCfg<float, 1> {
    static constexpr int BI = 256, BJ = 256, KC = 32, TM = 4, TN = 2, PROMOTE = KEVD_GRAM_PROMOTE;
    static constexpr int NT = 512, WGM = 2, WGN = 4, WG_PER_CU = 1, WPE = 2, PB = 16;
};
#elif KEVD_GRAM1_CFG == 3
template <> struct GramCfg<float, 1> {
    static constexpr int BI = 256, BJ = 256, KC = 16, TM = 4, TN = 2, PROMOTE = KEVD_GRAM_PROMOTE;
    static constexpr int NT = 512, WGM = 2, WGN = 4, WG_PER_CU = 1, WPE = 2, PB = 16;
};
#elif KEVD_GRAM1_CFG == 4
// three workgroups per CU: 16-row chunks and 8-value promotion batches keep
// a workgroup within 168 VGPRs and 54 KiB of LDS
template <> struct GramCfg<float, 1> {
    static constexpr int BI = 128, BJ = 128, KC = 16, TM = 2, TN = 2, PROMOTE = KEVD_GRAM_PROMOTE;
    static constexpr int NT = 256, WGM = 2, WGN = 2, WG_PER_CU = 3, WPE = 3, PB = 8;
};
#endif
template <int NG> struct GramCfg<double, NG> {
    static constexpr int BI = 128, BJ = 64, KC = 16, TM = 4, TN = 2, PROMOTE = 1 << 30;
    static constexpr int NT = 256, WGM = 2, WGN = 2, WG_PER_CU = 2, WPE = 2, PB = 16;
};

template <typename T, int NG>
struct GramTile {
    typedef GramCfg<T, NG> C;
    typedef Mfma<T> M;
    static constexpr int BI = C::BI, BJ = C::BJ, KC = C::KC, TM = C::TM, TN = C::TN, NT = C::NT, PB = C::PB;
    static constexpr int IL = 2;                            // interleaved tiles per load
    static constexpr int EPV = 16 / (int)sizeof(T);
    static constexpr int PI = BI + EPV, PJ = BJ + EPV;
    static constexpr int WOFF = KC * PI + KC * PJ;          // weights after the panels
    static constexpr int STAGE = WOFF + NG * KC;
    static constexpr int VI = BI / EPV, VJ = BJ / EPV;     // 16 B vectors per panel row
    static constexpr int NVI = (KC * VI + NT - 1) / NT, NVJ = (KC * VJ + NT - 1) / NT;
    static constexpr int WR = TM * M::MT, WC = TN * M::MT; // wave tile
    static_assert(C::WGM * WR == BI && C::WGN * WC == BJ && C::WGM * C::WGN * 64 == NT, "Gram wave layout");
    static_assert(NG * KC <= NT, "Gram weight staging");
    static_assert((KC * VI) % NT == 0 && (KC * VJ) % NT == 0, "Gram panel staging");
    typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type pair_t;

    typename M::acc_t acc[NG][TM][TN];
    // one chunk's global loads held in registers until stored to LDS
    struct Stage {
        uint4 ri[NVI], rj[NVJ];
        T wv;  // this thread's staged weight
    };
    int wm, wn, h, li;
    bool live = true;  // false: this wave's block lies above the diagonal (never read)

    __device__ inline void zero()
    {
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int v = 0; v < M::ACC; ++v) acc[g][a][b][v] = (T)0;
    }

    __device__ static inline void mask(uint4& u, int64_t col, int64_t r)
    {
        if (col + EPV > r) {
            T* e = reinterpret_cast<T*>(&u);
#pragma unroll
            for (int q = 0; q < EPV; ++q)
                if (col + q >= r) e[q] = (T)0;
        }
    }

    // F is read through a buffer resource over this split's rows [k_beg,
    // k_end) (num_records = their bytes, below 2^31): a row at or past k_end
    // reads as zero by the hardware's range check, so the panel loads carry
    // no predicate and no branch.  kro = first row of the chunk - k_beg.
    __device__ inline void gload(Stage& sg, __amdgpu_buffer_rsrc_t rs, int64_t ldf, int64_t ci, int64_t cj,
                                 int64_t kro, int64_t k_beg, int64_t k_end,
                                 const T* __restrict__ w1, const T* __restrict__ w2)
    {
        typedef __attribute__((ext_vector_type(4))) unsigned u4;
#pragma unroll
        for (int j = 0; j < NVI; ++j) {
            int lin = threadIdx.x + NT * j;
            int kr = lin / VI, cv = lin % VI;
            uint32_t off = (uint32_t)(((kro + kr) * ldf + ci + cv * EPV) * (int64_t)sizeof(T));
            u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            sg.ri[j] = make_uint4(v.x, v.y, v.z, v.w);
        }
#pragma unroll
        for (int j = 0; j < NVJ; ++j) {
            int lin = threadIdx.x + NT * j;
            int kr = lin / VJ, cv = lin % VJ;
            uint32_t off = (uint32_t)(((kro + kr) * ldf + cj + cv * EPV) * (int64_t)sizeof(T));
            u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            sg.rj[j] = make_uint4(v.x, v.y, v.z, v.w);
        }
        int64_t k0 = k_beg + kro;
        // threads [0, NG*KC) stage the chunk's row weights
        sg.wv = (T)0;
        if ((int)threadIdx.x < NG * KC) {
            int g = threadIdx.x / KC, kr = threadIdx.x % KC;
            int64_t row = k0 + kr;
            if (row < k_end) sg.wv = g == 0 ? w1[row] : w2[row];
        }
    }

    // Columns at or beyond r of a 16-byte vector that straddles r are zeroed
    // here rather than after the load, so that no load is waited for before
    // the chunk that needs it.
    __device__ inline void sstore(Stage& sg, T* st, int64_t ci, int64_t cj, int64_t r)
    {
#pragma unroll
        for (int j = 0; j < NVI; ++j) {
            int lin = threadIdx.x + NT * j;
            if (lin < KC * VI) {
                int kr = lin / VI, cv = lin % VI;
                mask(sg.ri[j], ci + cv * EPV, r);
                *reinterpret_cast<uint4*>(st + kr * PI + cv * EPV) = sg.ri[j];
            }
        }
#pragma unroll
        for (int j = 0; j < NVJ; ++j) {
            int lin = threadIdx.x + NT * j;
            if (lin < KC * VJ) {
                int kr = lin / VJ, cv = lin % VJ;
                mask(sg.rj[j], cj + cv * EPV, r);
                *reinterpret_cast<uint4*>(st + KC * PI + kr * PJ + cv * EPV) = sg.rj[j];
            }
        }
        if ((int)threadIdx.x < NG * KC) st[WOFF + threadIdx.x] = sg.wv;
    }

    // The fragments of K step s + 1 are read from LDS while step s
    // multiplies, and all weighted A operands of a step are formed before
    // its MFMAs, so no VALU write waits on an MFMA still reading a register.
    __device__ inline void compute(const T* cur)
    {
        if (!live) return;
        constexpr int NS = KC / M::NH;
        pair_t araw[2][TM / IL], b[2][TN / IL];
        T wk[2][NG];
        auto frag = [&](int q, int s) {
            int kr = s * M::NH + h;
#pragma unroll
            for (int g = 0; g < NG; ++g) wk[q][g] = cur[WOFF + g * KC + kr];
#pragma unroll
            for (int x = 0; x < TM / IL; ++x)
                araw[q][x] = *reinterpret_cast<const pair_t*>(cur + kr * PI + wm * WR + x * M::MT * IL + IL * li);
#pragma unroll
            for (int y = 0; y < TN / IL; ++y)
                b[q][y] = *reinterpret_cast<const pair_t*>(
                    cur + KC * PI + kr * PJ + wn * WC + y * M::MT * IL + IL * li);
        };
        frag(0, 0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int q = s & 1;
            if (s + 1 < NS) frag(q ^ 1, s + 1);
            T a[NG][TM];
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int x = 0; x < TM / IL; ++x) {
                    a[g][IL * x] = araw[q][x].x * wk[q][g];
                    a[g][IL * x + 1] = araw[q][x].y * wk[q][g];
                }
            // keep the next step's LDS reads and this step's operand scaling
            // ahead of its MFMAs (the scheduler would otherwise sink the
            // reads next to their uses and serialise them)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int x = 0; x < TM; ++x)
#pragma unroll
                    for (int y = 0; y < TN / IL; ++y) {
                        M::mma(acc[g][x][IL * y], a[g][x], b[q][y].x);
                        M::mma(acc[g][x][IL * y + 1], a[g][x], b[q][y].y);
                    }
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    // tile a = IL*x + q covers panel rows wm*WR + x*MT*IL + IL*i + q (i =
    // MFMA output row), likewise columns.  The partial tile is addressed
    // through a buffer resource: the lane part of the offset is one VGPR, the
    // compile-time part of each element's offset an SGPR constant, and the
    // read-modify-write runs PB values per batch so the loads overlap.
    __device__ inline void promote(double* out, bool first)
    {
        constexpr int NVAL = NG * TM * TN * M::ACC;
        __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(out, (short)0, NG * BI * BJ * (int)sizeof(double), 0x00020000);
        int voff = ((wm * WR + IL * M::row_lane(h)) * BJ + wn * WC + IL * li) * (int)sizeof(double);
        typedef __attribute__((ext_vector_type(2))) unsigned u2;
#pragma unroll
        for (int b0 = 0; b0 < NVAL; b0 += PB) {
            double x[PB];
#pragma unroll
            for (int q = 0; q < PB; ++q) {
                int e = b0 + q;
                int v = e % M::ACC, b = (e / M::ACC) % TN, a = (e / (M::ACC * TN)) % TM,
                    g = e / (M::ACC * TN * TM);
                x[q] = (double)acc[g][a][b][v];
            }
            if (!first) {
                // all PB loads issue before the first add uses one
                double y[PB];
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    int e = b0 + q;
                    int v = e % M::ACC, b = (e / M::ACC) % TN, a = (e / (M::ACC * TN)) % TM,
                        g = e / (M::ACC * TN * TM);
                    int row = (a / IL) * M::MT * IL + IL * M::row_off(v) + (a % IL);
                    int col = (b / IL) * M::MT * IL + (b % IL);
                    int soff = (g * BI * BJ + row * BJ + col) * (int)sizeof(double);
                    y[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < PB; ++q) x[q] += y[q];
            }
#pragma unroll
            for (int q = 0; q < PB; ++q) {
                int e = b0 + q;
                int v = e % M::ACC, b = (e / M::ACC) % TN, a = (e / (M::ACC * TN)) % TM,
                    g = e / (M::ACC * TN * TM);
                int row = (a / IL) * M::MT * IL + IL * M::row_off(v) + (a % IL);
                int col = (b / IL) * M::MT * IL + (b % IL);
                int soff = (g * BI * BJ + row * BJ + col) * (int)sizeof(double);
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x[q]), rs, voff, soff, 0);
            }
        }
        zero();
    }
};

template <typename T, int NG>
__global__ __launch_bounds__((GramCfg<T, NG>::NT))
__attribute__((amdgpu_waves_per_eu(GramCfg<T, NG>::WPE, GramCfg<T, NG>::WPE)))
void k_gram(const T* __restrict__ F, int64_t ldf, int64_t n, int64_t r, const T* __restrict__ w1,
            const T* __restrict__ w2, const int* __restrict__ tileI, const int* __restrict__ tileJ,
            int ntiles, int nitems, int64_t rows_per_split, double* part)
{
    typedef GramTile<T, NG> G;
    __shared__ __attribute__((aligned(16))) T lds[2 * G::STAGE];
    // Workgroups go to the 8 XCDs round-robin by id.  XCD x takes a
    // contiguous range of the split-major items (item = split * ntiles +
    // tile), so the workgroups resident on one XCD read the same rows of F
    // and its L2 serves the column panels they share.
    int id = blockIdx.x, xcd = id & 7, per = nitems >> 3, rem = nitems & 7;
    int item = xcd * per + (xcd < rem ? xcd : rem) + (id >> 3);
    int tile = item % ntiles, split = item / ntiles;
    int64_t ci = (int64_t)tileI[tile] * G::BI, cj = (int64_t)tileJ[tile] * G::BJ;
    int64_t k_beg = (int64_t)split * rows_per_split;
    int64_t k_end = k_beg + rows_per_split < n ? k_beg + rows_per_split : n;
    double* out = part + ((int64_t)split * ntiles + tile) * (int64_t)NG * G::BI * G::BJ;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(F + k_beg * ldf), (short)0,
        (int)(k_end > k_beg ? (k_end - k_beg) * ldf * (int64_t)sizeof(T) : 0), 0x00020000);

    G g;
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    g.wm = w / G::C::WGN;
    g.wn = w % G::C::WGN;
    g.h = lane / Mfma<T>::MT;
    g.li = lane % Mfma<T>::MT;
    // on tiles that straddle the diagonal, a wave block with every row index
    // below every column index only holds entries the reduction never reads
    g.live = ci + (int64_t)(g.wm + 1) * G::WR - 1 >= cj + (int64_t)g.wn * G::WC;
    g.zero();
    bool first = true;
    int64_t nk = k_end > k_beg ? (k_end - k_beg + G::KC - 1) / G::KC : 0;
    int since = 0;
    auto step = [&](const T* cur) {
        g.compute(cur);
        since += G::KC;
        if (since >= G::C::PROMOTE) {
            g.promote(out, first);
            first = false;
            since = 0;
        }
    };
    typename G::Stage s0;
    if (nk > 0) {
        g.gload(s0, rs, ldf, ci, cj, 0, k_beg, k_end, w1, w2);
        g.sstore(s0, lds, ci, cj, r);
    }
    if constexpr (NG == 1) {
        // Loads run two chunks ahead (the single Gram has the registers):
        // chunk c multiplies from stage c & 1 while c + 1 waits in one
        // register set and c + 2 loads into the other.
        typename G::Stage s1;
        if (nk > 1) g.gload(s1, rs, ldf, ci, cj, G::KC, k_beg, k_end, w1, w2);
        __syncthreads();
        for (int64_t c = 0; c < nk; c += 2) {
            if (c + 2 < nk) g.gload(s0, rs, ldf, ci, cj, (c + 2) * G::KC, k_beg, k_end, w1, w2);
            step(lds);
            if (c + 1 < nk) g.sstore(s1, lds + G::STAGE, ci, cj, r);
            __syncthreads();
            if (c + 1 >= nk) break;
            if (c + 3 < nk) g.gload(s1, rs, ldf, ci, cj, (c + 3) * G::KC, k_beg, k_end, w1, w2);
            step(lds + G::STAGE);
            if (c + 2 < nk) g.sstore(s0, lds, ci, cj, r);
            __syncthreads();
        }
    } else {
        __syncthreads();
        for (int64_t kc = 0; kc < nk; ++kc) {
            T* cur = lds + (kc & 1) * G::STAGE;
            T* nxt = lds + ((kc + 1) & 1) * G::STAGE;
            bool more = kc + 1 < nk;
            if (more) g.gload(s0, rs, ldf, ci, cj, (kc + 1) * G::KC, k_beg, k_end, w1, w2);
            step(cur);
            if (more) g.sstore(s0, nxt, ci, cj, r);
            __syncthreads();
        }
    }
    if (since > 0 || first) g.promote(out, first);
}

// a register set chosen at compile time
template <int N> struct RegSet { static constexpr int value = N; };

// ---------------------------------------------- weighted Gram, bf16x6
// G = F^T diag(w) F for FP32 F on the bf16 matrix cores (bf16x6 products,
// FP32 accumulation, FP64 promotion every PROMOTE rows as in k_gram, here
// 4096).
// Two tile shapes, chosen per problem by kevd_autotune_tiles:
//   256 x 256 on 8 waves (2 x 4) of 128 x 64, one workgroup per CU (the
//   default: half the panel re-reads of 128-wide tiles at large r);
//   128 x 128 on 4 waves (2 x 2) of 64 x 64, two workgroups per CU (less
//   waste in the partial edge and diagonal tiles at small r).
// 16 rows of F per chunk.  Thread t stages one column of one panel (threads
// 0..BI-1 the weighted A panel, BI..2BI-1 the B panel): it reads the
// column's 16 rows (one dword per row; a wave reads 256 contiguous bytes of
// each row), splits each value into bf16 hi, mid, lo and writes each part
// as two 16-byte halves of 8 rows.  LDS per stage: [panel][part][half]
// [column][8 rows], so an MFMA lane's fragment (8 rows of one column) is one
// 16-byte read and 32 lanes read or write 512 contiguous bytes.
template <int B_, int WGM_, int WGN_, int TM_, int TN_, int WPC>
struct GramX6Cfg {
    static constexpr int BI = B_, BJ = B_, KC = 16, NT = 2 * B_, WGM = WGM_, WGN = WGN_;
    static constexpr int TM = TM_, TN = TN_, WR = 32 * TM_, WC = 32 * TN_, PB = 16, PROMOTE = KEVD_GRAM_X6_PROMOTE;
    static constexpr int WG_PER_CU = WPC;
    static constexpr int HALF = B_ * 16;  // bytes of one half slab (B_ columns x 8 rows)
    static constexpr int PART = 2 * HALF;
    static constexpr int PANEL = 3 * PART;
    static constexpr int STAGE = 2 * PANEL;  // 48 KiB (256) / 24 KiB (128)
    static_assert(WGM * WR == BI && WGN * WC == BJ && WGM * WGN * 64 == NT, "Gram x6 wave layout");
};
typedef GramX6Cfg<256, 2, 4, 4, 2, 1> GramX6;
typedef GramX6Cfg<128, 2, 2, 2, 2, 2> GramX6s;

template <typename C>
__global__ __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_gram_x6(const float* __restrict__ F, int64_t ldf, int64_t n, int64_t r, const float* __restrict__ w,
               const int* __restrict__ tileI, const int* __restrict__ tileJ, int ntiles, int nitems,
               int64_t rows_per_split, double* part)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds[2 * C::STAGE];
    // the split-major item order and XCD assignment of k_gram
    int id = blockIdx.x, xcd = id & 7, per = nitems >> 3, rem = nitems & 7;
    int item = xcd * per + (xcd < rem ? xcd : rem) + (id >> 3);
    int tile = item % ntiles, split = item / ntiles;
    const int64_t ci = (int64_t)tileI[tile] * C::BI, cj = (int64_t)tileJ[tile] * C::BJ;
    const int64_t k_beg = (int64_t)split * rows_per_split;
    const int64_t k_end = k_beg + rows_per_split < n ? k_beg + rows_per_split : n;
    double* out = part + ((int64_t)split * ntiles + tile) * (int64_t)C::BI * C::BJ;
    // F and w over this split's rows: rows at or past k_end read as zero
    const int nb = k_end > k_beg ? (int)((k_end - k_beg) * ldf * 4) : 0;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(F + k_beg * ldf), (short)0, nb,
                                                                  0x00020000);
    __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(w + k_beg), (short)0, k_end > k_beg ? (int)((k_end - k_beg) * 4) : 0, 0x00020000);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int panel = t / C::BI;  // wave-uniform
    const int scol = t % C::BI;
    const int64_t gcol = (panel ? cj : ci) + scol;
    const bool colok = gcol < r;
    const uint32_t coff = (uint32_t)(gcol * 4);
    const uint32_t rowb = (uint32_t)(ldf * 4);
    const int wm = wave / C::WGN, wn = wave % C::WGN, li = lane & 31, hh = lane >> 5;
    // a wave block with every row index below every column index holds only
    // entries the reduction never reads
    const bool live = ci + (int64_t)(wm + 1) * C::WR - 1 >= cj + (int64_t)wn * C::WC;

    floatx16 acc[C::TM][C::TN];
    auto zero = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int a = 0; a < C::TM; ++a)
#pragma unroll
            for (int b = 0; b < C::TN; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
    };
    zero();

    // The whole offset goes in the VGPR operand: the range check that turns
    // rows at or past k_end into zeros applies to it.  Rows of F are loaded
    // two chunks ahead into two register sets (a load has two chunks'
    // products to land), the weights one chunk ahead.
    // (the register set is a constant, so v stays in registers)
    float v[2][C::KC], wv[C::KC];
    auto gload = [&](auto set, int64_t kro) __attribute__((always_inline)) {
        const uint32_t base = (uint32_t)kro * rowb + coff;
#pragma unroll
        for (int k = 0; k < C::KC; ++k)
            v[decltype(set)::value][k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, base + k * rowb, 0, 0));
    };
    auto wload = [&](int64_t kro) __attribute__((always_inline)) {
        if (panel == 0) {
            // the whole vector is bit-cast: clang reads element 0 for a
            // __builtin_bit_cast of a single ext-vector element (q.y)
            typedef __attribute__((ext_vector_type(4))) float f4;
#pragma unroll
            for (int j = 0; j < C::KC / 4; ++j) {
                f4 q = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rw, (uint32_t)kro * 4 + 16 * j, 0, 0));
                wv[4 * j] = q.x;
                wv[4 * j + 1] = q.y;
                wv[4 * j + 2] = q.z;
                wv[4 * j + 3] = q.w;
            }
        }
    };
    auto sstore = [&](auto set, unsigned char* st) __attribute__((always_inline)) {
        unsigned char* dst = st + panel * C::PANEL + scol * 16;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 p0, p1, p2;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float x = colok ? v[decltype(set)::value][8 * h + e] : 0.0f;
                if (panel == 0) x *= wv[8 * h + e];
                __bf16 a0, a1, a2;
                bf16_split3(x, a0, a1, a2);
                p0[e] = a0, p1[e] = a1, p2[e] = a2;
            }
            *reinterpret_cast<bf16x8*>(dst + h * C::HALF) = p0;
            *reinterpret_cast<bf16x8*>(dst + C::PART + h * C::HALF) = p1;
            *reinterpret_cast<bf16x8*>(dst + 2 * C::PART + h * C::HALF) = p2;
        }
    };
    auto compute = [&](const unsigned char* st) __attribute__((always_inline)) {
        if (!live) return;
        bf16x8 af[C::TM][3], bf[C::TN][3];
#pragma unroll
        for (int b = 0; b < C::TN; ++b)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                bf[b][q] = *reinterpret_cast<const bf16x8*>(st + C::PANEL + q * C::PART + hh * C::HALF +
                                                            (wn * C::WC + b * 32 + li) * 16);
#pragma unroll
        for (int a = 0; a < C::TM; ++a)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                af[a][q] = *reinterpret_cast<const bf16x8*>(st + q * C::PART + hh * C::HALF +
                                                            (wm * C::WR + a * 32 + li) * 16);
        // products outermost: eight independent accumulators between two
        // MFMAs into the same one
#pragma unroll
        for (int p = 0; p < X6_NPROD; ++p)
#pragma unroll
            for (int a = 0; a < C::TM; ++a)
#pragma unroll
                for (int b = 0; b < C::TN; ++b) mfma_bf16_32(acc[a][b], af[a][x6_a(p)], bf[b][x6_b(p)]);
    };
    // value v of tile (a, b): row wm*WR + 32 a + 8 (v / 4) + 4 hh + v % 4,
    // column wn*WC + 32 b + li; the partial tile is row-major BI x BJ FP64
    auto promote = [&](bool first) __attribute__((always_inline)) {
        __amdgpu_buffer_rsrc_t ro =
            __builtin_amdgcn_make_buffer_rsrc(out, (short)0, C::BI * C::BJ * (int)sizeof(double), 0x00020000);
        const int voff = ((wm * C::WR + 4 * hh) * C::BJ + wn * C::WC + li) * (int)sizeof(double);
        typedef __attribute__((ext_vector_type(2))) unsigned u2;
        constexpr int NVAL = C::TM * C::TN * 16;
#pragma unroll
        for (int b0 = 0; b0 < NVAL; b0 += C::PB) {
            double x[C::PB];
#pragma unroll
            for (int q = 0; q < C::PB; ++q) {
                int e = b0 + q, vv = e % 16, b = (e / 16) % C::TN, a = e / (16 * C::TN);
                x[q] = (double)acc[a][b][vv];
            }
            if (!first) {
                double y[C::PB];
#pragma unroll
                for (int q = 0; q < C::PB; ++q) {
                    int e = b0 + q, vv = e % 16, b = (e / 16) % C::TN, a = e / (16 * C::TN);
                    int soff = ((a * 32 + 8 * (vv >> 2) + (vv & 3)) * C::BJ + b * 32) * (int)sizeof(double);
                    y[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(ro, voff, soff, 0));
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < C::PB; ++q) x[q] += y[q];
            }
#pragma unroll
            for (int q = 0; q < C::PB; ++q) {
                int e = b0 + q, vv = e % 16, b = (e / 16) % C::TN, a = e / (16 * C::TN);
                int soff = ((a * 32 + 8 * (vv >> 2) + (vv & 3)) * C::BJ + b * 32) * (int)sizeof(double);
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x[q]), ro, voff, soff, 0);
            }
        }
        zero();
    };

    const int64_t nk = k_end > k_beg ? (k_end - k_beg + C::KC - 1) / C::KC : 0;
    bool first = true;
    int since = 0;
    // chunk j's rows sit in v[j & 1]; loads past the last chunk read zeros
    // (the range check) and are stored to the idle stage, never multiplied
    if (nk > 0) {
        wload(0);
        gload(RegSet<0>(), 0);
        gload(RegSet<1>(), C::KC);
        sstore(RegSet<0>(), lds);
        wload(C::KC);
        gload(RegSet<0>(), 2 * C::KC);
    }
    __syncthreads();
    // A workgroup's waves are placed on the CU's four SIMDs in turn, so
    // waves w and w + 4 share one.  With one workgroup per CU the barrier
    // lines up their phases (both multiply, then both split and store, and
    // the matrix core idles through the split); the second wave of each
    // pair splits and stores first, so each wave's split runs under the
    // other's products.
    const bool store_first = C::WG_PER_CU == 1 && KEVD_GRAM_X6_PINGPONG && wave >= C::NT / 128;
    // chunk c: multiply it, store chunk c + 1 (in vn), then load the
    // weights of c + 2 and the rows of c + 3 into vn.  Each order has a loop
    // of its own: in one loop with both orders the register-reuse waits
    // would take the other order's loads as outstanding and wait for all.
    auto run = [&](auto order) __attribute__((always_inline)) {
        auto iter = [&](int64_t c, auto vn) __attribute__((always_inline)) {
            unsigned char* cur = lds + (c & 1) * C::STAGE;
            unsigned char* nxt = lds + ((c + 1) & 1) * C::STAGE;
            if constexpr (decltype(order)::value) {
                sstore(vn, nxt);
                wload((c + 2) * C::KC);
                gload(vn, (c + 3) * C::KC);
                compute(cur);
            } else {
                compute(cur);
                sstore(vn, nxt);
                wload((c + 2) * C::KC);
                gload(vn, (c + 3) * C::KC);
            }
            since += C::KC;
            if (since >= C::PROMOTE) {
                promote(first);
                first = false;
                since = 0;
            }
            __syncthreads();
        };
        int64_t c = 0;
        for (; c + 1 < nk; c += 2) {
            iter(c, RegSet<1>());
            iter(c + 1, RegSet<0>());
        }
        if (c < nk) iter(c, RegSet<1>());
    };
    if (store_first)
        run(RegSet<1>());
    else
        run(RegSet<0>());
    if (since > 0 || first) promote(first);
}

template <typename T>
__global__ void k_cast(const double* x, int64_t n, T* out)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (T)x[i];
}

// G_g(i, j) = sum_split part[split][tile][g](i, j) for the lower triangle,
// mirrored; column-major r x r FP64.
template <int BI, int BJ, int NG>
__global__ __launch_bounds__(256) void k_gram_reduce(const double* part, int nsplit, int ntiles,
                                                     const int* tileI, const int* tileJ, int64_t r,
                                                     double* G1, double* G2)
{
    int tile = blockIdx.x;
    int64_t ci = (int64_t)tileI[tile] * BI, cj = (int64_t)tileJ[tile] * BJ;
    for (int e = blockIdx.y * 256 + threadIdx.x; e < BI * BJ; e += 256 * gridDim.y) {
        int row = e / BJ, col = e % BJ;
        int64_t i = ci + row, j = cj + col;
        if (i >= r || j >= r || i < j) continue;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            double s = 0.0;
            for (int sp = 0; sp < nsplit; ++sp)
                s += part[(((int64_t)sp * ntiles + tile) * NG + g) * BI * BJ + e];
            double* G = g == 0 ? G1 : G2;
            G[i + j * r] = s;
            G[j + i * r] = s;
        }
    }
}

// ------------------------------------------------------------- U GEMM

// U(i, c) = scale_i * (F_i . Y_c) for c < kout, FP32 row-major n x kout;
// column c == kout (if extra) goes to P1_i in FP64.  Y rows are the B
// operand (kc x r, leading dimension ldy).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ugemm(const T* __restrict__ F, int64_t ldf, int64_t n,
                                               int64_t r, const T* __restrict__ Y, int64_t ldy,
                                               int kc, int kout, const double* __restrict__ scale,
                                               float* U, double* P1)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM;
    __shared__ __attribute__((aligned(16))) T lds[G::LDS_ELEMS];
    __shared__ double scs[BM];  // the block's row scales
    G g;
    g.init();
    int64_t r0 = (int64_t)blockIdx.x * BM;
    int c0 = blockIdx.y * Cfg<T>::BN;
    for (int i = threadIdx.x; i < BM; i += 256) scs[i] = r0 + i < n ? scale[r0 + i] : 0.0;
    g.set_live_cols(kc - c0);
    g.mainloop_live(lds, F + r0 * ldf, ldf, n - r0, Y + (int64_t)c0 * ldy, ldy, kc - c0, r);
    __syncthreads();  // scs (the main loop's barriers publish it too, for r > 0)
    // U rows of this block through a buffer resource: a row past n falls
    // outside its range and is dropped, offsets are 32-bit (BM kout values)
    int64_t vrows = n - r0 < BM ? n - r0 : BM;
    __amdgpu_buffer_rsrc_t rsU =
        __builtin_amdgcn_make_buffer_rsrc(U + r0 * kout, (short)0, (int)(vrows * kout * 4), 0x00020000);
    double sc[G::NTM][G::M::ACC];
    int rl[G::NTM][G::M::ACC];
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < G::M::ACC; ++v) {
            rl[a][v] = g.row_of(a, v);
            sc[a][v] = scs[rl[a][v]];
        }
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                int col = c0 + g.col_of(b);
                double x = (double)g.acc[a][b][v] * sc[a][v];
                if (col < kout)
                    buffer_store((float)x, rsU, (uint32_t)((rl[a][v] * kout + col) * 4));
                else if (P1 && col < kc && r0 + rl[a][v] < n)
                    P1[r0 + rl[a][v]] = x;
            }
}

// The U product with an FP32 factor on the bf16 matrix cores: k_ugemm<float>
// with the bf16x6 main loop (bf16x6.h), for kc <= 128 output columns.  Yx
// holds the B operand's parts chunk-major (k_pack_y_x6).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ugemm_x6(
    const float* __restrict__ F, int64_t ldf, int64_t n, int64_t r, const __bf16* __restrict__ Yx, int kc, int kout,
    const double* __restrict__ scale, float* U, double* P1)
{
    typedef Cfg<float>::G G;
    constexpr int BM = Cfg<float>::BM;
    static_assert(X6Stage<G>::BYTES <= G::LDS_ELEMS * (int)sizeof(float), "the x6 stage fits");
    __shared__ __attribute__((aligned(16))) float lds[G::LDS_ELEMS];
    __shared__ double scs[BM];  // the block's row scales
    G g;
    g.init();
    int64_t r0 = (int64_t)blockIdx.x * BM;
    for (int i = threadIdx.x; i < BM; i += 256) scs[i] = r0 + i < n ? scale[r0 + i] : 0.0;
    g.set_live_cols(kc);
    x6_nt_mainloop_live<G>(g, reinterpret_cast<unsigned char*>(lds), F + r0 * ldf, ldf, n - r0, Yx, r);
    __syncthreads();  // scs
    int64_t vrows = n - r0 < BM ? n - r0 : BM;
    __amdgpu_buffer_rsrc_t rsU =
        __builtin_amdgcn_make_buffer_rsrc(U + r0 * kout, (short)0, (int)(vrows * kout * 4), 0x00020000);
    double sc[G::NTM][G::M::ACC];
    int rl[G::NTM][G::M::ACC];
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int v = 0; v < G::M::ACC; ++v) {
            rl[a][v] = g.row_of(a, v);
            sc[a][v] = scs[rl[a][v]];
        }
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                int col = g.col_of(b);
                double x = (double)g.acc[a][b][v] * sc[a][v];
                if (col < kout)
                    buffer_store((float)x, rsU, (uint32_t)((rl[a][v] * kout + col) * 4));
                else if (P1 && col < kc && r0 + rl[a][v] < n)
                    P1[r0 + rl[a][v]] = x;
            }
}

__global__ void k_trace_shift(double* A, int64_t r, double* tr_out)
{
    __shared__ double s[256];
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < r; i += 256) t += A[i + i * r];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    double tr = s[0];
    double sh = DBL_EPSILON * tr;
    for (int64_t i = threadIdx.x; i < r; i += 256) A[i + i * r] += sh;
    if (threadIdx.x == 0) {
        tr_out[0] = tr;
        tr_out[1] = sh;
    }
}

// Y (column-major r x k, FP64) -> rows of the GEMM B operand in T with
// leading dimension ldy; optionally scaled per column and with one extra
// row `extra` (length r).
template <typename T>
__global__ void k_pack_y(const double* Y, int64_t r, int k, const double* colscale,
                         const double* extra, T* out, int64_t ldy)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int c = blockIdx.y;
    if (i >= r) return;
    if (c < k) {
        double v = Y[i + (int64_t)c * r];
        if (colscale) v *= colscale[c];
        out[(int64_t)c * ldy + i] = (T)v;
    } else if (extra) {
        out[(int64_t)c * ldy + i] = (T)extra[i];
    }
}

// k_pack_y's B operand as bf16x6 parts for k_ugemm_x6: row c < 128 of the
// operand (column c of Y scaled by colscale[c], or extra for c = k), split
// into three bf16 parts, chunk-major in fragment order (X6_CHUNK elements
// per 32 values of r, x6_idx); values past r and rows past k (+1 with
// extra) are zero.  Thread per (value, row).
__global__ void k_pack_y_x6(const double* Y, int64_t r, int k, const double* colscale, const double* extra,
                            int64_t nchunk, __bf16* out)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int c = blockIdx.y;
    if (i >= nchunk * 32) return;
    float v = 0.0f;
    if (i < r) {
        if (c < k) {
            double x = Y[i + (int64_t)c * r];
            if (colscale) x *= colscale[c];
            v = (float)x;
        } else if (c == k && extra) {
            v = (float)extra[i];
        }
    }
    __bf16 h, md, lo;
    bf16_split3(v, h, md, lo);
    __bf16* e = out + (i / 32) * X6_CHUNK + x6_idx(c, (int)(i % 32));
    e[0] = h;
    e[X6_CHUNK / 3] = md;
    e[2 * X6_CHUNK / 3] = lo;
}

__global__ void k_rowsum_dev(const double* P1, int64_t n, double* bdev)
{
    __shared__ double s[256];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(P1[i] - 1.0));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bdev[blockIdx.x] = s[0];
}

__global__ void k_fill_d(double* p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_inv_sqrt_tail(const double* lam, int64_t r, int k, double* out)
{
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= k) return;
    double l = lam[r - k + c];
    out[c] = l > 0.0 ? 1.0 / sqrt(l) : 0.0;
}

// ------------------------------------------------- regenerated K panels

// K(p0 + i, S_j) for i < np, j < r, into out (row-major, leading dimension
// ldo; columns r .. ldo-1 zero): the distance product on MFMA with FP64
// norms and the downdate's Gaussian (gauss_val), exactly 1 on a point's own
// pivot.  One BM x BN tile per block.
template <typename T>
__global__ __launch_bounds__(256) void k_kpanel(const T* __restrict__ X, const double* __restrict__ nrm,
                                                int32_t dp, int64_t p0, int64_t np, const T* __restrict__ XS,
                                                const double* __restrict__ nrmS, const int64_t* __restrict__ piv,
                                                int64_t r, GaussScale gs, T* __restrict__ out, int64_t ldo)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    __shared__ __attribute__((aligned(16))) T lds[G::LDS_ELEMS];
    G g;
    g.init();
    const int64_t r0 = (int64_t)blockIdx.x * BM, c0 = (int64_t)blockIdx.y * BN;
    g.mainloop(lds, X + (p0 + r0) * dp, dp, np - r0, XS + c0 * dp, dp, r - c0, dp);
#pragma unroll
    for (int a = 0; a < G::NTM; ++a)
#pragma unroll
        for (int b = 0; b < G::NTN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                const int64_t row = r0 + g.row_of(a, v), col = c0 + g.col_of(b);
                if (row >= np || col >= ldo) continue;
                T kv = (T)0;
                if (col < r) {
                    double D = nrm[p0 + row] + nrmS[col] - 2.0 * (double)g.acc[a][b][v];
                    D = D > 0.0 ? D : 0.0;
                    kv = piv[col] == p0 + row ? (T)1 : gauss_val(D, gs, (T*)nullptr);
                }
                out[row * ldo + col] = kv;
            }
}

// y += x over count entries.
__global__ void k_add(double* y, const double* x, int64_t count)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) y[i] += x[i];
}

// ------------------------------------------------------------- host side

struct EvdBufs {
    double *part = nullptr, *c = nullptr, *t = nullptr, *y = nullptr, *y2 = nullptr;
    double *w1 = nullptr, *w2 = nullptr, *bmin = nullptr, *bdev = nullptr, *mm = nullptr;
};

// Rows [0, n) of an n x r row-major matrix with leading dimension ldf: the
// resident factor, or one regenerated panel of K(:,S).
template <typename T> struct Rows {
    const T* F;
    int64_t ldf, n, r;
};

// part[b][j] for the COLSUM_ROWS-row blocks of src (block b of src is
// part row b): column partials of w^T src (w null: ones).
template <typename T>
void colsum_part(kevd_ctx* ctx, const Rows<T>& src, const double* w, double* part, int64_t rpad,
                 int rows = COLSUM_ROWS)
{
    int nb = (int)((src.n + rows - 1) / rows);
    k_colsum<T><<<nb, 256, 0, ctx->stream>>>(src.F, src.ldf, src.n, src.r, w, part, rpad, rows);
}

template <typename T>
kevd_status colsum_rows(kevd_ctx* ctx, const Rows<T>& src, const double* w, double* c)
{
    // at least 4 blocks per CU (rows a multiple of 32, at most COLSUM_ROWS)
    int64_t rows = KEVD_ROUND_UP((src.n + 4 * (int64_t)ctx->num_cu - 1) / (4 * (int64_t)ctx->num_cu), 32);
    rows = std::max<int64_t>(32, std::min<int64_t>(COLSUM_ROWS, rows));
    int nb = (int)((src.n + rows - 1) / rows);
    int64_t rpad = KEVD_ROUND_UP(src.r, 4);
    double* part = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "e_cpart", (size_t)nb * rpad, &part));
    ProfScope ps(ctx, TC_MATVEC);
    colsum_part<T>(ctx, src, w, part, rpad, (int)rows);
    k_colsum_reduce<<<(unsigned)((src.r + 63) / 64), 1024, 0, ctx->stream>>>(part, nb, src.r, rpad, c);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

// Launch of k_rowdot over src with nb blocks (per-block min / max into
// bmin, bdev).
template <typename T>
void rowdot_launch(kevd_ctx* ctx, const Rows<T>& src, const double* c, double* y, int mode, const double* s,
                   double* s_new, double* bmin, double* bdev, int nb)
{
    if (src.r <= ROWDOT_LDS_MAX)
        k_rowdot<T, true><<<nb, 256, src.r * sizeof(double), ctx->stream>>>(src.F, src.ldf, src.n, src.r, c, y,
                                                                            bmin, mode, s, s_new, bdev);
    else
        k_rowdot<T, false><<<nb, 256, 0, ctx->stream>>>(src.F, src.ldf, src.n, src.r, c, y, bmin, mode, s, s_new,
                                                       bdev);
}

inline int rowdot_blocks(const kevd_ctx* ctx, int64_t n)
{
    return (int)std::min<int64_t>((n + 15) / 16, 8 * (int64_t)ctx->num_cu);
}

// y = F c; returns min(y) (and max dev in Sinkhorn mode) through host_mm.
template <typename T>
kevd_status rowdot_rows(kevd_ctx* ctx, const Rows<T>& src, const double* c, double* y, int mode, const double* s,
                   double* s_new, double host_mm[2])
{
    int nb = rowdot_blocks(ctx, src.n);
    double *bmin, *bdev, *mm;
    KEVD_TRY(ctx_buf_t(ctx, "e_bmin", nb, &bmin));
    KEVD_TRY(ctx_buf_t(ctx, "e_bdev", nb, &bdev));
    KEVD_TRY(ctx_buf_t(ctx, "e_mm", 2, &mm));
    {
        ProfScope ps(ctx, TC_MATVEC);
        rowdot_launch<T>(ctx, src, c, y, mode, s, s_new, bmin, bdev, nb);
        k_block_minmax<<<1, 256, 0, ctx->stream>>>(bmin, mode == 1 ? bdev : nullptr, nb, mm);
    }
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipMemcpyAsync(host_mm, mm, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

// G = F^T diag(w) F for FP32 F by k_gram_x6 and the split reduction.
template <typename C>
kevd_status gram_x6(kevd_ctx* ctx, const Rows<float>& src, const double* w, double* G, const char* wbuf)
{
    int64_t n = src.n, r = src.r;
    int ti_n = (int)((r + C::BI - 1) / C::BI), tj_n = (int)((r + C::BJ - 1) / C::BJ);
    std::vector<int> TI, TJ;
    for (int I = 0; I < ti_n; ++I)
        for (int J = 0; J < tj_n; ++J)
            if ((int64_t)I * C::BI + C::BI - 1 >= (int64_t)J * C::BJ) {
                TI.push_back(I);
                TJ.push_back(J);
            }
    int ntiles = (int)TI.size();
    // One workgroup per CU; at least ~2048 workgroups, the split count
    // chosen so that the grid fills whole waves of the device's CUs.  A
    // split's rows are addressed by 32-bit buffer offsets.
    const int64_t slots = C::WG_PER_CU * (int64_t)ctx->num_cu;
    const int64_t rp_max = ((int64_t)INT32_MAX / (src.ldf * 4)) / C::KC * C::KC - C::KC;
    if (rp_max < C::KC) return kevd_fail(ctx, KEVD_EINVAL, "Gram: factor rows too long for 32-bit offsets");
    int ns0 = (int)std::max<int64_t>({(int64_t)1, (2048 + ntiles - 1) / ntiles, (n + rp_max - 1) / rp_max});
    int nsplit = 1;
    int64_t rps = n;
    double best = -1.0;
    for (int ns = ns0; ns <= 2 * ns0; ++ns) {
        int64_t rp = KEVD_ROUND_UP((n + ns - 1) / ns, C::KC);
        if (rp < 2048) rp = 2048;
        if (rp > rp_max) rp = rp_max;
        int64_t nse = (n + rp - 1) / rp, wk = nse * ntiles;
        double eff = (double)wk / (double)(slots * ((wk + slots - 1) / slots));
        if (eff > best + 1e-9) {
            best = eff;
            nsplit = (int)nse;
            rps = rp;
        }
    }
    int nitems = ntiles * nsplit;
    int *dTI, *dTJ;
    double* part;
    float* wt = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, wbuf, n, &wt));
    k_cast<float><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(w, n, wt);
    KEVD_TRY(ctx_buf_t(ctx, "e_xTI", ntiles, &dTI));
    KEVD_TRY(ctx_buf_t(ctx, "e_xTJ", ntiles, &dTJ));
    KEVD_TRY(ctx_buf_t(ctx, "e_gpart", (size_t)nsplit * ntiles * C::BI * C::BJ, &part));
    KEVD_HIP(ctx, hipMemcpyAsync(dTI, TI.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(dTJ, TJ.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    {
        ProfScope ps(ctx, TC_GRAM);
        k_gram_x6<C><<<nitems, C::NT, 0, ctx->stream>>>((const float*)src.F, src.ldf, n, r, wt, dTI, dTJ, ntiles,
                                                      nitems, rps, part);
    }
    k_gram_reduce<C::BI, C::BJ, 1><<<dim3(ntiles, gram_reduce_slices(ctx, ntiles)), 256, 0, ctx->stream>>>(
        part, nsplit, ntiles, dTI, dTJ, r, G, nullptr);
    KEVD_HIP(ctx, hipGetLastError());
    ctx->tm.gram_flop += (double)n * (double)r * (double)(r + 1);
    return KEVD_OK;
}

template <typename T, int NG>
kevd_status gram_rows(kevd_ctx* ctx, const Rows<T>& src, const double* w1, const double* w2, double* G1, double* G2)
{
    if constexpr (std::is_same<T, float>::value) {
        if (use_bf16x6()) {
            if (ctx->gram_tile == 128) {
                KEVD_TRY(gram_x6<GramX6s>(ctx, src, w1, G1, "e_gw1"));
                if (NG > 1) KEVD_TRY(gram_x6<GramX6s>(ctx, src, w2, G2, "e_gw2"));
            } else {
                KEVD_TRY(gram_x6<GramX6>(ctx, src, w1, G1, "e_gw1"));
                if (NG > 1) KEVD_TRY(gram_x6<GramX6>(ctx, src, w2, G2, "e_gw2"));
            }
            return KEVD_OK;
        }
    }
    typedef GramCfg<T, NG> C;
    int64_t n = src.n, r = src.r;
    int ti_n = (int)((r + C::BI - 1) / C::BI), tj_n = (int)((r + C::BJ - 1) / C::BJ);
    std::vector<int> TI, TJ;
    for (int I = 0; I < ti_n; ++I)
        for (int J = 0; J < tj_n; ++J)
            if ((int64_t)I * C::BI + C::BI - 1 >= (int64_t)J * C::BJ) {
                TI.push_back(I);
                TJ.push_back(J);
            }
    int ntiles = (int)TI.size();
    // At least ~2048 workgroups, with the split count chosen so that the
    // grid fills whole waves of the device's resident slots.
    const int64_t slots = C::WG_PER_CU * (int64_t)ctx->num_cu;
    // A split's rows are addressed by 32-bit buffer offsets: at most 2^31
    // bytes of F per split.
    const int64_t rp_max = ((int64_t)INT32_MAX / (src.ldf * (int64_t)sizeof(T))) / C::KC * C::KC - C::KC;
    if (rp_max < C::KC) return kevd_fail(ctx, KEVD_EINVAL, "Gram: factor rows too long for 32-bit offsets");
    int ns0 = (int)std::max<int64_t>({(int64_t)1, (2048 + ntiles - 1) / ntiles, (n + rp_max - 1) / rp_max});
    int nsplit = 1;
    int64_t rps = n;
    double best = -1.0;
    for (int ns = ns0; ns <= 2 * ns0; ++ns) {
        int64_t rp = KEVD_ROUND_UP((n + ns - 1) / ns, C::KC);
        if (KEVD_GRAM_SPLIT_ROWS > 0) rp = KEVD_ROUND_UP(KEVD_GRAM_SPLIT_ROWS, C::KC);
        if (rp < 2048) rp = 2048;
        if (rp > rp_max) rp = rp_max;
        int64_t nse = (n + rp - 1) / rp, w = nse * ntiles;
        double eff = (double)w / (double)(slots * ((w + slots - 1) / slots));
        if (eff > best + 1e-9) {
            best = eff;
            nsplit = (int)nse;
            rps = rp;
        }
    }
    int nitems = ntiles * nsplit;
    int *dTI, *dTJ;
    double* part;
    T *wt1 = nullptr, *wt2 = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "e_gw1", n, &wt1));
    k_cast<T><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(w1, n, wt1);
    if (NG > 1) {
        KEVD_TRY(ctx_buf_t(ctx, "e_gw2", n, &wt2));
        k_cast<T><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(w2, n, wt2);
    }
    KEVD_TRY(ctx_buf_t(ctx, "e_gTI", ntiles, &dTI));
    KEVD_TRY(ctx_buf_t(ctx, "e_gTJ", ntiles, &dTJ));
    KEVD_TRY(ctx_buf_t(ctx, "e_gpart", (size_t)nsplit * ntiles * NG * C::BI * C::BJ, &part));
    KEVD_HIP(ctx, hipMemcpyAsync(dTI, TI.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(dTJ, TJ.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    {
        ProfScope ps(ctx, TC_GRAM);
        k_gram<T, NG><<<nitems, C::NT, 0, ctx->stream>>>(src.F, src.ldf, n, r, wt1, wt2, dTI, dTJ, ntiles, nitems,
                                                      rps, part);
    }
    k_gram_reduce<C::BI, C::BJ, NG><<<dim3(ntiles, gram_reduce_slices(ctx, ntiles)), 256, 0, ctx->stream>>>(
        part, nsplit, ntiles, dTI, dTJ, r, G1, G2);
    KEVD_HIP(ctx, hipGetLastError());
    ctx->tm.gram_flop += (double)NG * (double)n * (double)r * (double)(r + 1);
    return KEVD_OK;
}

// Y (r x k column-major FP64, scaled per column by colscale) and an
// optional extra column, packed as the B operand of the U product.
template <typename T>
kevd_status pack_y(kevd_ctx* ctx, int64_t r, const double* Y, int k, const double* colscale, const double* extra,
                   T** Yp_out, int64_t* ldy_out)
{
    int kc = k + (extra ? 1 : 0);
    int64_t ldy = KEVD_ROUND_UP(r, 64);
    T* Yp;
    KEVD_TRY(ctx_buf_t(ctx, "e_Ypack", (size_t)kc * ldy, &Yp));
    k_pack_y<T><<<dim3((unsigned)((r + 255) / 256), kc), 256, 0, ctx->stream>>>(Y, r, k, colscale, extra, Yp,
                                                                                ldy);
    KEVD_HIP(ctx, hipGetLastError());
    *Yp_out = Yp;
    *ldy_out = ldy;
    return KEVD_OK;
}

// U = diag(rowscale) src Yp (and P1 from the extra column) for src's rows;
// U and P1 point at src's first row.
template <typename T>
kevd_status ugemm_rows(kevd_ctx* ctx, const Rows<T>& src, const T* Yp, int64_t ldy, int kc, int k,
                       const double* rowscale, float* U, double* P1)
{
    dim3 grid((unsigned)((src.n + Cfg<T>::BM - 1) / Cfg<T>::BM), (unsigned)((kc + Cfg<T>::BN - 1) / Cfg<T>::BN));
    {
        ProfScope ps(ctx, TC_UGEMM);
        k_ugemm<T><<<grid, 256, 0, ctx->stream>>>(src.F, src.ldf, src.n, src.r, Yp, ldy, kc, k, rowscale, U, P1);
    }
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

// One fused Sinkhorn pass over src (see k_sinkhorn_rows / k_sinkhorn_pass):
// q = the next iterate's factor per row, the per-block partials of
// src^T q into part (row b of part for block b of src) and the per-block
// statistics into bstat.  Returns false when the rank is beyond the fused
// kernels' register tiling.
template <typename T>
bool sk_fused_ok(int64_t r)
{
    constexpr int V = 16 / (int)sizeof(T), VPT_MAX = sizeof(T) == 4 ? 4 : 8;
    int vpt = (int)(((r + V - 1) / V + 255) / 256);
    return vpt <= VPT_MAX;
}

template <typename T>
void sk_pass_launch(kevd_ctx* ctx, const Rows<T>& src, const double* u, const double* s, int damped, double* q,
                    double* part, int64_t rpad, double* bstat)
{
    constexpr int V = 16 / (int)sizeof(T), VPT_MAX = sizeof(T) == 4 ? 4 : 8;
    const int64_t r = src.r;
    const int nb = (int)((src.n + SK_ROWS - 1) / SK_ROWS);
    const int vpt = (int)(((r + V - 1) / V + 255) / 256);
#define KEVD_SK_ROWS(VQ)                                                                                   \
    k_sinkhorn_rows<T, VQ><<<nb, 256, 0, ctx->stream>>>(src.F, src.ldf, src.n, r, u, s, damped, q, part, rpad, \
                                                       bstat)
#define KEVD_SK_PASS(VP)                                                                                   \
    k_sinkhorn_pass<T, VP><<<nb, 256, 0, ctx->stream>>>(src.F, src.ldf, src.n, r, u, s, damped, q, part, rpad, \
                                                       bstat)
    const int vpq = (int)(((r + V - 1) / V + 63) / 64);
    if constexpr (sizeof(T) == 4) {
        if (vpq <= 8) {
            if (vpq <= 2)
                KEVD_SK_ROWS(2);
            else if (vpq <= 4)
                KEVD_SK_ROWS(4);
            else
                KEVD_SK_ROWS(8);
            return;
        }
    }
    if (vpt <= 1)
        KEVD_SK_PASS(1);
    else if (vpt <= 2)
        KEVD_SK_PASS(2);
    else if (vpt <= 4)
        KEVD_SK_PASS(4);
    else
        KEVD_SK_PASS(VPT_MAX);
#undef KEVD_SK_PASS
#undef KEVD_SK_ROWS
}

// The passes of a kEVD over the resident factor F (n x rank).
template <typename T> struct FactorOps {
    kevd_ctx* ctx;
    Rows<T> src;
    explicit FactorOps(kevd_ctx* c) : ctx(c), src{(const T*)c->F, c->ldf, c->n, c->rank} {}
    kevd_status begin() { return KEVD_OK; }
    // c = F^T w (w null: ones; from the downdate's partials when F is the
    // factor kevd_arpc left)
    kevd_status colsum(const double* w, double* c)
    {
        if (!w && ctx->cs_valid && ctx->cs_part && ctx->cs_nblk > 0) {
            ProfScope ps(ctx, TC_MATVEC);
            k_colsum_reduce<<<(unsigned)((src.r + 63) / 64), 1024, 0, ctx->stream>>>(ctx->cs_part, (int)ctx->cs_nblk,
                                                                                    src.r, ctx->ldf, c);
            KEVD_HIP(ctx, hipGetLastError());
            return KEVD_OK;
        }
        return colsum_rows<T>(ctx, src, w, c);
    }
    // y = F c (Sinkhorn mode 1 as k_rowdot); min(y), max dev to host
    kevd_status rowdot(const double* c, double* y, int mode, const double* s, double* s_new, double mm[2])
    {
        return rowdot_rows<T>(ctx, src, c, y, mode, s, s_new, mm);
    }
    kevd_status gram1(const double* w, double* G) { return gram_rows<T, 1>(ctx, src, w, nullptr, G, nullptr); }
    kevd_status gram2(const double* w1, const double* w2, double* A, double* B)
    {
        return gram_rows<T, 2>(ctx, src, w1, w2, A, B);
    }
    // U = diag(rowscale) F Y diag(colscale) (and P1 = diag(rowscale) F extra)
    kevd_status ugemm(const double* Y, int k, const double* colscale, const double* extra, const double* rowscale,
                      float* U, double* P1)
    {
        const int kc = k + (extra ? 1 : 0);
        if constexpr (std::is_same<T, float>::value) {
            if (use_bf16x6() && kc <= Cfg<float>::BN) {
                // the product on the bf16 matrix cores
                const int64_t nchunk = (src.r + 31) / 32;
                __bf16* Yx;
                KEVD_TRY(ctx_buf(ctx, "e_Ypackx6", (size_t)nchunk * X6_CHUNK * sizeof(__bf16), (void**)&Yx));
                k_pack_y_x6<<<dim3((unsigned)((nchunk * 32 + 255) / 256), Cfg<float>::BN), 256, 0, ctx->stream>>>(
                    Y, src.r, k, colscale, extra, nchunk, Yx);
                {
                    ProfScope ps(ctx, TC_UGEMM);
                    k_ugemm_x6<<<(unsigned)((src.n + Cfg<float>::BM - 1) / Cfg<float>::BM), 256, 0, ctx->stream>>>(
                        src.F, src.ldf, src.n, src.r, Yx, kc, k, rowscale, U, P1);
                }
                KEVD_HIP(ctx, hipGetLastError());
                return KEVD_OK;
            }
        }
        T* Yp;
        int64_t ldy;
        KEVD_TRY(pack_y<T>(ctx, src.r, Y, k, colscale, extra, &Yp, &ldy));
        return ugemm_rows<T>(ctx, src, Yp, ldy, k + (extra ? 1 : 0), k, rowscale, U, P1);
    }
    // one fused Sinkhorn pass: q, v = F^T q and the five statistics (st3)
    kevd_status skpass(const double* u, const double* s, int damped, double* q, double* v, double* st3)
    {
        const int nb = (int)((src.n + SK_ROWS - 1) / SK_ROWS);
        const int64_t rpad = KEVD_ROUND_UP(src.r, 4);
        double *part, *bstat;
        KEVD_TRY(ctx_buf_t(ctx, "e_skpart", (size_t)nb * rpad, &part));
        KEVD_TRY(ctx_buf_t(ctx, "e_skstat", (size_t)SK_NSTAT * nb, &bstat));
        {
            ProfScope ps(ctx, TC_MATVEC);
            sk_pass_launch<T>(ctx, src, u, s, damped, q, part, rpad, bstat);
            k_colsum_reduce<<<(unsigned)((src.r + 63) / 64), 1024, 0, ctx->stream>>>(part, nb, src.r, rpad, v);
            k_stat_reduce<<<1, 256, 0, ctx->stream>>>(bstat, nb, st3);
        }
        KEVD_HIP(ctx, hipGetLastError());
        return KEVD_OK;
    }
};

#define BLAS(call)                                                                  \
    do {                                                                            \
        rocblas_status b_ = (call);                                                 \
        if (b_ != rocblas_status_success)                                           \
            return kevd_fail(ctx, KEVD_EHIP, std::string(#call) + " failed: " +     \
                                                 rocblas_status_to_string(b_));     \
    } while (0)

// The passes of a kEVD with no factor held: the reference's LM form
// (ks_kevd.cu regenerates K(R,S) row panels in every pass, 755-800,
// 1091-1168, 1627-1690, 2229-2275).  With the pivots S and L^-1
// (K(S,S) = L L^T) kept by the LM form of ARPC, F = K(:,S) L^-T is never
// formed: every pass regenerates K(R,S) for PANEL_ROWS rows at a time
// (k_kpanel) and runs the factor-form kernels on the panel, and the
// r-vectors and r x r results move between the two bases with L^-1:
//   F^T w = L^-1 (K^T w),  F c = K (L^-T c),  F^T W F = L^-1 (K^T W K) L^-T,
//   F Y = K (L^-T Y).
// L^-1 is row-major lower, so read column-major it is Lt = L^-T (upper).
// Device memory beyond the points: one panel (PANEL_ROWS x rank) and the
// per-block partials, instead of the n x rank factor.
constexpr int64_t PANEL_ROWS = 65536;  // a multiple of every kernel's row block

template <typename T> struct PanelOps {
    kevd_ctx* ctx;
    int64_t n, r, ldo = 0, P = 0;
    const double* Lt = nullptr;  // L^-T, column-major upper, leading dimension ldl
    int64_t ldl = 0;
    T *panel = nullptr, *XS = nullptr;
    double *nrmS = nullptr, *tv = nullptr, *tY = nullptr, *tG = nullptr;
    int64_t* piv = nullptr;
    GaussScale gs{};
    explicit PanelOps(kevd_ctx* c) : ctx(c), n(c->n), r(c->rank) {}

    kevd_status begin()
    {
        if (!ctx->lm_linv || (int64_t)ctx->pivots.size() != r)
            return kevd_fail(ctx, KEVD_EINVAL, "evd: no pivots and L^-1 to regenerate the factor from");
        ldo = KEVD_ROUND_UP(r, 64);
        P = KEVD_ROUND_UP(std::min<int64_t>(n, PANEL_ROWS), 512);
        Lt = ctx->lm_linv;
        ldl = ctx->lm_ldl;
        KEVD_TRY(ctx_buf_t(ctx, "p_panel", (size_t)P * ldo, &panel));
        KEVD_TRY(ctx_buf_t(ctx, "p_XS", (size_t)r * ctx->dp, &XS));
        KEVD_TRY(ctx_buf_t(ctx, "p_nrmS", r, &nrmS));
        KEVD_TRY(ctx_buf_t(ctx, "p_piv", r, &piv));
        KEVD_TRY(ctx_buf_t(ctx, "p_tv", r, &tv));
        KEVD_HIP(ctx, hipMemcpyAsync(piv, ctx->pivots.data(), r * sizeof(int64_t), hipMemcpyHostToDevice,
                                     ctx->stream));
        KEVD_TRY(kblock_gather(ctx, piv, r, sizeof(T) == 4 ? KEVD_MIXED : KEVD_FP64, XS, nrmS));
        gs = GaussScale{-1.0 / ctx->factor_eps, -1.4426950408889634074 / ctx->factor_eps};
        return KEVD_OK;
    }
    int64_t panels() const { return (n + P - 1) / P; }
    Rows<T> gen(int64_t p0)
    {
        const int64_t np = std::min(P, n - p0);
        const T* X = sizeof(T) == 4 ? (const T*)ctx->X32 : (const T*)ctx->X64;
        const double* nrm = sizeof(T) == 4 ? ctx->nrm32 : ctx->nrm64;
        dim3 grid((unsigned)((np + Cfg<T>::BM - 1) / Cfg<T>::BM), (unsigned)((ldo + Cfg<T>::BN - 1) / Cfg<T>::BN));
        k_kpanel<T><<<grid, 256, 0, ctx->stream>>>(X, nrm, ctx->dp, p0, np, XS, nrmS, piv, r, gs, panel, ldo);
        return Rows<T>{panel, ldo, np, r};
    }
    // x := L^-1 x (to the factor's basis) or x := L^-T x (to the kernel's)
    kevd_status to_f(double* x)
    {
        rocblas_handle h = blas_handle(ctx);
        BLAS(rocblas_dtrmv(h, rocblas_fill_upper, rocblas_operation_transpose, rocblas_diagonal_non_unit,
                           (rocblas_int)r, Lt, (rocblas_int)ldl, x, 1));
        return KEVD_OK;
    }
    kevd_status to_k(double* x)
    {
        rocblas_handle h = blas_handle(ctx);
        BLAS(rocblas_dtrmv(h, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                           (rocblas_int)r, Lt, (rocblas_int)ldl, x, 1));
        return KEVD_OK;
    }
    kevd_status copy_to_k(const double* x, double* out)
    {
        KEVD_HIP(ctx, hipMemcpyAsync(out, x, r * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        return to_k(out);
    }

    kevd_status colsum(const double* w, double* c)
    {
        const int nbt = (int)((n + COLSUM_ROWS - 1) / COLSUM_ROWS);
        const int64_t rpad = KEVD_ROUND_UP(r, 4);
        double* part;
        KEVD_TRY(ctx_buf_t(ctx, "p_cpart", (size_t)nbt * rpad, &part));
        {
            ProfScope ps(ctx, TC_MATVEC);
            for (int64_t p0 = 0; p0 < n; p0 += P)
                colsum_part<T>(ctx, gen(p0), w ? w + p0 : nullptr, part + (p0 / COLSUM_ROWS) * rpad, rpad);
            k_colsum_reduce<<<(unsigned)((r + 63) / 64), 1024, 0, ctx->stream>>>(part, nbt, r, rpad, c);
        }
        KEVD_HIP(ctx, hipGetLastError());
        return to_f(c);
    }
    kevd_status rowdot(const double* c, double* y, int mode, const double* s, double* s_new, double mm[2])
    {
        KEVD_TRY(copy_to_k(c, tv));
        const int nbmax = rowdot_blocks(ctx, P);
        double *bmin, *bdev, *mmd;
        KEVD_TRY(ctx_buf_t(ctx, "p_bmin", (size_t)panels() * nbmax, &bmin));
        KEVD_TRY(ctx_buf_t(ctx, "p_bdev", (size_t)panels() * nbmax, &bdev));
        KEVD_TRY(ctx_buf_t(ctx, "e_mm", 2, &mmd));
        int off = 0;
        {
            ProfScope ps(ctx, TC_MATVEC);
            for (int64_t p0 = 0; p0 < n; p0 += P) {
                Rows<T> src = gen(p0);
                const int nb = rowdot_blocks(ctx, src.n);
                rowdot_launch<T>(ctx, src, tv, y + p0, mode, s ? s + p0 : nullptr, s_new ? s_new + p0 : nullptr,
                                 bmin + off, bdev + off, nb);
                off += nb;
            }
            k_block_minmax<<<1, 256, 0, ctx->stream>>>(bmin, mode == 1 ? bdev : nullptr, off, mmd);
        }
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_HIP(ctx, hipMemcpyAsync(mm, mmd, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return KEVD_OK;
    }
    // G = L^-1 (K^T diag(w) K) L^-T, the Gram accumulated over the panels
    template <int NG>
    kevd_status gram(const double* w1, const double* w2, double* G1, double* G2)
    {
        const size_t rr = (size_t)r * r;
        double *acc1, *acc2 = nullptr, *t1, *t2 = nullptr;
        KEVD_TRY(ctx_buf_t(ctx, "p_gacc1", rr, &acc1));
        KEVD_TRY(ctx_buf_t(ctx, "p_gt1", rr, &t1));
        if (NG > 1) {
            KEVD_TRY(ctx_buf_t(ctx, "p_gacc2", rr, &acc2));
            KEVD_TRY(ctx_buf_t(ctx, "p_gt2", rr, &t2));
        }
        const unsigned gr = (unsigned)((rr + 255) / 256);
        for (int64_t p0 = 0; p0 < n; p0 += P) {
            Rows<T> src = gen(p0);
            double* o1 = p0 == 0 ? acc1 : t1;
            double* o2 = p0 == 0 ? acc2 : t2;
            KEVD_TRY((gram_rows<T, NG>(ctx, src, w1 + p0, NG > 1 ? w2 + p0 : nullptr, o1, o2)));
            if (p0 > 0) {
                k_add<<<gr, 256, 0, ctx->stream>>>(acc1, t1, (int64_t)rr);
                if (NG > 1) k_add<<<gr, 256, 0, ctx->stream>>>(acc2, t2, (int64_t)rr);
            }
        }
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_TRY(sandwich(acc1, t1, G1));
        if (NG > 1) KEVD_TRY(sandwich(acc2, t2, G2));
        return KEVD_OK;
    }
    // G = L^-1 Ak L^-T = Lt^T Ak Lt (tmp: r x r scratch)
    kevd_status sandwich(const double* Ak, double* tmp, double* G)
    {
        rocblas_handle h = blas_handle(ctx);
        const double one = 1.0;
        const rocblas_int R = (rocblas_int)r, LD = (rocblas_int)ldl;
        ProfScope ps(ctx, TC_DENSE);
        BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, R, R, &one, Lt, LD, Ak, R, tmp, R));
        BLAS(rocblas_dtrmm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, R, R, &one, Lt, LD, tmp, R, G, R));
        return KEVD_OK;
    }
    kevd_status gram1(const double* w, double* G) { return gram<1>(w, nullptr, G, nullptr); }
    kevd_status gram2(const double* w1, const double* w2, double* A, double* B) { return gram<2>(w1, w2, A, B); }
    kevd_status ugemm(const double* Y, int k, const double* colscale, const double* extra, const double* rowscale,
                      float* U, double* P1)
    {
        rocblas_handle h = blas_handle(ctx);
        const double one = 1.0;
        double *Yk, *ek = nullptr;
        KEVD_TRY(ctx_buf_t(ctx, "p_Yk", (size_t)r * k, &Yk));
        BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, (rocblas_int)r, (rocblas_int)k, &one, Lt, (rocblas_int)ldl, Y,
                           (rocblas_int)r, Yk, (rocblas_int)r));  // Yk = L^-T Y
        if (extra) {
            KEVD_TRY(ctx_buf_t(ctx, "p_ek", r, &ek));
            KEVD_TRY(copy_to_k(extra, ek));
        }
        T* Yp;
        int64_t ldy;
        KEVD_TRY(pack_y<T>(ctx, r, Yk, k, colscale, ek, &Yp, &ldy));
        const int kc = k + (extra ? 1 : 0);
        for (int64_t p0 = 0; p0 < n; p0 += P)
            KEVD_TRY(ugemm_rows<T>(ctx, gen(p0), Yp, ldy, kc, k, rowscale + p0, U + p0 * k,
                                   P1 ? P1 + p0 : nullptr));
        return KEVD_OK;
    }
    kevd_status skpass(const double* u, const double* s, int damped, double* q, double* v, double* st3)
    {
        KEVD_TRY(copy_to_k(u, tv));
        const int nbt = (int)((n + SK_ROWS - 1) / SK_ROWS);
        const int64_t rpad = KEVD_ROUND_UP(r, 4);
        double *part, *bstat;
        KEVD_TRY(ctx_buf_t(ctx, "p_skpart", (size_t)nbt * rpad, &part));
        KEVD_TRY(ctx_buf_t(ctx, "p_skstat", (size_t)SK_NSTAT * nbt, &bstat));
        {
            ProfScope ps(ctx, TC_MATVEC);
            for (int64_t p0 = 0; p0 < n; p0 += P)
                sk_pass_launch<T>(ctx, gen(p0), tv, s + p0, damped, q + p0, part + (p0 / SK_ROWS) * rpad, rpad,
                                  bstat + SK_NSTAT * (p0 / SK_ROWS));
            k_colsum_reduce<<<(unsigned)((r + 63) / 64), 1024, 0, ctx->stream>>>(part, nbt, r, rpad, v);
            k_stat_reduce<<<1, 256, 0, ctx->stream>>>(bstat, nbt, st3);
        }
        KEVD_HIP(ctx, hipGetLastError());
        return to_f(v);
    }
};

// Symmetric eigensolve of the column-major r x r FP64 matrix W (lower
// triangle used); V overwrites W, lam ascending.
kevd_status eigh(kevd_ctx* ctx, double* W, int64_t r, double* lam, bool timed = true)
{
    rocblas_handle h = blas_handle(ctx);
    double* E;
    int* info;
    KEVD_TRY(ctx_buf_t(ctx, "e_E", r, &E));
    KEVD_TRY(ctx_buf_t(ctx, "e_info", 1, &info));
    {
        ProfScope ps(ctx, timed ? TC_EIG : -1);
        BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)r, W,
                              (rocblas_int)r, lam, E, info));
    }
    int hi = 0;
    KEVD_HIP(ctx, hipMemcpyAsync(&hi, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hi != 0) return kevd_fail(ctx, KEVD_EEIG, "syevd failed, info " + std::to_string(hi));
    return KEVD_OK;
}

// ------------------------------------------------ leading eigenpairs

__global__ void k_randn(double* Q, int64_t count, uint64_t seed)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    uint32_t o[4];
    kevd_philox((uint32_t)i, 0u, 7u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o);
    double u1 = kevd_u64_to_unit(((uint64_t)o[1] << 32) | o[0]);
    double u2 = (double)((((uint64_t)o[3] << 32) | o[2]) >> 11) * 0x1.0p-53;
    Q[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void k_symmetrize(double* H, int l)
{
    int i = blockIdx.x * 16 + (threadIdx.x >> 4), j = blockIdx.y * 16 + (threadIdx.x & 15);
    if (i < l && j < l && i < j) {
        double a = 0.5 * (H[i + (int64_t)j * l] + H[j + (int64_t)i * l]);
        H[i + (int64_t)j * l] = a;
        H[j + (int64_t)i * l] = a;
    }
}

// res2[c] = s_i^T (G s_i) - theta_i^2 for the leading k Ritz pairs, where
// G = Y^T Y and Y = W Q: the squared residual ||W x_i - theta_i x_i||^2 of
// the Ritz vector x_i = Q s_i (Q orthonormal, ||s_i|| = 1).
__global__ void k_ritz_res2(const double* S, const double* GS, const double* th, int l, int k,
                            double* res2)
{
    int c = blockIdx.x;
    int i = l - k + c;
    double acc = 0.0;
    for (int j = threadIdx.x; j < l; j += 256) acc += S[j + (int64_t)i * l] * GS[j + (int64_t)i * l];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) res2[c] = red[0] - th[i] * th[i];
}

__global__ void k_lam_tail(const double* th, int l, int64_t r, double* lam)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= r) return;
    lam[i] = i >= r - l ? th[i - (r - l)] : NAN;
}

// G += c tr(G) I for the l x l column-major Gram matrix of shifted
// Cholesky QR (one workgroup).
__global__ __launch_bounds__(256) void k_chol_shift(double* G, int l, double c)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < l; i += 256) acc += G[i + (int64_t)i * l];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    double s = c * red[0];
    for (int i = threadIdx.x; i < l; i += 256) G[i + (int64_t)i * l] += s;
}

// Upper Cholesky factor of the symmetric positive definite l x l matrix A
// (column-major, upper triangle read and overwritten with R, A = R^T R), in
// one workgroup, for l <= CHOL_LMAX.  Right-looking over 32-row panels:
//   the 32 x 32 diagonal block is factored by one wave in registers (lane c
//   owns column c); the pivot and row k of R reach every lane as scalar
//   operands (v_readlane), with no LDS round trip;
//   R12 = R11^-T A12 by forward substitution, one thread per trailing
//   column, also kept in LDS;
//   A22 -= R12^T R12 on the upper triangle, read-modify-write through L2.
// *info = 0, or the 1-based index of the first non-positive pivot (the
// factorisation stops there), as LAPACK potrf.
constexpr int CHOL_LMAX = 512;

// Lane `lane` (wave-uniform) of a double, broadcast through scalar registers.
__device__ inline double readlane_f64(double v, int lane)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
// A batched launch factors matrix b = blockIdx.x at A + b l^2 (info + b).
__global__ __launch_bounds__(1024) void k_potrf_upper(double* A, int l, int* info)
{
    A += (int64_t)blockIdx.x * l * l;
    info += blockIdx.x;
    __shared__ double D[32][33];
    __shared__ double P[32][CHOL_LMAX - 32];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j0 = 0; j0 < l; j0 += 32) {
        const int nb = l - j0 < 32 ? l - j0 : 32, j1 = j0 + nb, nt = l - j1;
        for (int e = tid; e < nb * nb; e += 1024) {
            int r = e % nb, c = e / nb;
            D[r][c] = A[(j0 + r) + (int64_t)(j0 + c) * l];
        }
        __syncthreads();
        if (tid < 64) {
            // one wave; lanes 32..63 mirror lanes 0..31 (c >= 32 owns nothing)
            const int c = tid;
            double col[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) col[r] = (r < nb && c < nb) ? D[r][c] : 0.0;
            int fail = 0;
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                if (k < nb && !fail) {
                    double akk = readlane_f64(col[k], k);
                    if (!(akk > 0.0)) {
                        fail = j0 + k + 1;
                    } else {
                        double rkk = sqrt(akk);
                        double rkc = c == k ? rkk : (c > k && c < 32 ? col[k] / rkk : 0.0);
                        col[k] = rkc;
#pragma unroll
                        for (int i = k + 1; i < 32; ++i) {
                            double rki = readlane_f64(rkc, i);
                            if (c >= i) col[i] -= rki * rkc;
                        }
                    }
                }
            }
            if (c == 0 && fail) bad = fail;
#pragma unroll
            for (int r = 0; r < 32; ++r)
                if (r <= c && c < nb) D[r][c] = col[r];
        }
        __syncthreads();
        if (bad) break;
        for (int e = tid; e < nb * nb; e += 1024) {
            int r = e % nb, c = e / nb;
            if (r <= c) A[(j0 + r) + (int64_t)(j0 + c) * l] = D[r][c];
        }
        // R12 = R11^-T A12: column c of the panel solves R11^T x = a
        if (tid < nt) {
            const int c = j1 + tid;
            double x[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) x[k] = k < nb ? A[(j0 + k) + (int64_t)c * l] : 0.0;
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                if (k < nb) {
                    double v = x[k];
#pragma unroll
                    for (int m = 0; m < 32; ++m)
                        if (m < k) v -= D[m][k] * x[m];
                    x[k] = v / D[k][k];
                }
            }
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (k < nb) {
                    A[(j0 + k) + (int64_t)c * l] = x[k];
                    P[k][tid] = x[k];
                }
        }
        __syncthreads();
        // A22 -= R12^T R12 on the upper triangle, in 4 x 4 register tiles
        // (tile rows ti <= tile columns tc): 8 LDS reads per 16 FMAs
        const int ntt = (nt + 3) / 4;
        for (int e = tid; e < ntt * ntt; e += 1024) {
            int ti = e % ntt, tc = e / ntt;
            if (ti > tc) continue;
            double acc[4][4] = {};
            for (int k = 0; k < nb; ++k) {
                double pi[4], pc[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    pi[q] = 4 * ti + q < nt ? P[k][4 * ti + q] : 0.0;
                    pc[q] = 4 * tc + q < nt ? P[k][4 * tc + q] : 0.0;
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] += pi[a] * pc[b];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                int cc = 4 * tc + b;
                if (cc >= nt) break;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    int ii = 4 * ti + a;
                    if (ii < nt && ii <= cc) A[(j1 + ii) + (int64_t)(j1 + cc) * l] -= acc[a][b];
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) *info = bad;
}

// The same factorisation for l <= CHOLR_LMAX with the whole upper triangle
// in registers: right-looking and unblocked, one barrier per column.  The
// 32 x 32 blocks (bi <= bj) of the upper triangle are dealt to the eight
// waves round-robin in row-major order (block b to wave b % 8); lane
// (pr, pc) of a wave holds the 4 x 4 patch rows 4 pr .. 4 pr + 3, columns
// 4 pc .. 4 pc + 3 of each of its blocks.  Step k reads row k of the partly
// updated matrix from LDS, published during step k - 1 as u_j = A(k, j) for
// j > k and 0 for j <= k (from the first column of k's block on, which
// covers every row and column a live block reads), with A(k, k) and its
// reciprocal (formed by the publishing lane) in their own slots.  Row k goes
// straight to A (A(k, k) and u_j, scaled into R(k, k) = sqrt(A(k, k)) and
// R(k, j) = u_j / R(k, k) after the last step), and every entry takes
// A(i, j) -= u_i (u_j / A(k, k)) (8 LDS reads, 4 multiplies and 16 FMAs per
// patch; rows and columns <= k see u = 0).  Then the owners of row k + 1
// publish it into the other LDS buffer.  A block is final once k passes its
// last row, so a wave's blocks (in order of their last row) retire in
// phases: phase p runs the steps up to block p's last row on blocks p ..,
// and every step is straight-line code (a branch around a block would make
// the compiler keep a second copy of its registers).  Entries outside l
// read as 0; the strict lower triangle of A is left as it was.  *info as
// k_potrf_upper (rows from the failing one on keep A).
constexpr int CHOLR_LMAX = 256, CHOLR_NT = 512, CHOLR_MAXB = 5;
__global__ __launch_bounds__(CHOLR_NT) void k_potrf_reg(double* A, int l, int* info)
{
    typedef __attribute__((ext_vector_type(2))) double d2;
    A += (int64_t)blockIdx.x * l * l;
    info += blockIdx.x;
    __shared__ __attribute__((aligned(16))) double buf[2][CHOLR_LMAX + 2];  // + a spare slot
    __shared__ double dbuf[2][2];
    const int tid = threadIdx.x, lane = tid & 63;
    // the wave index as a scalar: the block bookkeeping is wave-uniform
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pr = lane >> 3, pc = lane & 7;
    const int nb = (l + 31) / 32, nblk = nb * (nb + 1) / 2;
    const int nq = nblk > w ? (nblk - w + 7) / 8 : 0;  // blocks of this wave
    // slot s holds the wave's block s - (MAXB - nq) (none below 0, so the
    // slots run in order of retirement: empty ones first)
    int rb[CHOLR_MAXB], cb[CHOLR_MAXB], kend[CHOLR_MAXB];
    bool have[CHOLR_MAXB];
    double e[CHOLR_MAXB][4][4];
#pragma unroll
    for (int sl = 0; sl < CHOLR_MAXB; ++sl) {
        const int q = sl - (CHOLR_MAXB - nq);
        int b = w + 8 * q, bi = 0;
        have[sl] = q >= 0;
        if (!have[sl]) b = 0;
        while (bi < nb && b >= nb - bi) b -= nb - bi++;  // row-major over bi <= bj
        rb[sl] = have[sl] ? 32 * bi : 0;  // an empty slot is never stepped
        cb[sl] = have[sl] ? 32 * (bi + b) : 0;
        kend[sl] = have[sl] ? min(l, 32 * bi + 32) : 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int i = rb[sl] + 4 * pr + a, j = cb[sl] + 4 * pc + c;
                e[sl][a][c] = (have[sl] && i < l && j < l) ? A[i + (int64_t)j * l] : 0.0;
            }
    }
    for (int j = tid; j < 2 * (CHOLR_LMAX + 2); j += CHOLR_NT) (&buf[0][0])[j] = 0.0;
    __syncthreads();
    // row 0 into buffer 0 (its block row is the first one)
#pragma unroll
    for (int sl = 0; sl < CHOLR_MAXB; ++sl)
        if (have[sl] && rb[sl] == 0 && pr == 0)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int j = cb[sl] + 4 * pc + c;
                if (j == 0) {
                    dbuf[0][0] = e[sl][0][c];
                    dbuf[0][1] = 1.0 / e[sl][0][c];
                }
                if (j < l) buf[0][j] = j > 0 ? e[sl][0][c] : 0.0;
            }
    __syncthreads();
    int bad = 0, k = 0;
    // phase p < MAXB: the steps up to block p's last row on slots p ..;
    // phase MAXB: the rest of the steps (every wave takes every barrier)
#pragma unroll
    for (int p = 0; p <= CHOLR_MAXB; ++p) {
        const int k_end = p < CHOLR_MAXB ? kend[p] : l;
        for (; k < k_end; ++k) {
            const double* u = buf[k & 1];
            const double dk = dbuf[k & 1][0], invd = dbuf[k & 1][1];
            if (!(dk > 0.0)) {  // every thread reads the same pivot: a uniform exit
                bad = k + 1;
                goto done;
            }
            // row k of R before its scaling by 1 / R(k, k) (after the steps)
            if (tid < l - k) A[k + (int64_t)(k + tid) * l] = tid == 0 ? dk : u[k + tid];
            double* nxt = buf[(k + 1) & 1];
#pragma unroll
            for (int sl = p; sl < CHOLR_MAXB; ++sl) {
                const int r0 = rb[sl] + 4 * pr, c0 = cb[sl] + 4 * pc;
                const d2 rv0 = *reinterpret_cast<const d2*>(u + r0), rv1 = *reinterpret_cast<const d2*>(u + r0 + 2);
                const d2 cv0 = *reinterpret_cast<const d2*>(u + c0), cv1 = *reinterpret_cast<const d2*>(u + c0 + 2);
                const double ur[4] = {rv0.x, rv0.y, rv1.x, rv1.y};
                const double uc[4] = {cv0.x * invd, cv0.y * invd, cv1.x * invd, cv1.y * invd};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) e[sl][a][c] -= ur[a] * uc[c];
                // row k + 1: published, 0 up to its diagonal (a store past
                // l goes to the spare slot CHOLR_LMAX).  The patch row holding it, (k + 1) % 4, is the same for the whole
                // wave: a scalar switch picks it.
                if (have[sl] && rb[sl] <= k + 1 && k + 1 < rb[sl] + 32) {  // uniform
                    if (pr == ((k + 1 - rb[sl]) >> 2)) {
                        double v[4];
                        switch ((k + 1) & 3) {
                        case 0:
#pragma unroll
                            for (int c = 0; c < 4; ++c) v[c] = e[sl][0][c];
                            break;
                        case 1:
#pragma unroll
                            for (int c = 0; c < 4; ++c) v[c] = e[sl][1][c];
                            break;
                        case 2:
#pragma unroll
                            for (int c = 0; c < 4; ++c) v[c] = e[sl][2][c];
                            break;
                        default:
#pragma unroll
                            for (int c = 0; c < 4; ++c) v[c] = e[sl][3][c];
                        }
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const int j = c0 + c;
                            nxt[j < l ? j : CHOLR_LMAX] = j > k + 1 ? v[c] : 0.0;
                            if (j == k + 1) {  // the pivot and its reciprocal
                                dbuf[(k + 1) & 1][0] = v[c];
                                dbuf[(k + 1) & 1][1] = 1.0 / v[c];
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
done:
    // R(i, i) = sqrt(A(i, i)) and R(i, j) = A(i, j) / R(i, i) on the final
    // rows (all, or those above a failing pivot): thread (i, j0) scales row
    // i in the columns j >= i of parity j0, the column-major loads of a wave
    // contiguous
    const int kf = bad ? bad - 1 : l;
    double* rs = &buf[0][0];
    __syncthreads();
    for (int i = tid; i < kf; i += CHOLR_NT) rs[i] = sqrt(A[i + (int64_t)i * l]);
    __syncthreads();
    const int i = tid & (CHOLR_LMAX - 1), j0 = tid >> 8;
    if (i < kf) {
        const double d = rs[i], ri = 1.0 / d;
#pragma unroll 8
        for (int j = i + ((i ^ j0) & 1); j < l; j += 2) A[i + (int64_t)j * l] = j == i ? d : A[i + (int64_t)j * l] * ri;
    }
    if (tid == 0) *info = bad;
}

// KEVD_CHOL_PANEL=1 in the environment keeps the panel Cholesky
// (k_potrf_upper) for every block width, for comparison.
bool chol_panel_only()
{
    static const bool v = [] {
        const char* e = std::getenv("KEVD_CHOL_PANEL");
        return e && e[0] == '1';
    }();
    return v;
}

// Householder tridiagonalisation of the symmetric l x l FP64 matrix M
// (column-major in A, lda = l, l <= SYTRD_LMAX; both triangles read) in one
// workgroup, with LAPACK dsytrd's lower-storage output (Golub and Van Loan,
// Matrix Computations, 4th ed., Alg. 8.3.1): D, E the tridiagonal, and
// reflector j = I - tau_j v v^T with v(j+1) = 1 and v(j+2:l) stored in
// A(j+2:l, j), so rocSOLVER's dstedc and dormtr finish the eigensolve.
// Reflector j is dlarfg's: beta = -sign(alpha) ||(alpha, x)||,
// tau = (beta - alpha) / beta, v = x / (alpha - beta).  Then
// p = tau M v, w = p - (tau / 2)(p^T v) v and M -= v w^T + w v^T on the
// trailing block.  The lower block-triangle of M lives in registers: on
// 32 x 32 blocks (A, B), A >= B, thread (ty < 16, tx < 32) holds rows ty and
// ty + 16 of column tx of every block (diagonal blocks in full).  The
// symmetric product sums each stored entry into its row and, below the
// diagonal blocks, into its column; both partial sets are reduced through
// LDS in a fixed order (the row partials padded against bank conflicts).
// Every wave forms the reflector itself from the published column, so a
// column costs three barriers.
constexpr int SYTRD_LMAX = 256;
__device__ __attribute__((always_inline)) inline void sytrd_lower_body(double* A, int l, double* D, double* E,
                                                                       double* tau)
{
    constexpr int NBK = SYTRD_LMAX / 32, NBLK = NBK * (NBK + 1) / 2;
    __shared__ double colv[2][SYTRD_LMAX], pv[2][SYTRD_LMAX];
    __shared__ double rpart[32][SYTRD_LMAX + 1], cpart[16][SYTRD_LMAX];
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5, lane = tid & 63;
    double e[NBLK][2];
#pragma unroll
    for (int a = 0; a < NBK; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                int row = 32 * a + ty + 16 * hh, col = 32 * b + tx;
                e[a * (a + 1) / 2 + b][hh] = (row < l && col < l) ? A[row + (int64_t)col * l] : 0.0;
            }
    __syncthreads();  // every entry is in registers before A is written
    for (int i = 0; i + 1 < l; ++i) {
        const int q = i & 1, bi = i >> 5;
        // column i below the diagonal (and its diagonal entry) from its owners
        if (tx == (i & 31)) {
#pragma unroll
            for (int a = 0; a < NBK; ++a)
#pragma unroll
                for (int b = 0; b <= a; ++b)
                    if (b == bi) {
#pragma unroll
                        for (int hh = 0; hh < 2; ++hh) {
                            int row = 32 * a + ty + 16 * hh;
                            double v = e[a * (a + 1) / 2 + b][hh];
                            if (row > i && row < l) colv[q][row] = v;
                            if (row == i) D[i] = v;
                        }
                    }
        }
        __syncthreads();
        // the reflector, formed identically by every wave
        double s2 = 0.0;
        for (int row = i + 2 + lane; row < l; row += 64) s2 += colv[q][row] * colv[q][row];
        s2 = wave_sum(s2);
        const double alpha = colv[q][i + 1];
        double t = 0.0, scale = 0.0, beta = alpha;
        if (s2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + s2), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        if (tid == 0) {
            E[i] = beta;
            tau[i] = t;
            A[(i + 1) + (int64_t)i * l] = beta;
        }
        auto vval = [&](int x) { return x == i + 1 ? 1.0 : ((x > i + 1 && x < l) ? colv[q][x] * scale : 0.0); };
        for (int row = i + 2 + tid; row < l; row += 512) A[row + (int64_t)i * l] = colv[q][row] * scale;
        // p = M v on rows > i: row sums over the blocks left of the diagonal
        // (and the diagonal block), column sums over the blocks below it
        double vc[NBK], csum[NBK];
#pragma unroll
        for (int c = 0; c < NBK; ++c) {
            vc[c] = vval(32 * c + tx);
            csum[c] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < NBK; ++a) {
            double r0 = 0.0, r1 = 0.0;
            if (32 * a + 31 > i) {
                const double v0 = vval(32 * a + ty), v1 = vval(32 * a + ty + 16);
#pragma unroll
                for (int c = 0; c <= a; ++c) {
                    const double e0 = e[a * (a + 1) / 2 + c][0], e1 = e[a * (a + 1) / 2 + c][1];
                    r0 += e0 * vc[c];
                    r1 += e1 * vc[c];
                    if (c < a) csum[c] += e0 * v0 + e1 * v1;
                }
            }
            rpart[tx][32 * a + ty] = r0;
            rpart[tx][32 * a + ty + 16] = r1;
        }
#pragma unroll
        for (int c = 0; c < NBK; ++c) cpart[ty][32 * c + tx] = csum[c];
        __syncthreads();
        if (tid < SYTRD_LMAX) {
            const int x = tid;
            double sr = 0.0, sq = 0.0;
            if (x > i && x < l) {
#pragma unroll 8
                for (int j = 0; j < 32; ++j) sr += rpart[j][x];
#pragma unroll 8
                for (int j = 0; j < 16; ++j) sq += cpart[j][x];
            }
            pv[q][x] = (x > i && x < l) ? t * (sr + sq) : 0.0;
        }
        __syncthreads();
        // w = p - (tau / 2)(p^T v) v; every wave forms p^T v itself
        double dot = 0.0;
        for (int x = i + 1 + lane; x < l; x += 64) dot += pv[q][x] * vval(x);
        dot = wave_sum(dot);
        const double kc = -0.5 * t * dot;
        double wc[NBK];
#pragma unroll
        for (int c = 0; c < NBK; ++c) wc[c] = pv[q][32 * c + tx] + kc * vc[c];
#pragma unroll
        for (int a = 0; a < NBK; ++a) {
            if (32 * a + 31 > i) {
                double vr[2], wr[2];
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    vr[hh] = vval(32 * a + ty + 16 * hh);
                    wr[hh] = pv[q][32 * a + ty + 16 * hh] + kc * vr[hh];
                }
#pragma unroll
                for (int c = 0; c <= a; ++c)
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh)
                        e[a * (a + 1) / 2 + c][hh] -= vr[hh] * wc[c] + wr[hh] * vc[c];
            }
        }
    }
    // the last diagonal entry
    const int il = l - 1, bl = il >> 5;
    if (tx == (il & 31)) {
#pragma unroll
        for (int a = 0; a < NBK; ++a)
            if (a == bl) {
#pragma unroll
                for (int hh = 0; hh < 2; ++hh)
                    if (32 * a + ty + 16 * hh == il) D[il] = e[a * (a + 1) / 2 + a][hh];
            }
    }
}
__global__ __launch_bounds__(512) void k_sytrd_lower(double* A, int l, double* D, double* E, double* tau)
{
    sytrd_lower_body(A, l, D, E, tau);
}
// matrix b = blockIdx.x at A + b l^2 (D, E, tau + b l)
__global__ __launch_bounds__(512) void k_sytrd_lower_b(double* A, int l, double* D, double* E, double* tau)
{
    const int64_t b = blockIdx.x;
    sytrd_lower_body(A + b * l * l, l, D + b * l, E + b * l, tau + b * l);
}

// Eigenpairs of the symmetric tridiagonal matrices of a batch (matrix b:
// diagonal D + b l, sub-diagonal E + b l, order l <= TQL_LMAX), one
// workgroup each, by implicit symmetric QR steps with the Wilkinson shift
// (Golub and Van Loan, Matrix Computations, 4th ed., Alg. 8.3.2 and 8.3.3):
// each step takes the bottom unreduced block [p, q] (sub-diagonal entries
// below DBL_EPSILON (|a_i| + |a_i+1|) are set to zero first) and chases the
// bulge of a shifted Givens rotation down it.  Wave 0 forms the
// steps (the deflation scan by the whole wave, the rotations by lane 0); the
// next waves apply them to the rows of Z (held in LDS, one thread per row),
// one step behind, so a step costs one barrier.  The
// eigenvalues go to lam + b l ascending and the eigenvectors to the
// column-major l x l matrix Z + b l^2 in the same order; *info (+ b) = 1 if
// the steps exceed 30 l.
constexpr int TQL_LMAX = 128;
__device__ inline int tql_step(double* a, double* e, int l, double* c, double* s, int* lo)
{
    // (called by all 64 lanes of one wave; l <= 128) deflation: lane i tests
    // e[i] and e[i + 64], and two ballots give the bottom unreduced block
    const int lane = threadIdx.x & 63;
    uint64_t m[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = lane + 64 * h;
        bool nz = false;
        if (i + 1 < l) {
            double ei = e[i];
            if (fabs(ei) <= DBL_EPSILON * (fabs(a[i]) + fabs(a[i + 1]))) {
                e[i] = 0.0;
                ei = 0.0;
            }
            nz = ei != 0.0;
        }
        m[h] = __ballot(nz);
    }
    // q - 1: the last coupling left; p: one past the last zero below it
    const int top = m[1] ? 127 - __builtin_clzll(m[1]) : (m[0] ? 63 - __builtin_clzll(m[0]) : -1);
    if (top < 0) return 0;
    const uint64_t below0 = top >= 64 ? ~0ull : (top ? (1ull << top) - 1 : 0ull);
    const uint64_t below1 = top > 64 ? (1ull << (top - 64)) - 1 : 0ull;
    const uint64_t z0 = ~m[0] & below0, z1 = ~m[1] & below1;
    const int zt = z1 ? 127 - __builtin_clzll(z1) : (z0 ? 63 - __builtin_clzll(z0) : -1);
    const int p = zt + 1, q = top + 1;
    *lo = p;
    if (lane != 0) return q - p;
    // the Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to a_q
    const double d = 0.5 * (a[q - 1] - a[q]), bq = e[q - 1];
    const double mu = a[q] - bq * bq / (d + (d >= 0.0 ? 1.0 : -1.0) * sqrt(d * d + bq * bq));
    // the chain of rotations carries a[k] and e[k] (as updated by the
    // previous rotation) in registers; a[k + 1] and e[k + 1] are read as they
    // were, so the loads need not wait for the chain
    double x = a[p] - mu, z = e[p];
    double ak = a[p], ek = e[p], en = 0.0;
    for (int k = p; k < q; ++k) {
        const double ga = a[k + 1], e1 = k + 1 < q ? e[k + 1] : 0.0;
        // the rotation of rows/columns (k, k + 1) that zeroes z against x
        // (one division: c = x / r and s = -z / r through 1 / r)
        const double r = sqrt(x * x + z * z);
        const bool nzr = r > 0.0;
        const double ir = nzr ? 1.0 / r : 0.0;
        const double ck = nzr ? x * ir : 1.0, sk = nzr ? -z * ir : 0.0;
        if (k > p) e[k - 1] = r;
        const double al = ak, be = ek;
        a[k] = ck * ck * al - 2.0 * ck * sk * be + sk * sk * ga;
        ak = sk * sk * al + 2.0 * ck * sk * be + ck * ck * ga;
        en = ck * sk * (al - ga) + (ck * ck - sk * sk) * be;
        // the bulge at (k + 2, k) and the next rotation's pair
        z = -sk * e1;
        ek = ck * e1;
        x = en;
        c[k - p] = ck;
        s[k - p] = sk;
    }
    a[q] = ak;
    e[q - 1] = en;
    return q - p;
}

__global__ __launch_bounds__(256) void k_tridiag_eig_b(const double* __restrict__ D, const double* __restrict__ E,
                                                       int l, double* __restrict__ Z, double* __restrict__ lam,
                                                       int* __restrict__ info)
{
    const int64_t b = blockIdx.x;
    D += b * l;
    E += b * l;
    Z += b * l * l;
    lam += b * l;
    info += b;
    __shared__ double zs[TQL_LMAX][TQL_LMAX + 1];  // Z row-major, padded
    __shared__ double a[TQL_LMAX], e[TQL_LMAX], rc[2][TQL_LMAX], rs[2][TQL_LMAX];
    __shared__ int nrot[2], rlo[2], rank[TQL_LMAX];
    const int tid = threadIdx.x, row = tid - 64;
    for (int i = tid; i < l * l; i += 256) zs[i / l][i % l] = i / l == i % l ? 1.0 : 0.0;
    for (int i = tid; i < l; i += 256) {
        a[i] = D[i];
        e[i] = i + 1 < l ? E[i] : 0.0;
    }
    __syncthreads();
    if (tid < 64) {
        int lo = 0;
        const int nr = tql_step(a, e, l, rc[0], rs[0], &lo);
        if (tid == 0) {
            nrot[0] = nr;
            rlo[0] = lo;
        }
    }
    __syncthreads();
    int bad = 0;
    for (int t = 0;; ++t) {
        const int nr = nrot[t & 1], lo = rlo[t & 1];
        if (nr == 0) break;  // every thread reads the same count: a uniform exit
        if (t >= 30 * l) {
            bad = 1;
            break;
        }
        if (tid < 64) {
            // the next step, into the other buffer
            int lo2 = 0;
            const int nr2 = tql_step(a, e, l, rc[(t + 1) & 1], rs[(t + 1) & 1], &lo2);
            if (tid == 0) {
                nrot[(t + 1) & 1] = nr2;
                rlo[(t + 1) & 1] = lo2;
            }
        } else if (row >= 0 && row < l) {
            // Z <- Z R_k^T for the step's rotations, on this row
            const double* cc = rc[t & 1];
            const double* ss = rs[t & 1];
            double zk = zs[row][lo];
            for (int j = 0; j < nr; ++j) {
                const double zk1 = zs[row][lo + j + 1];
                zs[row][lo + j] = cc[j] * zk - ss[j] * zk1;
                zk = ss[j] * zk + cc[j] * zk1;
            }
            zs[row][lo + nr] = zk;
        }
        __syncthreads();
    }
    // ascending order: the rank of each eigenvalue (ties by index)
    for (int i = tid; i < l; i += 256) {
        int rk = 0;
        for (int j = 0; j < l; ++j) rk += (a[j] < a[i] || (a[j] == a[i] && j < i)) ? 1 : 0;
        rank[i] = rk;
        lam[rk] = a[i];
    }
    __syncthreads();
    for (int r = tid; r < l; r += 256)
        for (int i = 0; i < l; ++i) Z[r + (int64_t)rank[i] * l] = zs[r][i];
    if (tid == 0) *info = bad;
}

// Z <- Q Z for the orthogonal Q = H_0 H_1 ... H_{l-2} of k_sytrd_lower's
// reflectors (H_i = I - tau_i v v^T, v(i+1) = 1, v(i+2:l) = A(i+2:l, i)), as
// LAPACK dormtr('L', 'L', 'N'), for each l x l column-major Z of a batch
// (l <= TQL_LMAX), one workgroup each: thread j holds column j, the
// reflectors applied last to first.
__global__ __launch_bounds__(256) void k_ormtr_b(const double* __restrict__ A, const double* __restrict__ tau, int l,
                                                 double* __restrict__ Z)
{
    const int64_t b = blockIdx.x;
    A += b * l * l;
    tau += b * l;
    Z += b * l * l;
    __shared__ double zs[TQL_LMAX][TQL_LMAX + 1];
    __shared__ double v[TQL_LMAX];
    const int tid = threadIdx.x;
    for (int i = tid; i < l * l; i += 256) zs[i % l][i / l] = Z[i];
    for (int i = l - 2; i >= 0; --i) {
        __syncthreads();  // the previous reflector's updates are done with v
        for (int r = tid; r < l; r += 256) v[r] = r == i + 1 ? 1.0 : (r > i + 1 ? A[r + (int64_t)i * l] : 0.0);
        __syncthreads();
        if (tid < l) {
            double w = 0.0;
            for (int r = i + 1; r < l; ++r) w += v[r] * zs[r][tid];
            w *= tau[i];
            for (int r = i + 1; r < l; ++r) zs[r][tid] -= v[r] * w;
        }
    }
    __syncthreads();
    for (int i = tid; i < l * l; i += 256) Z[i] = zs[i % l][i / l];
}


// The same product for one l x l column-major Z with l <= SYTRD_LMAX (the
// leading solve's Rayleigh-Ritz): the columns are independent, so each wave
// takes one (lane holding rows lane + 64 m) and applies the reflectors last
// to first, two per round: their two dot products with the column and their
// own product are reduced by one set of shuffles (lane 0's sums broadcast),
// and the next pair's columns load while the current pair is applied.  No
// block barrier is taken.
constexpr int ORMTR_WAVES = 4;
__global__ __launch_bounds__(64 * ORMTR_WAVES) void k_ormtr_cols(const double* __restrict__ A,
                                                                  const double* __restrict__ tau, int l,
                                                                  double* __restrict__ Z)
{
    constexpr int RM = SYTRD_LMAX / 64;
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * ORMTR_WAVES + (threadIdx.x >> 6);
    if (j >= l) return;  // a whole wave
    double* zc = Z + (int64_t)j * l;
    double z[RM], vn[RM];
#pragma unroll
    for (int m = 0; m < RM; ++m) {
        const int r = lane + 64 * m;
        z[m] = r < l ? zc[r] : 0.0;
    }
    // v of reflector i: 1 at row i + 1, A(r, i) below it, 0 above
    auto load_v = [&](int i, double* v) {
#pragma unroll
        for (int m = 0; m < RM; ++m) {
            const int r = lane + 64 * m;
            v[m] = r == i + 1 ? 1.0 : (r > i + 1 && r < l ? A[r + (int64_t)i * l] : 0.0);
        }
    };
    // two reflectors per round, i then i - 1: with a = v_i^T z,
    // b = v_{i-1}^T z and c = v_{i-1}^T v_i reduced by the same shuffles,
    // H_{i-1} H_i z = z - al v_i - be v_{i-1}, al = tau_i a,
    // be = tau_{i-1} (b - al c)
    double vn0[RM];
    int i = l - 2;
    if (i >= 1) {
        load_v(i, vn);
        load_v(i - 1, vn0);
    }
    for (; i >= 1; i -= 2) {
        double v1[RM], v0[RM];
#pragma unroll
        for (int m = 0; m < RM; ++m) {
            v1[m] = vn[m];
            v0[m] = vn0[m];
        }
        const double t1 = tau[i], t0 = tau[i - 1];
        if (i >= 3) {
            load_v(i - 2, vn);
            load_v(i - 3, vn0);
        }
        double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
        for (int m = 0; m < RM; ++m) {
            a += v1[m] * z[m];
            b += v0[m] * z[m];
            c += v0[m] * v1[m];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
        }
        const double al = __shfl(a, 0, 64) * t1;
        const double be = (__shfl(b, 0, 64) - al * __shfl(c, 0, 64)) * t0;
#pragma unroll
        for (int m = 0; m < RM; ++m) z[m] = z[m] - al * v1[m] - be * v0[m];
    }
    if (i == 0) {  // an odd count leaves reflector 0
        load_v(0, vn);
        double a = 0.0;
#pragma unroll
        for (int m = 0; m < RM; ++m) a += vn[m] * z[m];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
        const double al = __shfl(a, 0, 64) * tau[0];
#pragma unroll
        for (int m = 0; m < RM; ++m) z[m] -= al * vn[m];
    }
#pragma unroll
    for (int m = 0; m < RM; ++m) {
        const int r = lane + 64 * m;
        if (r < l) zc[r] = z[m];
    }
}

// KEVD_ROCSOLVER_ORMTR=1 in the environment keeps rocSOLVER's dormtr for the
// leading solve's back-transformation, for comparison.
bool rocsolver_ormtr()
{
    static const bool v = [] {
        const char* e = std::getenv("KEVD_ROCSOLVER_ORMTR");
        return e && e[0] == '1';
    }();
    return v;
}

// C = A^T B for column-major r x l blocks A, B (lda = ldb = r; C is l x l,
// ldc = l), FP64.  With l x l outputs and K = r the plain GEMM runs on a
// handful of workgroups; here K is split into GRAM_TN_SPLIT strided-batched
// GEMMs (the remainder rows added into the first partial) and the partials
// are summed in a fixed order.
constexpr int GRAM_TN_SPLIT = 8;
__global__ void k_sum_parts(const double* __restrict__ part, int ns, int64_t count, double* __restrict__ out)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double s = 0.0;
    for (int b = 0; b < ns; ++b) s += part[(int64_t)b * count + e];
    out[e] = s;
}

kevd_status gram_tn(kevd_ctx* ctx, const double* A, const double* B, int64_t r, int l, double* C)
{
    rocblas_handle h = blas_handle(ctx);
    const rocblas_int R = (rocblas_int)r, Lb = (rocblas_int)l;
    double one = 1.0, zero = 0.0;
    const int64_t kc = r / GRAM_TN_SPLIT, rem = r - GRAM_TN_SPLIT * kc;
    if (kc < 64) {
        BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb, R, &one, A, R, B, R,
                           &zero, C, Lb));
        return KEVD_OK;
    }
    const int64_t cnt = (int64_t)l * l;
    double* part;
    KEVD_TRY(ctx_buf_t(ctx, "q_gram_part", (size_t)GRAM_TN_SPLIT * cnt, &part));
    BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb,
                                       (rocblas_int)kc, &one, A, R, kc, B, R, kc, &zero, part, Lb, cnt,
                                       GRAM_TN_SPLIT));
    if (rem > 0)
        BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb, (rocblas_int)rem, &one,
                           A + GRAM_TN_SPLIT * kc, R, B + GRAM_TN_SPLIT * kc, R, &one, part, Lb));
    k_sum_parts<<<(unsigned)((cnt + 255) / 256), 256, 0, ctx->stream>>>(part, GRAM_TN_SPLIT, cnt, C);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

// Orthonormal basis of the columns of the r x l block Y, in place.
// Householder: rocSOLVER geqrf + orgqr.  Otherwise shifted Cholesky QR 3
// (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput.
// 42 (2020) A477): R = chol(Y^T Y + s I), Y <- Y R^-1 with
// s = 11 (r l + l (l + 1)) u ||Y||_F^2, then two plain Cholesky QR passes.
// It is stable for cond(Y) up to O(1/u) and costs three small GEMM +
// potrf + trsm sweeps instead of l sequential Householder steps.  The
// potrf statuses go to info[0..2]; a nonzero one means the basis is not
// to be trusted and the caller falls back to Householder.
kevd_status orth_basis(kevd_ctx* ctx, double* Y, int64_t r, int l, bool householder, double* G,
                       double* tau, int* info, int passes = 3)
{
    rocblas_handle h = blas_handle(ctx);
    const rocblas_int R = (rocblas_int)r, Lb = (rocblas_int)l;
    if (householder) {
        BLAS(rocsolver_dgeqrf(h, R, Lb, Y, R, tau));
        BLAS(rocsolver_dorgqr(h, R, Lb, Lb, Y, R, tau));
        return KEVD_OK;
    }
    double one = 1.0;
    const double c = 11.0 * ((double)r * l + (double)l * (l + 1)) * 0x1.0p-53;
    for (int pass = 0; pass < passes; ++pass) {
        KEVD_TRY(gram_tn(ctx, Y, Y, r, l, G));
        if (pass == 0) k_chol_shift<<<1, 256, 0, ctx->stream>>>(G, l, c);
        if (l <= CHOLR_LMAX && !chol_panel_only())
            k_potrf_reg<<<1, CHOLR_NT, 0, ctx->stream>>>(G, l, info + pass);
        else if (l <= CHOL_LMAX)
            k_potrf_upper<<<1, 1024, 0, ctx->stream>>>(G, l, info + pass);
        else
            BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, Lb, G, Lb, info + pass));
        BLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, R, Lb, &one, G, Lb, Y, R));
    }
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

// Leading k eigenpairs of the symmetric r x r FP64 matrix W (column-major).
// lam (device, r entries) gets eigenvalues ascending, *Vk an r x k block of
// eigenvectors in the same order (the last column is the leading one).
// all = true (or r < 2 l): the full divide-and-conquer eigensolve (every
// eigenvalue, as the reference's syevd).  Otherwise block subspace
// iteration on l = max(k + 64, min(4k, 256)) columns (rounded up to 32):
// Q <- orth(W Q), each step reducing the unwanted
// components of eigenvector i by lambda_{l+1} / lambda_i.  Rayleigh-Ritz
// (H = Q^T W Q = S diag(theta) S^T) runs after the first few steps (where
// the context's last solve converged, at least 4) and then after as many
// steps as the slowest Ritz pair needs, until the leading k residuals
// ||W x - theta x|| <= 1e-6 theta + 1e-14 theta_max.  The error of a Ritz
// value is of the order of the squared relative residual, so the stopping
// rule gives ~1e-12 relative eigenvalues.
kevd_status eig_batch_join(kevd_ctx* ctx, double* W, int64_t r, int k, double* lam, double** Vk, int* iters,
                           bool* handled);

kevd_status eig_leading(kevd_ctx* ctx, double* W, int64_t r, int k, bool all, double* lam,
                        double** Vk, int* iters)
{
    // block: k + 64 columns at least, 4k (at most 256) when that is larger
    int l = (int)KEVD_ROUND_UP(std::max(k + 64, std::min(4 * k, 256)), 32);
    *iters = 0;
    if (all || r < 2 * (int64_t)l) {
        KEVD_TRY(eigh(ctx, W, r, lam));
        *Vk = W + (r - k) * r;
        return KEVD_OK;
    }
    if (ctx->eig_batch) {
        bool handled = false;
        kevd_status bs = eig_batch_join(ctx, W, r, k, lam, Vk, iters, &handled);
        if (handled) return bs;
    }
    rocblas_handle h = blas_handle(ctx);
    double *Q, *Y, *H, *G, *GS, *th, *tau, *res2, *E, *X;
    int* info;
    KEVD_TRY(ctx_buf_t(ctx, "q_Q", (size_t)r * l, &Q));
    KEVD_TRY(ctx_buf_t(ctx, "q_Y", (size_t)r * l, &Y));
    KEVD_TRY(ctx_buf_t(ctx, "q_H", (size_t)l * l, &H));
    KEVD_TRY(ctx_buf_t(ctx, "q_G", (size_t)l * l, &G));
    KEVD_TRY(ctx_buf_t(ctx, "q_GS", (size_t)l * l, &GS));
    KEVD_TRY(ctx_buf_t(ctx, "q_th", l, &th));
    KEVD_TRY(ctx_buf_t(ctx, "q_tau", l, &tau));
    KEVD_TRY(ctx_buf_t(ctx, "q_res2", k, &res2));
    KEVD_TRY(ctx_buf_t(ctx, "q_E", l, &E));
    KEVD_TRY(ctx_buf_t(ctx, "q_X", (size_t)r * k, &X));
    // info[0]: Rayleigh-Ritz syevd; info[1 + 3 j + pass]: Cholesky QR
    // passes of the j-th orthonormalisation since the last check
    const int max_orth = 8;
    KEVD_TRY(ctx_buf_t(ctx, "q_info", 1 + 3 * max_orth, &info));
    ProfScope ps(ctx, TC_EIG);
    const rocblas_int R = (rocblas_int)r, Lb = (rocblas_int)l;
    double one = 1.0, zero = 0.0;
    std::vector<double> hres((size_t)k), hth((size_t)l);
    std::vector<int> hinfo(1 + 3 * max_orth);
    // A Rayleigh-Ritz check (an l x l syevd, latency-bound) costs several
    // iterations, so the first one goes where the last solve on this context
    // converged (at least step 4): repeated solves on similar matrices then
    // check once, and a guess past convergence costs only a few iterations.
    const int maxit = 16;
    const int first_check = std::max(4, std::min(ctx->eig_converged_at, maxit - 1));
    for (int householder = 0; householder < 2; ++householder) {
        KEVD_HIP(ctx, hipMemsetAsync(info, 0, (1 + 3 * max_orth) * sizeof(int), ctx->stream));
        k_randn<<<(unsigned)((r * l + 255) / 256), 256, 0, ctx->stream>>>(Q, r * l, 0x6b657664ull);
        KEVD_HIP(ctx, hipGetLastError());
        int north = 0;
        // The Gaussian start needs no orthonormalisation: with r >= 2l its
        // condition number is at most about (sqrt(r) + sqrt(l)) /
        // (sqrt(r) - sqrt(l)) <= 5.8, so W Q is as well conditioned as the
        // products of later steps, and only span(Q) matters.
        bool conv = false, broken = false;
        int it = 0, next_check = first_check;
        for (it = 1; it <= maxit; ++it) {
            BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, R, Lb, R, &one, W, R, Q, R,
                               &zero, Y, R));
            if (it < next_check && north < max_orth) {
                // the iteration only follows span(Q), so the next product
                // needs a well-conditioned basis, not an orthonormal one: one
                // shifted Cholesky QR pass (cond(Q) = 1 + O(cond(Y)^2 u)); the
                // basis that Rayleigh-Ritz reads gets a second, plain pass,
                // which leaves it orthonormal to O(u)
                bool before_check = it + 1 >= next_check || north + 1 >= max_orth;
                KEVD_TRY(orth_basis(ctx, Y, r, l, householder, G, tau, info + 1 + 3 * north++,
                                    before_check ? 2 : 1));
                ctx->tm.eig_orth_passes += before_check ? 2 : 1;
                std::swap(Q, Y);
                continue;
            }
            KEVD_TRY(gram_tn(ctx, Q, Y, r, l, H));
            KEVD_TRY(gram_tn(ctx, Y, Y, r, l, G));
            k_symmetrize<<<dim3((l + 15) / 16, (l + 15) / 16), 256, 0, ctx->stream>>>(H, l);
            if (l <= SYTRD_LMAX) {
                // one-workgroup tridiagonalisation, rocSOLVER's divide and
                // conquer on T, and the back-transformation by one wave per
                // column; the eigenvectors land in GS, which becomes H
                k_sytrd_lower<<<1, 512, 0, ctx->stream>>>(H, l, th, E, tau);
                KEVD_HIP(ctx, hipGetLastError());
                BLAS(rocsolver_dstedc(h, rocblas_evect_tridiagonal, Lb, th, E, GS, Lb, info));
                if (rocsolver_ormtr()) {
                    BLAS(rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, Lb, Lb,
                                          H, Lb, tau, GS, Lb));
                } else {
                    k_ormtr_cols<<<(l + ORMTR_WAVES - 1) / ORMTR_WAVES, 64 * ORMTR_WAVES, 0, ctx->stream>>>(H, tau,
                                                                                                         l, GS);
                    KEVD_HIP(ctx, hipGetLastError());
                }
                std::swap(H, GS);
            } else {
                BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, Lb, H, Lb, th, E, info));
            }
            BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, Lb, Lb, Lb, &one, G, Lb, H,
                               Lb, &zero, GS, Lb));
            k_ritz_res2<<<k, 256, 0, ctx->stream>>>(H, GS, th, l, k, res2);
            KEVD_HIP(ctx, hipGetLastError());
            KEVD_HIP(ctx, hipMemcpyAsync(hinfo.data(), info, hinfo.size() * sizeof(int), hipMemcpyDeviceToHost,
                                         ctx->stream));
            KEVD_HIP(ctx, hipMemcpyAsync(hres.data(), res2, k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KEVD_HIP(ctx, hipMemcpyAsync(hth.data(), th, l * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (int j = 1; j < 1 + 3 * north; ++j) broken |= hinfo[j] != 0;
            if (broken) break;
            if (hinfo[0] != 0)
                return kevd_fail(ctx, KEVD_EEIG, "Rayleigh-Ritz eigensolve failed, info " + std::to_string(hinfo[0]));
            double thmax = std::fabs(hth[l - 1]);
            conv = true;
            // Steps until the next check: the residual of Ritz pair c shrinks
            // by about rho_c = theta_min / theta_c per step (theta_min, the
            // smallest Ritz value of the block, stands in for lambda_{l+1}).
            int need = 1;
            for (int c = 0; c < k; ++c) {
                double t = hth[l - k + c];
                double res = std::sqrt(std::max(hres[c], 0.0));
                double tolc = 1e-6 * std::fabs(t) + 1e-14 * thmax;
                if (res <= tolc) continue;
                conv = false;
                double rho = std::fabs(hth[0]) / std::max(std::fabs(t), 1e-300);
                int m = rho > 0.0 && rho < 0.9 ? (int)std::ceil(std::log(tolc / res) / std::log(rho)) : 1;
                need = std::max(need, std::min(m, 6));
            }
            if (conv || it == maxit) break;
            // the last step is always a check; a check on the next step needs
            // the basis orthonormal to O(u) (two passes), as before the first
            next_check = std::min(it + need, maxit);
            KEVD_HIP(ctx, hipMemsetAsync(info, 0, (1 + 3 * max_orth) * sizeof(int), ctx->stream));
            north = 0;
            KEVD_TRY(orth_basis(ctx, Y, r, l, householder, G, tau, info + 1 + 3 * north++,
                                next_check == it + 1 ? 2 : 1));
            ctx->tm.eig_orth_passes += next_check == it + 1 ? 2 : 1;
            std::swap(Q, Y);
        }
        *iters = it;
        ctx->tm.eig_iters = it;
        ctx->tm.eig_householder = householder;
        if (broken) continue;  // a Cholesky QR pass lost definiteness: redo with Householder
        if (!conv) {
            // The step limit came first (a slowly decaying spectrum: the
            // rate lambda_{l+1} / lambda_k close to 1).  The full solve gives
            // every pair to working accuracy; W is unchanged by the
            // iteration, which only read it.
            ctx->tm.eig_full_fallback = 1;
            KEVD_TRY(eigh(ctx, W, r, lam, false));
            *Vk = W + (r - k) * r;
            return KEVD_OK;
        }
        ctx->eig_converged_at = it;
        // Ritz vectors of the leading k: X = Q S(:, l-k:l)
        BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, R, (rocblas_int)k, Lb, &one, Q, R,
                           H + (int64_t)(l - k) * l, Lb, &zero, X, R));
        k_lam_tail<<<(unsigned)((r + 255) / 256), 256, 0, ctx->stream>>>(th, l, r, lam);
        KEVD_HIP(ctx, hipGetLastError());
        *Vk = X;
        return KEVD_OK;
    }
    return kevd_fail(ctx, KEVD_EEIG, "subspace iteration: orthonormalisation failed");
}


// ------------------------------------------- batched leading eigenpairs
// The ensemble's problems each solve an r_p x r_p leading-eigenpair problem
// whose cost is latency: 6-9 subspace steps of small GEMMs, Cholesky QR
// passes and l x l Rayleigh-Ritz eigensolves, 3.4-4.3 ms whatever r_p
// (48 problems: 182 ms in sequence, about 45 ms on 4 hardware queues).
// Here the problems that reach the subspace iteration meet, and one solve
// runs the same iteration for all of them with strided-batched rocBLAS /
// rocSOLVER calls on copies of W_p zero-padded to the largest order R (the
// padding adds eigenvalue 0, below every leading eigenvalue of the
// positive semidefinite W_p, and leaves the Ritz vectors' padded rows 0).
// The stopping rule is eig_leading's, applied to every problem; problems
// not converged at the step limit, or a Cholesky QR that loses
// definiteness, fall back to each problem's own solve.

__global__ void k_pad_copy(const double* const* __restrict__ Wp, const int64_t* __restrict__ rs, int64_t R,
                           double* __restrict__ Wb)
{
    const int p = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= R * R) return;
    const int64_t r = rs[p], i = e % R, j = e / R;
    Wb[(int64_t)p * R * R + e] = i < r && j < r ? Wp[p][i + j * r] : 0.0;
}

// strict lower triangle of each l x l matrix := 0 (grid: tiles, problems)
__global__ void k_zero_lower_b(double* G, int l)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)l * l) return;
    if (e % l > e / l) G[(int64_t)blockIdx.y * l * l + e] = 0.0;
}

__global__ __launch_bounds__(256) void k_chol_shift_b(double* G, int l, double c)
{
    double* g = G + (int64_t)blockIdx.x * l * l;
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < l; i += 256) acc += g[i + (int64_t)i * l];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    double sh = c * red[0];
    for (int i = threadIdx.x; i < l; i += 256) g[i + (int64_t)i * l] += sh;
}

__global__ void k_symmetrize_b(double* H, int l)
{
    double* h = H + (int64_t)blockIdx.z * l * l;
    int i = blockIdx.x * 16 + (threadIdx.x >> 4), j = blockIdx.y * 16 + (threadIdx.x & 15);
    if (i < l && j < l && i < j) {
        double a = 0.5 * (h[i + (int64_t)j * l] + h[j + (int64_t)i * l]);
        h[i + (int64_t)j * l] = a;
        h[j + (int64_t)i * l] = a;
    }
}

__global__ void k_ritz_res2_b(const double* S, const double* GS, const double* th, int l, int k, double* res2)
{
    const int p = blockIdx.y, c = blockIdx.x, i = l - k + c;
    const double* s = S + (int64_t)p * l * l;
    const double* gs = GS + (int64_t)p * l * l;
    double acc = 0.0;
    for (int j = threadIdx.x; j < l; j += 256) acc += s[j + (int64_t)i * l] * gs[j + (int64_t)i * l];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) res2[(int64_t)p * k + c] = red[0] - th[(int64_t)p * l + i] * th[(int64_t)p * l + i];
}

}  // namespace

struct EigBatch {
    kevd_ctx* m;
    int k, l;
    int expected;  // slots that have neither joined nor left yet, plus those joined
    int arrived = 0;
    bool done = false;
    std::mutex mu;
    std::condition_variable cv;
    struct Slot {
        kevd_ctx* ctx = nullptr;
        double* W = nullptr;
        int64_t r = 0;
        double* lam = nullptr;
        double* X = nullptr;
        bool joined = false, left = false, ok = false;
        int iters = 0;
    };
    std::vector<Slot> slots;
};

namespace {

// The batched solve over the joined slots (called with the batch's mutex
// held, every joined problem waiting).  Marks slot.ok for the problems it
// solved; the others run their own solve.
kevd_status eig_batch_run(EigBatch& B)
{
    kevd_ctx* m = B.m;
    kevd_ctx* ctx = m;  // for BLAS()
    std::vector<int> J;
    for (int i = 0; i < (int)B.slots.size(); ++i)
        if (B.slots[i].joined) J.push_back(i);
    const int P = (int)J.size();
    if (P < 2) return KEVD_OK;  // nothing to batch: the one problem solves alone
    TraceRange tr("kEVD batched eigensolve");
    const int l = B.l, k = B.k;
    int64_t R = 0;
    for (int i : J) R = std::max(R, B.slots[i].r);
    KEVD_HIP(m, hipSetDevice(m->device));
    rocblas_handle h = blas_handle(m);
    if (!h) return kevd_fail(m, KEVD_EHIP, "rocblas handle creation failed");
    const int max_orth = 8;
    double *Wb, *Q, *Y, *H, *G, *GS, *th, *E, *res2, *Xb;
    int64_t* rs;
    const double** Wp;
    int* info;
    const int ninfo = (1 + 3 * max_orth) * P;
    KEVD_TRY(ctx_buf_t(m, "eb_W", (size_t)P * R * R, &Wb));
    KEVD_TRY(ctx_buf_t(m, "eb_Q", (size_t)P * R * l, &Q));
    KEVD_TRY(ctx_buf_t(m, "eb_Y", (size_t)P * R * l, &Y));
    KEVD_TRY(ctx_buf_t(m, "eb_H", (size_t)P * l * l, &H));
    KEVD_TRY(ctx_buf_t(m, "eb_G", (size_t)P * l * l, &G));
    KEVD_TRY(ctx_buf_t(m, "eb_GS", (size_t)P * l * l, &GS));
    KEVD_TRY(ctx_buf_t(m, "eb_th", (size_t)P * l, &th));
    KEVD_TRY(ctx_buf_t(m, "eb_E", (size_t)P * l, &E));
    KEVD_TRY(ctx_buf_t(m, "eb_res2", (size_t)P * k, &res2));
    KEVD_TRY(ctx_buf_t(m, "eb_X", (size_t)P * R * k, &Xb));
    KEVD_TRY(ctx_buf_t(m, "eb_rs", P, &rs));
    KEVD_TRY(ctx_buf(m, "eb_Wp", P * sizeof(double*), (void**)&Wp));
    KEVD_TRY(ctx_buf_t(m, "eb_info", ninfo, &info));
    double* T;
    int *tinfo, *nsw;
    KEVD_TRY(ctx_buf_t(m, "eb_T", (size_t)P * R * l, &T));
    KEVD_TRY(ctx_buf_t(m, "eb_tinfo", P, &tinfo));
    KEVD_TRY(ctx_buf_t(m, "eb_nsw", P, &nsw));
    double* tau;
    KEVD_TRY(ctx_buf_t(m, "eb_tau", (size_t)P * l, &tau));
    std::vector<int64_t> hr(P);
    std::vector<const double*> hw(P);
    for (int q = 0; q < P; ++q) hr[q] = B.slots[J[q]].r, hw[q] = B.slots[J[q]].W;
    hipStream_t s = m->stream;
    KEVD_HIP(m, hipMemcpyAsync(rs, hr.data(), P * sizeof(int64_t), hipMemcpyHostToDevice, s));
    KEVD_HIP(m, hipMemcpyAsync(Wp, hw.data(), P * sizeof(double*), hipMemcpyHostToDevice, s));
    k_pad_copy<<<dim3((unsigned)((R * R + 255) / 256), P), 256, 0, s>>>(Wp, rs, R, Wb);
    k_randn<<<(unsigned)((P * R * l + 255) / 256), 256, 0, s>>>(Q, P * R * l, 0x6b657664ull);
    KEVD_HIP(m, hipGetLastError());
    const rocblas_int Rr = (rocblas_int)R, Lb = (rocblas_int)l, Pb = (rocblas_int)P;
    const rocblas_stride sRR = R * R, sRL = R * l, sLL = (rocblas_stride)l * l, sL = l;
    double one = 1.0, zero = 0.0;
    const double cshift = 11.0 * ((double)R * l + (double)l * (l + 1)) * 0x1.0p-53;
    // passes shifted (first) then plain Cholesky QR; info of pass j at info + j * P
    auto orth = [&](int slot0, int passes) -> kevd_status {
        for (int pass = 0; pass < passes; ++pass) {
            BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb, Rr,
                                               &one, Y, Rr, sRL, Y, Rr, sRL, &zero, G, Lb, sLL, Pb));
            if (pass == 0) k_chol_shift_b<<<P, 256, 0, s>>>(G, l, cshift);
            // one workgroup per problem (rocSOLVER's batched potrf and rocBLAS's
            // batched trsm run as hundreds of per-matrix kernels), then
            // Y <- Y R^-1 as a batched GEMM with the inverted factor
            if (l <= CHOLR_LMAX && !chol_panel_only())
                k_potrf_reg<<<P, CHOLR_NT, 0, s>>>(G, l, info + (int64_t)(slot0 + pass) * P);
            else
                k_potrf_upper<<<P, 1024, 0, s>>>(G, l, info + (int64_t)(slot0 + pass) * P);
            BLAS(rocsolver_dtrtri_strided_batched(h, rocblas_fill_upper, rocblas_diagonal_non_unit, Lb, G, Lb, sLL,
                                                  tinfo, Pb));
            // the GEMM reads the whole matrix: the strict lower triangle still
            // holds the Gram
            k_zero_lower_b<<<dim3((unsigned)(((int64_t)l * l + 255) / 256), P), 256, 0, s>>>(G, l);
            BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, Rr, Lb, Lb, &one, Y,
                                               Rr, sRL, G, Lb, sLL, &zero, T, Rr, sRL, Pb));
            std::swap(Y, T);
        }
        KEVD_HIP(m, hipGetLastError());
        return KEVD_OK;
    };
    std::vector<int> hinfo(ninfo);
    std::vector<double> hres((size_t)P * k), hth((size_t)P * l);
    // dead: a Cholesky QR pass or the Rayleigh-Ritz eigensolve of that
    // problem failed; the batch elements are independent, so the others go on
    std::vector<char> conv(P, 0), dead(P, 0);
    KEVD_HIP(m, hipMemsetAsync(info, 0, ninfo * sizeof(int), s));
    const int maxit = 16;
    // as the single solve: the first Rayleigh-Ritz check (a batched Jacobi
    // eigensolve, about 9 ms for 48 blocks of 128) goes where the last
    // batched solve on this master context converged, at least step 4
    int next_check = std::max(4, std::min(m->ens_eig_converged_at, maxit - 1)), north = 0, it = 0;
    bool all_conv = false;
    for (it = 1; it <= maxit; ++it) {
        BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, Rr, Lb, Rr, &one, Wb,
                                           Rr, sRR, Q, Rr, sRL, &zero, Y, Rr, sRL, Pb));
        if (it < next_check && north < max_orth) {
            bool before_check = it + 1 >= next_check || north + 1 >= max_orth;
            KEVD_TRY(orth(1 + 3 * north++, before_check ? 2 : 1));
            std::swap(Q, Y);
            continue;
        }
        BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb, Rr, &one,
                                           Q, Rr, sRL, Y, Rr, sRL, &zero, H, Lb, sLL, Pb));
        BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, Lb, Lb, Rr, &one,
                                           Y, Rr, sRL, Y, Rr, sRL, &zero, G, Lb, sLL, Pb));
        k_symmetrize_b<<<dim3((l + 15) / 16, (l + 15) / 16, P), 256, 0, s>>>(H, l);
        if (l <= TQL_LMAX) {
            // per block: the one-workgroup tridiagonalisation, implicit QR
            // steps on the tridiagonal and the back-transformation, each a
            // launch over all problems (eigenvalues ascending in th, the
            // eigenvectors into GS, which becomes H)
            k_sytrd_lower_b<<<P, 512, 0, s>>>(H, l, th, E, tau);
            k_tridiag_eig_b<<<P, 256, 0, s>>>(th, E, l, GS, th, info);
            k_ormtr_b<<<P, 256, 0, s>>>(H, tau, l, GS);
            KEVD_HIP(m, hipGetLastError());
            std::swap(H, GS);
        } else {
            // batched Jacobi (eigenvalues ascending; its batched divide and
            // conquer runs as thousands of per-matrix kernels), converged once
            // the off-diagonal norm is 1e-14 of the matrix's (a tolerance of
            // 0 asks for machine precision, which it need not reach)
            BLAS(rocsolver_dsyevj_strided_batched(h, rocblas_esort_ascending, rocblas_evect_original,
                                                  rocblas_fill_lower, Lb, H, Lb, sLL, 1e-14, E, 100, nsw, th, sL,
                                                  info, Pb));
        }
        BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, Lb, Lb, Lb, &one, G, Lb,
                                           sLL, H, Lb, sLL, &zero, GS, Lb, sLL, Pb));
        k_ritz_res2_b<<<dim3(k, P), 256, 0, s>>>(H, GS, th, l, k, res2);
        KEVD_HIP(m, hipGetLastError());
        KEVD_HIP(m, hipMemcpyAsync(hinfo.data(), info, ninfo * sizeof(int), hipMemcpyDeviceToHost, s));
        KEVD_HIP(m, hipMemcpyAsync(hres.data(), res2, hres.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        KEVD_HIP(m, hipMemcpyAsync(hth.data(), th, hth.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        KEVD_HIP(m, hipStreamSynchronize(s));
        for (int j = 0; j < (1 + 3 * north) * P; ++j)
            if (hinfo[j] != 0) dead[j % P] = 1;
        all_conv = true;
        int need = 1;
        for (int q = 0; q < P; ++q) {
            if (dead[q]) {
                conv[q] = 0;
                continue;
            }
            const double* tq = hth.data() + (size_t)q * l;
            const double thmax = std::fabs(tq[l - 1]);
            bool cq = true;
            for (int c = 0; c < k; ++c) {
                double t = tq[l - k + c], res = std::sqrt(std::max(hres[(size_t)q * k + c], 0.0));
                double tolc = 1e-6 * std::fabs(t) + 1e-14 * thmax;
                if (res <= tolc) continue;
                cq = false;
                double rho = std::fabs(tq[0]) / std::max(std::fabs(t), 1e-300);
                int mm = rho > 0.0 && rho < 0.9 ? (int)std::ceil(std::log(tolc / res) / std::log(rho)) : 1;
                need = std::max(need, std::min(mm, 6));
            }
            conv[q] = cq;
            all_conv &= cq;
        }
        if (all_conv || it == maxit) break;
        next_check = std::min(it + need, maxit);
        KEVD_HIP(m, hipMemsetAsync(info, 0, ninfo * sizeof(int), s));
        north = 0;
        KEVD_TRY(orth(1 + 3 * north++, next_check == it + 1 ? 2 : 1));
        std::swap(Q, Y);
    }
    // Ritz vectors of the leading k: X = Q S(:, l-k:l), then each problem's
    // first r_p rows and its eigenvalue tail
    BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, Rr, (rocblas_int)k, Lb,
                                       &one, Q, Rr, sRL, H + (int64_t)(l - k) * l, Lb, sLL, &zero, Xb, Rr,
                                       (rocblas_stride)R * k, Pb));
    for (int q = 0; q < P; ++q) {
        if (!conv[q]) continue;
        EigBatch::Slot& sl = B.slots[J[q]];
        KEVD_HIP(m, hipMemcpy2DAsync(sl.X, sl.r * sizeof(double), Xb + (int64_t)q * R * k, R * sizeof(double),
                                     sl.r * sizeof(double), k, hipMemcpyDeviceToDevice, s));
        k_lam_tail<<<(unsigned)((sl.r + 255) / 256), 256, 0, s>>>(th + (int64_t)q * l, l, sl.r, sl.lam);
    }
    KEVD_HIP(m, hipGetLastError());
    KEVD_HIP(m, hipStreamSynchronize(s));
    for (int q = 0; q < P; ++q) {
        B.slots[J[q]].ok = conv[q] != 0;
        B.slots[J[q]].iters = it;
    }
    if (all_conv) m->ens_eig_converged_at = it;
    return KEVD_OK;
}

// with the mutex held: run once every slot has joined or left
void eig_batch_maybe_run(EigBatch& B)
{
    if (B.done || B.arrived < B.expected) return;
    if (B.arrived > 0 && eig_batch_run(B) != KEVD_OK)
        for (auto& sl : B.slots) sl.ok = false;  // the batch failed: every problem solves alone
    B.done = true;
    B.cv.notify_all();
}

kevd_status eig_batch_join(kevd_ctx* ctx, double* W, int64_t r, int k, double* lam, double** Vk, int* iters,
                           bool* handled)
{
    EigBatch& B = *ctx->eig_batch;
    *handled = false;
    const int slot = ctx->eig_batch_slot;
    if (slot < 0 || slot >= (int)B.slots.size() || k != B.k) return KEVD_OK;
    double* X;
    KEVD_TRY(ctx_buf_t(ctx, "q_X", (size_t)r * k, &X));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));  // W complete before another stream reads it
    std::unique_lock<std::mutex> g(B.mu);
    if (B.done) return KEVD_OK;  // too late for the batch: solve alone
    EigBatch::Slot& sl = B.slots[slot];
    sl.ctx = ctx, sl.W = W, sl.r = r, sl.lam = lam, sl.X = X;
    sl.joined = true;
    ++B.arrived;
    eig_batch_maybe_run(B);
    B.cv.wait(g, [&] { return B.done; });
    if (!sl.ok) return KEVD_OK;  // not converged, or the batch broke down: solve alone
    *handled = true;
    *Vk = X;
    *iters = sl.iters;
    ctx->tm.eig_iters = sl.iters;
    ctx->tm.eig_batched = 1;
    return KEVD_OK;
}

}  // namespace

EigBatch* eig_batch_create(kevd_ctx* master, int nslots, int k_out)
{
    EigBatch* b = new EigBatch();
    b->m = master;
    b->k = k_out;
    b->l = (int)KEVD_ROUND_UP(std::max(k_out + 64, std::min(4 * k_out, 256)), 32);
    b->expected = nslots;
    b->slots.resize(nslots);
    return b;
}

void eig_batch_leave(EigBatch* b, int slot)
{
    std::lock_guard<std::mutex> g(b->mu);
    EigBatch::Slot& sl = b->slots[slot];
    if (sl.joined || sl.left) return;
    sl.left = true;
    --b->expected;
    eig_batch_maybe_run(*b);
}

void eig_batch_destroy(EigBatch* b) { delete b; }

namespace {

// ------------------------------------------------ intermediate record
// The reference's dump datasets (ks_kevd.cu:531-542, 809-815, 935-941,
// 1183-1218, 1274-1280, 1562-1576, 1705-1744, 1945-1969, 2164-2170), in its
// basis: the kernel columns K(:,S), with K(S,S) = L L^T, where this kEVD
// works with F = K(:,S) L^-T.  From t = F^T 1, u = F^T d~^-1 and the F-basis
// Grams A = F^T D~^-2 F = L^-1 Atil L^-T, B = F^T Q~^-1 F = L^-1 Btil L^-T:
//   c1 = L t, p1 = L^-T t, cq = L u, pq = L^-T u,
//   Atil = L A L^T (lower triangle, strict upper zero: the reference's syrk
//   fills only the lower one), R = chol(A + sI) upper (strict lower zero),
//   G = R L^-1, Btil = L B L^T, W = R B R^T, Z = L^-T R^-1 V_k,
// and dtil = d~, qtil = q~, V, lam as computed.  Matrices are column-major.
// Every entry is copied to the host when it is formed (synchronously).

// out(i, j) = i <= j ? Lc(j, i) : 0: the upper factor R = Lc^T from the
// lower Cholesky factor in the lower triangle of Lc.
__global__ __launch_bounds__(256) void k_rec_upper(const double* Lc, int64_t r, double* out)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= r * r) return;
    int64_t i = e % r, j = e / r;
    out[e] = i <= j ? Lc[j + i * r] : 0.0;
}

__global__ __launch_bounds__(256) void k_fill1(double* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 1.0;
}

__global__ __launch_bounds__(256) void k_rec_tril(double* M, int64_t r)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= r * r) return;
    if (e % r < e / r) M[e] = 0.0;
}

struct Record {
    kevd_ctx* ctx = nullptr;
    rocblas_handle h = nullptr;
    int64_t r = 0;
    double *L = nullptr, *M = nullptr, *M2 = nullptr, *v = nullptr;
    bool on = false;

    kevd_status begin(kevd_ctx* c, rocblas_handle hd)
    {
        ctx = c, h = hd, r = c->rank;
        on = c->dump;
        if (!on) return KEVD_OK;
        KEVD_TRY(ctx_buf_t(ctx, "rec_L", (size_t)r * r, &L));
        KEVD_TRY(ctx_buf_t(ctx, "rec_M", (size_t)r * r, &M));
        KEVD_TRY(ctx_buf_t(ctx, "rec_M2", (size_t)r * r, &M2));
        KEVD_TRY(ctx_buf_t(ctx, "rec_v", r, &v));
        return pivot_cholesky(ctx, L);
    }
    kevd_status put(const char* name, const double* dev, size_t count)
    {
        if (!on) return KEVD_OK;
        std::vector<double>& out = ctx->record[name];
        out.resize(count);
        KEVD_HIP(ctx, hipMemcpyAsync(out.data(), dev, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return KEVD_OK;
    }
    // name = L x (lower = true) or L^-T x
    kevd_status vec(const char* name, const double* x, bool times_l)
    {
        if (!on) return KEVD_OK;
        KEVD_HIP(ctx, hipMemcpyAsync(v, x, r * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        rocblas_status st = times_l ? rocblas_dtrmv(h, rocblas_fill_lower, rocblas_operation_none,
                                                    rocblas_diagonal_non_unit, (rocblas_int)r, L, (rocblas_int)r, v, 1)
                                    : rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_transpose,
                                                    rocblas_diagonal_non_unit, (rocblas_int)r, L, (rocblas_int)r, v, 1);
        if (st != rocblas_status_success) return kevd_fail(ctx, KEVD_EHIP, "record: triangular vector op failed");
        return put(name, v, r);
    }
    // name = L X L^T, lower triangle only when tril
    kevd_status sandwich(const char* name, const double* X, bool tril)
    {
        if (!on) return KEVD_OK;
        const double one = 1.0;
        const rocblas_int R = (rocblas_int)r;
        if (rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                          R, R, &one, L, R, X, R, M, R) != rocblas_status_success ||
            rocblas_dtrmm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, R, R, &one, L, R, M, R, M2, R) != rocblas_status_success)
            return kevd_fail(ctx, KEVD_EHIP, "record: triangular product failed");
        if (tril) k_rec_tril<<<(unsigned)((r * r + 255) / 256), 256, 0, ctx->stream>>>(M2, r);
        KEVD_HIP(ctx, hipGetLastError());
        return put(name, M2, (size_t)r * r);
    }
    // R from the lower Cholesky factor Lc of A + sI, and G = R L^-1
    kevd_status chol(const double* Lc)
    {
        if (!on) return KEVD_OK;
        const double one = 1.0;
        const rocblas_int R = (rocblas_int)r;
        k_rec_upper<<<(unsigned)((r * r + 255) / 256), 256, 0, ctx->stream>>>(Lc, r, M);
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_TRY(put("R", M, (size_t)r * r));
        if (rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                          R, R, &one, L, R, M, R) != rocblas_status_success)
            return kevd_fail(ctx, KEVD_EHIP, "record: triangular solve failed");
        return put("G", M, (size_t)r * r);
    }
    // Z = L^-T Y for the r x k block Y = R^-1 V_k
    kevd_status z(const double* Y, int k)
    {
        if (!on) return KEVD_OK;
        const double one = 1.0;
        KEVD_HIP(ctx, hipMemcpyAsync(M, Y, (size_t)r * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        if (rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, (rocblas_int)r, (rocblas_int)k, &one, L, (rocblas_int)r, M,
                          (rocblas_int)r) != rocblas_status_success)
            return kevd_fail(ctx, KEVD_EHIP, "record: triangular solve failed");
        return put("Z", M, (size_t)r * k);
    }
};

template <typename T, typename Ops>
kevd_status evd_ref(kevd_ctx* ctx, Ops& ops, const kevd_evd_params* p, kevd_evd_result* res)
{
    int64_t n = ctx->n, r = ctx->rank;
    int k = p->k_out;
    rocblas_handle h = blas_handle(ctx);
    if (!h) return kevd_fail(ctx, KEVD_EHIP, "rocblas handle creation failed");
    double *t, *dt, *cq, *qt, *w1, *w2, *A, *B, *W, *lam, *v2, *P1, *trs;
    int* info;
    KEVD_TRY(ctx_buf_t(ctx, "e_t", r, &t));
    KEVD_TRY(ctx_buf_t(ctx, "e_dt", n, &dt));
    KEVD_TRY(ctx_buf_t(ctx, "e_cq", r, &cq));
    KEVD_TRY(ctx_buf_t(ctx, "e_qt", n, &qt));
    KEVD_TRY(ctx_buf_t(ctx, "e_w1", n, &w1));
    KEVD_TRY(ctx_buf_t(ctx, "e_w2", n, &w2));
    KEVD_TRY(ctx_buf_t(ctx, "e_A", (size_t)r * r, &A));
    KEVD_TRY(ctx_buf_t(ctx, "e_B", (size_t)r * r, &B));
    KEVD_TRY(ctx_buf_t(ctx, "e_W", (size_t)r * r, &W));
    KEVD_TRY(ctx_buf_t(ctx, "e_lam", r, &lam));
    KEVD_TRY(ctx_buf_t(ctx, "e_v2", r, &v2));
    KEVD_TRY(ctx_buf_t(ctx, "e_P1", n, &P1));
    KEVD_TRY(ctx_buf_t(ctx, "e_trs", 2, &trs));
    KEVD_TRY(ctx_buf_t(ctx, "e_pinfo", 1, &info));
    unsigned gn = (unsigned)((n + 255) / 256);
    double mm[2];
    Record rec;
    KEVD_TRY(rec.begin(ctx, h));
    KEVD_TRY(ops.colsum(nullptr, t));  // t = F^T 1
    KEVD_TRY(ops.rowdot(t, dt, 0, nullptr, nullptr, mm));  // d~ = F t
    res->min_degree = mm[0];
    KEVD_TRY(rec.vec("c1", t, true));
    KEVD_TRY(rec.vec("p1", t, false));
    KEVD_TRY(rec.put("dtil", dt, n));
    if (!(mm[0] > 0.0))
        return kevd_fail(ctx, KEVD_ENONPOS_DEGREE, "non-positive degree d~ (min " + std::to_string(mm[0]) + ")");
    k_weights<<<gn, 256, 0, ctx->stream>>>(dt, n, 2, w1);  // 1/d~
    KEVD_TRY(ops.colsum(w1, cq));                      // cq = F^T d~^-1
    KEVD_TRY(ops.rowdot(cq, qt, 0, nullptr, nullptr, mm));  // q~ = F cq
    KEVD_TRY(rec.vec("cq", cq, true));
    KEVD_TRY(rec.vec("pq", cq, false));
    KEVD_TRY(rec.put("qtil", qt, n));
    if (!(mm[0] > 0.0))
        return kevd_fail(ctx, KEVD_ENONPOS_Q, "non-positive q~ (min " + std::to_string(mm[0]) + ")");
    k_weights<<<gn, 256, 0, ctx->stream>>>(dt, n, 0, w1);  // 1/d~^2
    k_weights<<<gn, 256, 0, ctx->stream>>>(qt, n, 1, w2);  // 1/q~
    KEVD_HIP(ctx, hipGetLastError());
    {
        TraceRange tr("kEVD Grams A, B");
        KEVD_TRY(ops.gram2(w1, w2, A, B));
    }
    KEVD_TRY(rec.sandwich("Atil", A, true));
    KEVD_TRY(rec.put("A", A, (size_t)r * r));
    KEVD_TRY(rec.sandwich("Btil", B, false));
    // dense phase 1: A + sI = L L^T (R = L^T)
    double trsh[2];
    {
        ProfScope ps(ctx, TC_DENSE);
        k_trace_shift<<<1, 256, 0, ctx->stream>>>(A, r, trs);
        BLAS(rocsolver_dpotrf(h, rocblas_fill_lower, (rocblas_int)r, A, (rocblas_int)r, info));
    }
    int pinfo = 0;
    KEVD_HIP(ctx, hipMemcpyAsync(trsh, trs, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(&pinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    res->shift = trsh[1];
    if (pinfo != 0)
        return kevd_fail(ctx, KEVD_EPOTRF, "Cholesky of A + sI failed at minor " + std::to_string(pinfo));
    KEVD_TRY(rec.chol(A));
    // dense phase 2: W = R B R^T = L^T B L
    double one = 1.0, zero = 0.0;
    {
        ProfScope ps(ctx, TC_DENSE);
        BLAS(rocblas_dtrmm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                           rocblas_diagonal_non_unit, (rocblas_int)r, (rocblas_int)r, &one, A,
                           (rocblas_int)r, B, (rocblas_int)r, W, (rocblas_int)r));  // W = B L
        // v2 = B cq for the row-sum check, before B is reused
        BLAS(rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)r, (rocblas_int)r, &one, B,
                           (rocblas_int)r, cq, 1, &zero, v2, 1));
        BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, (rocblas_int)r, (rocblas_int)r, &one, A,
                           (rocblas_int)r, W, (rocblas_int)r, B, (rocblas_int)r));  // B = L^T W
    }
    KEVD_TRY(rec.put("W", B, (size_t)r * r));
    double* Vk = nullptr;
    int eit = 0;
    {
        TraceRange tr("kEVD eigensolve");
        KEVD_TRY(eig_leading(ctx, B, r, k, p->all_eigenvalues != 0, lam, &Vk, &eit));
    }
    // all r eigenvectors when the solve was a full one (they sit in B)
    const bool full_v = Vk == B + (r - k) * r;
    KEVD_TRY(rec.put("V", full_v ? B : Vk, (size_t)r * (full_v ? r : k)));
    KEVD_TRY(rec.put("lam", lam, r));
    // Y = L^-T V_k  (the trailing k eigenvectors, ascending)
    {
        ProfScope ps(ctx, TC_DENSE);
        BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, (rocblas_int)r, (rocblas_int)k, &one, A,
                           (rocblas_int)r, Vk, (rocblas_int)r));
    }
    KEVD_TRY(rec.z(Vk, k));
    // U = D~^-1 F Y, and P1 = D~^-1 F (B cq)
    k_weights<<<gn, 256, 0, ctx->stream>>>(dt, n, 2, w1);
    KEVD_TRY(ops.ugemm(Vk, k, nullptr, v2, w1, ctx->U, P1));
    int nbd = 1024;
    double* bdev;
    KEVD_TRY(ctx_buf_t(ctx, "e_rsdev", nbd, &bdev));
    k_rowsum_dev<<<nbd, 256, 0, ctx->stream>>>(P1, n, bdev);
    double* mmd;
    KEVD_TRY(ctx_buf_t(ctx, "e_mm2", 2, &mmd));
    k_block_minmax<<<1, 256, 0, ctx->stream>>>(bdev, bdev, nbd, mmd);
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipMemcpyAsync(mm, mmd, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(res->lam, lam, r * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    res->rowsum_dev = mm[1];
    return KEVD_OK;
}

// Step schedule of the fused Sinkhorn iteration: the coefficients of
// s_{j+1} = a1 s_j + a2 s_{j-1} + a3 T(s_j) after the pass that measured
// s_j.  In the error of s the map T acts as -P on the modes below the
// scale, with the spectrum of P in [0, lam_2] (P is positive semidefinite).
// The first NPLAIN steps are plain (s <- T(s)); then the Chebyshev
// semi-iteration for I + P on [1, 1 + lam] (Golub and Varga, Numer. Math.
// 3 (1961) 147; the recurrence of Saad, Iterative Methods for Sparse
// Linear Systems, 2nd ed., Alg. 12.1) contracts the error by about
// 1 / (sig + sqrt(sig^2 - 1)) per step, sig = (2 + lam) / lam.  lam starts
// at 0.3; when the RMS deviation falls more slowly than that by a factor
// 1.5 over two steps, the mode that decays at the observed rate sits at
// lam' = lam (x + 1) / 2 with x from the Chebyshev growth outside the
// interval, and the recurrence restarts with 1.1 lam' (at least 1.25 lam,
// at most 0.995).
struct SkCheb {
    static constexpr int NPLAIN = 2;
    double lam = 0.3;
    int j = 0;  // steps decided
    int k = 0;  // steps of the current Chebyshev run
    double rho = 0.0;
    std::vector<double> e2;  // RMS deviation of every measured iterate

    void next(double err, double* a1, double* a2, double* a3)
    {
        e2.push_back(err);
        ++j;
        if (j <= NPLAIN) {
            *a1 = 0.0, *a2 = 0.0, *a3 = 1.0;
            return;
        }
        double th = 1.0 + 0.5 * lam, de = 0.5 * lam, sig = th / de;
        size_t m = e2.size();
        if (k >= 3 && e2[m - 3] > 0.0) {
            double g = sig + std::sqrt(sig * sig - 1.0);
            double q = std::sqrt(e2[m - 1] / e2[m - 3]);
            if (q > 1.5 / g) {
                double z = q * g, x = 0.5 * (z + 1.0 / z);
                lam = std::min(0.995, std::max(1.25 * lam, 1.1 * lam * (x + 1.0) * 0.5));
                k = 0;
                th = 1.0 + 0.5 * lam, de = 0.5 * lam, sig = th / de;
            }
        }
        double alpha, beta;
        if (k == 0) {
            rho = 1.0 / sig;
            alpha = 0.0, beta = 1.0 / th;
        } else {
            double rn = 1.0 / (2.0 * sig - rho);
            alpha = rn * rho, beta = 2.0 * rn / de;
            rho = rn;
        }
        ++k;
        *a1 = 1.0 - beta + alpha, *a2 = -alpha, *a3 = beta;
    }
};

// ------------------------------------------- batched Sinkhorn (ensembles)
// One Sinkhorn iteration of many problems as one launch per stage: the
// fused pass over every problem's row blocks (one launch per register
// tiling VPQ), the column-partial and statistics reductions, and the next
// iterates.  Each block runs the single-problem kernels' bodies on its
// problem, so every problem's iterates are those of its own kevd_evd.
struct SkDesc {
    const float* F;
    int64_t ldf, n, r, rpad;
    const double *u, *s;
    double *q, *part, *bstat, *v, *st;
    int damped, nb;
};
// out = a x + b y + c z over count entries (k_lincomb3's arithmetic)
struct LcDesc {
    double a, b, c;
    const double *x, *y, *z;
    double* out;
    int64_t count;
};

template <int VPQ>
__global__ __launch_bounds__(256) void k_ens_skpass(const SkDesc* __restrict__ D, const int2* __restrict__ items)
{
    const int2 it = items[blockIdx.x];
    const SkDesc& d = D[it.x];
    sinkhorn_rows_body<float, VPQ>(d.F, d.ldf, d.n, d.r, d.u, d.s, d.damped, d.q, d.part, d.rpad, d.bstat, it.y);
}

// items: (problem, 64-column block) for the partials; block b < nprob of
// the second launch shape reduces problem b's statistics
__global__ __launch_bounds__(1024) void k_ens_colsum_reduce(const SkDesc* __restrict__ D, const int2* __restrict__ items)
{
    const int2 it = items[blockIdx.x];
    const SkDesc& d = D[it.x];
    colsum_reduce_body(d.part, d.nb, d.r, d.rpad, d.v, it.y);
}

__global__ __launch_bounds__(256) void k_ens_stat_reduce(const SkDesc* __restrict__ D)
{
    const SkDesc& d = D[blockIdx.x];
    stat_reduce_body(d.bstat, d.nb, d.st);
}

// blockIdx.y: the descriptor; a grid-stride loop over its entries
__global__ __launch_bounds__(256) void k_ens_lincomb(const LcDesc* __restrict__ L)
{
    const LcDesc& d = L[blockIdx.y];
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < d.count; j += (int64_t)gridDim.x * 256) {
        double v = d.a != 0.0 ? d.a * d.x[j] : 0.0;
        if (d.b != 0.0) v += d.b * d.y[j];
        if (d.c != 0.0) v += d.c * d.z[j];
        d.out[j] = v;
    }
}

template <typename T, typename Ops>
kevd_status evd_sinkhorn(kevd_ctx* ctx, Ops& ops, const kevd_evd_params* p, kevd_evd_result* res)
{
    int64_t n = ctx->n, r = ctx->rank;
    int k = p->k_out;
    rocblas_handle h = blas_handle(ctx);
    if (!h) return kevd_fail(ctx, KEVD_EHIP, "rocblas handle creation failed");
    double *s, *s2, *u, *Ks, *w1, *G, *lam, *isq;
    KEVD_TRY(ctx_buf_t(ctx, "e_s", n, &s));
    KEVD_TRY(ctx_buf_t(ctx, "e_s2", n, &s2));
    KEVD_TRY(ctx_buf_t(ctx, "e_u", r, &u));
    KEVD_TRY(ctx_buf_t(ctx, "e_Ks", n, &Ks));
    KEVD_TRY(ctx_buf_t(ctx, "e_w1", n, &w1));
    KEVD_TRY(ctx_buf_t(ctx, "e_A", (size_t)r * r, &G));
    KEVD_TRY(ctx_buf_t(ctx, "e_lam", r, &lam));
    KEVD_TRY(ctx_buf_t(ctx, "e_isq", k, &isq));
    unsigned gn = (unsigned)((n + 255) / 256);
    double mm[2];
    int it = 0;
    double dev = INFINITY, smin_fused = NAN;
    int maxit = p->sinkhorn_max_iter > 0 ? p->sinkhorn_max_iter : 1000;
    double tol = p->sinkhorn_tol > 0.0 ? p->sinkhorn_tol : 1e-12;
    // kevd_ensemble_evd's batched iteration may have scaled this problem
    // already (the same passes and schedule): take its result
    const bool pre = ctx->sk_pre.valid;
    ctx->sk_pre.valid = false;
    if (pre) {
        ctx->tm.sinkhorn_batched = 1;
        s = ctx->sk_pre.s;
        it = ctx->sk_pre.iters;
        dev = ctx->sk_pre.dev;
        smin_fused = ctx->sk_pre.smin;
    } else {
        // start from s = 1 / sqrt(mean(K^ 1)) so that s K^ s is O(1); with
        // u = F^T 1, mean(K^ 1) = 1^T F F^T 1 / n = |u|^2 / n.
        KEVD_TRY(ops.colsum(nullptr, u));
        std::vector<double> hu((size_t)r);
        KEVD_HIP(ctx, hipMemcpyAsync(hu.data(), u, r * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double mean = 0.0;
        for (double v : hu) mean += v * v;
        mean /= (double)n;
        if (!(mean > 0.0)) return kevd_fail(ctx, KEVD_ENONPOS_DEGREE, "non-positive row sum of K^");
        if (sk_fused_ok<T>(r)) {
            // One fused pass per iteration.  Near the fixed point s* the error
            // e in s = s* (1 + e) maps under s <- 1 / (K^ s) to -P e, with
            // P = S* K^ S* bistochastic and positive semidefinite (K^ = F F^T).
            // Its eigenvalue 1 (the common scale of s) is removed exactly by
            // T(s) = sqrt(mean(s .* K^ s)) / (K^ s), which is invariant under
            // s -> c s; every other mode of the map lies in [-lam_2, 0].  The
            // steps are SkCheb's: two plain steps s <- T(s), then the Chebyshev
            // semi-iteration for the spectrum [0, lam_2] (three-term recurrence
            // on s and, by linearity, on u = F^T s).  Three rises of the
            // deviation in a row switch to the damped update s <- sqrt(s / K^ s)
            // (rate <= 1/2).
            double* st3;
            KEVD_TRY(ctx_buf_t(ctx, "e_skst3", SK_NSTAT, &st3));
            // iterate s_j in Sb[j % 3] (built by k_lincomb3 before its pass),
            // q_j in q; u_j = F^T s_j in ucur, u_{j-1} in uprev, v_j = F^T q_j
            // in uv
            double *s3, *q, *uprev, *uv, *un;
            KEVD_TRY(ctx_buf_t(ctx, "e_s3", n, &s3));
            KEVD_TRY(ctx_buf_t(ctx, "e_q0", n, &q));
            KEVD_TRY(ctx_buf_t(ctx, "e_u2", r, &uprev));
            KEVD_TRY(ctx_buf_t(ctx, "e_uv", r, &uv));
            KEVD_TRY(ctx_buf_t(ctx, "e_un", r, &un));
            double* Sb[3] = {s, s2, s3};
            double* ucur = u;
            unsigned gr = (unsigned)((r + 255) / 256);
            // s_0 = c0 * 1 and u_0 = c0 F^T 1
            double c0 = 1.0 / std::sqrt(mean);
            k_fill_d<<<gn, 256, 0, ctx->stream>>>(Sb[0], n, c0);
            k_lincomb3<<<gr, 256, 0, ctx->stream>>>(r, c0, u, 0.0, nullptr, 0.0, nullptr, ucur);
            int damped = 0;
            SkCheb cheb;
            int rises = 0;
            double prev = INFINITY, st[SK_NSTAT];
            int j = 0;
            for (it = 1; it <= maxit; ++it, ++j) {
                KEVD_TRY(ops.skpass(ucur, Sb[j % 3], damped, q, uv, st3));
                KEVD_HIP(ctx, hipGetLastError());
                KEVD_HIP(ctx, hipMemcpyAsync(st, st3, SK_NSTAT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
                KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
                if (!(st[0] > 0.0)) {
                    // the first pass sees K^ 1 up to a positive constant: a degree
                    if (it == 1) return kevd_fail(ctx, KEVD_ENONPOS_DEGREE, "non-positive row sum of K^");
                    return kevd_fail(ctx, KEVD_ENONPOS_Q, "Sinkhorn: non-positive K^ s");
                }
                dev = st[1];
                smin_fused = st[3];
                s = Sb[j % 3];  // the iterate this pass measured
                if (dev <= tol || it == maxit) break;
                bool was_damped = damped != 0;
                rises = dev > prev ? rises + 1 : 0;
                prev = dev;
                // s_{j+1} = a1 s_j + a2 s_{j-1} + a3 q_j
                double a1, a2, a3;
                if (was_damped) {
                    a1 = 0.0, a2 = 0.0, a3 = 1.0;
                } else {
                    double m = st[2] / (double)n;
                    cheb.next(std::sqrt(st[4] / (double)n), &a1, &a2, &a3);
                    a3 *= std::sqrt(m);
                }
                k_lincomb3<<<gr, 256, 0, ctx->stream>>>(r, a1, ucur, a2, uprev, a3, uv, un);
                k_lincomb3<<<gn, 256, 0, ctx->stream>>>(n, a1, Sb[j % 3], a2, Sb[(j + 2) % 3], a3, q, Sb[(j + 1) % 3]);
                std::swap(uprev, ucur);  // uprev = u_j
                std::swap(ucur, un);     // ucur = u_{j+1}; un free
                damped = (was_damped || rises >= 3) ? 1 : 0;
            }
        } else {
            // rank beyond the register tiling: two passes per damped iteration
            k_fill_d<<<gn, 256, 0, ctx->stream>>>(s, n, 1.0 / std::sqrt(mean));
            for (it = 1; it <= maxit; ++it) {
                KEVD_TRY(ops.colsum(s, u));
                KEVD_TRY(ops.rowdot(u, Ks, 1, s, s2, mm));
                if (!(mm[0] > 0.0)) {
                    if (it == 1) return kevd_fail(ctx, KEVD_ENONPOS_DEGREE, "non-positive row sum of K^");
                    return kevd_fail(ctx, KEVD_ENONPOS_Q, "Sinkhorn: non-positive K^ s");
                }
                dev = mm[1];
                if (dev <= tol || it == maxit) break;
                std::swap(s, s2);
            }
        }
    }
    res->sinkhorn_iters = it > maxit ? maxit : it;
    res->rowsum_dev = dev;
    res->shift = 0.0;
    // s holds the scaling whose deviation was measured last
    k_weights<<<gn, 256, 0, ctx->stream>>>(s, n, 3, w1);  // s^2
    KEVD_HIP(ctx, hipGetLastError());
    {
        TraceRange tr("kEVD Gram F^T S^2 F");
        KEVD_TRY(ops.gram1(w1, G));
    }
    double* Vk = nullptr;
    int eit = 0;
    {
        TraceRange tr("kEVD eigensolve");
        KEVD_TRY(eig_leading(ctx, G, r, k, p->all_eigenvalues != 0, lam, &Vk, &eit));
    }
    k_inv_sqrt_tail<<<(k + 255) / 256, 256, 0, ctx->stream>>>(lam, r, k, isq);
    KEVD_TRY(ops.ugemm(Vk, k, isq, nullptr, s, ctx->U, nullptr));
    KEVD_HIP(ctx, hipMemcpyAsync(res->lam, lam, r * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (std::isnan(smin_fused)) {
        // two-pass path: min(s) by a block-partial reduction
        int nbm = (int)std::min<int64_t>((n + 255) / 256, 4096);
        double *pm, *smm;
        KEVD_TRY(ctx_buf_t(ctx, "e_spm", nbm, &pm));
        KEVD_TRY(ctx_buf_t(ctx, "e_smm", 2, &smm));
        k_min_partial<<<nbm, 256, 0, ctx->stream>>>(s, n, pm);
        k_block_minmax<<<1, 256, 0, ctx->stream>>>(pm, nullptr, nbm, smm);
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_HIP(ctx, hipMemcpyAsync(mm, smm, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        smin_fused = mm[0];
    } else {
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    res->min_degree = smin_fused;
    return KEVD_OK;
}

template <typename T, typename Ops>
kevd_status run_norm(kevd_ctx* ctx, Ops& ops, const kevd_evd_params* p, kevd_evd_result* res)
{
    KEVD_TRY(ops.begin());
    return p->norm == KEVD_NORM_REF ? evd_ref<T>(ctx, ops, p, res) : evd_sinkhorn<T>(ctx, ops, p, res);
}

template <typename T>
kevd_status run_evd(kevd_ctx* ctx, bool panels, const kevd_evd_params* p, kevd_evd_result* res)
{
    if constexpr (sizeof(T) == 8) {
        if (panels) {
            PanelOps<T> ops(ctx);
            return run_norm<T>(ctx, ops, p, res);
        }
    }
    FactorOps<T> ops(ctx);
    return run_norm<T>(ctx, ops, p, res);
}

}  // namespace

// The batched Sinkhorn iteration of kevd_ensemble_evd (declared in
// internal.h): evd_sinkhorn's fused schedule (SkCheb, the damped switch
// after three rises, the stop at tol or the step limit) run for every
// problem in lockstep, one launch per stage for all problems and one host
// synchronisation per iteration for all of them instead of one per problem
// and iteration.  The iterates live in each problem's own named buffers,
// under the names evd_sinkhorn uses, and the scaling is left in sk_pre for
// the problem's kevd_evd, which then skips its own iteration.
kevd_status ens_sinkhorn(kevd_ctx** ctxs, int P, const kevd_evd_params* p, int* batched)
{
    *batched = 0;
    kevd_ctx* m = ctxs[0];
    for (int i = 0; i < P; ++i) ctxs[i]->sk_pre.valid = false;
    if (p->norm != KEVD_NORM_SINKHORN || P < 2) return KEVD_OK;
    for (int i = 0; i < P; ++i) {
        kevd_ctx* c = ctxs[i];
        const int64_t r = c->rank;
        // the fused FP32 pass and F^T 1 from the downdate's partials
        if (!c->have_factor || c->fprec != KEVD_MIXED || !c->cs_valid || !c->cs_part || c->cs_nblk < 1 ||
            c->device != m->device || r < 1 || ((r + 3) / 4 + 63) / 64 > 8 || c->n < 1)
            return KEVD_OK;
    }
    TraceRange tr("kEVD batched Sinkhorn");
    KEVD_HIP(m, hipSetDevice(m->device));
    for (int i = 0; i < P; ++i) KEVD_HIP(ctxs[i], hipStreamSynchronize(ctxs[i]->stream));
    hipStream_t st = m->stream;
    const int maxit = p->sinkhorn_max_iter > 0 ? p->sinkhorn_max_iter : 1000;
    const double tol = p->sinkhorn_tol > 0.0 ? p->sinkhorn_tol : 1e-12;
    struct Prob {
        int64_t n, r, rpad;
        int nb, grp;  // row blocks; register tiling VPQ = 2 << grp
        double *Sb[3], *q, *ucur, *uprev, *uv, *un, *part, *bstat;
        int j = 0, damped = 0, rises = 0;
        double prev = INFINITY;
        SkCheb cheb;
        bool active = true, done = false;
        double dev = INFINITY, smin = NAN, *sfin = nullptr;
        int iters = 0;
    };
    std::vector<Prob> pr((size_t)P);
    int64_t nitems = 0, ncitems = 0;
    for (int i = 0; i < P; ++i) {
        kevd_ctx* c = ctxs[i];
        Prob& q = pr[i];
        q.n = c->n;
        q.r = c->rank;
        q.rpad = KEVD_ROUND_UP(q.r, 4);
        q.nb = (int)((q.n + SK_ROWS - 1) / SK_ROWS);
        const int vpq = (int)(((q.r + 3) / 4 + 63) / 64);
        q.grp = vpq <= 2 ? 0 : (vpq <= 4 ? 1 : 2);
        KEVD_TRY(ctx_buf_t(c, "e_s", q.n, &q.Sb[0]));
        KEVD_TRY(ctx_buf_t(c, "e_s2", q.n, &q.Sb[1]));
        KEVD_TRY(ctx_buf_t(c, "e_s3", q.n, &q.Sb[2]));
        KEVD_TRY(ctx_buf_t(c, "e_q0", q.n, &q.q));
        KEVD_TRY(ctx_buf_t(c, "e_u", q.r, &q.ucur));
        KEVD_TRY(ctx_buf_t(c, "e_u2", q.r, &q.uprev));
        KEVD_TRY(ctx_buf_t(c, "e_uv", q.r, &q.uv));
        KEVD_TRY(ctx_buf_t(c, "e_un", q.r, &q.un));
        KEVD_TRY(ctx_buf_t(c, "e_skpart", (size_t)q.nb * q.rpad, &q.part));
        KEVD_TRY(ctx_buf_t(c, "e_skstat", (size_t)SK_NSTAT * q.nb, &q.bstat));
        nitems += q.nb;
        ncitems += (q.r + 63) / 64;
    }
    // device and pinned staging: descriptors, work items, statistics
    const size_t b_desc = (size_t)P * sizeof(SkDesc), b_items = (size_t)nitems * sizeof(int2),
                 b_citems = (size_t)ncitems * sizeof(int2), b_lc = (size_t)2 * P * sizeof(LcDesc),
                 b_stat = (size_t)P * SK_NSTAT * sizeof(double);
    const size_t o_items = KEVD_ROUND_UP(b_desc, 256), o_citems = o_items + KEVD_ROUND_UP(b_items, 256),
                 o_lc = o_citems + KEVD_ROUND_UP(b_citems, 256), o_stat = o_lc + KEVD_ROUND_UP(b_lc, 256),
                 total = o_stat + KEVD_ROUND_UP(b_stat, 256);
    unsigned char *dst, *hst;
    KEVD_TRY(ctx_buf(m, "es_stage", total, (void**)&dst));
    KEVD_TRY(ctx_pinned(m, total, (void**)&hst));
    double* dstat = reinterpret_cast<double*>(dst + o_stat);
    const double* hstat = reinterpret_cast<const double*>(hst + o_stat);
    // u_0 = F^T 1 from the downdate's partials; mean(K^ 1) = |u_0|^2 / n
    std::vector<std::vector<double>> hu((size_t)P);
    for (int i = 0; i < P; ++i) {
        kevd_ctx* c = ctxs[i];
        Prob& q = pr[i];
        k_colsum_reduce<<<(unsigned)((q.r + 63) / 64), 1024, 0, st>>>(c->cs_part, (int)c->cs_nblk, q.r, c->ldf,
                                                                       q.ucur);
        hu[i].resize((size_t)q.r);
        KEVD_HIP(m, hipMemcpyAsync(hu[i].data(), q.ucur, q.r * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    KEVD_HIP(m, hipStreamSynchronize(st));
    for (int i = 0; i < P; ++i) {
        Prob& q = pr[i];
        double mean = 0.0;
        for (double v : hu[i]) mean += v * v;
        mean /= (double)q.n;
        if (!(mean > 0.0)) {  // its own kEVD reports the failure
            q.active = false;
            continue;
        }
        const double c0 = 1.0 / std::sqrt(mean);
        k_fill_d<<<(unsigned)((q.n + 255) / 256), 256, 0, st>>>(q.Sb[0], q.n, c0);
        k_lincomb3<<<(unsigned)((q.r + 255) / 256), 256, 0, st>>>(q.r, c0, q.ucur, 0.0, nullptr, 0.0, nullptr,
                                                                   q.ucur);
    }
    KEVD_HIP(m, hipGetLastError());
    std::vector<int> act;
    for (int it = 1; it <= maxit; ++it) {
        act.clear();
        for (int i = 0; i < P; ++i)
            if (pr[i].active) act.push_back(i);
        if (act.empty()) break;
        const int na = (int)act.size();
        SkDesc* hd = reinterpret_cast<SkDesc*>(hst);
        int2* hi = reinterpret_cast<int2*>(hst + o_items);
        int2* hci = reinterpret_cast<int2*>(hst + o_citems);
        int64_t ni[3] = {0, 0, 0}, nci = 0;
        for (int a = 0; a < na; ++a) {
            const Prob& q = pr[act[a]];
            hd[a] = SkDesc{(const float*)ctxs[act[a]]->F, ctxs[act[a]]->ldf, q.n, q.r, q.rpad, q.ucur, q.Sb[q.j % 3],
                           q.q, q.part, q.bstat, q.uv, dstat + (int64_t)SK_NSTAT * a, q.damped, q.nb};
            for (int c = 0; c < (int)((q.r + 63) / 64); ++c) hci[nci++] = make_int2(a, c);
        }
        // row-block items grouped by register tiling, heaviest group first
        int64_t off[3] = {0, 0, 0};
        for (int a = 0; a < na; ++a) ni[pr[act[a]].grp] += pr[act[a]].nb;
        off[1] = ni[0];
        off[2] = ni[0] + ni[1];
        int64_t fill[3] = {off[0], off[1], off[2]};
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < pr[act[a]].nb; ++b) hi[fill[pr[act[a]].grp]++] = make_int2(a, b);
        KEVD_HIP(m, hipMemcpyAsync(dst, hst, o_lc, hipMemcpyHostToDevice, st));
        const SkDesc* dd = reinterpret_cast<const SkDesc*>(dst);
        const int2* di = reinterpret_cast<const int2*>(dst + o_items);
        if (ni[0]) k_ens_skpass<2><<<(unsigned)ni[0], 256, 0, st>>>(dd, di + off[0]);
        if (ni[1]) k_ens_skpass<4><<<(unsigned)ni[1], 256, 0, st>>>(dd, di + off[1]);
        if (ni[2]) k_ens_skpass<8><<<(unsigned)ni[2], 256, 0, st>>>(dd, di + off[2]);
        k_ens_colsum_reduce<<<(unsigned)nci, 1024, 0, st>>>(dd, reinterpret_cast<const int2*>(dst + o_citems));
        k_ens_stat_reduce<<<(unsigned)na, 256, 0, st>>>(dd);
        KEVD_HIP(m, hipGetLastError());
        KEVD_HIP(m, hipMemcpyAsync(hst + o_stat, dstat, (size_t)na * SK_NSTAT * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
        KEVD_HIP(m, hipStreamSynchronize(st));
        // each problem's step, as evd_sinkhorn's loop takes it
        LcDesc* hl = reinterpret_cast<LcDesc*>(hst + o_lc);
        int nl = 0;
        int64_t lmax = 0;
        for (int a = 0; a < na; ++a) {
            Prob& q = pr[act[a]];
            const double* sv = hstat + (int64_t)SK_NSTAT * a;
            if (!(sv[0] > 0.0)) {  // its own kEVD reports the failure
                q.active = false;
                continue;
            }
            q.dev = sv[1];
            q.smin = sv[3];
            if (q.dev <= tol || it == maxit) {
                q.active = false;
                q.done = true;
                q.sfin = q.Sb[q.j % 3];
                q.iters = it;
                continue;
            }
            const bool was_damped = q.damped != 0;
            q.rises = q.dev > q.prev ? q.rises + 1 : 0;
            q.prev = q.dev;
            double a1, a2, a3;
            if (was_damped) {
                a1 = 0.0, a2 = 0.0, a3 = 1.0;
            } else {
                const double mq = sv[2] / (double)q.n;
                q.cheb.next(std::sqrt(sv[4] / (double)q.n), &a1, &a2, &a3);
                a3 *= std::sqrt(mq);
            }
            hl[nl++] = LcDesc{a1, a2, a3, q.ucur, q.uprev, q.uv, q.un, q.r};
            hl[nl++] = LcDesc{a1, a2, a3, q.Sb[q.j % 3], q.Sb[(q.j + 2) % 3], q.q, q.Sb[(q.j + 1) % 3], q.n};
            lmax = std::max(lmax, q.n);
            std::swap(q.uprev, q.ucur);
            std::swap(q.ucur, q.un);
            q.damped = (was_damped || q.rises >= 3) ? 1 : 0;
            ++q.j;
        }
        if (nl > 0) {
            KEVD_HIP(m, hipMemcpyAsync(dst + o_lc, hst + o_lc, (size_t)nl * sizeof(LcDesc), hipMemcpyHostToDevice,
                                       st));
            const unsigned gx = (unsigned)std::min<int64_t>((lmax + 255) / 256, 1024);
            k_ens_lincomb<<<dim3(gx, (unsigned)nl), 256, 0, st>>>(reinterpret_cast<const LcDesc*>(dst + o_lc));
            KEVD_HIP(m, hipGetLastError());
        }
    }
    KEVD_HIP(m, hipStreamSynchronize(st));
    for (int i = 0; i < P; ++i) {
        if (!pr[i].done) continue;
        kevd_ctx::SkPre& sp = ctxs[i]->sk_pre;
        sp.valid = true;
        sp.s = pr[i].sfin;
        sp.iters = pr[i].iters;
        sp.dev = pr[i].dev;
        sp.smin = pr[i].smin;
        ++*batched;
    }
    return KEVD_OK;
}

extern "C" kevd_status kevd_evd(kevd_ctx* ctx, const kevd_evd_params* p, kevd_evd_result* res)
{
    TraceRange trace_range("kevd_evd");
    if (!ctx || !p || !res) return kevd_fail(ctx, KEVD_EINVAL, "evd: null argument");
    // no factor but the LM form's pivots and L^-1: regenerate K(R,S) panels
    const bool panels = !ctx->have_factor && ctx->lm && ctx->lm_linv;
    if (!(ctx->have_factor || panels) || ctx->rank < 1) return kevd_fail(ctx, KEVD_EINVAL, "evd: no factor");
    if (p->k_out < 1 || p->k_out > ctx->rank || p->k_out > 1024)
        return kevd_fail(ctx, KEVD_EINVAL, "evd: need 1 <= k_out <= min(rank, 1024)");
    if (p->norm != KEVD_NORM_REF && p->norm != KEVD_NORM_SINKHORN)
        return kevd_fail(ctx, KEVD_EINVAL, "evd: unknown normalisation");
    if (!res->lam) return kevd_fail(ctx, KEVD_EINVAL, "evd: lam output missing");
    if (p->prec != KEVD_FP64 && p->prec != KEVD_MIXED) return kevd_fail(ctx, KEVD_EINVAL, "evd: precision");
    if (!panels && p->prec != ctx->fprec)
        return kevd_fail(ctx, KEVD_EINVAL, "evd: precision must match the factor's precision");
    // Without a factor every pass works in the kernel's basis, F = K(:,S) L^-T:
    // the rounding of K and of its Gram is carried through L^-1, whose norm
    // reaches tol^-1/2, so FP32 kernel values would cost up to tol^-1 in the
    // r x r matrices (measured: top eigenvalues off by O(1)).  FP64 only.
    if (panels && p->prec != KEVD_FP64)
        return kevd_fail(ctx, KEVD_EINVAL,
                         "evd: the kEVD without a factor (after the LM form of ARPC) runs in FP64 only");
    if (ctx->rank > 46340) return kevd_fail(ctx, KEVD_EINVAL, "evd: rank too large for the r x r phase");
    ctx->record.clear();
    if (ctx->dump && p->norm == KEVD_NORM_REF && ctx->pivots.size() != (size_t)ctx->rank)
        return kevd_fail(ctx, KEVD_EINVAL,
                         "evd: the intermediate record needs the pivot Cholesky factor, so a factor from kevd_arpc");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->tm.evd_ms = 0;
    ctx->tm.gram_ms = 0;
    ctx->tm.gram_launches = 0;
    ctx->tm.gram_flop = 0;
    ctx->tm.eig_ms = 0;
    ctx->tm.matvec_ms = 0;
    ctx->tm.dense_ms = 0;
    ctx->tm.u_ms = 0;
    ctx->tm.eig_iters = 0;
    ctx->tm.eig_householder = 0;
    ctx->tm.eig_full_fallback = 0;
    ctx->tm.eig_orth_passes = 0;
    ctx->tm.eig_batched = 0;
    ctx->tm.sinkhorn_batched = 0;
    prof_reset(ctx);
    size_t ubytes = (size_t)ctx->n * p->k_out * sizeof(float);
    if (ctx->U_bytes < ubytes) {
        if (ctx->U) KEVD_HIP(ctx, hipFree(ctx->U));
        ctx->U = nullptr;
        KEVD_HIP(ctx, hipMalloc(&ctx->U, ubytes));
        ctx->U_bytes = ubytes;
    }
    ctx->k_out = p->k_out;
    res->sinkhorn_iters = 0;
    const bool f32 = !panels && ctx->fprec == KEVD_MIXED;
    kevd_status s = f32 ? run_evd<float>(ctx, panels, p, res) : run_evd<double>(ctx, panels, p, res);
    if (s != KEVD_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        prof_reset(ctx);
        return s;
    }
    if (res->U)
        KEVD_HIP(ctx, hipMemcpyAsync(res->U, ctx->U, ubytes, hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    auto t1 = std::chrono::steady_clock::now();
    ctx->tm.evd_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    prof_resolve(ctx);
    return KEVD_OK;
}

extern "C" kevd_status kevd_get_u(kevd_ctx* ctx, float* U_host)
{
    if (!ctx || !U_host || !ctx->U || ctx->k_out < 1)
        return kevd_fail(ctx, KEVD_EINVAL, "get_u: no eigenvectors computed");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    KEVD_HIP(ctx, hipMemcpyAsync(U_host, ctx->U, (size_t)ctx->n * ctx->k_out * sizeof(float),
                                 hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

extern "C" kevd_status kevd_set_dump(kevd_ctx* ctx, int32_t on)
{
    if (!ctx) return KEVD_EINVAL;
    ctx->dump = on != 0;
    if (!ctx->dump) ctx->record.clear();
    return KEVD_OK;
}

extern "C" kevd_status kevd_get_intermediate(kevd_ctx* ctx, const char* name, double* out, int64_t* count)
{
    if (!ctx || !name || !count) return kevd_fail(ctx, KEVD_EINVAL, "get_intermediate: null argument");
    auto it = ctx->record.find(name);
    if (it == ctx->record.end())
        return kevd_fail(ctx, KEVD_EINVAL, std::string("get_intermediate: no recorded '") + name + "'");
    const int64_t m = (int64_t)it->second.size();
    if (out) {
        if (*count < m)
            return kevd_fail(ctx, KEVD_EINVAL, std::string("get_intermediate: '") + name + "' holds " +
                                                   std::to_string(m) + " values");
        std::memcpy(out, it->second.data(), (size_t)m * sizeof(double));
    }
    *count = m;
    return KEVD_OK;
}

namespace {
std::mutex tile_mu;
std::map<std::tuple<int, int64_t, int64_t, int64_t>, int> tile_cache;
}  // namespace

extern "C" kevd_status kevd_autotune_tiles(kevd_ctx* ctx, int32_t* gram_tile, double* gram_ms)
{
    TraceRange trace_range("kevd_autotune_tiles");
    if (!ctx || !gram_tile) return kevd_fail(ctx, KEVD_EINVAL, "autotune_tiles: null argument");
    if (!ctx->have_factor || ctx->rank < 1) return kevd_fail(ctx, KEVD_EINVAL, "autotune_tiles: no factor");
    if (ctx->fprec != KEVD_MIXED || !use_bf16x6()) {
        // only the bf16x6 Gram of an FP32 factor has a tile choice
        ctx->gram_tile = 0;
        *gram_tile = 0;
        return KEVD_OK;
    }
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    const auto key = std::make_tuple(ctx->device, ctx->n, ctx->rank, ctx->ldf);
    if (!gram_ms) {
        std::lock_guard<std::mutex> g(tile_mu);
        auto it = tile_cache.find(key);
        if (it != tile_cache.end()) {
            ctx->gram_tile = it->second;
            *gram_tile = it->second;
            return KEVD_OK;
        }
    }
    const int64_t n = ctx->n, r = ctx->rank;
    double *w, *G;
    KEVD_TRY(ctx_buf_t(ctx, "t_w", n, &w));
    KEVD_TRY(ctx_buf_t(ctx, "t_G", (size_t)r * r, &G));
    k_fill1<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(w, n);
    KEVD_HIP(ctx, hipGetLastError());
    const Rows<float> src{(const float*)ctx->F, ctx->ldf, n, r};
    hipEvent_t a, b;
    KEVD_HIP(ctx, hipEventCreate(&a));
    KEVD_HIP(ctx, hipEventCreate(&b));
    double t[2] = {1e30, 1e30};
    kevd_status st = KEVD_OK;
    for (int c = 0; c < 2 && st == KEVD_OK; ++c) {
        for (int rep = 0; rep < 3 && st == KEVD_OK; ++rep) {  // the first run warms up
            if (hipEventRecord(a, ctx->stream) != hipSuccess) st = KEVD_EHIP;
            if (st == KEVD_OK)
                st = c == 0 ? gram_x6<GramX6>(ctx, src, w, G, "t_gw") : gram_x6<GramX6s>(ctx, src, w, G, "t_gw");
            float ms = 0.f;
            if (st == KEVD_OK && hipEventRecord(b, ctx->stream) != hipSuccess) st = KEVD_EHIP;
            if (st == KEVD_OK && hipEventSynchronize(b) != hipSuccess) st = KEVD_EHIP;
            if (st == KEVD_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) st = KEVD_EHIP;
            if (st == KEVD_OK && rep > 0) t[c] = std::min(t[c], (double)ms);
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st == KEVD_EHIP) return kevd_fail(ctx, st, "autotune_tiles: timing failed");
    if (st != KEVD_OK) return st;
    const int choice = t[1] < t[0] ? 128 : 256;
    {
        std::lock_guard<std::mutex> g(tile_mu);
        tile_cache[key] = choice;
    }
    ctx->gram_tile = choice;
    *gram_tile = choice;
    if (gram_ms) {
        gram_ms[0] = t[0];
        gram_ms[1] = t[1];
    }
    return KEVD_OK;
}

extern "C" kevd_status kevd_set_gram_tile(kevd_ctx* ctx, int32_t gram_tile)
{
    if (!ctx) return KEVD_EINVAL;
    if (gram_tile != 0 && gram_tile != 128 && gram_tile != 256)
        return kevd_fail(ctx, KEVD_EINVAL, "set_gram_tile: 0, 128 or 256");
    ctx->gram_tile = gram_tile;
    return KEVD_OK;
}
