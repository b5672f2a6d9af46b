// bandwidth.hip -- fused Gaussian-kernel bandwidth sweep.
//
// S(eps_j) = sum_{i in rows} sum_{c < n} exp(-||x_i - x_c||^2 / eps_j) for all
// eps_j from one launch.  The reference materialises the distance matrix in
// HBM (dmat_comp, ks_bandwidth.cu:181-232) and re-reads it once per
// bandwidth (kernelsum_comp 276-325, launched Nbw times at 566-572); here a
// block computes a 128 x 128 distance tile on MFMA (|x|^2 + |y|^2 - 2 x.y,
// FP64 norms), keeps it in registers and evaluates every eps_j against it.
//
// FP32: the exponent arguments d * c_j (c_j = -log2(e) / eps_j) are MFMA
// outer products, so the vector ALU issues only v_exp_f32 and one add per
// evaluation.  A v_mfma_f32_32x32x2_f32 takes one distance register of the
// wave as its A operand (A[i][k] = lane i + 32 k) and 32 coefficients as B,
// with B zero in the other half of the lanes, so C[i][j] = d(lane i [+32])
// * c_j: 32 distances x 32 eps per instruction.  Lane l receives only
// column j = l % 32 of C, i.e. one eps per lane and set of 32, so each lane
// sums its own eps over its 16 outputs; per tile the FP32 partials of the
// two lanes of an eps are added into FP64 sums in LDS, and the waves and
// blocks are reduced once at the end in fixed order.
// FP64: the same tile, exp(d * c_j) per value, FP64 sums per chunk of 16 eps.
#include <cmath>
#include <vector>

#include "gemm.h"
#include "internal.h"

namespace {

constexpr int NEPS_MAX = 128;
constexpr int ECH = 16;      // FP64: eps values per register chunk
constexpr int EPS_SET = 32;  // FP32: eps per outer-product set
constexpr int NSET_MAX = NEPS_MAX / EPS_SET;

// Distance of tile entry (a, b, v) of g, as T for the exp stage: FP64
// norms minus twice the MFMA dot product, clamped at zero, exactly zero for
// a row's own column; entries outside the rows / columns get `pad`.
template <typename G, typename T>
__device__ inline T tile_dist(const G& g, int a, int b, int v, int64_t r0, int64_t c0, int64_t nrows,
                              int64_t cend, const double* nR, const double* nrm, const int64_t* ridx, T pad)
{
    int64_t row = r0 + g.row_of(a, v);
    int64_t col = c0 + g.col_of(b);
    if (row >= nrows || col >= cend) return pad;
    double Dd = nR[row] + nrm[col] - 2.0 * (double)g.acc[a][b][v];
    Dd = Dd > 0.0 ? Dd : 0.0;
    if (ridx[row] == col) Dd = 0.0;
    return (T)Dd;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bw_sweep(const T* __restrict__ XR, const double* __restrict__ nR,
                                                  const int64_t* __restrict__ ridx, int64_t nrows,
                                                  const T* __restrict__ X, const double* __restrict__ nrm,
                                                  int64_t n, int32_t dp, const T* __restrict__ coef,
                                                  int neps, int64_t cols_per_block, double* part)
{
    typedef typename Cfg<T>::G G;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    constexpr int NV = G::NTM * G::NTN * G::M::ACC;  // distances per lane
    __shared__ __attribute__((aligned(16))) T lds[G::LDS_ELEMS];
    __shared__ double acc64[4][NEPS_MAX];
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = lane; j < NEPS_MAX; j += 64) acc64[w][j] = 0.0;
    int64_t r0 = (int64_t)blockIdx.x * BM;
    int64_t cbeg = (int64_t)blockIdx.y * cols_per_block;
    int64_t cend = cbeg + cols_per_block < n ? cbeg + cols_per_block : n;
    G g;
    for (int64_t c0 = cbeg; c0 < cend; c0 += BN) {
        g.init();
        g.mainloop(lds, XR + r0 * dp, dp, nrows - r0, X + c0 * dp, dp, cend - c0, dp);
        T D[NV];
#pragma unroll
        for (int a = 0; a < G::NTM; ++a)
#pragma unroll
            for (int b = 0; b < G::NTN; ++b)
#pragma unroll
                for (int v = 0; v < G::M::ACC; ++v)
                    D[(a * G::NTN + b) * G::M::ACC + v] =
                        tile_dist(g, a, b, v, r0, c0, nrows, cend, nR, nrm, ridx, (T)INFINITY);
        for (int e0 = 0; e0 < neps; e0 += ECH) {
            T s[ECH];
#pragma unroll
            for (int j = 0; j < ECH; ++j) s[j] = (T)0;
            T c[ECH];
#pragma unroll
            for (int j = 0; j < ECH; ++j) c[j] = coef[e0 + j];  // coef padded to a multiple of ECH
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                T d = D[q];
#pragma unroll
                for (int j = 0; j < ECH; ++j) s[j] += exp(d * c[j]);  // c = -1 / eps
            }
#pragma unroll
            for (int j = 0; j < ECH; ++j) {
                double x = (double)s[j];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
                if (lane == 0) acc64[w][e0 + j] += x;
            }
        }
        // the next mainloop overwrites the LDS stage; its first barrier
        // orders that after every wave's last read of this tile
    }
    __syncthreads();
    double* out = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NEPS_MAX;
    for (int j = threadIdx.x; j < NEPS_MAX; j += 256)
        out[j] = (acc64[0][j] + acc64[1][j]) + (acc64[2][j] + acc64[3][j]);
}

// bcoef: per set s of 32 eps, two 64-lane B operands: [2s] holds c_{32s+l}
// in lanes l < 32 and 0 above, [2s+1] holds 0 below and c_{32s+l-32} above.
// Padding distances are FLT_MAX: times a zero B entry that is exactly 0, and
// times any |c| >= 1e-30 (eps <= 1.4e30, checked on the host) it rounds to
// -inf or below -3e8, whose exp2 is 0.
__global__ __launch_bounds__(256) void k_bw_sweep_f32(const float* __restrict__ XR, const double* __restrict__ nR,
                                                      const int64_t* __restrict__ ridx, int64_t nrows,
                                                      const float* __restrict__ X, const double* __restrict__ nrm,
                                                      int64_t n, int32_t dp, const float* __restrict__ bcoef,
                                                      int nset, int64_t cols_per_block, double* part)
{
    typedef Cfg<float>::G G;
    constexpr int BM = Cfg<float>::BM, BN = Cfg<float>::BN;
    constexpr int NV = G::NTM * G::NTN * G::M::ACC;  // distances per lane
    __shared__ __attribute__((aligned(16))) float lds[G::LDS_ELEMS];
    // FP64 sums of eps 32 s + m per wave; 4 KB, so that two blocks fit a CU
    __shared__ double lacc[4][NSET_MAX][EPS_SET];
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane < EPS_SET)
        for (int s = 0; s < NSET_MAX; ++s) lacc[w][s][lane] = 0.0;
    int64_t r0 = (int64_t)blockIdx.x * BM;
    int64_t cbeg = (int64_t)blockIdx.y * cols_per_block;
    int64_t cend = cbeg + cols_per_block < n ? cbeg + cols_per_block : n;
    const floatx16 zero = {};
    G g;
    for (int64_t c0 = cbeg; c0 < cend; c0 += BN) {
        g.init();
        g.mainloop(lds, XR + r0 * dp, dp, nrows - r0, X + c0 * dp, dp, cend - c0, dp);
        float D[NV];
#pragma unroll
        for (int a = 0; a < G::NTM; ++a)
#pragma unroll
            for (int b = 0; b < G::NTN; ++b)
#pragma unroll
                for (int v = 0; v < G::M::ACC; ++v)
                    D[(a * G::NTN + b) * G::M::ACC + v] =
                        tile_dist(g, a, b, v, r0, c0, nrows, cend, nR, nrm, ridx, 3.402823466e38f);
        for (int s = 0; s < nset; ++s) {
            const float blo = bcoef[(2 * s) * 64 + lane], bhi = bcoef[(2 * s + 1) * 64 + lane];
            float ps0 = 0.f, ps1 = 0.f, ps2 = 0.f, ps3 = 0.f;
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                floatx16 lo = __builtin_amdgcn_mfma_f32_32x32x2f32(D[q], blo, zero, 0, 0, 0);
                floatx16 hi = __builtin_amdgcn_mfma_f32_32x32x2f32(D[q], bhi, zero, 0, 0, 0);
#pragma unroll
                for (int u = 0; u < 16; u += 4) {
                    ps0 += __builtin_amdgcn_exp2f(lo[u]);
                    ps1 += __builtin_amdgcn_exp2f(lo[u + 1]);
                    ps2 += __builtin_amdgcn_exp2f(lo[u + 2]);
                    ps3 += __builtin_amdgcn_exp2f(lo[u + 3]);
                    ps0 += __builtin_amdgcn_exp2f(hi[u]);
                    ps1 += __builtin_amdgcn_exp2f(hi[u + 1]);
                    ps2 += __builtin_amdgcn_exp2f(hi[u + 2]);
                    ps3 += __builtin_amdgcn_exp2f(hi[u + 3]);
                }
            }
            // lanes m and m + 32 hold the same eps
            double x = (double)((ps0 + ps1) + (ps2 + ps3));
            x += __shfl_xor(x, 32, 64);
            if (lane < EPS_SET) lacc[w][s][lane] += x;
        }
        // the next mainloop overwrites the LDS stage; its first barrier
        // orders that after every wave's last read of this tile
    }
    __syncthreads();
    double* out = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NEPS_MAX;
    for (int j = threadIdx.x; j < NEPS_MAX; j += 256) {
        int s = j / EPS_SET, m = j % EPS_SET;
        out[j] = s < nset ? (lacc[0][s][m] + lacc[1][s][m]) + (lacc[2][s][m] + lacc[3][s][m]) : 0.0;
    }
}

// FP32 sweep for dp <= 64 (dp = DPT): the row block's A fragments stay in
// registers for the whole launch, and each 128-column tile of X is loaded
// into registers one tile ahead (issued before the exp stage of the current
// tile, stored to the other LDS buffer after it), so the distance product
// costs LDS reads and 4 DPT MFMAs per wave instead of a pipelined GEMM with
// exposed global latency.  Distances, layout and the exp stage are those of
// k_bw_sweep_f32 (16x16x4 tiles: wave w owns rows 32 w .. 32 w + 31, lane
// (li, h) = (lane % 16, lane / 16) holds rows 16 t + 4 h + v and columns
// 16 y + li of the tile).
template <int DPT>
__global__ __launch_bounds__(256) void k_bw_sweep_f32r(const float* __restrict__ XR, const double* __restrict__ nR,
                                                       const int64_t* __restrict__ ridx, int64_t nrows,
                                                       const float* __restrict__ X, const double* __restrict__ nrm,
                                                       int64_t n, const float* __restrict__ bcoef, int nset,
                                                       int64_t cols_per_block, double* part)
{
    constexpr int BM = 128, BN = 128, PB = DPT + 4, NS = DPT / 16;
    constexpr int Q4 = DPT / 4;               // float4 per point
    constexpr int NLD = BN * Q4 / 256;        // float4 loads per thread per tile
    constexpr int NV = 2 * 8 * 4;             // distances per lane
    static_assert(DPT % 16 == 0 && (BN * Q4) % 256 == 0, "tile shape");
    __shared__ __attribute__((aligned(16))) float Bs[2][BN * PB];
    __shared__ double lacc[4][NSET_MAX][EPS_SET];
    __shared__ float bsh[NSET_MAX * 2][64];  // B operands of the products (LDS: no vmcnt waits)
    __shared__ double rnorm[BM];             // the block's row norms and own columns
    __shared__ int64_t rself[BM];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, h = lane >> 4;
    if (lane < EPS_SET)
        for (int s = 0; s < NSET_MAX; ++s) lacc[w][s][lane] = 0.0;
    for (int i = threadIdx.x; i < nset * 2 * 64; i += 256) bsh[i / 64][i % 64] = bcoef[i];
    const int64_t r0 = (int64_t)blockIdx.x * BM;
    if (threadIdx.x < BM) {
        int64_t row = r0 + threadIdx.x;
        bool ok = row < nrows;
        rnorm[threadIdx.x] = ok ? nR[row] : 0.0;
        rself[threadIdx.x] = ok ? ridx[row] : -1;
    }
    const int64_t cbeg = (int64_t)blockIdx.y * cols_per_block;
    const int64_t cend = cbeg + cols_per_block < n ? cbeg + cols_per_block : n;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // Rows and columns past the ends are read from the last valid row /
    // column (clamped addresses, no selects that would force a wait on the
    // loads); the distance epilogue masks them.
    float4 a[2][NS];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        int64_t row = r0 + w * 32 + t * 16 + li;
        row = row < nrows ? row : nrows - 1;
#pragma unroll
        for (int s = 0; s < NS; ++s) a[t][s] = *reinterpret_cast<const float4*>(XR + row * DPT + s * 16 + h * 4);
    }
    // next tile's points: float4 f = threadIdx.x + 256 i of the tile is
    // point f / Q4, components 4 (f % Q4) ..
    floatx4 pre[NLD];
#define BW_GLOAD(C0)                                                                              \
    _Pragma("unroll") for (int i = 0; i < NLD; ++i)                                               \
    {                                                                                             \
        int f = threadIdx.x + 256 * i, k4 = f % Q4;                                               \
        int64_t col = (C0) + f / Q4;                                                              \
        col = col < cend ? col : cend - 1;                                                        \
        pre[i] = *reinterpret_cast<const floatx4*>(X + col * DPT + k4 * 4);                       \
    }
#define BW_SSTORE(BUF)                                                                            \
    _Pragma("unroll") for (int i = 0; i < NLD; ++i)                                               \
    {                                                                                             \
        int f = threadIdx.x + 256 * i, col = f / Q4, k4 = f % Q4;                                 \
        *reinterpret_cast<floatx4*>((BUF) + col * PB + k4 * 4) = pre[i];                          \
    }
    if (cbeg < cend) {
        BW_GLOAD(cbeg)
        BW_SSTORE(Bs[0])
    }
    const floatx16 zero = {};
    int buf = 0;
    for (int64_t c0 = cbeg; c0 < cend; c0 += BN, buf ^= 1) {
        // tile c0 is in Bs[buf]; every wave has left the previous tile, so
        // Bs[buf ^ 1] is free for the next one
        __syncthreads();
        const bool more = c0 + BN < cend;
        floatx4 acc[2][8];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int y = 0; y < 8; ++y) acc[t][y] = floatx4{0.f, 0.f, 0.f, 0.f};
        const float* B = Bs[buf];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            float4 b[8];
#pragma unroll
            for (int y = 0; y < 8; ++y)
                b[y] = *reinterpret_cast<const float4*>(B + (y * 16 + li) * PB + s * 16 + h * 4);
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int y = 0; y < 8; ++y)
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        Mfma16f::mma(acc[t][y], Mfma16f::get(a[t][s], j), Mfma16f::get(b[y], j));
        }
        // distances: FP64 norms of the lane's 8 columns (clamped, loaded
        // together) and rows (LDS); padding as in k_bw_sweep_f32
        double ncol[8];
#pragma unroll
        for (int y = 0; y < 8; ++y) {
            int64_t col = c0 + y * 16 + li;
            ncol[y] = nrm[col < cend ? col : cend - 1];
        }
        float D[NV];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rl = w * 32 + t * 16 + 4 * h + v;
                const double nr = rnorm[rl];
                const int64_t self = rself[rl];
                const bool rok = r0 + rl < nrows;
#pragma unroll
                for (int y = 0; y < 8; ++y) {
                    int64_t col = c0 + y * 16 + li;
                    double Dd = nr + ncol[y] - 2.0 * (double)acc[t][y][v];
                    Dd = Dd > 0.0 ? Dd : 0.0;
                    Dd = self == col ? 0.0 : Dd;
                    D[(t * 8 + y) * 4 + v] = rok && col < cend ? (float)Dd : 3.402823466e38f;
                }
            }
        // the next tile's loads come after the epilogue's, so that no wait
        // on those (in-order counter) waits on the prefetch too
        if (more) {
            BW_GLOAD(c0 + BN)
        }
        for (int s = 0; s < nset; ++s) {
            const float blo = bsh[2 * s][lane], bhi = bsh[2 * s + 1][lane];
            float ps0 = 0.f, ps1 = 0.f, ps2 = 0.f, ps3 = 0.f;
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                floatx16 lo = __builtin_amdgcn_mfma_f32_32x32x2f32(D[q], blo, zero, 0, 0, 0);
                floatx16 hi = __builtin_amdgcn_mfma_f32_32x32x2f32(D[q], bhi, zero, 0, 0, 0);
#pragma unroll
                for (int u = 0; u < 16; u += 4) {
                    ps0 += __builtin_amdgcn_exp2f(lo[u]);
                    ps1 += __builtin_amdgcn_exp2f(lo[u + 1]);
                    ps2 += __builtin_amdgcn_exp2f(lo[u + 2]);
                    ps3 += __builtin_amdgcn_exp2f(lo[u + 3]);
                    ps0 += __builtin_amdgcn_exp2f(hi[u]);
                    ps1 += __builtin_amdgcn_exp2f(hi[u + 1]);
                    ps2 += __builtin_amdgcn_exp2f(hi[u + 2]);
                    ps3 += __builtin_amdgcn_exp2f(hi[u + 3]);
                }
            }
            double x = (double)((ps0 + ps1) + (ps2 + ps3));
            x += __shfl_xor(x, 32, 64);
            if (lane < EPS_SET) lacc[w][s][lane] += x;
        }
        if (more) {
            BW_SSTORE(Bs[buf ^ 1])
        }
    }
#undef BW_GLOAD
#undef BW_SSTORE
    __syncthreads();
    double* out = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NEPS_MAX;
    for (int j = threadIdx.x; j < NEPS_MAX; j += 256) {
        int s = j / EPS_SET, m = j % EPS_SET;
        out[j] = s < nset ? (lacc[0][s][m] + lacc[1][s][m]) + (lacc[2][s][m] + lacc[3][s][m]) : 0.0;
    }
}

// S[j] = sum over blocks of part[b][j]: one workgroup per eps, a strided
// sum per thread and a fixed-shape tree, so the result does not depend on
// scheduling.
__global__ __launch_bounds__(256) void k_bw_reduce(const double* part, int nblk, double* S)
{
    __shared__ double red[256];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += part[(int64_t)b * NEPS_MAX + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) S[j] = red[0];
}

// Launch shape and buffer sizes of one sweep.
struct SweepPlan {
    int nset = 0, nep = 0;
    size_t ncoef = 0, esz = 0;
    int64_t rt = 0, ncs = 0, cpb = 0;
    int nblk = 0;
};

template <typename T>
SweepPlan plan_sweep(int64_t n, int64_t nrows, int neps)
{
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    const bool f32 = sizeof(T) == 4;
    SweepPlan P;
    P.esz = sizeof(T);
    P.nep = (neps + ECH - 1) / ECH * ECH;
    P.nset = (neps + EPS_SET - 1) / EPS_SET;
    // FP64: the coefficients padded to whole chunks of ECH; FP32: the B
    // operands of the outer-product sets (see k_bw_sweep_f32)
    P.ncoef = f32 ? (size_t)P.nset * 2 * 64 : (size_t)P.nep;
    P.rt = (nrows + BM - 1) / BM;
    int64_t ctiles = (n + BN - 1) / BN;
    int64_t want = 4096 / (P.rt > 0 ? P.rt : 1);
    if (want < 1) want = 1;
    if (want > ctiles) want = ctiles;
    P.cpb = ((ctiles + want - 1) / want) * BN;
    P.ncs = (n + P.cpb - 1) / P.cpb;
    P.nblk = (int)(P.rt * P.ncs);
    return P;
}

// Device buffers of one sweep: the rows' indices, the gathered rows and
// their norms, the eps coefficients, the per-block partial sums and S.
struct SweepBufs {
    int64_t* ri = nullptr;
    void* XR = nullptr;
    double* nR = nullptr;
    void* coef = nullptr;
    double* part = nullptr;
    double* S = nullptr;
};

// Carves SweepBufs out of one region at 256-byte aligned offsets; returns
// the bytes used (base may be null to only measure).
size_t carve_sweep(const SweepPlan& P, int64_t nrows, int32_t dp, char* base, SweepBufs* B)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += KEVD_ROUND_UP(bytes, 256);
        return p;
    };
    SweepBufs b;
    b.ri = (int64_t*)take((size_t)nrows * sizeof(int64_t));
    b.XR = take((size_t)nrows * dp * P.esz);
    b.nR = (double*)take((size_t)nrows * sizeof(double));
    b.coef = take(P.ncoef * P.esz);
    b.part = (double*)take((size_t)P.nblk * NEPS_MAX * sizeof(double));
    b.S = (double*)take(NEPS_MAX * sizeof(double));
    if (B) *B = b;
    return off;
}

// One sweep on the context stream: rows and eps from the host, S to the
// host.  Every buffer is written before it is read (the gather fills XR and
// nR, the coefficient copy covers all ncoef entries, the sweep kernel writes
// every partial), so none needs clearing.
template <typename T>
kevd_status sweep(kevd_ctx* ctx, const SweepPlan& P, const SweepBufs& B, const int64_t* rows, int64_t nrows,
                  const double* eps, int neps, double* S_out)
{
    int64_t n = ctx->n;
    int32_t dp = ctx->dp;
    const bool f32 = sizeof(T) == 4;
    KEVD_HIP(ctx, hipMemcpyAsync(B.ri, rows, nrows * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KEVD_TRY(kblock_gather(ctx, B.ri, nrows, f32 ? KEVD_MIXED : KEVD_FP64, B.XR, B.nR));
    std::vector<T> hc(P.ncoef, (T)0);
    for (int j = 0; j < neps; ++j) {
        if (f32) {
            float c = (float)(-1.4426950408889634074 / eps[j]);
            int st = j / EPS_SET, m = j % EPS_SET;
            hc[(size_t)(2 * st) * 64 + m] = (T)c;
            hc[(size_t)(2 * st + 1) * 64 + 32 + m] = (T)c;
        } else {
            hc[j] = (T)(-1.0 / eps[j]);
        }
    }
    KEVD_HIP(ctx, hipMemcpyAsync(B.coef, hc.data(), P.ncoef * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    const T* X = f32 ? (const T*)ctx->X32 : (const T*)ctx->X64;
    const double* nrm = f32 ? ctx->nrm32 : ctx->nrm64;
    const T* XR = (const T*)B.XR;
    const T* coef = (const T*)B.coef;
    const int64_t* ri = B.ri;
    const double* nR = B.nR;
    dim3 grid((unsigned)P.rt, (unsigned)P.ncs);
    if constexpr (sizeof(T) == 4) {
        if (dp == 32)
            k_bw_sweep_f32r<32><<<grid, 256, 0, ctx->stream>>>(XR, nR, ri, nrows, X, nrm, n, coef, P.nset, P.cpb,
                                                               B.part);
        else if (dp == 64)
            k_bw_sweep_f32r<64><<<grid, 256, 0, ctx->stream>>>(XR, nR, ri, nrows, X, nrm, n, coef, P.nset, P.cpb,
                                                               B.part);
        else
            k_bw_sweep_f32<<<grid, 256, 0, ctx->stream>>>(XR, nR, ri, nrows, X, nrm, n, dp, coef, P.nset, P.cpb,
                                                          B.part);
    } else {
        k_bw_sweep<T><<<grid, 256, 0, ctx->stream>>>(XR, nR, ri, nrows, X, nrm, n, dp, coef, P.nep, P.cpb,
                                                     B.part);
    }
    k_bw_reduce<<<neps, 256, 0, ctx->stream>>>(B.part, P.nblk, B.S);
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipMemcpyAsync(S_out, B.S, neps * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

kevd_status sweep_validate(kevd_ctx* ctx, const int64_t* rows, int64_t nrows, const double* eps, int32_t neps,
                           kevd_precision prec, const double* S_out)
{
    if (!ctx || ctx->n < 1) return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: no points");
    if (!rows || nrows < 1 || !eps || neps < 1 || neps > NEPS_MAX || !S_out)
        return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: need rows, 1 <= neps <= 128 and outputs");
    for (int32_t j = 0; j < neps; ++j)
        if (!(eps[j] > 0.0) || !(eps[j] <= 1e30))
            return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: eps values must lie in (0, 1e30]");
    for (int64_t i = 0; i < nrows; ++i)
        if (rows[i] < 0 || rows[i] >= ctx->n)
            return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: row index out of range");
    if (prec != KEVD_FP64 && prec != KEVD_MIXED) return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: precision");
    return KEVD_OK;
}

template <typename T>
kevd_status sweep_ctx_bufs(kevd_ctx* ctx, const int64_t* rows, int64_t nrows, const double* eps, int neps,
                           double* S_out)
{
    SweepPlan P = plan_sweep<T>(ctx->n, nrows, neps);
    SweepBufs B;
    KEVD_TRY(ctx_buf_t(ctx, "bw_ri", nrows, &B.ri));
    KEVD_TRY(ctx_buf(ctx, "bw_XR", (size_t)nrows * ctx->dp * sizeof(T), &B.XR));
    KEVD_TRY(ctx_buf_t(ctx, "bw_nR", nrows, &B.nR));
    KEVD_TRY(ctx_buf(ctx, "bw_coef", P.ncoef * sizeof(T), &B.coef));
    KEVD_TRY(ctx_buf_t(ctx, "bw_part", (size_t)P.nblk * NEPS_MAX, &B.part));
    KEVD_TRY(ctx_buf_t(ctx, "bw_S", NEPS_MAX, &B.S));
    return sweep<T>(ctx, P, B, rows, nrows, eps, neps, S_out);
}

template <typename T>
kevd_status sweep_scratch(kevd_ctx* ctx, const int64_t* rows, int64_t nrows, const double* eps, int neps,
                          double* S_out)
{
    SweepPlan P = plan_sweep<T>(ctx->n, nrows, neps);
    size_t need = carve_sweep(P, nrows, ctx->dp, nullptr, nullptr);
    if (need > KEVD_SCRATCH_BYTES)
        return kevd_fail(ctx, KEVD_EINVAL,
                         "bandwidth (allocation-free): " + std::to_string(nrows) + " rows need " +
                             std::to_string(need) + " bytes of scratch, more than the fixed " +
                             std::to_string(KEVD_SCRATCH_BYTES));
    if (!ctx->scratch) return kevd_fail(ctx, KEVD_EHIP, "bandwidth (allocation-free): no scratch address");
    SweepBufs B;
    carve_sweep(P, nrows, ctx->dp, (char*)ctx->scratch, &B);
    return sweep<T>(ctx, P, B, rows, nrows, eps, neps, S_out);
}

}  // namespace

extern "C" kevd_status kevd_bandwidth_sweep(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                            const double* eps, int32_t neps, kevd_precision prec,
                                            double* S_out)
{
    TraceRange trace_range("kevd_bandwidth_sweep");
    KEVD_TRY(sweep_validate(ctx, rows, nrows, eps, neps, prec, S_out));
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    return prec == KEVD_MIXED ? sweep_ctx_bufs<float>(ctx, rows, nrows, eps, neps, S_out)
                              : sweep_ctx_bufs<double>(ctx, rows, nrows, eps, neps, S_out);
}

extern "C" kevd_status kevd_bandwidth_sweep_noalloc(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                                    const double* eps, int32_t neps, kevd_precision prec,
                                                    double* S_out)
{
    TraceRange trace_range("kevd_bandwidth_sweep_noalloc");
    KEVD_TRY(sweep_validate(ctx, rows, nrows, eps, neps, prec, S_out));
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    return prec == KEVD_MIXED ? sweep_scratch<float>(ctx, rows, nrows, eps, neps, S_out)
                              : sweep_scratch<double>(ctx, rows, nrows, eps, neps, S_out);
}

// ks_bandwidth.cu:609-642: log of the sums, centred-difference slope over
// log eps, first maximum of the slope, eopt = eps[idmax+1], mdim = 2 slope.
// The reference takes log(2 S_lower + 1) with S_lower the strict lower
// triangle over NM; for full row sums that is log(S / rows), which differs
// from log(S) by a constant and leaves the slope unchanged.
extern "C" kevd_status kevd_bandwidth_post(const double* eps, const double* S, int32_t neps,
                                           double* logS, double* slope, double* eopt, double* mdim)
{
    if (!eps || !S || neps < 3 || !logS || !slope || !eopt || !mdim) return KEVD_EINVAL;
    for (int32_t i = 0; i < neps; ++i) {
        if (!(S[i] > 0.0) || !(eps[i] > 0.0)) return KEVD_EINVAL;
        logS[i] = std::log(S[i]);
    }
    for (int32_t i = 1; i < neps - 1; ++i)
        slope[i - 1] = (logS[i + 1] - logS[i - 1]) / (std::log(eps[i + 1]) - std::log(eps[i - 1]));
    int32_t idmax = 0;
    for (int32_t i = 1; i < neps - 2; ++i)
        if (slope[i] > slope[idmax]) idmax = i;
    *eopt = eps[idmax + 1];
    *mdim = 2.0 * slope[idmax];
    return KEVD_OK;
}
