// bandwidth.hip -- fused Gaussian-kernel bandwidth sweep.
//
// S(eps_j) = sum_{i in rows} sum_{c < n} exp(-||x_i - x_c||^2 / eps_j) for all
// eps_j from one launch.  The reference materialises the distance matrix in
// HBM (dmat_comp, ks_bandwidth.cu:181-232) and re-reads it once per
// bandwidth (kernelsum_comp 276-325, launched Nbw times at 566-572); here a
// block computes a 128 x 128 distance tile on MFMA (|x|^2 + |y|^2 - 2 x.y,
// FP64 norms), keeps it in registers and evaluates every eps_j against it.
// Per (tile, 16 eps) the lane partials are summed across the wave and added
// into FP64 accumulators; blocks write FP64 partials reduced in fixed order.
#include <cmath>
#include <vector>

#include "gemm.h"
#include "internal.h"

namespace {

constexpr int NEPS_MAX = 128;
constexpr int ECH = 16;  // eps values per register chunk

template <typename T>
__device__ inline T kexp(T d, T c);
template <>
__device__ inline float kexp<float>(float d, float c)
{
    // c = -log2(e) / eps; v_exp_f32 computes 2^x
    return __builtin_amdgcn_exp2f(d * c);
}
template <>
__device__ inline double kexp<double>(double d, double c)
{
    // c = -1 / eps
    return exp(d * c);
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
                for (int v = 0; v < G::M::ACC; ++v) {
                    int64_t row = r0 + g.row_of(a, v);
                    int64_t col = c0 + g.col_of(b);
                    T dv = (T)INFINITY;  // exp(-inf) = 0 for padding
                    if (row < nrows && col < cend) {
                        double Dd = nR[row] + nrm[col] - 2.0 * (double)g.acc[a][b][v];
                        Dd = Dd > 0.0 ? Dd : 0.0;
                        if (ridx[row] == col) Dd = 0.0;
                        dv = (T)Dd;
                    }
                    D[(a * G::NTN + b) * G::M::ACC + v] = dv;
                }
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
                for (int j = 0; j < ECH; ++j) s[j] += kexp<T>(d, c[j]);
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

__global__ void k_bw_reduce(const double* part, int nblk, double* S)
{
    int j = threadIdx.x;
    if (j >= NEPS_MAX) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * NEPS_MAX + j];
    S[j] = s;
}

template <typename T>
kevd_status sweep(kevd_ctx* ctx, const int64_t* ri, int64_t nrows, const double* eps, int neps,
                  double* S_out)
{
    int64_t n = ctx->n;
    int32_t dp = ctx->dp;
    constexpr int BM = Cfg<T>::BM, BN = Cfg<T>::BN;
    bool f32 = sizeof(T) == 4;
    T* XR;
    double* nR;
    T* coef;
    double *part, *S;
    KEVD_TRY(ctx_buf_t(ctx, "bw_XR", (size_t)nrows * dp, &XR));
    KEVD_TRY(ctx_buf_t(ctx, "bw_nR", nrows, &nR));
    KEVD_TRY(kblock_gather(ctx, ri, nrows, f32 ? KEVD_MIXED : KEVD_FP64, XR, nR));
    int nep = (neps + ECH - 1) / ECH * ECH;
    std::vector<T> hc((size_t)nep, (T)0);
    for (int j = 0; j < neps; ++j)
        hc[j] = f32 ? (T)(-1.4426950408889634074 / eps[j]) : (T)(-1.0 / eps[j]);
    KEVD_TRY(ctx_buf_t(ctx, "bw_coef", nep, &coef));
    KEVD_HIP(ctx, hipMemcpyAsync(coef, hc.data(), nep * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    int64_t rt = (nrows + BM - 1) / BM;
    int64_t ctiles = (n + BN - 1) / BN;
    int64_t want = 4096 / (rt > 0 ? rt : 1);
    if (want < 1) want = 1;
    if (want > ctiles) want = ctiles;
    int64_t cpb = ((ctiles + want - 1) / want) * BN;
    int64_t ncs = (n + cpb - 1) / cpb;
    int nblk = (int)(rt * ncs);
    KEVD_TRY(ctx_buf_t(ctx, "bw_part", (size_t)nblk * NEPS_MAX, &part));
    KEVD_TRY(ctx_buf_t(ctx, "bw_S", NEPS_MAX, &S));
    const T* X = f32 ? (const T*)ctx->X32 : (const T*)ctx->X64;
    const double* nrm = f32 ? ctx->nrm32 : ctx->nrm64;
    k_bw_sweep<T><<<dim3((unsigned)rt, (unsigned)ncs), 256, 0, ctx->stream>>>(XR, nR, ri, nrows, X, nrm, n, dp,
                                                                            coef, nep, cpb, part);
    k_bw_reduce<<<1, NEPS_MAX, 0, ctx->stream>>>(part, nblk, S);
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipMemcpyAsync(S_out, S, neps * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

}  // namespace

extern "C" kevd_status kevd_bandwidth_sweep(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                            const double* eps, int32_t neps, kevd_precision prec,
                                            double* S_out)
{
    if (!ctx || ctx->n < 1) return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: no points");
    if (!rows || nrows < 1 || !eps || neps < 1 || neps > NEPS_MAX || !S_out)
        return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: need rows, 1 <= neps <= 128 and outputs");
    for (int32_t j = 0; j < neps; ++j)
        if (!(eps[j] > 0.0) || !std::isfinite(eps[j]))
            return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: eps values must be positive and finite");
    for (int64_t i = 0; i < nrows; ++i)
        if (rows[i] < 0 || rows[i] >= ctx->n)
            return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: row index out of range");
    if (prec != KEVD_FP64 && prec != KEVD_MIXED) return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: precision");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    int64_t* ri;
    KEVD_TRY(ctx_buf_t(ctx, "bw_ri", nrows, &ri));
    KEVD_HIP(ctx, hipMemcpyAsync(ri, rows, nrows * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    return prec == KEVD_MIXED ? sweep<float>(ctx, ri, nrows, eps, neps, S_out)
                              : sweep<double>(ctx, ri, nrows, eps, neps, S_out);
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
