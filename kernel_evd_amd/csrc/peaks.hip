// peaks.hip -- box-measured throughput ceilings for the roofline fractions.
//
// The spec-sheet peaks (FP32 matrix 157.3 TFLOP/s, FP64 matrix 78.6, HBM
// 8 TB/s) are what a part can issue on paper; each probe here runs the one
// instruction or access pattern a hot kernel is bound by, across every CU,
// and reports what this device sustains.  The reference sizes its launches
// from the occupancy API (ks_bandwidth.cu:436-474) and reports no ceilings;
// this has no counterpart there.
//
//   MFMA      waves issue chains of independent MFMAs on register operands
//             (no memory traffic); 2 M N K flop per instruction: FP32,
//             FP64 and the bf16 forms the bf16x6 GEMMs use.
//   exp       16 independent v_exp_f32 recurrences per lane, x <- 2^-x (the
//             negation is a free source modifier, the value stays in (0, 1]).
//   HBM copy  a 16-byte-per-lane grid-stride copy of a buffer far larger
//             than the L2 / MALL; read + written bytes per second.
//   HBM read  the same grid-stride pattern reading only (an XOR of the
//             words per lane is written once), the ceiling of the
//             read-dominated streaming passes.
#include "bf16x6.h"
#include "internal.h"

namespace {

constexpr int PK_THREADS = 256;
constexpr int PK_CHAINS = 16;

__global__ __launch_bounds__(PK_THREADS) void k_peak_f32_16(int iters, float* out)
{
    floatx4 acc[PK_CHAINS];
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    float a = 1e-3f * (float)(threadIdx.x & 63), b = 1e-3f;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < PK_CHAINS; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_f32_32(int iters, float* out)
{
    constexpr int CH = PK_CHAINS / 4;  // 16 accumulators each: 64 VGPRs
    floatx16 acc[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[j][v] = 0.f;
    float a = 1e-3f * (float)(threadIdx.x & 63), b = 1e-3f;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < CH; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[j], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < CH; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) s += acc[j][v];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

// bf16 operands as the bf16x6 GEMMs issue them (v_mfma_f32_16x16x32_bf16 in
// the downdate, v_mfma_f32_32x32x16_bf16 in the Gram)
__global__ __launch_bounds__(PK_THREADS) void k_peak_bf16_16(int iters, float* out)
{
    floatx4 acc[PK_CHAINS];
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        a[e] = (__bf16)(1e-3f * (float)((threadIdx.x + e) & 63));
        b[e] = (__bf16)1e-3f;
    }
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < PK_CHAINS; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_bf16_32(int iters, float* out)
{
    constexpr int CH = PK_CHAINS / 4;
    floatx16 acc[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[j][v] = 0.f;
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        a[e] = (__bf16)(1e-3f * (float)((threadIdx.x + e) & 63));
        b[e] = (__bf16)1e-3f;
    }
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < CH; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < CH; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) s += acc[j][v];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_f64(int iters, float* out)
{
    doublex4 acc[PK_CHAINS / 2];
#pragma unroll
    for (int j = 0; j < PK_CHAINS / 2; ++j) acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
    double a = 1e-3 * (double)(threadIdx.x & 63), b = 1e-3;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < PK_CHAINS / 2; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < PK_CHAINS / 2; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = (float)s;
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_exp(int iters, float* out)
{
    float v[PK_CHAINS];
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) v[j] = 0.05f * (float)j + 1e-4f * (float)(threadIdx.x & 63);
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int j = 0; j < PK_CHAINS; ++j) v[j] = __builtin_amdgcn_exp2f(-v[j]);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < PK_CHAINS; ++j) s += v[j];
    out[blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          int64_t n16)
{
    int64_t stride = (int64_t)gridDim.x * PK_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * PK_THREADS + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(PK_THREADS) void k_peak_read(const uint4* __restrict__ src, int64_t n16, uint32_t* out)
{
    int64_t stride = (int64_t)gridDim.x * PK_THREADS;
    uint32_t x = 0;
    for (int64_t i = (int64_t)blockIdx.x * PK_THREADS + threadIdx.x; i < n16; i += stride) {
        uint4 v = src[i];
        x ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    out[blockIdx.x * PK_THREADS + threadIdx.x] = x;
}

__global__ void k_peak_fill(uint4* p, int64_t n16)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = uint4{(uint32_t)i, 1u, 2u, 3u};
}

// Best of `reps` timed launches (after one untimed) of `launch` on the
// context stream, in seconds.
template <typename F>
kevd_status best_time(kevd_ctx* ctx, int reps, F launch, double* sec)
{
    hipEvent_t a, b;
    KEVD_HIP(ctx, hipEventCreate(&a));
    KEVD_HIP(ctx, hipEventCreate(&b));
    launch();
    KEVD_HIP(ctx, hipGetLastError());
    double best = 1e30;
    kevd_status st = KEVD_OK;
    for (int r = 0; r < reps && st == KEVD_OK; ++r) {
        float ms = 0.f;
        if (hipEventRecord(a, ctx->stream) != hipSuccess) st = KEVD_EHIP;
        launch();
        if (st == KEVD_OK && hipEventRecord(b, ctx->stream) != hipSuccess) st = KEVD_EHIP;
        if (st == KEVD_OK && hipEventSynchronize(b) != hipSuccess) st = KEVD_EHIP;
        if (st == KEVD_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) st = KEVD_EHIP;
        if (st == KEVD_OK && ms > 0.f && ms * 1e-3 < best) best = ms * 1e-3;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != KEVD_OK) return kevd_fail(ctx, st, "peaks: event timing failed");
    KEVD_HIP(ctx, hipGetLastError());
    *sec = best;
    return KEVD_OK;
}

}  // namespace

extern "C" kevd_status kevd_measure_peaks(kevd_ctx* ctx, int64_t copy_bytes, kevd_peaks* out)
{
    if (!ctx || !out) return kevd_fail(ctx, KEVD_EINVAL, "peaks: need a context and an output");
    if (copy_bytes < (int64_t)(16 << 20) || copy_bytes > ((int64_t)64 << 30))
        return kevd_fail(ctx, KEVD_EINVAL, "peaks: copy_bytes must lie in [16 MiB, 64 GiB]");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    KEVD_TRY(ctx_check_arch(ctx));
    *out = kevd_peaks{};
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const int nblk = 8 * ctx->num_cu;
    const double waves = 4.0 * nblk;
    float* sink;
    KEVD_TRY(ctx_buf_t(ctx, "peak_sink", (size_t)nblk * PK_THREADS, &sink));
    hipStream_t s = ctx->stream;
    double t;
    const int it32 = 6000, it64 = 3000, itexp = 20000;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_f32_16<<<nblk, PK_THREADS, 0, s>>>(it32, sink); }, &t));
    out->mfma_f32_tflops = waves * it32 * PK_CHAINS * (2.0 * 16 * 16 * 4) / t * 1e-12;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_f32_32<<<nblk, PK_THREADS, 0, s>>>(it32, sink); }, &t));
    out->mfma_f32_32_tflops = waves * it32 * (PK_CHAINS / 4) * (2.0 * 32 * 32 * 2) / t * 1e-12;
    const int itb = 12000;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_bf16_16<<<nblk, PK_THREADS, 0, s>>>(itb, sink); }, &t));
    out->mfma_bf16_tflops = waves * itb * PK_CHAINS * (2.0 * 16 * 16 * 32) / t * 1e-12;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_bf16_32<<<nblk, PK_THREADS, 0, s>>>(itb, sink); }, &t));
    out->mfma_bf16_32_tflops = waves * itb * (PK_CHAINS / 4) * (2.0 * 32 * 32 * 16) / t * 1e-12;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_f64<<<nblk, PK_THREADS, 0, s>>>(it64, sink); }, &t));
    out->mfma_f64_tflops = waves * it64 * (PK_CHAINS / 2) * (2.0 * 16 * 16 * 4) / t * 1e-12;
    KEVD_TRY(best_time(ctx, 3, [&] { k_peak_exp<<<nblk, PK_THREADS, 0, s>>>(itexp, sink); }, &t));
    out->exp_f32_per_s = waves * 64.0 * itexp * PK_CHAINS / t;

    int64_t n16 = copy_bytes / 16;
    void *src = nullptr, *dst = nullptr;
    KEVD_HIP(ctx, hipMalloc(&src, (size_t)n16 * 16));
    if (hipMalloc(&dst, (size_t)n16 * 16) != hipSuccess) {
        (void)hipFree(src);
        return kevd_fail(ctx, KEVD_ENOMEM, "peaks: cannot allocate the copy buffers");
    }
    const int cblk = 4 * ctx->num_cu;
    k_peak_fill<<<cblk, 256, 0, s>>>((uint4*)src, n16);
    kevd_status st = best_time(ctx, 5, [&] {
        k_peak_copy<<<cblk, PK_THREADS, 0, s>>>((const uint4*)src, (uint4*)dst, n16);
    }, &t);
    double tr = 0.0;
    if (st == KEVD_OK) {
        out->hbm_copy_gbs = 2.0 * (double)n16 * 16.0 / t * 1e-9;
        // sink: cblk * 256 words, inside the peak_sink buffer (nblk >= cblk)
        st = best_time(ctx, 5, [&] {
            k_peak_read<<<cblk, PK_THREADS, 0, s>>>((const uint4*)src, n16, (uint32_t*)sink);
        }, &tr);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (st != KEVD_OK) return st;
    out->hbm_read_gbs = (double)n16 * 16.0 / tr * 1e-9;
    return KEVD_OK;
}
