// ctx.hip -- context lifetime, point upload (centring + norms), the median
// bandwidth estimator and the standalone kernel-block entry point.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "gemm.h"
#include "internal.h"

TraceRange::TraceRange(const char* name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

// The allocation-free entry points keep their intermediates here: a global of
// the code object, so every device the library is loaded on holds its own copy
// at its own address.  Host code never passes this name to a kernel; it uses
// the per-device address resolved below (kevd_ctx::scratch).
static __device__ __align__(256) unsigned char scratch[KEVD_SCRATCH_BYTES];

// Device address of `scratch` on the current device, resolved once per device.
static kevd_status scratch_address(int device, void** out)
{
    static std::mutex mu;
    static std::map<int, void*> per_device;
    std::lock_guard<std::mutex> g(mu);
    auto it = per_device.find(device);
    if (it == per_device.end()) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return KEVD_EHIP;
        it = per_device.emplace(device, p).first;
    }
    *out = it->second;
    return KEVD_OK;
}

// ------------------------------------------------------------ errors, bufs

bool use_bf16x6()
{
    static const bool on = [] {
        const char* e = std::getenv("KEVD_FP32_MFMA");
        return !(e && e[0] == '1');
    }();
    return on;
}

kevd_status kevd_fail(kevd_ctx* ctx, kevd_status s, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return s;
}

kevd_status kevd_hip_fail(kevd_ctx* ctx, hipError_t e, const char* what, const char* file,
                          int line)
{
    std::string m = std::string(what) + " -> " + hipGetErrorName(e) + " (" +
                    hipGetErrorString(e) + ") at " + file + ":" + std::to_string(line);
    return kevd_fail(ctx, e == hipErrorOutOfMemory ? KEVD_ENOMEM : KEVD_EHIP, m);
}

kevd_status ctx_buf(kevd_ctx* ctx, const char* name, size_t bytes, void** out)
{
    DevBuf& b = ctx->bufs[name];
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (b.p) {
            KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
            KEVD_HIP(ctx, hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        KEVD_HIP(ctx, hipMalloc(&b.p, bytes));
        b.bytes = bytes;
    }
    *out = b.p;
    return KEVD_OK;
}

kevd_status ctx_pinned(kevd_ctx* ctx, size_t bytes, void** out)
{
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) KEVD_HIP(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        KEVD_HIP(ctx, hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
        ctx->pinned_bytes = bytes;
    }
    *out = ctx->pinned;
    return KEVD_OK;
}

// -------------------------------------------------------------- profiling

static hipEvent_t pool_event(kevd_ctx* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

ProfScope::ProfScope(kevd_ctx* c, int category) : ctx(c), cat(category)
{
    // a negative category times nothing (a call nested in a timed scope)
    if (!ctx->profiling || cat < 0) return;
    a = pool_event(ctx);
    b = pool_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
}

ProfScope::~ProfScope()
{
    if (!ctx->profiling || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->launches.push_back({cat, a, b});
}

void prof_reset(kevd_ctx* ctx)
{
    for (auto& l : ctx->launches) {
        ctx->event_pool.push_back(l.start);
        ctx->event_pool.push_back(l.stop);
    }
    ctx->launches.clear();
}

void prof_resolve(kevd_ctx* ctx)
{
    if (!ctx->profiling) return;
    double ms[TC_COUNT] = {};
    int64_t cnt[TC_COUNT] = {};
    for (auto& l : ctx->launches) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, l.start, l.stop) == hipSuccess) {
            ms[l.cat] += t;
            cnt[l.cat] += 1;
        }
    }
    ctx->tm.downdate_ms += ms[TC_DOWNDATE];
    ctx->tm.downdate_launches += cnt[TC_DOWNDATE];
    ctx->tm.gram_ms += ms[TC_GRAM];
    ctx->tm.gram_launches += cnt[TC_GRAM];
    ctx->tm.eig_ms += ms[TC_EIG];
    ctx->tm.sample_ms += ms[TC_SAMPLE];
    ctx->tm.rejection_ms += ms[TC_REJECT];
    ctx->tm.matvec_ms += ms[TC_MATVEC];
    ctx->tm.dense_ms += ms[TC_DENSE];
    ctx->tm.u_ms += ms[TC_UGEMM];
    prof_reset(ctx);
}

kevd_status ctx_check_arch(kevd_ctx* ctx)
{
    hipDeviceProp_t prop;
    KEVD_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return kevd_fail(ctx, KEVD_EINVAL,
                         std::string("libkevd is built for gfx950 only; device is ") +
                             prop.gcnArchName);
    ctx->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    return KEVD_OK;
}

// ------------------------------------------------------------------- API

extern "C" const char* kevd_status_string(kevd_status s)
{
    switch (s) {
    case KEVD_OK: return "ok";
    case KEVD_EINVAL: return "invalid argument";
    case KEVD_ENONPOS_DEGREE: return "non-positive degree";
    case KEVD_EPOTRF: return "Cholesky of the shifted Gram failed";
    case KEVD_ENONPOS_Q: return "non-positive q";
    case KEVD_EEIG: return "eigensolver failed";
    case KEVD_EHIP: return "HIP error";
    case KEVD_ENOMEM: return "out of memory";
    }
    return "unknown status";
}

extern "C" kevd_status kevd_create(int device, kevd_ctx** out)
{
    if (!out) return KEVD_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return KEVD_EINVAL;
    kevd_ctx* ctx = new kevd_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return KEVD_EHIP;
    }
    kevd_status s = ctx_check_arch(ctx);
    if (s == KEVD_OK) s = scratch_address(device, &ctx->scratch);
    if (s != KEVD_OK) {
        delete ctx;
        return s;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return KEVD_EHIP;
    }
    *out = ctx;
    return KEVD_OK;
}

void blas_destroy(kevd_ctx* ctx);

extern "C" void kevd_destroy(kevd_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    blas_destroy(ctx);
    prof_reset(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto& kv : ctx->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    void* ptrs[] = {ctx->X32, ctx->X64, ctx->nrm32, ctx->nrm64, ctx->F, ctx->U};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* kevd_last_error(const kevd_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" kevd_status kevd_set_profiling(kevd_ctx* ctx, int enable)
{
    if (!ctx) return KEVD_EINVAL;
    ctx->profiling = enable != 0;
    return KEVD_OK;
}

extern "C" kevd_status kevd_get_timings(const kevd_ctx* ctx, kevd_timings* out)
{
    if (!ctx || !out) return KEVD_EINVAL;
    *out = ctx->tm;
    return KEVD_OK;
}

// --------------------------------------------------------------- points

// Column means in a fixed reduction order (deterministic).
__global__ __launch_bounds__(256) void k_colmean(const double* X64raw, const float* X32raw,
                                                 int64_t n, int32_t dim, double* mean)
{
    int j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256)
        s += X64raw ? X64raw[i * dim + j] : (double)X32raw[i * dim + j];
    __shared__ double part[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) mean[j] = ((part[0] + part[1]) + (part[2] + part[3])) / (double)n;
}

// Centre, pad to dp, write FP64 and FP32 copies and both squared norms.
// One wave per row.
__global__ __launch_bounds__(256) void k_center(const double* X64raw, const float* X32raw,
                                                int64_t n, int32_t dim, int32_t dp,
                                                const double* mean, double* X64, float* X32,
                                                double* nrm64, double* nrm32)
{
    int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s64 = 0.0, s32 = 0.0;
    for (int j = lane; j < dp; j += 64) {
        double v = 0.0;
        if (j < dim) {
            double raw = X64raw ? X64raw[row * dim + j] : (double)X32raw[row * dim + j];
            v = raw - mean[j];
        }
        float f = (float)v;
        X64[row * dp + j] = v;
        X32[row * dp + j] = f;
        s64 += v * v;
        s32 += (double)f * (double)f;
    }
    s64 = wave_sum(s64);
    s32 = wave_sum(s32);
    if (lane == 0) {
        nrm64[row] = s64;
        nrm32[row] = s32;
    }
}

static kevd_status set_points_impl(kevd_ctx* ctx, const double* X64h, const float* X32h,
                                   int64_t n, int32_t dim)
{
    if (!ctx || n < 2 || dim < 1 || dim > 4096 || (!X64h && !X32h))
        return kevd_fail(ctx, KEVD_EINVAL, "set_points: need n >= 2, 1 <= dim <= 4096");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    void* olds[] = {ctx->X32, ctx->X64, ctx->nrm32, ctx->nrm64, ctx->F};
    for (void* p : olds)
        if (p) KEVD_HIP(ctx, hipFree(p));
    ctx->X32 = nullptr; ctx->X64 = nullptr; ctx->nrm32 = nullptr; ctx->nrm64 = nullptr;
    ctx->F = nullptr; ctx->F_bytes = 0;
    ctx->have_factor = false;
    ctx->cs_valid = false;
    ctx->lm = false;
    ctx->rank = 0;
    ctx->pivots.clear();

    int32_t dp = KEVD_ROUND_UP(dim, 32);
    ctx->n = n;
    ctx->dim = dim;
    ctx->dp = dp;
    KEVD_HIP(ctx, hipMalloc(&ctx->X64, (size_t)n * dp * sizeof(double)));
    KEVD_HIP(ctx, hipMalloc(&ctx->X32, (size_t)n * dp * sizeof(float)));
    KEVD_HIP(ctx, hipMalloc(&ctx->nrm64, (size_t)n * sizeof(double)));
    KEVD_HIP(ctx, hipMalloc(&ctx->nrm32, (size_t)n * sizeof(double)));

    size_t esz = X64h ? sizeof(double) : sizeof(float);
    void* raw = nullptr;
    KEVD_TRY(ctx_buf(ctx, "raw_points", (size_t)n * dim * esz, &raw));
    double* mean = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "col_mean", dim, &mean));
    KEVD_HIP(ctx, hipMemcpyAsync(raw, X64h ? (const void*)X64h : (const void*)X32h,
                                 (size_t)n * dim * esz, hipMemcpyHostToDevice, ctx->stream));
    const double* r64 = X64h ? (const double*)raw : nullptr;
    const float* r32 = X64h ? nullptr : (const float*)raw;
    k_colmean<<<dim, 256, 0, ctx->stream>>>(r64, r32, n, dim, mean);
    KEVD_HIP(ctx, hipGetLastError());
    k_center<<<(unsigned)((n + 3) / 4), 256, 0, ctx->stream>>>(r64, r32, n, dim, dp, mean,
                                                               ctx->X64, ctx->X32, ctx->nrm64,
                                                               ctx->nrm32);
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

extern "C" kevd_status kevd_set_points(kevd_ctx* ctx, const double* X, int64_t n, int32_t dim)
{
    TraceRange trace_range("kevd_set_points");
    return set_points_impl(ctx, X, nullptr, n, dim);
}

extern "C" kevd_status kevd_set_points_f32(kevd_ctx* ctx, const float* X, int64_t n,
                                           int32_t dim)
{
    TraceRange trace_range("kevd_set_points_f32");
    return set_points_impl(ctx, nullptr, X, n, dim);
}

// --------------------------------------------------------------- median

// Squared distance of pair t: either the t-th pair (i > j) of the strict
// lower triangle, or a Philox pair with i != j.  Direct differences, FP64.
__global__ __launch_bounds__(256) void k_pair_sqdist(const double* X64, int64_t n, int32_t dp,
                                                     int64_t m, int exact, uint64_t seed,
                                                     double* out)
{
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    int64_t i, j;
    if (exact) {
        // largest i with i(i-1)/2 <= t
        i = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
        while (i * (i - 1) / 2 > t) --i;
        while ((i + 1) * i / 2 <= t) ++i;
        j = t - i * (i - 1) / 2;
    } else {
        uint32_t o[4];
        kevd_philox((uint32_t)t, 0u, KEVD_STREAM_MEDIAN, 0u, (uint32_t)seed,
                    (uint32_t)(seed >> 32), o);
        i = (int64_t)(((uint64_t)o[0] * (uint64_t)n) >> 32);
        j = (int64_t)(((uint64_t)o[1] * (uint64_t)(n - 1)) >> 32);
        if (j >= i) ++j;
    }
    const double* a = X64 + i * dp;
    const double* b = X64 + j * dp;
    double s = 0.0;
    for (int k = 0; k < dp; ++k) {
        double d = a[k] - b[k];
        s += d * d;
    }
    out[t] = s;
}

extern "C" kevd_status kevd_median_sqdist(kevd_ctx* ctx, uint64_t seed, int64_t nsample,
                                          double* eps_out)
{
    TraceRange trace_range("kevd_median_sqdist");
    if (!ctx || !eps_out || ctx->n < 2) return kevd_fail(ctx, KEVD_EINVAL, "median: no points");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    int64_t n = ctx->n;
    int64_t npairs = n * (n - 1) / 2;
    int exact = (nsample <= 0 || nsample >= npairs);
    if (exact && n > 16384)
        return kevd_fail(ctx, KEVD_EINVAL, "median: exact mode needs n <= 16384; give nsample");
    if (!exact && n >= ((int64_t)1 << 32))
        return kevd_fail(ctx, KEVD_EINVAL, "median: sampled mode needs n < 2^32");
    int64_t m = exact ? npairs : nsample;
    double* dv = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "median_d", m, &dv));
    k_pair_sqdist<<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(ctx->X64, n, ctx->dp,
                                                                       m, exact, seed, dv);
    KEVD_HIP(ctx, hipGetLastError());
    std::vector<double> h((size_t)m);
    KEVD_HIP(ctx, hipMemcpyAsync(h.data(), dv, (size_t)m * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t mid = (size_t)(m / 2);
    std::nth_element(h.begin(), h.begin() + mid, h.end());
    double hi = h[mid];
    if (m & 1) {
        *eps_out = hi;
    } else {
        double lo = *std::max_element(h.begin(), h.begin() + mid);
        *eps_out = 0.5 * (lo + hi);
    }
    return KEVD_OK;
}

// ---------------------------------------------------------- kernel block

// Gather rows idx of the centred points (FP32 or FP64 copy) and their norms.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_rows(const T* X, const double* nrm, int32_t dp,
                                                     const int64_t* idx, int64_t count,
                                                     T* out, double* nrm_out)
{
    int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (r >= count) return;
    int64_t src = idx[r];
    for (int j = lane; j < dp; j += 64) out[r * dp + j] = X[src * dp + j];
    if (lane == 0) nrm_out[r] = nrm[src];
}

kevd_status kblock_gather(kevd_ctx* ctx, const int64_t* idx_dev, int64_t count,
                          kevd_precision prec, void* Xout, double* nrm_out)
{
    if (count <= 0) return KEVD_OK;
    unsigned g = (unsigned)((count + 3) / 4);
    if (prec == KEVD_MIXED)
        k_gather_rows<float><<<g, 256, 0, ctx->stream>>>(ctx->X32, ctx->nrm32, ctx->dp, idx_dev,
                                                         count, (float*)Xout, nrm_out);
    else
        k_gather_rows<double><<<g, 256, 0, ctx->stream>>>(ctx->X64, ctx->nrm64, ctx->dp, idx_dev,
                                                          count, (double*)Xout, nrm_out);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

// K(rows, cols) = exp(-(|x|^2 + |y|^2 - 2 x.y) / eps) with the x.y product
// on MFMA (dist_exp): one BM x BN output tile per block.  The diagonal
// (same global index) is exactly 1.
template <typename T>
__global__ __launch_bounds__(256) void k_kblock(const T* XR, const double* nR, const int64_t* ridx,
                                                int64_t nr, const T* XC, const double* nC,
                                                const int64_t* cidx, int64_t nc, int32_t dp,
                                                GaussScale gs, double* out)
{
    typedef typename Cfg<T>::G G;
    __shared__ __attribute__((aligned(16))) T lds[G::LDS_ELEMS];
    G g;
    g.init();
    int64_t r0 = (int64_t)blockIdx.x * Cfg<T>::BM;
    int64_t c0 = (int64_t)blockIdx.y * Cfg<T>::BN;
    g.mainloop(lds, XR + r0 * dp, dp, nr - r0, XC + c0 * dp, dp, nc - c0, dp);
    constexpr int TM = G::NTM, TN = G::NTN;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int v = 0; v < G::M::ACC; ++v) {
                int64_t row = r0 + g.row_of(a, v);
                int64_t col = c0 + g.col_of(b);
                if (row < nr && col < nc) {
                    double D = nR[row] + nC[col] - 2.0 * (double)g.acc[a][b][v];
                    D = D > 0.0 ? D : 0.0;
                    double K = (ridx[row] == cidx[col]) ? 1.0 : (double)gauss_val(D, gs, (T*)nullptr);
                    out[row * nc + col] = K;
                }
            }
}

extern "C" kevd_status kevd_kernel_block(kevd_ctx* ctx, const int64_t* rows, int64_t nr,
                                         const int64_t* cols, int64_t nc, double eps,
                                         kevd_precision prec, double* K_out)
{
    TraceRange trace_range("kevd_kernel_block");
    if (!ctx || ctx->n < 1) return kevd_fail(ctx, KEVD_EINVAL, "kernel_block: no points");
    if (nr < 0 || nc < 0 || !(eps > 0.0) || (nr && !rows) || (nc && !cols) ||
        ((nr && nc) && !K_out))
        return kevd_fail(ctx, KEVD_EINVAL, "kernel_block: bad arguments");
    if (nr == 0 || nc == 0) return KEVD_OK;
    for (int64_t i = 0; i < nr; ++i)
        if (rows[i] < 0 || rows[i] >= ctx->n)
            return kevd_fail(ctx, KEVD_EINVAL, "kernel_block: row index out of range");
    for (int64_t i = 0; i < nc; ++i)
        if (cols[i] < 0 || cols[i] >= ctx->n)
            return kevd_fail(ctx, KEVD_EINVAL, "kernel_block: column index out of range");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    size_t esz = prec == KEVD_MIXED ? sizeof(float) : sizeof(double);
    int64_t *ri = nullptr, *ci = nullptr;
    void *XR = nullptr, *XC = nullptr;
    double *nR = nullptr, *nC = nullptr, *out = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "kb_ri", nr, &ri));
    KEVD_TRY(ctx_buf_t(ctx, "kb_ci", nc, &ci));
    KEVD_TRY(ctx_buf(ctx, "kb_XR", (size_t)nr * ctx->dp * esz, &XR));
    KEVD_TRY(ctx_buf(ctx, "kb_XC", (size_t)nc * ctx->dp * esz, &XC));
    KEVD_TRY(ctx_buf_t(ctx, "kb_nR", nr, &nR));
    KEVD_TRY(ctx_buf_t(ctx, "kb_nC", nc, &nC));
    KEVD_TRY(ctx_buf_t(ctx, "kb_out", (size_t)nr * nc, &out));
    KEVD_HIP(ctx, hipMemcpyAsync(ri, rows, nr * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(ci, cols, nc * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KEVD_TRY(kblock_gather(ctx, ri, nr, prec, XR, nR));
    KEVD_TRY(kblock_gather(ctx, ci, nc, prec, XC, nC));
    GaussScale gs{-1.0 / eps, -1.4426950408889634074 / eps};
    if (prec == KEVD_MIXED) {
        dim3 grid((unsigned)((nr + Cfg<float>::BM - 1) / Cfg<float>::BM),
                  (unsigned)((nc + Cfg<float>::BN - 1) / Cfg<float>::BN));
        k_kblock<float><<<grid, 256, 0, ctx->stream>>>((const float*)XR, nR, ri, nr,
                                                       (const float*)XC, nC, ci, nc, ctx->dp, gs,
                                                       out);
    } else {
        dim3 grid((unsigned)((nr + Cfg<double>::BM - 1) / Cfg<double>::BM),
                  (unsigned)((nc + Cfg<double>::BN - 1) / Cfg<double>::BN));
        k_kblock<double><<<grid, 256, 0, ctx->stream>>>((const double*)XR, nR, ri, nr,
                                                        (const double*)XC, nC, ci, nc, ctx->dp,
                                                        gs, out);
    }
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipMemcpyAsync(K_out, out, (size_t)nr * nc * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}
