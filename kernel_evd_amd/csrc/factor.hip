// factor.hip -- moving the factor F across the boundary, the reference's
// L^-1 export, and the rocBLAS/rocSOLVER handle used by the r x r dense
// phase.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "internal.h"

rocblas_handle blas_handle(kevd_ctx* ctx)
{
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
        rocblas_set_stream(h, ctx->stream);
        ctx->blas = h;
    }
    return (rocblas_handle)ctx->blas;
}

void blas_destroy(kevd_ctx* ctx)
{
    if (ctx->blas) rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert_rows(const S* src, int64_t lds, D* dst, int64_t ldd,
                                                      int64_t n, int64_t cols)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t total = n * cols;
    if (i >= total) return;
    int64_t r = i / cols, c = i % cols;
    dst[r * ldd + c] = (D)src[r * lds + c];
}

template <typename S, typename D>
static kevd_status convert(kevd_ctx* ctx, const S* src, int64_t lds, D* dst, int64_t ldd, int64_t n,
                           int64_t cols)
{
    int64_t total = n * cols;
    if (total == 0) return KEVD_OK;
    k_convert_rows<S, D><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(src, lds, dst, ldd,
                                                                                 n, cols);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

extern "C" kevd_status kevd_get_factor(kevd_ctx* ctx, void* F_host, int64_t ld, kevd_precision as)
{
    if (!ctx || !F_host || !ctx->have_factor || ld < ctx->rank)
        return kevd_fail(ctx, KEVD_EINVAL, "get_factor: no factor or ld < rank");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    size_t esz = as == KEVD_MIXED ? sizeof(float) : sizeof(double);
    size_t fsz = ctx->fprec == KEVD_MIXED ? sizeof(float) : sizeof(double);
    const void* src = ctx->F;
    int64_t lsrc = ctx->ldf;
    if (as != ctx->fprec) {
        void* tmp = nullptr;
        KEVD_TRY(ctx_buf(ctx, "f_conv", (size_t)ctx->n * ctx->rank * esz, &tmp));
        if (as == KEVD_MIXED)
            KEVD_TRY(convert(ctx, (const double*)ctx->F, ctx->ldf, (float*)tmp, ctx->rank, ctx->n, ctx->rank));
        else
            KEVD_TRY(convert(ctx, (const float*)ctx->F, ctx->ldf, (double*)tmp, ctx->rank, ctx->n, ctx->rank));
        src = tmp;
        lsrc = ctx->rank;
        fsz = esz;
    }
    KEVD_HIP(ctx, hipMemcpy2DAsync(F_host, ld * esz, src, lsrc * fsz, ctx->rank * esz, ctx->n,
                                   hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KEVD_OK;
}

extern "C" kevd_status kevd_set_factor(kevd_ctx* ctx, const void* F_host, int64_t n, int64_t r,
                                       int64_t ld, kevd_precision prec)
{
    if (!ctx || !F_host || n != ctx->n || r < 1 || r > n || ld < r ||
        (prec != KEVD_FP64 && prec != KEVD_MIXED))
        return kevd_fail(ctx, KEVD_EINVAL, "set_factor: need n == points, 1 <= r <= n, ld >= r");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t esz = prec == KEVD_MIXED ? sizeof(float) : sizeof(double);
    int64_t ldf = kevd_factor_ld(r);
    size_t bytes = (size_t)n * ldf * esz;
    if (ctx->F && ctx->F_bytes < bytes) {
        KEVD_HIP(ctx, hipFree(ctx->F));
        ctx->F = nullptr;
        ctx->F_bytes = 0;
    }
    if (!ctx->F) {
        KEVD_HIP(ctx, hipMalloc(&ctx->F, bytes));
        ctx->F_bytes = bytes;
    }
    KEVD_HIP(ctx, hipMemsetAsync(ctx->F, 0, bytes, ctx->stream));
    KEVD_HIP(ctx, hipMemcpy2DAsync(ctx->F, ldf * esz, F_host, ld * esz, r * esz, n,
                                   hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->ldf = ldf;
    ctx->rank = r;
    ctx->fprec = prec;
    ctx->have_factor = true;
    ctx->cs_valid = false;  // F^T 1 partials belong to kevd_arpc's factors only
    ctx->lm = false;
    ctx->pivots.clear();
    return KEVD_OK;
}

// L(a, c) = F(S_a, c) for c <= a, column-major r x r in FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_L(const T* F, int64_t ldf, const int64_t* piv,
                                                  int64_t r, double* L)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= r * r) return;
    int64_t c = i / r, a = i % r;
    L[a + c * r] = c <= a ? (double)F[piv[a] * ldf + c] : 0.0;
}

// Column-major r x r copy of the LM form's L^-1 (row-major, ld ldl).
__global__ __launch_bounds__(256) void k_linv_colmajor(const double* Linv, int64_t ldl, int64_t r, double* out)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= r * r) return;
    int64_t c = i / r, a = i % r;
    out[a + c * r] = c <= a ? Linv[a * ldl + c] : 0.0;
}

// K(S,S) = L L^T for the pivots S of the held factor: L lower triangular,
// column-major r x r FP64 with the strict upper triangle zero.  The F form
// holds L as its pivot rows F(S,:); the LM form holds L^-1, inverted here.
kevd_status pivot_cholesky(kevd_ctx* ctx, double* L)
{
    int64_t r = ctx->rank;
    bool lm = ctx->lm && ctx->lm_linv;
    if (r <= 0 || ctx->pivots.size() != (size_t)r || (!lm && !ctx->have_factor))
        return kevd_fail(ctx, KEVD_EINVAL, "the pivot Cholesky factor needs a factor built by kevd_arpc");
    unsigned g = (unsigned)((r * r + 255) / 256);
    if (lm) {
        int* info = nullptr;
        KEVD_TRY(ctx_buf_t(ctx, "l_info", 1, &info));
        k_linv_colmajor<<<g, 256, 0, ctx->stream>>>(ctx->lm_linv, ctx->lm_ldl, r, L);
        KEVD_HIP(ctx, hipGetLastError());
        rocblas_handle h = blas_handle(ctx);
        if (!h) return kevd_fail(ctx, KEVD_EHIP, "rocblas handle creation failed");
        if (rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)r, L,
                             (rocblas_int)r, info) != rocblas_status_success)
            return kevd_fail(ctx, KEVD_EHIP, "rocsolver_dtrtri failed");
        int hinfo = 0;
        KEVD_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hinfo != 0) return kevd_fail(ctx, KEVD_EINVAL, "the kept L^-1 is singular");
        return KEVD_OK;
    }
    int64_t* piv = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "l_piv", r, &piv));
    KEVD_HIP(ctx, hipMemcpyAsync(piv, ctx->pivots.data(), r * sizeof(int64_t), hipMemcpyHostToDevice,
                                 ctx->stream));
    if (ctx->fprec == KEVD_MIXED)
        k_gather_L<float><<<g, 256, 0, ctx->stream>>>((const float*)ctx->F, ctx->ldf, piv, r, L);
    else
        k_gather_L<double><<<g, 256, 0, ctx->stream>>>((const double*)ctx->F, ctx->ldf, piv, r, L);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

extern "C" kevd_status kevd_get_linv(kevd_ctx* ctx, double* Linv_host)
{
    if (ctx && Linv_host && ctx->lm && ctx->lm_linv && ctx->pivots.size() == (size_t)ctx->rank && ctx->rank > 0) {
        // the LM form kept L^-1 itself
        KEVD_HIP(ctx, hipSetDevice(ctx->device));
        int64_t r = ctx->rank;
        double* out = nullptr;
        KEVD_TRY(ctx_buf_t(ctx, "l_L", (size_t)r * r, &out));
        k_linv_colmajor<<<(unsigned)((r * r + 255) / 256), 256, 0, ctx->stream>>>(ctx->lm_linv, ctx->lm_ldl, r, out);
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_HIP(ctx, hipMemcpyAsync(Linv_host, out, (size_t)r * r * sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream));
        KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return KEVD_OK;
    }
    if (!ctx || !Linv_host || !ctx->have_factor || ctx->pivots.size() != (size_t)ctx->rank)
        return kevd_fail(ctx, KEVD_EINVAL, "get_linv: needs a factor built by kevd_arpc");
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    int64_t r = ctx->rank;
    int64_t* piv = nullptr;
    double* L = nullptr;
    int* info = nullptr;
    KEVD_TRY(ctx_buf_t(ctx, "l_piv", r, &piv));
    KEVD_TRY(ctx_buf_t(ctx, "l_L", (size_t)r * r, &L));
    KEVD_TRY(ctx_buf_t(ctx, "l_info", 1, &info));
    KEVD_HIP(ctx, hipMemcpyAsync(piv, ctx->pivots.data(), r * sizeof(int64_t), hipMemcpyHostToDevice,
                                 ctx->stream));
    unsigned g = (unsigned)((r * r + 255) / 256);
    if (ctx->fprec == KEVD_MIXED)
        k_gather_L<float><<<g, 256, 0, ctx->stream>>>((const float*)ctx->F, ctx->ldf, piv, r, L);
    else
        k_gather_L<double><<<g, 256, 0, ctx->stream>>>((const double*)ctx->F, ctx->ldf, piv, r, L);
    KEVD_HIP(ctx, hipGetLastError());
    rocblas_handle h = blas_handle(ctx);
    if (!h) return kevd_fail(ctx, KEVD_EHIP, "rocblas handle creation failed");
    if (rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)r, L,
                         (rocblas_int)r, info) != rocblas_status_success)
        return kevd_fail(ctx, KEVD_EHIP, "rocsolver_dtrtri failed");
    int hinfo = 0;
    KEVD_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KEVD_HIP(ctx, hipMemcpyAsync(Linv_host, L, (size_t)r * r * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hinfo != 0) return kevd_fail(ctx, KEVD_EINVAL, "get_linv: F(S,:) is singular");
    return KEVD_OK;
}
