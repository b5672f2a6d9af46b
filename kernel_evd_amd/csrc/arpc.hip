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

#include "arpc_kernels.h"

namespace {

using namespace arpck;

// Single-problem kernels: thin wrappers over the shared round bodies.
__global__ __launch_bounds__(256) void k_reduce_d(const double* __restrict__ d, int64_t n, double* pmax,
                                                  double* psum, const ArpcState* st, int force)
{
    reduce_d_body(d, n, pmax, psum, st, force, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_weights_scan(const double* __restrict__ d, int64_t n,
                                                      const double* __restrict__ pmax, int nb, int shift_base,
                                                      uint64_t* prefix, uint64_t* btot, const ArpcState* st)
{
    weights_scan_body(d, n, pmax, nb, shift_base, prefix, btot, st, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_draw(const double* __restrict__ psum, const uint64_t* __restrict__ btot,
                                              int nb, const uint64_t* __restrict__ prefix, int64_t n, double trK,
                                              uint64_t seed, int b, int64_t* Sprime, ArpcState* st, double* err,
                                              const int64_t* inject, int64_t n_inject)
{
    draw_body(psum, btot, nb, prefix, n, trK, seed, b, Sprime, st, err, inject, n_inject);
}

template <typename T>
__global__ __launch_bounds__(256) void k_form_H(const double* __restrict__ X64, int32_t dp, const T* __restrict__ F,
                                                int64_t ldf, const int64_t* __restrict__ Sprime,
                                                const ArpcState* st, double neg_inv_eps, double* H)
{
    form_H_body<T>(X64, dp, F, ldf, Sprime, st, neg_inv_eps, H, blockIdx.x, blockIdx.y);
}

template <typename T>
__global__ __launch_bounds__(256) void k_rejchol(const double* __restrict__ H, ArpcState* st, uint64_t seed,
                                                 double tol, int inject, const int64_t* __restrict__ Sprime,
                                                 int64_t* Si, T* LinvT, double* Linv64)
{
    rejchol_body<T>(H, st, seed, tol, inject, Sprime, Si, LinvT, Linv64);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_pivots(const T* __restrict__ Xsrc, const double* __restrict__ nrm,
                                                       int32_t dp, const T* __restrict__ F, int64_t ldf,
                                                       const int64_t* __restrict__ Si, const ArpcState* st,
                                                       T* XS, double* nrmS, T* FS)
{
    gather_pivots_body<T>(Xsrc, nrm, dp, F, ldf, Si, st, XS, nrmS, FS, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_downdate(
    const T* __restrict__ X, int32_t dp, const double* __restrict__ nrm, T* __restrict__ F, int64_t ldf,
    double* __restrict__ d, int64_t n, const T* __restrict__ XS, const double* __restrict__ nrmS,
    const int64_t* __restrict__ Si, const T* __restrict__ FS, const T* __restrict__ LinvT, const ArpcState* st,
    GaussScale gs)
{
    typedef DowndateLds<T> DL;
    __shared__ __attribute__((aligned(16))) T lds[DL::ELEMS];
    __shared__ double rs[DL::WN][DL::BM];
    __shared__ DowndateEp<T> ep;
    downdate_body<T>(X, dp, nrm, F, ldf, d, n, XS, nrmS, Si, FS, LinvT, st, gs, blockIdx.x, lds, rs, &ep);
}

__global__ __launch_bounds__(BMAX) void k_commit(ArpcState* st, const int64_t* __restrict__ Si, int64_t* pivots,
                                                 int32_t* accepted)
{
    commit_body(st, Si, pivots, accepted);
}

__global__ void k_final_err(const double* psum, int nb, double trK, const ArpcState* st, double* err)
{
    final_err_body(psum, nb, trK, st, err);
}

__global__ void k_fill(double* p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
kevd_status run_rounds(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, int64_t n_inject,
                       const ArpcBufs& B)
{
    int64_t n = ctx->n;
    int nb = B.nb;
    double *pmax = B.pmax, *psum = B.psum, *H = B.H, *nrmS = B.nrmS, *Linv64 = B.Linv64;
    uint64_t *prefix = B.prefix, *btot = B.btot;
    int64_t *Sprime = B.Sprime, *Si = B.Si;
    T *XS = (T*)B.XS, *FS = (T*)B.FS, *LinvT = (T*)B.LinvT;
    const int64_t* inject_dev = B.inj;
    ArpcState* st = B.st;
    int64_t* pivots_dev = B.piv;
    int32_t* acc_dev = B.acc;
    double* err_dev = B.err;
    double* d = B.d;

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
            k_draw<<<1, 256, 0, s>>>(psum, btot, nb, prefix, n, trK, p->seed, p->block, Sprime, st, err_dev,
                                     inject_dev, n_inject);
        }
        {
            ProfScope ps(ctx, TC_REJECT);
            k_form_H<T><<<dim3(BMAX / 16, BMAX / 16), 256, 0, s>>>(ctx->X64, ctx->dp, F, ctx->ldf,
                                                                  Sprime, st, gs.neg_inv_eps, H);
            k_rejchol<T><<<1, 256, 0, s>>>(H, st, p->seed, p->tol, inject_dev != nullptr, Sprime, Si, LinvT,
                                           Linv64);
            k_gather_pivots<T><<<BMAX, 256, 0, s>>>(Xsrc, nrm, ctx->dp, F, ctx->ldf, Si, st, XS,
                                                    nrmS, FS);
        }
        {
            ProfScope ps(ctx, TC_DOWNDATE);
            k_downdate<T><<<gdown, 256, 0, s>>>(Xsrc, ctx->dp, nrm, F, ctx->ldf, d, n, XS, nrmS, Si,
                                                FS, LinvT, st, gs);
        }
        k_commit<<<1, BMAX, 0, s>>>(st, Si, pivots_dev, acc_dev);
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

namespace arpck {

kevd_status arpc_validate(kevd_ctx* ctx, const kevd_arpc_params* p, const kevd_arpc_result* res,
                          int* max_rounds_out)
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
    *max_rounds_out = max_rounds;
    return KEVD_OK;
}

kevd_status arpc_prepare(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, ArpcBufs* B)
{
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
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

    int64_t n = ctx->n;
    KEVD_TRY(ctx_buf_t(ctx, "a_d", n, &B->d));
    KEVD_TRY(ctx_buf_t(ctx, "a_state", 1, &B->st));
    KEVD_TRY(ctx_buf_t(ctx, "a_piv", p->rank, &B->piv));
    KEVD_TRY(ctx_buf_t(ctx, "a_acc", max_rounds, &B->acc));
    KEVD_TRY(ctx_buf_t(ctx, "a_err", max_rounds, &B->err));
    B->inj = nullptr;
    if (p->inject_pivots) {
        KEVD_TRY(ctx_buf_t(ctx, "a_inj", p->n_inject, &B->inj));
        KEVD_HIP(ctx, hipMemcpyAsync(B->inj, p->inject_pivots, p->n_inject * sizeof(int64_t),
                                     hipMemcpyHostToDevice, ctx->stream));
    }
    int nb = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    B->nb = nb;
    KEVD_TRY(ctx_buf_t(ctx, "a_pmax", nb, &B->pmax));
    KEVD_TRY(ctx_buf_t(ctx, "a_psum", nb, &B->psum));
    KEVD_TRY(ctx_buf_t(ctx, "a_prefix", n, &B->prefix));
    KEVD_TRY(ctx_buf_t(ctx, "a_btot", nb, &B->btot));
    KEVD_TRY(ctx_buf_t(ctx, "a_Sprime", BMAX, &B->Sprime));
    KEVD_TRY(ctx_buf_t(ctx, "a_Si", BMAX, &B->Si));
    KEVD_TRY(ctx_buf_t(ctx, "a_H", BMAX * BMAX, &B->H));
    KEVD_TRY(ctx_buf_t(ctx, "a_Linv64", BMAX * BMAX, &B->Linv64));
    KEVD_TRY(ctx_buf_t(ctx, "a_nrmS", BMAX, &B->nrmS));
    KEVD_TRY(ctx_buf(ctx, "a_XS", (size_t)BMAX * ctx->dp * esz, &B->XS));
    KEVD_TRY(ctx_buf(ctx, "a_FS", (size_t)BMAX * ldf * esz, &B->FS));
    KEVD_TRY(ctx_buf(ctx, "a_LinvT", (size_t)BMAX * BMAX * esz, &B->LinvT));

    ArpcState init{};
    init.rank = p->rank;
    KEVD_HIP(ctx, hipMemcpyAsync(B->st, &init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
    KEVD_HIP(ctx, hipMemsetAsync(B->acc, 0, max_rounds * sizeof(int32_t), ctx->stream));
    KEVD_HIP(ctx, hipMemsetAsync(B->err, 0, max_rounds * sizeof(double), ctx->stream));
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(B->d, n, 1.0);
    KEVD_HIP(ctx, hipGetLastError());
    return KEVD_OK;
}

kevd_status arpc_finish(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, const ArpcBufs& B,
                        kevd_arpc_result* res)
{
    ArpcState fin{};
    KEVD_HIP(ctx, hipMemcpy(&fin, B.st, sizeof(fin), hipMemcpyDeviceToHost));
    res->rank = fin.Nr;
    res->rounds = fin.round;
    KEVD_HIP(ctx, hipMemcpy(res->pivots, B.piv, fin.Nr * sizeof(int64_t), hipMemcpyDeviceToHost));
    KEVD_HIP(ctx, hipMemcpy(res->accepted, B.acc, max_rounds * sizeof(int32_t), hipMemcpyDeviceToHost));
    KEVD_HIP(ctx, hipMemcpy(res->trace_err, B.err, max_rounds * sizeof(double), hipMemcpyDeviceToHost));
    ctx->rank = fin.Nr;
    ctx->pivots.assign(res->pivots, res->pivots + fin.Nr);
    ctx->have_factor = fin.Nr > 0;
    ctx->factor_eps = p->eps;
    // Algorithmic flop of the downdate: per round 2 n k Nr (overlap) +
    // 2 n k dim (distances) + n k^2 (triangular solve) + 2 n k (norms).
    double fl = 0.0, nr = 0.0;
    for (int r = 0; r < fin.round; ++r) {
        double k = res->accepted[r];
        fl += (double)ctx->n * k * (2.0 * nr + 2.0 * ctx->dim + k + 2.0);
        nr += k;
    }
    ctx->tm.downdate_flop = fl;
    return KEVD_OK;
}

}  // namespace arpck

extern "C" kevd_status kevd_arpc(kevd_ctx* ctx, const kevd_arpc_params* p, kevd_arpc_result* res)
{
    int max_rounds = 0;
    KEVD_TRY(arpck::arpc_validate(ctx, p, res, &max_rounds));
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->tm = kevd_timings{};
    prof_reset(ctx);
    arpck::ArpcBufs B;
    KEVD_TRY(arpck::arpc_prepare(ctx, p, max_rounds, &B));
    kevd_status s = p->prec == KEVD_MIXED ? run_rounds<float>(ctx, p, max_rounds, p->n_inject, B)
                                          : run_rounds<double>(ctx, p, max_rounds, p->n_inject, B);
    if (s != KEVD_OK) return s;
    KEVD_TRY(arpck::arpc_finish(ctx, p, max_rounds, B, res));
    auto t1 = std::chrono::steady_clock::now();
    ctx->tm.arpc_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    prof_resolve(ctx);
    return KEVD_OK;
}
