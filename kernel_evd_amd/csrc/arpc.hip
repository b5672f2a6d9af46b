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
//
// With params.regenerate the rounds run in the reference's own LM form (no
// factor stored; F(S',:), the overlap and L^-1 regenerated or maintained
// from the points, see arpc_kernels.h): run_rounds_lm below.
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include <rocblas/rocblas.h>

#include "arpc_kernels.h"

rocblas_handle blas_handle(kevd_ctx* ctx);

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

// A run that has used its round budget stops here (stop reason 8), so a
// replayed block of rounds may run past the budget harmlessly.
__global__ __launch_bounds__(256) void k_draw(const double* __restrict__ psum, const double* __restrict__ pmax,
                                              const uint64_t* __restrict__ btot, int nb,
                                              const uint64_t* __restrict__ prefix, int64_t n, double trK,
                                              uint64_t seed, int b, double tol, int64_t* Sprime, ArpcState* st,
                                              double* err, const int64_t* inject, int64_t n_inject, int max_rounds)
{
    if (threadIdx.x == 0 && !st->done && st->round >= max_rounds) {
        st->done = 1;
        st->stop_reason = 8;
    }
    __syncthreads();
    draw_body(psum, pmax, btot, nb, prefix, n, trK, seed, b, tol, Sprime, st, err, inject, n_inject);
}

template <typename T>
__global__ __launch_bounds__(FORMH_THREADS) void k_form_H(const double* __restrict__ X64, int32_t dp, const T* __restrict__ F,
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

template <typename T, bool X6>
__global__ __launch_bounds__(256) void k_gather_pivots(const T* __restrict__ Xsrc, const double* __restrict__ nrm,
                                                       int32_t dp, const T* __restrict__ F, int64_t ldf,
                                                       const int64_t* __restrict__ Si, const ArpcState* st,
                                                       T* XS, double* nrmS, T* FS, T* LinvT)
{
    gather_pivots_body<T, X6>(Xsrc, nrm, dp, F, ldf, Si, st, XS, nrmS, FS, LinvT, blockIdx.x);
}

template <typename T, bool X6>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_downdate(
    const T* __restrict__ X, int32_t dp, const double* __restrict__ nrm, T* __restrict__ F, int64_t ldf,
    double* __restrict__ d, int64_t n, const T* __restrict__ XS, const double* __restrict__ nrmS,
    const int64_t* __restrict__ Si, const T* __restrict__ FS, const T* __restrict__ LinvT, const ArpcState* st,
    GaussScale gs, double* __restrict__ cspart)
{
    typedef DowndateLds<T> DL;
    __shared__ __attribute__((aligned(16))) T lds[DL::ELEMS];
    __shared__ double rs[DL::WN][DL::BM];
    __shared__ DowndateEp<T> ep;
    downdate_body<T, X6>(X, dp, nrm, F, ldf, d, n, XS, nrmS, Si, FS, LinvT, st, gs, blockIdx.x, lds, rs, &ep,
                         cspart);
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

// ---------------------------------------------------------- LM variant

__global__ __launch_bounds__(256) void k_lm_kprop(const double* __restrict__ X64, int32_t dp,
                                                  const int64_t* __restrict__ Sprime, const int64_t* __restrict__ piv,
                                                  const ArpcState* st, double neg_inv_eps, double* P, int64_t ldp)
{
    lm_kprop_body(X64, dp, Sprime, piv, st, neg_inv_eps, P, ldp, blockIdx.y,
                  (int64_t)blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(FORMH_THREADS) void k_lm_form_H(const double* __restrict__ X64, int32_t dp,
                                                   const double* __restrict__ P2, int64_t ldp,
                                                   const int64_t* __restrict__ Sprime, const ArpcState* st,
                                                   double neg_inv_eps, double* H)
{
    form_H_body<double>(X64, dp, P2, ldp, Sprime, st, neg_inv_eps, H, blockIdx.x, blockIdx.y, true);
}

template <typename T>
__global__ __launch_bounds__(256) void k_lm_rejchol(const double* __restrict__ H, ArpcState* st, uint64_t seed,
                                                    double tol, int inject, const int64_t* __restrict__ Sprime,
                                                    int64_t* Si, T* LinvT, double* Linv64, int32_t* Spos)
{
    rejchol_body<T>(H, st, seed, tol, inject, Sprime, Si, LinvT, Linv64, Spos);
}

// X(S_i,:) and its norms for the prologue (rows past k zero), and the
// accepted rows of F(S',:) plus the pivot-point append.
template <typename T>
__global__ __launch_bounds__(256) void k_lm_gather(const double* __restrict__ P2, int64_t ldp,
                                                   const int32_t* __restrict__ Spos, const T* __restrict__ Xsrc,
                                                   const double* __restrict__ nrm, int32_t dp,
                                                   const int64_t* __restrict__ Si, const ArpcState* st, double* FSd,
                                                   T* XP, double* nrmP, T* XS, double* nrmS)
{
    if (st->done) return;
    int c = blockIdx.x;
    int64_t s = c < st->k ? Si[c] : -1;
    for (int j = threadIdx.x; j < dp; j += 256) XS[c * dp + j] = s >= 0 ? Xsrc[s * dp + j] : (T)0;
    if (threadIdx.x == 0) nrmS[c] = s >= 0 ? nrm[s] : 0.0;
    lm_gather_body<T>(P2, ldp, Spos, Xsrc, nrm, dp, Si, st, FSd, XP, nrmP, c);
}

__global__ __launch_bounds__(256) void k_lm_linv_append(const double* __restrict__ Lc,
                                                        const double* __restrict__ MTd, int64_t ldp,
                                                        const ArpcState* st, double* Linv, int64_t ldl)
{
    lm_linv_append_body(Lc, MTd, ldp, st, Linv, ldl, blockIdx.y, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_lm_cast(const double* __restrict__ src, float* dst, int64_t count)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = (float)src[i];
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_downdate_lm(
    const T* __restrict__ X, int32_t dp, const double* __restrict__ nrm, double* __restrict__ d, int64_t n,
    const T* __restrict__ XS, const double* __restrict__ nrmS, const int64_t* __restrict__ Si,
    const T* __restrict__ XP, const double* __restrict__ nrmP, const int64_t* __restrict__ piv,
    const T* __restrict__ MT, int64_t ldm, const T* __restrict__ LinvT, const ArpcState* st, GaussScale gs)
{
    typedef DowndateLds<T> DL;
    __shared__ __attribute__((aligned(16))) T lds[DL::ELEMS];
    __shared__ double rs[DL::WN][DL::BM];
    __shared__ DowndateEp<T> ep;
    downdate_lm_body<T>(X, dp, nrm, d, n, XS, nrmS, Si, XP, nrmP, piv, MT, ldm, LinvT, st, gs, blockIdx.x, lds,
                        rs, &ep);
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
    const bool x6 = std::is_same<T, float>::value && use_bf16x6();
    auto round = [&]() {
        {
            ProfScope ps(ctx, TC_SAMPLE);
            k_reduce_d<<<nb, 256, 0, s>>>(d, n, pmax, psum, st, 0);
            k_weights_scan<<<nb, 256, 0, s>>>(d, n, pmax, nb, shift_base, prefix, btot, st);
            k_draw<<<1, 256, 0, s>>>(psum, pmax, btot, nb, prefix, n, trK, p->seed, p->block, p->tol, Sprime, st,
                                     err_dev, inject_dev, n_inject, max_rounds);
        }
        {
            ProfScope ps(ctx, TC_REJECT);
            k_form_H<T><<<dim3(BMAX / 16, BMAX / 16), FORMH_THREADS, 0, s>>>(ctx->X64, ctx->dp, F, ctx->ldf,
                                                                  Sprime, st, gs.neg_inv_eps, H);
            k_rejchol<T><<<1, 256, 0, s>>>(H, st, p->seed, p->tol, inject_dev != nullptr, Sprime, Si, LinvT,
                                           Linv64);
            if (x6)
                k_gather_pivots<T, true><<<BMAX, 256, 0, s>>>(Xsrc, nrm, ctx->dp, F, ctx->ldf, Si, st, XS,
                                                              nrmS, FS, LinvT);
            else
                k_gather_pivots<T, false><<<BMAX, 256, 0, s>>>(Xsrc, nrm, ctx->dp, F, ctx->ldf, Si, st, XS,
                                                               nrmS, FS, LinvT);
        }
        {
            ProfScope ps(ctx, TC_DOWNDATE);
            if (x6)
                k_downdate<T, true><<<gdown, 256, 0, s>>>(Xsrc, ctx->dp, nrm, F, ctx->ldf, d, n, XS, nrmS, Si,
                                                          FS, LinvT, st, gs, B.cspart);
            else
                k_downdate<T, false><<<gdown, 256, 0, s>>>(Xsrc, ctx->dp, nrm, F, ctx->ldf, d, n, XS, nrmS, Si,
                                                           FS, LinvT, st, gs, B.cspart);
        }
        k_commit<<<1, BMAX, 0, s>>>(st, Si, pivots_dev, acc_dev);
    };
    // Without per-kernel timing the rounds are captured as a hipGraph of
    // rpg consecutive rounds and replayed: one launch (and one status copy)
    // per rpg rounds.  Every kernel returns at once after the device state
    // says done, and the draw enforces the round budget, so replays past the
    // end cost a few microseconds each.  Polling runs two launches behind.
    const int est = (int)((p->rank + p->block - 1) / p->block);
    const int rpg = ctx->profiling ? 1 : std::max(1, std::min(8, est / 8));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = !ctx->profiling && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (use_graph) {
        for (int i = 0; i < rpg; ++i) round();
        use_graph = hipStreamEndCapture(s, &graph) == hipSuccess && graph != nullptr;
        if (use_graph) use_graph = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        (void)hipGetLastError();
    }
    const int step = use_graph ? rpg : 1;
    for (int rd = 0, li = 0; rd < max_rounds; rd += step, ++li) {
        hipError_t le = hipSuccess;
        if (use_graph)
            le = hipGraphLaunch(exec, s);
        else
            round();
        if (le == hipSuccess) le = hipGetLastError();
        if (le != hipSuccess) {
            status = kevd_hip_fail(ctx, le, "ARPC round launch", __FILE__, __LINE__);
            break;
        }
        int slot = li & 3;
        hipError_t ce = hipMemcpyAsync(&host_st[slot], st, sizeof(ArpcState), hipMemcpyDeviceToHost, s);
        if (ce == hipSuccess) ce = hipEventRecord(ev[slot], s);
        if (ce != hipSuccess) {
            status = kevd_hip_fail(ctx, ce, "ARPC status copy", __FILE__, __LINE__);
            break;
        }
        if (li >= 2) {
            int old = (li - 2) & 3;
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
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return status;
}

#define LM_BLAS(call)                                                                 \
    do {                                                                              \
        rocblas_status b_ = (call);                                                   \
        if (b_ != rocblas_status_success)                                             \
            return kevd_fail(ctx, KEVD_EHIP, std::string(#call) + " failed: " +       \
                                                 rocblas_status_to_string(b_));       \
    } while (0)

// The rounds of the LM form.  Each round needs the rank so far on the host
// (it sizes the FP64 triangular products), so the state is read back after
// every round, as the reference's round loop does.
template <typename T>
kevd_status run_rounds_lm(kevd_ctx* ctx, const kevd_arpc_params* p, int max_rounds, int64_t n_inject,
                          const ArpcBufs& B)
{
    int64_t n = ctx->n;
    int nb = B.nb;
    const int32_t dp = ctx->dp;
    const T* Xsrc = std::is_same<T, float>::value ? (const T*)ctx->X32 : (const T*)ctx->X64;
    const double* nrm = std::is_same<T, float>::value ? ctx->nrm32 : ctx->nrm64;
    GaussScale gs{-1.0 / p->eps, -1.4426950408889634074 / p->eps};
    int shift_base = 62 - ceil_log2(n);
    double trK = (double)n;
    hipStream_t s = ctx->stream;
    rocblas_handle h = blas_handle(ctx);
    if (!h) return kevd_fail(ctx, KEVD_EHIP, "rocblas handle creation failed");

    // row length of every r-long row: room for a whole last pivot chunk
    const int64_t ldp = KEVD_ROUND_UP(p->rank + LM_CH, 64);
    const int64_t ldl = ldp;
    double *P, *P2, *FSd, *MTd, *Linv, *nrmP;
    T *XP, *MT;
    int32_t* Spos;
    KEVD_TRY(ctx_buf_t(ctx, "lm_P", (size_t)BMAX * ldp, &P));
    KEVD_TRY(ctx_buf_t(ctx, "lm_P2", (size_t)BMAX * ldp, &P2));
    KEVD_TRY(ctx_buf_t(ctx, "lm_FSd", (size_t)BMAX * ldp, &FSd));
    KEVD_TRY(ctx_buf_t(ctx, "lm_MTd", (size_t)BMAX * ldp, &MTd));
    KEVD_TRY(ctx_buf_t(ctx, "lm_Linv", (size_t)ldl * ldl, &Linv));
    KEVD_TRY(ctx_buf_t(ctx, "lm_nrmP", (size_t)ldp, &nrmP));
    KEVD_TRY(ctx_buf_t(ctx, "lm_XP", (size_t)ldp * dp, &XP));
    KEVD_TRY(ctx_buf_t(ctx, "lm_Spos", BMAX, &Spos));
    if (std::is_same<T, float>::value) KEVD_TRY(ctx_buf_t(ctx, "lm_MT", (size_t)BMAX * ldp, &MT));
    else MT = (T*)MTd;
    // zero: rows past k and columns past Nr must read as exact zeros
    for (double* z : {P, P2, FSd, MTd, Linv}) {
        size_t cnt = z == Linv ? (size_t)ldl * ldl : (size_t)BMAX * ldp;
        KEVD_HIP(ctx, hipMemsetAsync(z, 0, cnt * sizeof(double), s));
    }
    KEVD_HIP(ctx, hipMemsetAsync(XP, 0, (size_t)ldp * dp * sizeof(T), s));
    KEVD_HIP(ctx, hipMemsetAsync(nrmP, 0, (size_t)ldp * sizeof(double), s));
    if ((void*)MT != (void*)MTd) KEVD_HIP(ctx, hipMemsetAsync(MT, 0, (size_t)BMAX * ldp * sizeof(T), s));

    ArpcState* host_st = nullptr;
    KEVD_TRY(ctx_pinned(ctx, sizeof(ArpcState), (void**)&host_st));
    const double one = 1.0;
    unsigned gdown = (unsigned)((n + Cfg<T>::BM - 1) / Cfg<T>::BM);
    int64_t Nr = 0;
    for (int rd = 0; rd < max_rounds; ++rd) {
        {
            ProfScope ps(ctx, TC_SAMPLE);
            k_reduce_d<<<nb, 256, 0, s>>>(B.d, n, B.pmax, B.psum, B.st, 0);
            k_weights_scan<<<nb, 256, 0, s>>>(B.d, n, B.pmax, nb, shift_base, B.prefix, B.btot, B.st);
            k_draw<<<1, 256, 0, s>>>(B.psum, B.pmax, B.btot, nb, B.prefix, n, trK, p->seed, p->block, p->tol,
                                     B.Sprime, B.st, B.err, B.inj, n_inject, max_rounds);
        }
        {
            ProfScope ps(ctx, TC_REJECT);
            if (Nr > 0) {
                // F(S',:) = K(S',S) L^-T in FP64 (column-major view: P2^T = L^-1 P^T)
                k_lm_kprop<<<dim3((unsigned)((Nr + 255) / 256), BMAX), 256, 0, s>>>(
                    ctx->X64, dp, B.Sprime, B.piv, B.st, gs.neg_inv_eps, P, ldp);
                LM_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                      rocblas_diagonal_non_unit, (rocblas_int)Nr, BMAX, &one, Linv,
                                      (rocblas_int)ldl, P, (rocblas_int)ldp, P2, (rocblas_int)ldp));
            }
            k_lm_form_H<<<dim3(BMAX / 16, BMAX / 16), FORMH_THREADS, 0, s>>>(ctx->X64, dp, P2, ldp, B.Sprime, B.st,
                                                                 gs.neg_inv_eps, B.H);
            k_lm_rejchol<T><<<1, 256, 0, s>>>(B.H, B.st, p->seed, p->tol, B.inj != nullptr, B.Sprime, B.Si,
                                              (T*)B.LinvT, B.Linv64, Spos);
            k_lm_gather<T><<<BMAX, 256, 0, s>>>(P2, ldp, Spos, Xsrc, nrm, dp, B.Si, B.st, FSd, XP, nrmP,
                                                (T*)B.XS, B.nrmS);
            if (Nr > 0) {
                // M^T = F_{S_i} L^-1 (column-major view: MTd^T = L^-T FSd^T)
                LM_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                      rocblas_diagonal_non_unit, (rocblas_int)Nr, BMAX, &one, Linv,
                                      (rocblas_int)ldl, FSd, (rocblas_int)ldp, MTd, (rocblas_int)ldp));
                if ((void*)MT != (void*)MTd)
                    k_lm_cast<<<(unsigned)((BMAX * ldp + 255) / 256), 256, 0, s>>>(MTd, (float*)MT, BMAX * ldp);
            }
            k_lm_linv_append<<<dim3((unsigned)((Nr + BMAX + 255) / 256), BMAX), 256, 0, s>>>(B.Linv64, MTd, ldp,
                                                                                            B.st, Linv, ldl);
        }
        {
            ProfScope ps(ctx, TC_DOWNDATE);
            k_downdate_lm<T><<<gdown, 256, 0, s>>>(Xsrc, dp, nrm, B.d, n, (const T*)B.XS, B.nrmS, B.Si, XP, nrmP,
                                                   B.piv, MT, ldp, (const T*)B.LinvT, B.st, gs);
        }
        k_commit<<<1, BMAX, 0, s>>>(B.st, B.Si, B.piv, B.acc);
        KEVD_HIP(ctx, hipGetLastError());
        KEVD_HIP(ctx, hipMemcpyAsync(host_st, B.st, sizeof(ArpcState), hipMemcpyDeviceToHost, s));
        KEVD_HIP(ctx, hipStreamSynchronize(s));
        Nr = host_st->Nr;
        if (host_st->done) break;
    }
    k_reduce_d<<<nb, 256, 0, s>>>(B.d, n, B.pmax, B.psum, B.st, 1);
    k_final_err<<<1, 64, 0, s>>>(B.psum, nb, trK, B.st, B.err);
    KEVD_HIP(ctx, hipGetLastError());
    KEVD_HIP(ctx, hipStreamSynchronize(s));
    ctx->lm_linv = Linv;
    ctx->lm_ldl = ldl;
    return KEVD_OK;
}

#undef LM_BLAS

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
    // factor storage (none in the LM form: that is its point)
    int64_t ldf = kevd_factor_ld(p->rank);
    size_t esz = p->prec == KEVD_MIXED ? sizeof(float) : sizeof(double);
    size_t fbytes = (size_t)ctx->n * ldf * esz;
    KEVD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!p->regenerate) {
        if (ctx->F && ctx->F_bytes < fbytes) {
            KEVD_HIP(ctx, hipFree(ctx->F));
            ctx->F = nullptr;
            ctx->F_bytes = 0;
        }
        if (!ctx->F) {
            KEVD_HIP(ctx, hipMalloc(&ctx->F, fbytes));
            ctx->F_bytes = fbytes;
        }
    }
    ctx->have_factor = false;
    ctx->cs_valid = false;
    ctx->lm = false;
    ctx->lm_linv = nullptr;
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
    // mixed F form: the pivot points' and pivot rows' bf16x6 parts follow them
    KEVD_TRY(ctx_buf(ctx, "a_XS",
                     !p->regenerate && p->prec == KEVD_MIXED ? x6_xs_bytes(ctx->dp) : (size_t)BMAX * ctx->dp * esz,
                     &B->XS));
    KEVD_TRY(ctx_buf(ctx, "a_FS",
                     p->regenerate ? (size_t)BMAX * 64 * esz
                                   : (p->prec == KEVD_MIXED ? x6_fs_bytes(ldf) : (size_t)BMAX * ldf * esz),
                     &B->FS));
    KEVD_TRY(ctx_buf(ctx, "a_LinvT",
                     !p->regenerate && p->prec == KEVD_MIXED ? x6_linv_bytes() : (size_t)BMAX * BMAX * esz,
                     &B->LinvT));
    // F form: the downdate leaves F^T 1 as per-row-block partials
    B->cspart = nullptr;
    if (!p->regenerate) {
        const int64_t bm = p->prec == KEVD_MIXED ? Cfg<float>::BM : Cfg<double>::BM;
        KEVD_TRY(ctx_buf_t(ctx, "a_cspart", (size_t)((n + bm - 1) / bm) * ldf, &B->cspart));
        ctx->cs_nblk = (n + bm - 1) / bm;
    }

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
    ctx->cs_part = B.cspart;
    ctx->cs_valid = B.cspart != nullptr && fin.Nr > 0;
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
    TraceRange trace_range("kevd_arpc");
    int max_rounds = 0;
    KEVD_TRY(arpck::arpc_validate(ctx, p, res, &max_rounds));
    KEVD_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->tm = kevd_timings{};
    prof_reset(ctx);
    arpck::ArpcBufs B;
    KEVD_TRY(arpck::arpc_prepare(ctx, p, max_rounds, &B));
    kevd_status s;
    if (p->regenerate)
        s = p->prec == KEVD_MIXED ? run_rounds_lm<float>(ctx, p, max_rounds, p->n_inject, B)
                                  : run_rounds_lm<double>(ctx, p, max_rounds, p->n_inject, B);
    else
        s = p->prec == KEVD_MIXED ? run_rounds<float>(ctx, p, max_rounds, p->n_inject, B)
                                  : run_rounds<double>(ctx, p, max_rounds, p->n_inject, B);
    if (s != KEVD_OK) return s;
    KEVD_TRY(arpck::arpc_finish(ctx, p, max_rounds, B, res));
    if (p->regenerate) {
        // the LM form keeps S and L^-1, not F
        ctx->have_factor = false;
        ctx->lm = true;
    }
    auto t1 = std::chrono::steady_clock::now();
    ctx->tm.arpc_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    prof_resolve(ctx);
    return KEVD_OK;
}
