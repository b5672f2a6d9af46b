// internal.h -- the context object behind the kevd_ctx handle and host-side
// helpers shared by the translation units of libkevd.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "common.h"

// Size of the fixed device scratch array behind the allocation-free entry
// points (one array per device, held until the process ends).
constexpr size_t KEVD_SCRATCH_BYTES = KEVD_SCRATCH_MAX_BYTES;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// One device-side timing record: a named category and an event pair.
struct TimedLaunch {
    int cat;
    hipEvent_t start, stop;
};

enum TimeCat {
    TC_DOWNDATE = 0, TC_GRAM, TC_EIG, TC_SAMPLE, TC_REJECT, TC_MATVEC, TC_DENSE, TC_UGEMM,
    TC_COUNT
};

struct kevd_ctx {
    int device = 0;
    int num_cu = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    std::string err;

    // Point set, centred, resident in HBM.  dp = dim rounded up to 32 (zero
    // padded) so every distance GEMM chunk is whole.
    int64_t n = 0;
    int32_t dim = 0;
    int32_t dp = 0;
    float* X32 = nullptr;      // n x dp, FP32 copy of the centred points
    double* X64 = nullptr;     // n x dp, FP64 centred points
    double* nrm32 = nullptr;   // ||X32_i||^2 accumulated in FP64
    double* nrm64 = nullptr;   // ||X64_i||^2

    // Factor F (n x rank, row-major, leading dimension ldf) in fprec.
    void* F = nullptr;
    size_t F_bytes = 0;
    kevd_precision fprec = KEVD_FP64;
    int64_t rank = 0;
    int64_t ldf = 0;
    bool have_factor = false;
    std::vector<int64_t> pivots;  // acceptance order, when F came from kevd_arpc
    // F^T 1 of the factor kevd_arpc built, as the downdate's per-row-block
    // partials (cs_nblk x ldf FP64, a named buffer); cs_valid while F is
    // that factor.
    double* cs_part = nullptr;
    int64_t cs_nblk = 0;
    bool cs_valid = false;
    double factor_eps = 0.0;
    // The last kevd_arpc ran in the LM form: no F, L^-1 kept instead
    // (row-major lower, FP64, leading dimension lm_ldl; a named buffer).
    bool lm = false;
    double* lm_linv = nullptr;
    int64_t lm_ldl = 0;

    // Subspace-iteration step at which the last leading-eigenpair solve on
    // this context converged (0: none yet); the next solve runs its first
    // Rayleigh-Ritz check there.
    int eig_converged_at = 0;
    // The same for the ensemble's batched solve, kept on its master context
    // (the first context of kevd_ensemble_evd).
    int ens_eig_converged_at = 0;

    // Tile of the bf16x6 Gram (256 or 128; 0: 256), set by kevd_autotune_tiles.
    int gram_tile = 0;

    // Set by kevd_ensemble_evd: the leading-eigenpair solve joins one batched
    // solve over the ensemble's problems (slot = this problem's index).
    struct EigBatch* eig_batch = nullptr;
    int eig_batch_slot = -1;

    // Set by kevd_ensemble_evd when its batched Sinkhorn iteration has
    // scaled this problem: the next Sinkhorn kEVD takes the scaling s (one
    // of its named buffers, n entries), the iteration count, the final
    // deviation and min(s) instead of iterating itself.
    struct SkPre {
        bool valid = false;
        double* s = nullptr;
        int iters = 0;
        double dev = 0.0, smin = 0.0;
    } sk_pre;

    // Device address of this device's fixed scratch array (KEVD_SCRATCH_BYTES),
    // resolved when the context is created; shared by every context on the
    // device.
    void* scratch = nullptr;

    // The kEVD's intermediate record (kevd_set_dump): the reference's dump
    // datasets by name, FP64 on the host, column-major for matrices.
    bool dump = false;
    std::map<std::string, std::vector<double>> record;

    // kEVD output held on the device.
    float* U = nullptr;
    size_t U_bytes = 0;
    int32_t k_out = 0;

    // Named grow-only scratch buffers.
    std::map<std::string, DevBuf> bufs;

    // Pinned host staging for small device->host status copies.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;

    void* blas = nullptr;         // rocblas_handle, created on first use

    bool profiling = false;
    std::vector<TimedLaunch> launches;
    std::vector<hipEvent_t> event_pool;
    kevd_timings tm{};
    double flop_acc[TC_COUNT] = {};
};

// Grow-only named device buffer.
kevd_status ctx_buf(kevd_ctx* ctx, const char* name, size_t bytes, void** out);
template <typename P>
inline kevd_status ctx_buf_t(kevd_ctx* ctx, const char* name, size_t count, P** out)
{
    void* v = nullptr;
    kevd_status s = ctx_buf(ctx, name, count * sizeof(P), &v);
    *out = reinterpret_cast<P*>(v);
    return s;
}
kevd_status ctx_pinned(kevd_ctx* ctx, size_t bytes, void** out);

// A named roctx range over a host-side scope (an entry point or one of its
// phases), for timeline tools (rocprofv3 --marker-trace).  The ranges mark
// the host's enqueueing; device time is in the kernel trace.
struct TraceRange {
    explicit TraceRange(const char* name);
    ~TraceRange();
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

// Profiling: bracket a launch with an event pair on the context stream when
// profiling is on; resolve the pairs into ctx->tm after the stream syncs.
struct ProfScope {
    kevd_ctx* ctx;
    int cat;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(kevd_ctx* c, int category);
    ~ProfScope();
};
void prof_reset(kevd_ctx* ctx);
void prof_resolve(kevd_ctx* ctx);

// The gfx950 MFMA paths need a gfx950 device; reject anything else loudly.
kevd_status ctx_check_arch(kevd_ctx* ctx);

// The FP32 GEMMs of the mixed path (the downdate's overlap, the kEVD Grams)
// run with bf16x6 products on the bf16 matrix cores; KEVD_FP32_MFMA=1 in the
// environment selects plain FP32 MFMA instead (read once per process).
bool use_bf16x6();

// Cross-unit entry points.
kevd_status pivot_cholesky(kevd_ctx* ctx, double* L);

// Batched leading-eigenpair solve for an ensemble (evd.hip): every problem
// whose kEVD reaches the subspace iteration joins, the last one to arrive
// (or to leave without joining) runs one batched solve for all of them on
// the master context, and each continues with its own eigenvectors.
// nslots problems, k_out leading pairs each.  A problem that ends without
// joining must leave (eig_batch_leave is a no-op for one that joined).
struct EigBatch;
EigBatch* eig_batch_create(kevd_ctx* master, int nslots, int k_out);
// The Sinkhorn iterations of an ensemble's kEVDs as batched launches on the
// master context's stream (FP32 factors from kevd_arpc, rank <= 2048); sets
// each scaled problem's sk_pre.  Problems it cannot take keep their own
// iteration.  *batched: the number it scaled.
kevd_status ens_sinkhorn(kevd_ctx** ctxs, int nprob, const kevd_evd_params* p, int* batched);
void eig_batch_leave(EigBatch* b, int slot);
void eig_batch_destroy(EigBatch* b);
kevd_status kblock_gather(kevd_ctx* ctx, const int64_t* idx_dev, int64_t count,
                          kevd_precision prec, void* Xout, double* nrm_out);
