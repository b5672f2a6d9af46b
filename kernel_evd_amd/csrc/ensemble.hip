// ensemble.hip -- many independent kernel problems at once on one GPU.
//
// An ensemble is an array of ordinary contexts, one per problem (points set
// and eps chosen through the single-problem calls).  kevd_ensemble_arpc
// runs every problem's accelerated RPCholesky together:
//   * each stage of a round is ONE launch over all problems: the scan
//     stages take (problem, chunk) pairs from a chunk table, the pivot-block
//     stages one workgroup (or tile grid) per problem;
//   * the downdate is a persistent kernel: a device-wide grid of 2
//     workgroups per CU takes (problem, row block) items from an atomic
//     work counter, heaviest problems first, so problems of very different
//     sizes share the machine without a tail per problem;
//   * the round's launches are captured once into a hipGraph and replayed,
//     one graph launch per round; all per-round values (round index,
//     accepted count, rank so far, stop flags) live in device memory, so the
//     graph needs no host arguments and the host only polls an all-done flag
//     two rounds behind.
// Every problem runs the same round bodies as kevd_arpc (arpc_kernels.h),
// so its pivots and factor are bit-identical to a single-problem run.
// kevd_ensemble_evd runs the per-problem kEVDs concurrently on a pool of
// host threads, each problem on its own context stream (the r x r phases are
// latency-bound rocSOLVER sequences, which overlap across streams).
// kevd_autotune_block picks a problem's block size by timing candidates.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "arpc_kernels.h"

namespace {

using namespace arpck;

// Device descriptor of one problem's ARPC run.
template <typename T>
struct EnsProb {
    const T* X;
    const double* X64;
    const double* nrm;
    T* F;
    double* d;
    ArpcState* st;
    double *pmax, *psum, *H, *nrmS, *Linv64, *err;
    uint64_t *prefix, *btot;
    int64_t *Sprime, *Si, *piv;
    int32_t* acc;
    T *XS, *FS, *LinvT;
    double* cspart;  // F^T 1 partials
    int64_t n, ldf;
    double trK, tol;
    uint64_t seed;
    GaussScale gs;
    int32_t dp, nb, shift_base, block, max_rounds, pad;
};

template <typename T>
__global__ __launch_bounds__(256) void k_ens_reduce_d(const EnsProb<T>* P, const int2* chunks, int force)
{
    int2 c = chunks[blockIdx.x];
    const EnsProb<T>& q = P[c.x];
    reduce_d_body(q.d, q.n, q.pmax, q.psum, q.st, force, c.y);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ens_weights_scan(const EnsProb<T>* P, const int2* chunks)
{
    int2 c = chunks[blockIdx.x];
    const EnsProb<T>& q = P[c.x];
    weights_scan_body(q.d, q.n, q.pmax, q.nb, q.shift_base, q.prefix, q.btot, q.st, c.y);
}

// One workgroup per problem.  A problem that has used its round budget
// stops here (stop reason 8), as the single-problem host loop would.
template <typename T>
__global__ __launch_bounds__(256) void k_ens_draw(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.x];
    if (threadIdx.x == 0 && !q.st->done && q.st->round >= q.max_rounds) {
        q.st->done = 1;
        q.st->stop_reason = 8;
    }
    __syncthreads();
    draw_body(q.psum, q.pmax, q.btot, q.nb, q.prefix, q.n, q.trK, q.seed, q.block, q.tol, q.Sprime, q.st, q.err,
              nullptr, 0);
}

template <typename T>
__global__ __launch_bounds__(FORMH_THREADS) void k_ens_form_H(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.z];
    form_H_body<T>(q.X64, q.dp, q.F, q.ldf, q.Sprime, q.st, q.gs.neg_inv_eps, q.H, blockIdx.x, blockIdx.y);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ens_rejchol(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.x];
    rejchol_body<T>(q.H, q.st, q.seed, q.tol, 0, q.Sprime, q.Si, q.LinvT, q.Linv64);
}

template <typename T, bool X6>
__global__ __launch_bounds__(256) void k_ens_gather(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.y];
    gather_pivots_body<T, X6>(q.X, q.nrm, q.dp, q.F, q.ldf, q.Si, q.st, q.XS, q.nrmS, q.FS, q.LinvT, blockIdx.x);
}

__global__ void k_ens_zero(int* ctr)
{
    if (threadIdx.x == 0) *ctr = 0;
}

// Persistent downdate: each workgroup takes the next (problem, row block)
// item until the list is exhausted.  Items of finished problems (or of a
// round that accepted nothing) return at once.  The loop's two barriers keep
// the item slot consistent whether or not the body itself synchronised.
template <typename T, bool X6>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ens_downdate(
    const EnsProb<T>* P, const int2* items, int nitems, int* ctr)
{
    typedef DowndateLds<T> DL;
    __shared__ __attribute__((aligned(16))) T lds[DL::ELEMS];
    __shared__ double rs[DL::WN][DL::BM];
    __shared__ DowndateEp<T> ep;
    __shared__ int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(ctr, 1);
        __syncthreads();
        int it = item_s;
        __syncthreads();
        if (it >= nitems) break;
        int2 w = items[it];
        const EnsProb<T>& q = P[w.x];
        downdate_body<T, X6>(q.X, q.dp, q.nrm, q.F, q.ldf, q.d, q.n, q.XS, q.nrmS, q.Si, q.FS, q.LinvT, q.st,
                             q.gs, w.y, lds, rs, &ep, q.cspart);
    }
}

template <typename T>
__global__ __launch_bounds__(BMAX) void k_ens_commit(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.x];
    commit_body(q.st, q.Si, q.piv, q.acc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ens_all_done(const EnsProb<T>* P, int nprob, int* flag)
{
    __shared__ int all;
    if (threadIdx.x == 0) all = 1;
    __syncthreads();
    for (int i = threadIdx.x; i < nprob; i += 256)
        if (!P[i].st->done) all = 0;
    __syncthreads();
    if (threadIdx.x == 0) *flag = all;
}

template <typename T>
__global__ void k_ens_final_err(const EnsProb<T>* P)
{
    const EnsProb<T>& q = P[blockIdx.x];
    final_err_body(q.psum, q.nb, q.trK, q.st, q.err);
}

template <typename T>
kevd_status ens_rounds(kevd_ctx* m, kevd_ctx** ctxs, int nprob, const kevd_arpc_params* params,
                       const std::vector<ArpcBufs>& B, const std::vector<int>& maxr, int* used_graph)
{
    hipStream_t s = m->stream;
    // descriptors
    std::vector<EnsProb<T>> hp((size_t)nprob);
    std::vector<int2> chunks, items;
    for (int i = 0; i < nprob; ++i) {
        kevd_ctx* c = ctxs[i];
        const kevd_arpc_params& p = params[i];
        EnsProb<T>& q = hp[i];
        q.X = std::is_same<T, float>::value ? (const T*)c->X32 : (const T*)c->X64;
        q.X64 = c->X64;
        q.nrm = std::is_same<T, float>::value ? c->nrm32 : c->nrm64;
        q.F = (T*)c->F;
        q.d = B[i].d;
        q.st = B[i].st;
        q.pmax = B[i].pmax;
        q.psum = B[i].psum;
        q.H = B[i].H;
        q.nrmS = B[i].nrmS;
        q.Linv64 = B[i].Linv64;
        q.err = B[i].err;
        q.prefix = B[i].prefix;
        q.btot = B[i].btot;
        q.Sprime = B[i].Sprime;
        q.Si = B[i].Si;
        q.piv = B[i].piv;
        q.acc = B[i].acc;
        q.XS = (T*)B[i].XS;
        q.cspart = B[i].cspart;
        q.FS = (T*)B[i].FS;
        q.LinvT = (T*)B[i].LinvT;
        q.n = c->n;
        q.ldf = c->ldf;
        q.trK = (double)c->n;
        q.tol = p.tol;
        q.seed = p.seed;
        q.gs = GaussScale{-1.0 / p.eps, -1.4426950408889634074 / p.eps};
        q.dp = c->dp;
        q.nb = B[i].nb;
        q.shift_base = 62 - ceil_log2(c->n);
        q.block = p.block;
        q.max_rounds = maxr[i];
        q.pad = 0;
        for (int j = 0; j < B[i].nb; ++j) chunks.push_back(make_int2(i, j));
    }
    // downdate items, heaviest problems first (row blocks x target rank)
    std::vector<int> order((size_t)nprob);
    for (int i = 0; i < nprob; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double)ctxs[a]->n * params[a].rank > (double)ctxs[b]->n * params[b].rank;
    });
    constexpr int BM = Cfg<T>::BM;
    for (int i : order) {
        int64_t nblk = (ctxs[i]->n + BM - 1) / BM;
        for (int64_t j = 0; j < nblk; ++j) items.push_back(make_int2(i, (int)j));
    }
    if (items.size() > (size_t)INT32_MAX || chunks.size() > (size_t)INT32_MAX)
        return kevd_fail(m, KEVD_EINVAL, "ensemble: too many work items");
    EnsProb<T>* dP;
    int2 *dchunks, *ditems;
    int *ctr, *flag;
    KEVD_TRY(ctx_buf(m, "ens_desc", hp.size() * sizeof(EnsProb<T>), (void**)&dP));
    KEVD_TRY(ctx_buf_t(m, "ens_chunks", chunks.size(), &dchunks));
    KEVD_TRY(ctx_buf_t(m, "ens_items", items.size(), &ditems));
    KEVD_TRY(ctx_buf_t(m, "ens_ctr", 1, &ctr));
    KEVD_TRY(ctx_buf_t(m, "ens_flag", 1, &flag));
    KEVD_HIP(m, hipMemcpyAsync(dP, hp.data(), hp.size() * sizeof(EnsProb<T>), hipMemcpyHostToDevice, s));
    KEVD_HIP(m, hipMemcpyAsync(dchunks, chunks.data(), chunks.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    KEVD_HIP(m, hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    KEVD_HIP(m, hipStreamSynchronize(s));

    const unsigned nch = (unsigned)chunks.size();
    const int nit = (int)items.size();
    const unsigned gdown = (unsigned)std::min<int64_t>(nit, 2 * (int64_t)m->num_cu);
    const bool x6 = std::is_same<T, float>::value && use_bf16x6();
    auto round = [&]() {
        k_ens_reduce_d<T><<<nch, 256, 0, s>>>(dP, dchunks, 0);
        k_ens_weights_scan<T><<<nch, 256, 0, s>>>(dP, dchunks);
        k_ens_draw<T><<<nprob, 256, 0, s>>>(dP);
        k_ens_form_H<T><<<dim3(BMAX / 16, BMAX / 16, nprob), FORMH_THREADS, 0, s>>>(dP);
        k_ens_rejchol<T><<<nprob, 256, 0, s>>>(dP);
        k_ens_zero<<<1, 64, 0, s>>>(ctr);
        if (x6) {
            k_ens_gather<T, true><<<dim3(BMAX, nprob), 256, 0, s>>>(dP);
            k_ens_downdate<T, true><<<gdown, 256, 0, s>>>(dP, ditems, nit, ctr);
        } else {
            k_ens_gather<T, false><<<dim3(BMAX, nprob), 256, 0, s>>>(dP);
            k_ens_downdate<T, false><<<gdown, 256, 0, s>>>(dP, ditems, nit, ctr);
        }
        k_ens_commit<T><<<nprob, BMAX, 0, s>>>(dP);
        k_ens_all_done<T><<<1, 256, 0, s>>>(dP, nprob, flag);
    };
    // capture one round; fall back to direct launches if capture fails
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (use_graph) {
        round();
        use_graph = hipStreamEndCapture(s, &graph) == hipSuccess && graph != nullptr;
        if (use_graph) use_graph = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        (void)hipGetLastError();
    }
    int max_all = *std::max_element(maxr.begin(), maxr.end());
    int* hflag = nullptr;
    KEVD_TRY(ctx_pinned(m, 4 * sizeof(int), (void**)&hflag));
    for (int i = 0; i < 4; ++i) hflag[i] = 0;
    hipEvent_t ev[4];
    for (int i = 0; i < 4; ++i) KEVD_HIP(m, hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    kevd_status status = KEVD_OK;
    for (int rd = 0; rd < max_all; ++rd) {
        hipError_t e = hipSuccess;
        if (use_graph)
            e = hipGraphLaunch(exec, s);
        else
            round();
        if (e == hipSuccess) e = hipGetLastError();
        int slot = rd & 3;
        if (e == hipSuccess) e = hipMemcpyAsync(&hflag[slot], flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(ev[slot], s);
        if (e != hipSuccess) {
            status = kevd_hip_fail(m, e, "ensemble round", __FILE__, __LINE__);
            break;
        }
        if (rd >= 2) {
            int old = (rd - 2) & 3;
            hipError_t we = hipEventSynchronize(ev[old]);
            if (we != hipSuccess) {
                status = kevd_hip_fail(m, we, "ensemble status wait", __FILE__, __LINE__);
                break;
            }
            if (hflag[old]) break;
        }
    }
    if (status == KEVD_OK) {
        k_ens_reduce_d<T><<<nch, 256, 0, s>>>(dP, dchunks, 1);
        k_ens_final_err<T><<<nprob, 64, 0, s>>>(dP);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) status = kevd_hip_fail(m, le, "ensemble final error", __FILE__, __LINE__);
    }
    hipError_t se = hipStreamSynchronize(s);
    if (status == KEVD_OK && se != hipSuccess) status = kevd_hip_fail(m, se, "ensemble sync", __FILE__, __LINE__);
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    *used_graph = use_graph ? 1 : 0;
    return status;
}

// Autotuner cache: (device, n, dp, rank, precision, tol, eps, candidates) -> block size.
std::mutex tune_mu;
// key: device, n, padded dim, rank, precision, tol, eps, sorted candidate set
typedef std::tuple<int, int64_t, int32_t, int64_t, int, double, double, std::vector<int32_t>> TuneKey;
std::map<TuneKey, int32_t> tune_cache;

}  // namespace

extern "C" kevd_status kevd_ensemble_arpc(kevd_ctx** ctxs, int32_t nprob, const kevd_arpc_params* params,
                                          kevd_arpc_result* results, int32_t* used_graph)
{
    TraceRange trace_range("kevd_ensemble_arpc");
    if (!ctxs || nprob < 1 || !params || !results || !ctxs[0]) return KEVD_EINVAL;
    kevd_ctx* m = ctxs[0];
    for (int i = 0; i < nprob; ++i) {
        if (!ctxs[i]) return kevd_fail(m, KEVD_EINVAL, "ensemble: null context " + std::to_string(i));
        if (ctxs[i]->device != m->device)
            return kevd_fail(m, KEVD_EINVAL, "ensemble: every context must be on the first one's device");
        if (params[i].prec != params[0].prec)
            return kevd_fail(m, KEVD_EINVAL, "ensemble: every problem must use the same precision");
        if (params[i].inject_pivots)
            return kevd_fail(m, KEVD_EINVAL, "ensemble: injected pivots are a single-problem feature");
        if (params[i].regenerate)
            return kevd_fail(m, KEVD_EINVAL, "ensemble: the LM form (regenerate) is a single-problem feature");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return kevd_fail(m, KEVD_EINVAL, "ensemble: a context appears twice");
    }
    std::vector<int> maxr((size_t)nprob);
    for (int i = 0; i < nprob; ++i) {
        if (arpc_validate(ctxs[i], &params[i], &results[i], &maxr[i]) != KEVD_OK)
            return kevd_fail(m, KEVD_EINVAL, "ensemble problem " + std::to_string(i) + ": " + ctxs[i]->err);
    }
    KEVD_HIP(m, hipSetDevice(m->device));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<ArpcBufs> B((size_t)nprob);
    for (int i = 0; i < nprob; ++i) {
        ctxs[i]->tm = kevd_timings{};
        prof_reset(ctxs[i]);
        if (arpc_prepare(ctxs[i], &params[i], maxr[i], &B[i]) != KEVD_OK)
            return kevd_fail(m, KEVD_EHIP, "ensemble problem " + std::to_string(i) + ": " + ctxs[i]->err);
    }
    for (int i = 0; i < nprob; ++i) KEVD_HIP(m, hipStreamSynchronize(ctxs[i]->stream));
    int graph = 0;
    kevd_status s = params[0].prec == KEVD_MIXED ? ens_rounds<float>(m, ctxs, nprob, params, B, maxr, &graph)
                                                 : ens_rounds<double>(m, ctxs, nprob, params, B, maxr, &graph);
    if (s != KEVD_OK) return s;
    if (used_graph) *used_graph = graph;
    auto t1 = std::chrono::steady_clock::now();
    double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    for (int i = 0; i < nprob; ++i) {
        if (arpc_finish(ctxs[i], &params[i], maxr[i], B[i], &results[i]) != KEVD_OK)
            return kevd_fail(m, KEVD_EHIP, "ensemble problem " + std::to_string(i) + ": " + ctxs[i]->err);
        ctxs[i]->tm.arpc_ms = ms;
    }
    return KEVD_OK;
}

extern "C" kevd_status kevd_ensemble_evd(kevd_ctx** ctxs, int32_t nprob, const kevd_evd_params* p,
                                         kevd_evd_result* results, int32_t workers)
{
    TraceRange trace_range("kevd_ensemble_evd");
    if (!ctxs || nprob < 1 || !p || !results || !ctxs[0]) return KEVD_EINVAL;
    kevd_ctx* m = ctxs[0];
    for (int i = 0; i < nprob; ++i)
        if (!ctxs[i]) return kevd_fail(m, KEVD_EINVAL, "ensemble: null context " + std::to_string(i));
    int nw = workers > 0 ? workers : 8;
    nw = std::min(nw, (int)nprob);
    // The leading-eigenpair solves run as one batched solve (every problem
    // meets there, so each needs its own thread); up to EIG_BATCH_MAX
    // problems, leading pairs only.  On the 48-problem ensemble the kEVDs
    // take 80-82 ms against 89 ms with each problem's own solve on 8
    // threads.  KEVD_ENS_EIG_BATCH=0 in the environment keeps the own solves.
    constexpr int EIG_BATCH_MAX = 128;
    const char* want = std::getenv("KEVD_ENS_EIG_BATCH");
    const bool batch =
        !p->all_eigenvalues && nprob >= 2 && nprob <= EIG_BATCH_MAX && !(want && want[0] == '0');
    EigBatch* eb = batch ? eig_batch_create(m, nprob, p->k_out) : nullptr;
    if (eb) {
        nw = nprob;
        for (int i = 0; i < nprob; ++i) {
            ctxs[i]->eig_batch = eb;
            ctxs[i]->eig_batch_slot = i;
        }
    }
    // The Sinkhorn iterations first, batched over the problems on this
    // thread (KEVD_ENS_SK_BATCH=0 keeps each problem's own iteration); each
    // problem's kevd_evd then starts from its scaling.
    const char* want_sk = std::getenv("KEVD_ENS_SK_BATCH");
    int sk_batched = 0;
    if (!(want_sk && want_sk[0] == '0')) {
        kevd_status bs = ens_sinkhorn(ctxs, nprob, p, &sk_batched);
        if (bs != KEVD_OK) {
            for (int i = 0; i < nprob; ++i) ctxs[i]->sk_pre.valid = false;
            if (eb) {
                for (int i = 0; i < nprob; ++i) {
                    ctxs[i]->eig_batch = nullptr;
                    ctxs[i]->eig_batch_slot = -1;
                }
                eig_batch_destroy(eb);
            }
            return bs;
        }
    }
    std::vector<kevd_status> st((size_t)nprob, KEVD_OK);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= nprob) break;
            st[i] = kevd_evd(ctxs[i], p, &results[i]);
            if (eb) eig_batch_leave(eb, i);  // a problem that never joined (error, full solve)
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    // a scaling its problem's kevd_evd did not take (it failed before) is
    // not left for a later call
    for (int i = 0; i < nprob; ++i) ctxs[i]->sk_pre.valid = false;
    if (eb) {
        for (int i = 0; i < nprob; ++i) {
            ctxs[i]->eig_batch = nullptr;
            ctxs[i]->eig_batch_slot = -1;
        }
        eig_batch_destroy(eb);
    }
    for (int i = 0; i < nprob; ++i)
        if (st[i] != KEVD_OK) {
            std::string msg = "ensemble problem " + std::to_string(i) + ": " + ctxs[i]->err;
            return kevd_fail(m, st[i], msg);
        }
    return KEVD_OK;
}

extern "C" kevd_status kevd_autotune_block(kevd_ctx* ctx, const kevd_arpc_params* base,
                                           const int32_t* candidates, int32_t ncand, int32_t* best_block,
                                           double* cols_per_s)
{
    TraceRange trace_range("kevd_autotune_block");
    if (!ctx || !base || !candidates || ncand < 1 || !best_block)
        return kevd_fail(ctx, KEVD_EINVAL, "autotune: null argument or no candidates");
    std::vector<int32_t> cand_set(candidates, candidates + ncand);
    std::sort(cand_set.begin(), cand_set.end());
    cand_set.erase(std::unique(cand_set.begin(), cand_set.end()), cand_set.end());
    TuneKey key(ctx->device, ctx->n, ctx->dp, base->rank, (int)base->prec, base->tol, base->eps, cand_set);
    if (!cols_per_s) {
        std::lock_guard<std::mutex> g(tune_mu);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            *best_block = it->second;
            return KEVD_OK;
        }
    }
    int32_t best = candidates[0];
    double best_rate = -1.0;
    for (int c = 0; c < ncand; ++c) {
        int32_t b = candidates[c];
        if (b < 1 || b > BMAX) return kevd_fail(ctx, KEVD_EINVAL, "autotune: candidate block outside [1, 128]");
        kevd_arpc_params p = *base;
        p.block = b;
        p.inject_pivots = nullptr;
        p.n_inject = 0;
        p.max_rounds = 0;
        int mr = (int)(4 * ((p.rank + b - 1) / b) + 64);
        std::vector<int64_t> piv((size_t)p.rank);
        std::vector<int32_t> acc((size_t)mr);
        std::vector<double> err((size_t)mr);
        kevd_arpc_result r{};
        r.pivots = piv.data();
        r.accepted = acc.data();
        r.trace_err = err.data();
        r.rounds_cap = mr;
        double best_ms = INFINITY;
        int64_t cols = 0;
        // the first run also sizes the buffers; the faster of two counts
        for (int rep = 0; rep < 2; ++rep) {
            KEVD_TRY(kevd_arpc(ctx, &p, &r));
            best_ms = std::min(best_ms, ctx->tm.arpc_ms);
            cols = r.rank;
        }
        double rate = best_ms > 0.0 ? (double)cols / (best_ms * 1e-3) : 0.0;
        if (cols_per_s) cols_per_s[c] = rate;
        if (rate > best_rate) {
            best_rate = rate;
            best = b;
        }
    }
    *best_block = best;
    std::lock_guard<std::mutex> g(tune_mu);
    tune_cache[key] = best;
    return KEVD_OK;
}
