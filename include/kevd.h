/*
 * kevd.h -- C ABI of libkevd, the MI355X (gfx950) implementation of
 * accelerated randomly pivoted Cholesky (ARPC), the bistochastic kernel
 * EVD built on its factor, and the Gaussian-kernel bandwidth sweep.
 *
 * The reference exposes these stages as three standalone MPI programs
 * whose parameters are compile-time constants and whose interfaces are
 * HDF5 files (ks_arpclm.cu:566-1152, ks_kevd.cu:419-2403,
 * ks_bandwidth.cu:339-667).  Each entry point below replaces one stage or
 * one step of a stage; the comment above it cites what it replaces.
 *
 * Conventions
 *   - All device memory is owned by the context.  Host arrays belong to the
 *     caller and are copied in or out; capacities are the caller's job.
 *   - Every call returns a kevd_status.  No call exits the process.  The
 *     drivers map statuses to the reference's exit codes (2 = non-positive
 *     degree, 3 = potrf failure, 4 = non-positive q, 5 = eigensolver).
 *   - One context is used by one host thread at a time.  Work is ordered on
 *     the context's own HIP stream; every call returns with its host outputs
 *     complete.  Several contexts may share one device.
 *   - Point sets are row-major n x dim.  The kernel is
 *     K(i,j) = exp(-||x_i - x_j||^2 / eps).  The reference writes the same
 *     kernel as exp(-sum(dx^2) / (Nq * bw)) (ks_arpclm.cu:277), so
 *     eps = Nq * bw.
 */
#ifndef KEVD_H
#define KEVD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kevd_ctx kevd_ctx;

/* Status codes.  2..5 carry the meaning of the reference's MPI_Abort codes
 * (ks_kevd.cu:1170-1181 -> 2, 1510-1520 -> 3, 1692-1703 -> 4,
 * 1904-1913 -> 5). */
typedef enum {
    KEVD_OK = 0,
    KEVD_EINVAL = 1,             /* bad argument or call order            */
    KEVD_ENONPOS_DEGREE = 2,     /* a degree d~_i <= 0                    */
    KEVD_EPOTRF = 3,             /* Cholesky of the shifted Gram failed   */
    KEVD_ENONPOS_Q = 4,          /* a q~_i <= 0 (or a Sinkhorn failure)   */
    KEVD_EEIG = 5,               /* symmetric eigensolver failed          */
    KEVD_EHIP = 6,               /* HIP runtime or library error          */
    KEVD_ENOMEM = 7              /* device or host allocation failed      */
} kevd_status;

typedef enum {
    KEVD_FP64 = 0,   /* every step in FP64 (the plain form)                 */
    KEVD_MIXED = 1   /* FP32 MFMA distances / exp / downdate / Gram, FP64
                        pivot block, residual diagonal, scalings, eigensolve */
} kevd_precision;

/* ---------------------------------------------------------------- context */

/* Replaces the per-rank device setup of ks_arpclm.cu:683-689 (stream and
 * handle creation) for one GPU. */
kevd_status kevd_create(int device, kevd_ctx** out);
void        kevd_destroy(kevd_ctx* ctx);
const char* kevd_last_error(const kevd_ctx* ctx);
const char* kevd_status_string(kevd_status s);

/* Upload a point set (row-major n x dim).  Replaces the h5read of /u and
 * its cudaMemcpy (ks_arpclm.cu:652-655, 693-696).  The points are centred
 * on the device (the kernel is translation invariant); FP64 and FP32 copies
 * plus FP64 squared norms are kept resident.  Any factor held by the
 * context is dropped. */
kevd_status kevd_set_points(kevd_ctx* ctx, const double* X, int64_t n, int32_t dim);
kevd_status kevd_set_points_f32(kevd_ctx* ctx, const float* X, int64_t n, int32_t dim);

/* Median of the squared pairwise distances ||x_i - x_j||^2, i != j, in
 * FP64.  If nsample <= 0 or nsample >= n(n-1)/2 the median is exact over
 * all pairs i < j (allowed for n <= 16384); otherwise it is the median of
 * nsample pairs drawn with Philox4x32-10 from `seed`.  An even count takes
 * the mean of the two central order statistics.  The reference has no
 * median code: it hard-codes bw = 0.8 (ks_arpclm.cu:602). */
kevd_status kevd_median_sqdist(kevd_ctx* ctx, uint64_t seed, int64_t nsample,
                               double* eps_out);

/* Kernel block K(rows, cols), written row-major nr x nc into K_out (host,
 * FP64).  Replaces compute_kblock<T> (ks_arpclm.cu:242-280,
 * ks_kevd.cu:302-340).  The mixed precision runs the distance product on
 * FP32 MFMA with FP64 norms; FP64 runs it on FP64 MFMA. */
kevd_status kevd_kernel_block(kevd_ctx* ctx, const int64_t* rows, int64_t nr,
                              const int64_t* cols, int64_t nc, double eps,
                              kevd_precision prec, double* K_out);

/* ------------------------------------------------------------------ ARPC */

typedef struct {
    double eps;               /* kernel scale                               */
    int64_t rank;             /* target rank r (the factor has r columns)   */
    int32_t block;            /* proposals per round b, 1..128              */
    uint64_t seed;            /* Philox key                                 */
    double tol;               /* rejection tolerance on H_ii (ref 1e-10)    */
    kevd_precision prec;
    int32_t max_rounds;       /* guard; <= 0 means 4*ceil(r/b)+64           */
    const int64_t* inject_pivots; /* optional fixed pivot list (host)       */
    int64_t n_inject;         /* its length (>= rank when given)            */
    int32_t regenerate;       /* 1: the LM form (below); 0: store F         */
} kevd_arpc_params;

typedef struct {
    int64_t rank;             /* realised rank (<= params.rank)             */
    int32_t rounds;           /* rounds run                                 */
    int64_t* pivots;          /* out, capacity >= params.rank, acceptance order */
    int32_t* accepted;        /* out, per round, capacity >= max_rounds     */
    double* trace_err;        /* out, per round, tr(K - F F^T) / tr(K)      */
    int32_t rounds_cap;       /* capacity of accepted / trace_err           */
} kevd_arpc_result;

/* Accelerated randomly pivoted Cholesky.  Replaces the round loop of
 * ks_arpclm.cu:793-1046 (propose_block 193-224, the H overlap 833-874,
 * rejection_sample 300-408, the Step-3 row loop 929-995 and compute_dsum
 * 536-564).  The factor F (n x rank, K ~ F F^T) stays on the device.
 * Rounds run until the rank reaches params.rank (the last accepted block is
 * truncated), the residual mass is zero, or max_rounds is reached. */
kevd_status kevd_arpc(kevd_ctx* ctx, const kevd_arpc_params* p, kevd_arpc_result* res);
/* regenerate = 1 runs the reference's own LM form (ks_arpclm.cu:793-1046:
 * L^-1 kept, K(R,S) regenerated from the points every round, M and the
 * appended L^-1 block row of :907-914 and :480-534) instead of storing F:
 * device memory O(n + r^2) instead of n * r, for point sets whose factor does
 * not fit.  Same rounds, sampling and rejection; pivots, accepted counts and
 * trace errors come out as usual, L^-1 through kevd_get_linv; the context
 * then holds no factor (kevd_get_factor and kevd_evd report EINVAL until a
 * factor is set, e.g. by kevd_arpc with the pivots injected). */

/* Copy the factor to the host: n x rank row-major with leading dimension ld
 * (>= rank), as FP64 (`as` = KEVD_FP64) or FP32 (`as` = KEVD_MIXED). */
kevd_status kevd_get_factor(kevd_ctx* ctx, void* F_host, int64_t ld, kevd_precision as);

/* Install an external factor (n x r row-major, leading dimension ld) in the
 * given precision, for kevd_evd.  n must match the point set. */
kevd_status kevd_set_factor(kevd_ctx* ctx, const void* F_host, int64_t n, int64_t r,
                            int64_t ld, kevd_precision prec);

/* The inverse Cholesky factor of K(S,S) in the reference's meaning
 * (ks_arpclm.cu:1048-1096): r x r lower triangular, column-major, FP64,
 * with K(S,S) = L L^T and pivots in acceptance order.  Here L = F(S,:). */
kevd_status kevd_get_linv(kevd_ctx* ctx, double* Linv_host);

/* ------------------------------------------------------------------ kEVD */

typedef enum {
    KEVD_NORM_REF = 0,       /* two-step d~, q~ form of ks_kevd.cu            */
    KEVD_NORM_SINKHORN = 1   /* symmetric Sinkhorn scaling s .* (K^ s) = 1    */
} kevd_norm;

typedef struct {
    int32_t k_out;            /* leading eigenvectors returned              */
    int32_t norm;             /* kevd_norm                                  */
    double sinkhorn_tol;      /* max row-sum deviation target (Sinkhorn)    */
    int32_t sinkhorn_max_iter;
    kevd_precision prec;
    int32_t all_eigenvalues;  /* 1: all r eigenvalues (full symmetric
                                 eigensolve, the reference's /lam);
                                 0: the leading ones only, by block subspace
                                 iteration; lam entries not computed are NaN */
} kevd_evd_params;

typedef struct {
    double* lam;              /* out, rank values, ascending (FP64)         */
    float* U;                 /* out, n x k_out row-major, columns ascending,
                                 NULL to skip the copy                      */
    double rowsum_dev;        /* max_i |(P 1)_i - 1|                        */
    double shift;             /* the Cholesky shift s = DBL_EPSILON tr(A)   */
    double min_degree;        /* min d~ (ref) or min s (Sinkhorn)           */
    int32_t sinkhorn_iters;
} kevd_evd_result;

/* Bistochastic kernel EVD on the resident factor.  Replaces the four
 * streaming passes and two dense phases of ks_kevd.cu:755-2308 with the
 * F-form: d~ = F(F^T 1), q~ = F(F^T d~^-1), A = F^T D~^-2 F,
 * B = F^T Q~^-1 F, R = chol(A + sI), W = R B R^T, W = V diag(lam) V^T,
 * U = D~^-1 F R^-1 V_k. */
kevd_status kevd_evd(kevd_ctx* ctx, const kevd_evd_params* p, kevd_evd_result* res);

/* U on the device after kevd_evd, as FP32 n x k_out row-major. */
kevd_status kevd_get_u(kevd_ctx* ctx, float* U_host);

/* The intermediate record of the kEVD: the reference's dump_h5 switch and
 * its kevd_dump file's datasets (ks_kevd.cu:484, 531-542, 809-815, 935-941,
 * 1183-1218, 1274-1280, 1562-1576, 1705-1744, 1945-1969, 2164-2170).  With
 * on = 1, every later kevd_evd with KEVD_NORM_REF keeps on the host, in the
 * reference's basis K(:,S) with K(S,S) = L L^T (the factor must come from
 * kevd_arpc, else kevd_evd returns KEVD_EINVAL):
 *   c1, p1, cq, pq (r), dtil, qtil (n), lam (r),
 *   Atil (lower triangle, strict upper zero), A, R (upper, strict lower
 *   zero), G, Btil, W (r x r), V (r x r after a full eigensolve, else the
 *   r x k_out leading block), Z (r x k_out),
 * FP64, matrices column-major (the reference's flat order).  U is
 * kevd_get_u.  Each entry is copied to the host as it is formed, which
 * synchronises the stream: the timings of a recorded run include it.
 * on = 0 stops recording and frees the record.  A Sinkhorn run records
 * nothing. */
kevd_status kevd_set_dump(kevd_ctx* ctx, int32_t on);

/* One recorded array by its dataset name.  *count: in, the capacity of out;
 * out, the number of values.  out = NULL only reports the number.  An
 * unknown or unrecorded name, or too small a capacity, is KEVD_EINVAL. */
kevd_status kevd_get_intermediate(kevd_ctx* ctx, const char* name, double* out, int64_t* count);

/* ------------------------------------------------------------- bandwidth */

/* S(eps_j) = sum_{i in rows} sum_{all j'} exp(-||x_i - x_j'||^2 / eps_j),
 * FP64, all neps values from one fused launch.  Replaces dmat_comp and the
 * per-bandwidth kernelsum_comp launches (ks_bandwidth.cu:181-325, 537-576).
 * The distance tile runs on MFMA (FP32 with FP64 norms for KEVD_MIXED). */
kevd_status kevd_bandwidth_sweep(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                 const double* eps, int32_t neps,
                                 kevd_precision prec, double* S_out);

/* kevd_bandwidth_sweep without any device allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and kernels as kevd_bandwidth_sweep; every intermediate
 * (row indices, gathered rows, eps coefficients, per-block partial sums)
 * lives in a fixed scratch array of KEVD_SCRATCH_MAX_BYTES per device, which
 * the library holds until the process ends.  The points must be resident
 * (kevd_set_points).  A call whose intermediates exceed the array returns
 * KEVD_EINVAL with the byte count (about 0.5 KB per row at dim 64 in FP64,
 * plus 1 KB per partial-sum block: some 50,000 rows).  Every context on a
 * device shares that device's array: calls that overlap in time on one
 * device race and must be serialised by the caller. */
#define KEVD_SCRATCH_MAX_BYTES (32u << 20)
kevd_status kevd_bandwidth_sweep_noalloc(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                         const double* eps, int32_t neps,
                                         kevd_precision prec, double* S_out);

/* Post-processing of a sweep, ks_bandwidth.cu:609-642: logS_i = log(S_i),
 * the centred-difference slope over log eps, the first maximum of the slope,
 * eopt = eps[idmax+1] and mdim = 2 * slope[idmax].  slope has neps-2
 * entries.  Pure host arithmetic. */
kevd_status kevd_bandwidth_post(const double* eps, const double* S, int32_t neps,
                                double* logS, double* slope, double* eopt,
                                double* mdim);

/* --------------------------------------------------------------- KS data */

typedef struct {
    double L;                 /* domain length (configs: 22, 44, 88)       */
    int32_t M;                /* grid points (even)                        */
    double dt;                /* ETDRK4 step                               */
    int64_t n_discard;        /* steps discarded before the first sample   */
    int32_t stride;           /* steps between samples                     */
    uint64_t seed;            /* initial condition 0.1 * N(0,1) per point  */
    double ic_scale;          /* 0.1                                       */
} kevd_ks_params;

/* Kuramoto-Sivashinsky snapshots u_t = -u u_x - u_xx - u_xxxx on a periodic
 * grid, ETDRK4 with 3/2 dealiasing and contour-integral coefficients
 * (ks_datagen.py:5-89).  Writes n snapshots, row-major n x M, FP64.
 * Host-side generator (the trajectory is sequential in time). */
kevd_status kevd_ks_generate(const kevd_ks_params* p, int64_t n, double* out);

/* Delay embedding in the reference's point order (ks_arpclm.cu:263-268):
 * from a series of ndata snapshots (row-major ndata x M, FP64) build
 * N*M points of dimension nq, N = ndata - nq + 1, point g*N + t being the
 * window series[t .. t+nq-1][g].  Host-side. */
kevd_status kevd_delay_embed(const double* series, int64_t ndata, int32_t M, int32_t nq,
                             double* out);

/* --------------------------------------------------------------- timing */

typedef struct {
    double arpc_ms;           /* wall time of the last kevd_arpc          */
    double evd_ms;            /* wall time of the last kevd_evd           */
    double downdate_ms;       /* summed device time of the downdate kernel */
    int64_t downdate_launches;
    double downdate_flop;     /* algorithmic flop of those launches        */
    double gram_ms;           /* summed device time of the weighted Gram   */
    int64_t gram_launches;
    double gram_flop;
    double eig_ms;            /* device time of the r x r eigensolve       */
    double sample_ms;         /* pivot sampling kernels                    */
    double rejection_ms;      /* H formation + rejection Cholesky          */
    double matvec_ms;         /* streaming matvec passes                   */
    double dense_ms;          /* r x r dense phase other than eigensolve   */
    double u_ms;              /* U GEMM                                    */
    int32_t eig_iters;        /* subspace iterations of the leading-eigenpair
                                 solve (0: full divide-and-conquer)         */
    int32_t eig_householder;  /* 1 if that solve fell back to Householder QR */
    int32_t eig_full_fallback; /* 1 if subspace iteration had not converged
                                  after its step limit and the full
                                  divide-and-conquer solve gave the pairs  */
    int32_t eig_batched;      /* 1 if the pairs came from the ensemble's
                                 batched solve                              */
    int32_t eig_orth_passes;  /* Cholesky QR passes of the subspace
                                 iteration (0 for the other solves)         */
    int32_t sinkhorn_batched; /* 1 if the Sinkhorn iteration ran in the
                                 ensemble's batched launches               */
} kevd_timings;

/* Enable per-kernel HIP event timing on the context's stream (adds one
 * event pair per launch) and read the last call's figures. */
kevd_status kevd_set_profiling(kevd_ctx* ctx, int enable);
kevd_status kevd_get_timings(const kevd_ctx* ctx, kevd_timings* out);

/* ------------------------------------------------------------ ensembles
 * Many independent problems on one GPU.  Each problem is an ordinary
 * context (points and eps set through the calls above); all contexts must be
 * on the same device.  The reference runs one problem per program launch
 * (ks_arpclm.cu main, ks_kevd.cu main); these calls batch that loop.
 *
 * kevd_ensemble_arpc: every problem's ARPC (as kevd_arpc, params[i] and
 * results[i] per problem, one precision for all, no injected pivots) runs
 * as batched rounds: one launch per round stage for all problems, the
 * downdate as a persistent work-queue kernel, the round captured once as a
 * hipGraph and replayed.  Pivots and factors equal kevd_arpc's bit for bit.
 * Each context then holds its factor, as after kevd_arpc.  *used_graph
 * (optional) is 1 when the rounds ran as graph launches.  Errors are
 * reported on ctxs[0]. */
kevd_status kevd_ensemble_arpc(kevd_ctx** ctxs, int32_t nprob, const kevd_arpc_params* params,
                               kevd_arpc_result* results, int32_t* used_graph);
/* kevd_ensemble_evd: kevd_evd on every context with the same parameters,
 * each problem on its own context stream.  With leading eigenpairs only
 * (all_eigenvalues = 0) and 2..128 problems, every problem runs on its own
 * host thread and the leading-eigenpair solves of all problems run as one
 * batched subspace iteration (strided-batched GEMMs, one-workgroup Cholesky
 * per problem, batched Jacobi Rayleigh-Ritz, over W zero-padded to the
 * largest rank); a problem the batch does not converge solves alone.
 * Otherwise, or with KEVD_ENS_EIG_BATCH=0 in the environment, the problems
 * run on `workers` host threads (<= 0: 8) with their own solves.
 * results[i] as kevd_evd's; errors are reported on ctxs[0]. */
kevd_status kevd_ensemble_evd(kevd_ctx** ctxs, int32_t nprob, const kevd_evd_params* p,
                              kevd_evd_result* results, int32_t workers);
/* kevd_autotune_tiles: the output tile of the bf16x6 Gram (256 x 256 on 8
 * waves, one workgroup per CU, or 128 x 128 on 4 waves, two per CU) for
 * this context's FP32 factor, by timing both on the factor itself (one
 * warm-up and two timed runs each, the faster counts).  Later kevd_evd
 * calls on the context use the choice.  Without gram_ms the choice is
 * cached per (device, n, rank, leading dimension) for the process and a
 * cached answer is returned without measuring; with it (2 entries), both
 * are measured and their times written (256 first).  An FP64 factor, or
 * KEVD_FP32_MFMA=1, has no choice: *gram_tile = 0.  The tuning tile shapes
 * have no counterpart in the reference, which sizes launches with the
 * occupancy API (ks_bandwidth.cu:436-474). */
kevd_status kevd_autotune_tiles(kevd_ctx* ctx, int32_t* gram_tile, double* gram_ms);
/* Set the bf16x6 Gram's tile for this context without measuring: 0 (the
 * default, 256), 128 or 256; anything else is KEVD_EINVAL. */
kevd_status kevd_set_gram_tile(kevd_ctx* ctx, int32_t gram_tile);

/* kevd_autotune_block: the block size among candidates[0..ncand) with the
 * highest factor-columns-per-second of kevd_arpc on this context's points
 * (base supplies eps, rank, seed, tol, precision; two timed runs per
 * candidate, the faster one counts).  Without cols_per_s the choice is
 * cached per (device, n, padded dim, rank, precision, tol, eps, candidate
 * set) for the process
 * and a cached answer is returned without measuring; with it, every
 * candidate is measured and its rate written.  The run leaves the context
 * holding the last candidate's factor. */
kevd_status kevd_autotune_block(kevd_ctx* ctx, const kevd_arpc_params* base, const int32_t* candidates,
                                int32_t ncand, int32_t* best_block, double* cols_per_s);

/* ---------------------------------------------------------------- peaks
 * Box-measured ceilings for roofline fractions (no counterpart in the
 * reference, which reports none): the rate of each instruction or access
 * pattern a hot kernel is bound by, sustained over every CU of the
 * context's device.  copy_bytes in [16 MiB, 64 GiB] sizes the HBM copy
 * and read (two buffers of that size are allocated and freed inside the
 * call). */
typedef struct {
    double mfma_f32_tflops;     /* v_mfma_f32_16x16x4_f32, TFLOP/s          */
    double mfma_f32_32_tflops;  /* v_mfma_f32_32x32x2_f32, TFLOP/s          */
    double mfma_f64_tflops;     /* v_mfma_f64_16x16x4_f64, TFLOP/s          */
    double exp_f32_per_s;       /* v_exp_f32 results per second            */
    double hbm_copy_gbs;        /* device copy, read + written GB/s        */
    double hbm_read_gbs;        /* streaming read of one buffer, GB/s      */
    double mfma_bf16_tflops;    /* v_mfma_f32_16x16x32_bf16, TFLOP/s        */
    double mfma_bf16_32_tflops; /* v_mfma_f32_32x32x16_bf16, TFLOP/s        */
} kevd_peaks;
kevd_status kevd_measure_peaks(kevd_ctx* ctx, int64_t copy_bytes, kevd_peaks* out);

#ifdef __cplusplus
}
#endif

#endif /* KEVD_H */
