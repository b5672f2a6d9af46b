// gemm.h -- LDS-staged MFMA block GEMM C += A B^T used by every GEMM-shaped
// step of libkevd (distance products, the Cholesky downdate, the pivot-block
// residual, the U product).  A and B are row-major with K contiguous, which
// is how point sets X, the factor F and gathered pivot rows F(S,:) are laid
// out in HBM, so no operand is ever transposed in memory.
//
// Block of 256 threads = 4 waves arranged WM x WN; each wave owns TM x TN
// MFMA tiles of MT x MT.  K advances in chunks of KC through a double
// buffered LDS stage (one barrier per chunk); global loads of chunk c+1 are
// in flight while chunk c is multiplied.
#pragma once

#include "common.h"

// One element stored through a buffer resource (byte offset off); a store
// outside the resource's range is dropped by the hardware.
__device__ inline void buffer_store(float v, __amdgpu_buffer_rsrc_t rs, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off, 0, 0);
}
__device__ inline void buffer_store(double v, __amdgpu_buffer_rsrc_t rs, uint32_t off)
{
    typedef __attribute__((ext_vector_type(2))) unsigned u2;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, off, 0, 0);
}

template <typename T, int BM, int BN, int KC, int WM, int WN, int TM, int TN, typename MF = Mfma<T>>
struct NTGemm {
    typedef MF M;
    typedef typename M::acc_t acc_t;
    typedef typename M::vec_t vec_t;
    static constexpr int MT = M::MT;
    static constexpr int KSUB = M::NH * M::VEC;            // K values per sub-step
    static constexpr int NTM = TM, NTN = TN, NBM = BM, NBN = BN, NKC = KC;
    static constexpr int EPV = 16 / (int)sizeof(T);       // elements per 16 B
    static constexpr int PITCH = KC + EPV;                // LDS row pitch (elements)
    static constexpr int VPR = KC / EPV;                  // 16 B vectors per row
    static constexpr int RPP = 256 / VPR;                 // rows per load pass
    static constexpr int NVA = BM / RPP;
    static constexpr int NVB = BN / RPP;
    static constexpr int STAGE = (BM + BN) * PITCH;
    static constexpr int LDS_ELEMS = 2 * STAGE;
    static_assert(WM * WN == 4, "4 waves");
    static_assert(WM * TM * MT == BM && WN * TN * MT == BN, "tile shape");
    static_assert(KC % KSUB == 0 && BM % RPP == 0 && BN % RPP == 0, "load shape");

    acc_t acc[TM][TN];
    int wm, wn, h, li;
    int nact = TN;  // column tiles of this wave that hold live columns (wave-uniform)

    __device__ inline void init()
    {
        int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        wm = w / WN;
        wn = w % WN;
        h = lane / MT;
        li = lane % MT;
        zero();
    }

    __device__ inline void zero()
    {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int v = 0; v < M::ACC; ++v) acc[a][b][v] = (T)0;
    }

    // Row / column (within the BM x BN block tile) of accumulator value v of
    // tile (a, b) held by this lane.
    __device__ inline int row_of(int a, int v) const
    {
        return wm * TM * MT + a * MT + M::row_lane(h) + M::row_off(v);
    }
    __device__ inline int col_of(int b) const { return wn * TN * MT + b * MT + li; }

    // Load rows [0, nvalid) x columns [k0, k0 + KC) of src into registers;
    // rows past nvalid and columns past kvalid read as zero (so stale or
    // uninitialised memory beyond the live columns never reaches the MFMA).
    template <int NV>
    __device__ static inline void gload(uint4 (&r)[NV], const T* src, int64_t ld,
                                        int64_t nvalid, int64_t k0, int64_t kvalid)
    {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            int lin = threadIdx.x + 256 * j;
            int row = lin / VPR, cv = lin % VPR;
            int64_t kk = k0 + cv * EPV;
            if (row < nvalid && kk < kvalid) {
                r[j] = *reinterpret_cast<const uint4*>(src + (int64_t)row * ld + kk);
                if (kk + EPV > kvalid) {
                    T* e = reinterpret_cast<T*>(&r[j]);
#pragma unroll
                    for (int q = 0; q < EPV; ++q)
                        if (kk + q >= kvalid) e[q] = (T)0;
                }
            } else {
                r[j] = make_uint4(0, 0, 0, 0);
            }
        }
    }

    template <int NV>
    __device__ static inline void sstore(const uint4 (&r)[NV], T* lds)
    {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            int lin = threadIdx.x + 256 * j;
            int row = lin / VPR, cv = lin % VPR;
            *reinterpret_cast<uint4*>(lds + row * PITCH + cv * EPV) = r[j];
        }
    }

    // Restrict the MFMAs to the first `live_cols` block columns: column tiles
    // past them are skipped (their accumulators stay zero).
    __device__ inline void set_live_cols(int live_cols)
    {
        int first = wn * TN * MT;
        int t = (live_cols - first + MT - 1) / MT;
        nact = t < 0 ? 0 : (t > TN ? TN : t);
    }

    // acc += As Bs^T over one KC chunk held in LDS.
    __device__ inline void compute(const T* As, const T* Bs)
    {
#pragma unroll
        for (int s = 0; s < KC / KSUB; ++s) {
            vec_t a[TM], b[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[t] = *reinterpret_cast<const vec_t*>(
                    As + (wm * TM * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
#pragma unroll
            for (int t = 0; t < TN; ++t)
                if (t < nact)
                    b[t] = *reinterpret_cast<const vec_t*>(
                        Bs + (wn * TN * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
#pragma unroll
            for (int y = 0; y < TN; ++y)
                if (y < nact)
#pragma unroll
                    for (int j = 0; j < M::VEC; ++j)
#pragma unroll
                        for (int x = 0; x < TM; ++x) M::mma(acc[x][y], M::get(a[x], j), M::get(b[y], j));
        }
    }

    // acc += A[0:BM, 0:K] B[0:BN, 0:K]^T.  A points at the block's first row
    // with leading dimension lda and arows valid rows (the rest read as 0);
    // likewise B.  Columns past K read as 0; both leading dimensions must
    // cover K rounded up to KC.  Ends with a barrier, so the LDS can be
    // reused by the caller.
    __device__ inline void mainloop(T* lds, const T* A, int64_t lda, int64_t arows,
                                    const T* B, int64_t ldb, int64_t brows, int64_t K)
    {
        int64_t nk = (K + KC - 1) / KC;
        if (nk <= 0) return;
        uint4 ra[NVA], rb[NVB];
        gload<NVA>(ra, A, lda, arows, 0, K);
        gload<NVB>(rb, B, ldb, brows, 0, K);
        sstore<NVA>(ra, lds);
        sstore<NVB>(rb, lds + BM * PITCH);
        __syncthreads();
        for (int64_t kc = 0; kc < nk; ++kc) {
            T* cur = lds + (kc & 1) * STAGE;
            T* nxt = lds + ((kc + 1) & 1) * STAGE;
            bool more = kc + 1 < nk;
            if (more) {
                gload<NVA>(ra, A, lda, arows, (kc + 1) * KC, K);
                gload<NVB>(rb, B, ldb, brows, (kc + 1) * KC, K);
            }
            compute(cur, cur + BM * PITCH);
            if (more) {
                sstore<NVA>(ra, nxt);
                sstore<NVB>(rb, nxt + BM * PITCH);
            }
            __syncthreads();
        }
    }

    // Like gload, for chunks that lie wholly below kvalid, through a buffer
    // resource over the operand's valid rows (rows_rsrc): a row past them
    // reads as zero by the hardware's range check, so the loads carry no
    // predicate, no branch and no 64-bit address arithmetic.
    __device__ static inline __amdgpu_buffer_rsrc_t rows_rsrc(const T* src, int64_t ld, int64_t nvalid, int nmax)
    {
        int64_t rows = nvalid < nmax ? (nvalid > 0 ? nvalid : 0) : nmax;
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(src), (short)0,
                                                 (int)(rows * ld * (int64_t)sizeof(T)), 0x00020000);
    }
    template <int NV>
    __device__ static inline void gload_full(uint4 (&r)[NV], __amdgpu_buffer_rsrc_t rs, int64_t ld, int64_t k0)
    {
        typedef __attribute__((ext_vector_type(4))) unsigned u4;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            int lin = threadIdx.x + 256 * j;
            int row = lin / VPR, cv = lin % VPR;
            u4 v = __builtin_amdgcn_raw_buffer_load_b128(
                rs, (uint32_t)(((int64_t)row * ld + k0 + cv * EPV) * (int64_t)sizeof(T)), 0, 0);
            r[j] = make_uint4(v.x, v.y, v.z, v.w);
        }
    }

    // acc += As Bs^T over one KC chunk for NACT live column tiles (a
    // compile-time count, so no branch splits the MFMA stream).  The operand
    // fragments of sub-step s + 1 are read from LDS while sub-step s
    // multiplies, and the MFMAs cycle over the accumulators so consecutive
    // ones are independent.
    template <int NACT>
    __device__ inline void compute_n(const T* As, const T* Bs)
    {
        constexpr int NS = KC / KSUB;
        // double-buffer the fragments only while they are small (<= 24
        // VGPRs per set); a wide sub-step (16x16 tiles: 64 MFMAs) hides the
        // LDS latency of the next one's reads on its own
        constexpr bool DB = (TM + NACT) * (int)sizeof(vec_t) / 4 <= 24;
        if constexpr (!DB) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                vec_t a[TM], b[NACT];
#pragma unroll
                for (int t = 0; t < TM; ++t)
                    a[t] = *reinterpret_cast<const vec_t*>(
                        As + (wm * TM * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
#pragma unroll
                for (int t = 0; t < NACT; ++t)
                    b[t] = *reinterpret_cast<const vec_t*>(
                        Bs + (wn * TN * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
#pragma unroll
                for (int j = 0; j < M::VEC; ++j)
#pragma unroll
                    for (int y = 0; y < NACT; ++y)
#pragma unroll
                        for (int x = 0; x < TM; ++x) M::mma(acc[x][y], M::get(a[x], j), M::get(b[y], j));
                // keep the next sub-step's reads below this one's MFMAs (hoisted,
                // both sets would be live at once)
                __builtin_amdgcn_sched_barrier(0);
            }
            return;
        }
        vec_t a[2][TM], b[2][NACT];
        auto frag = [&](int q, int s) {
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[q][t] = *reinterpret_cast<const vec_t*>(
                    As + (wm * TM * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
#pragma unroll
            for (int t = 0; t < NACT; ++t)
                b[q][t] = *reinterpret_cast<const vec_t*>(
                    Bs + (wn * TN * MT + t * MT + li) * PITCH + s * KSUB + h * M::VEC);
        };
        frag(0, 0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s + 1 < NS) frag((s + 1) & 1, s + 1);
#pragma unroll
            for (int j = 0; j < M::VEC; ++j)
#pragma unroll
                for (int y = 0; y < NACT; ++y)
#pragma unroll
                    for (int x = 0; x < TM; ++x)
                        M::mma(acc[x][y], M::get(a[s & 1][x], j), M::get(b[s & 1][y], j));
        }
    }

    // mainloop for NACT live column tiles.  Global loads run two chunks
    // ahead: while chunk c is multiplied from LDS stage c&1, chunk c+1 waits
    // in one register set (stored to the other stage after the multiply)
    // and chunk c+2 is loading into the other, so a load has two chunks of
    // MFMA work to arrive.  Chunks past the last full one mask K.
    template <int NACT>
    __device__ inline void mainloop_n(T* lds, const T* A, int64_t lda, int64_t arows,
                                      const T* B, int64_t ldb, int64_t brows, int64_t K)
    {
        int64_t nk = (K + KC - 1) / KC, nfull = K / KC;
        if (nk <= 0) return;
        // A block is at most BM x lda (BN x ldb) elements: 32-bit offsets
        const __amdgpu_buffer_rsrc_t rsA = rows_rsrc(A, lda, arows, BM), rsB = rows_rsrc(B, ldb, brows, BN);
        auto load = [&](uint4 (&ra)[NVA], uint4 (&rb)[NVB], int64_t c) {
            if (c < nfull) {
                gload_full<NVA>(ra, rsA, lda, c * KC);
                gload_full<NVB>(rb, rsB, ldb, c * KC);
            } else {
                gload<NVA>(ra, A, lda, arows, c * KC, K);
                gload<NVB>(rb, B, ldb, brows, c * KC, K);
            }
        };
        // wide fragment sets (16x16 tiles) leave no room for a second
        // register set: loads run one chunk ahead
        constexpr bool DEEP = (TM + NACT) * (int)sizeof(vec_t) / 4 <= 24;
        if constexpr (!DEEP) {
            uint4 ra[NVA], rb[NVB];
            load(ra, rb, 0);
            sstore<NVA>(ra, lds);
            sstore<NVB>(rb, lds + BM * PITCH);
            __syncthreads();
            for (int64_t c = 0; c < nk; ++c) {
                T* cur = lds + (c & 1) * STAGE;
                T* nxt = lds + ((c + 1) & 1) * STAGE;
                bool more = c + 1 < nk;
                if (more) load(ra, rb, c + 1);
                compute_n<NACT>(cur, cur + BM * PITCH);
                if (more) {
                    sstore<NVA>(ra, nxt);
                    sstore<NVB>(rb, nxt + BM * PITCH);
                }
                __syncthreads();
            }
            return;
        }
        uint4 ra0[NVA], rb0[NVB], ra1[NVA], rb1[NVB];
        load(ra0, rb0, 0);
        sstore<NVA>(ra0, lds);
        sstore<NVB>(rb0, lds + BM * PITCH);
        if (nk > 1) load(ra1, rb1, 1);
        __syncthreads();
        for (int64_t c = 0; c < nk; c += 2) {
            // even chunk c: LDS stage 0, c+1 in set 1, c+2 loads into set 0
            if (c + 2 < nk) load(ra0, rb0, c + 2);
            compute_n<NACT>(lds, lds + BM * PITCH);
            if (c + 1 < nk) {
                sstore<NVA>(ra1, lds + STAGE);
                sstore<NVB>(rb1, lds + STAGE + BM * PITCH);
            }
            __syncthreads();
            if (c + 1 >= nk) break;
            // odd chunk c+1: LDS stage 1, c+2 in set 0, c+3 loads into set 1
            if (c + 3 < nk) load(ra1, rb1, c + 3);
            compute_n<NACT>(lds + STAGE, lds + STAGE + BM * PITCH);
            if (c + 2 < nk) {
                sstore<NVA>(ra0, lds);
                sstore<NVB>(rb0, lds + BM * PITCH);
            }
            __syncthreads();
        }
    }

    // mainloop_n dispatched on the wave's live column tiles (nact).
    template <int N = TN>
    __device__ inline void mainloop_live(T* lds, const T* A, int64_t lda, int64_t arows,
                                         const T* B, int64_t ldb, int64_t brows, int64_t K)
    {
        static_assert(WN == 1, "nact must be block-uniform: every wave owns all block columns");
        if (nact == N)
            mainloop_n<N>(lds, A, lda, arows, B, ldb, brows, K);
        else if constexpr (N > 1)
            mainloop_live<N - 1>(lds, A, lda, arows, B, ldb, brows, K);
    }

    // acc += Ts Bg^T as lds_global, for a B operand that is lower triangular
    // (Bg[j][m] = 0 for m > j, row j = output column j), NACT live column
    // tiles and K <= NACT * MT.  Column tile y only needs K < MT (y + 1), so
    // sub-step s (KSUB K values) runs the tiles y >= s / (MT / KSUB): about half
    // the MFMAs of the full product.  B fragments are read from global
    // memory one sub-step ahead.
    template <int NACT>
    __device__ inline void lds_global_tri(const T* Ts, int tp, const T* Bg, int64_t ldb, int K)
    {
        constexpr int SPT = MT / KSUB, NSM = NACT * SPT;
        static_assert(SPT >= 1, "a column tile spans at least one sub-step");
        const int ns = K / KSUB;
        vec_t b[2][NACT];
        auto bload = [&](int q, int s) {
#pragma unroll
            for (int t = 0; t < NACT; ++t)
                if (t >= s / SPT)
                    b[q][t] = *reinterpret_cast<const vec_t*>(
                        Bg + (int64_t)(wn * TN * MT + t * MT + li) * ldb + s * KSUB + h * M::VEC);
        };
        if (ns > 0) bload(0, 0);
#pragma unroll
        for (int s = 0; s < NSM; ++s) {
            if (s < ns) {
                if (s + 1 < ns) bload((s + 1) & 1, s + 1);
                vec_t a[TM];
#pragma unroll
                for (int t = 0; t < TM; ++t)
                    a[t] = *reinterpret_cast<const vec_t*>(
                        Ts + (wm * TM * MT + t * MT + li) * tp + s * KSUB + h * M::VEC);
#pragma unroll
                for (int j = 0; j < M::VEC; ++j)
#pragma unroll
                    for (int y = s / SPT; y < NACT; ++y)
#pragma unroll
                        for (int x = 0; x < TM; ++x)
                            M::mma(acc[x][y], M::get(a[x], j), M::get(b[s & 1][y], j));
            }
        }
    }

    template <int N = TN>
    __device__ inline void lds_global_tri_live(const T* Ts, int tp, const T* Bg, int64_t ldb, int K)
    {
        static_assert(WN == 1, "nact must be block-uniform: every wave owns all block columns");
        if (nact == N)
            lds_global_tri<N>(Ts, tp, Bg, ldb, K);
        else if constexpr (N > 1)
            lds_global_tri_live<N - 1>(Ts, tp, Bg, ldb, K);
    }

    // acc += Ts[0:BM, 0:K] Bg[0:BN, 0:K]^T with Ts in LDS (row pitch tp) and
    // Bg read straight from global memory (a small operand every block
    // shares, served from L2).  K multiple of 8.  The B operand of sub-step
    // s + 1 is loaded while sub-step s multiplies, hiding the L2 latency.
    __device__ inline void lds_global(const T* Ts, int tp, const T* Bg, int64_t ldb, int K)
    {
        int ns = K / KSUB;
        if (ns <= 0) return;
        vec_t b[TN], bn[TN];
#pragma unroll
        for (int t = 0; t < TN; ++t)
            if (t < nact)
                b[t] = *reinterpret_cast<const vec_t*>(Bg + (int64_t)(wn * TN * MT + t * MT + li) * ldb + h * M::VEC);
        for (int s = 0; s < ns; ++s) {
            if (s + 1 < ns) {
#pragma unroll
                for (int t = 0; t < TN; ++t)
                    if (t < nact)
                        bn[t] = *reinterpret_cast<const vec_t*>(
                            Bg + (int64_t)(wn * TN * MT + t * MT + li) * ldb + (s + 1) * KSUB + h * M::VEC);
            }
            vec_t a[TM];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[t] = *reinterpret_cast<const vec_t*>(
                    Ts + (wm * TM * MT + t * MT + li) * tp + s * KSUB + h * M::VEC);
#pragma unroll
            for (int y = 0; y < TN; ++y)
                if (y < nact)
#pragma unroll
                    for (int j = 0; j < M::VEC; ++j)
#pragma unroll
                        for (int x = 0; x < TM; ++x) M::mma(acc[x][y], M::get(a[x], j), M::get(b[y], j));
#pragma unroll
            for (int t = 0; t < TN; ++t) b[t] = bn[t];
        }
    }
};

// Tile configurations.  The four waves split the block rows and each owns
// all BN = 128 columns, so a narrower live column range (fewer accepted
// pivots in a round, k_out + 1 < 128 eigenvectors) drops whole MFMA column
// tiles evenly on every wave.  FP32: 128 x 128 block, waves of 32 x 128
// (4 tiles of 32x32).  FP64: 64 x 128 block, waves of 16 x 128 (8 tiles of
// 16x16).
template <typename T> struct Cfg;
template <> struct Cfg<float> {
    typedef NTGemm<float, 128, 128, 32, 4, 1, 2, 8, Mfma16f> G;
    static constexpr int BM = 128, BN = 128, KC = 32;
};
template <> struct Cfg<double> {
    typedef NTGemm<double, 64, 128, 16, 4, 1, 1, 8> G;
    static constexpr int BM = 64, BN = 128, KC = 16;
};

// exp(-D / eps) for the Gaussian kernel.  FP64: libm-accurate exp.  FP32:
// the argument a = -D log2(e) / eps is formed in FP64 and split into an
// integer part and a fraction f in [-1/2, 1/2]; 2^f runs on v_exp_f32 and
// the integer part is an exact exponent shift, so a large |a| costs no
// accuracy (the FP32 rounding of a itself would cost |a| ulps).  Below
// a = -160 the value is 0, chosen by a select (no branch, so a caller's
// unrolled loop stays one basic block); the clamp keeps the integer
// conversion in range there.
struct GaussScale {
    double neg_inv_eps;        // -1 / eps
    double neg_log2e_inv_eps;  // -log2(e) / eps
};
__device__ inline double gauss_val(double D, const GaussScale& g, double*)
{
    return exp(D * g.neg_inv_eps);
}
__device__ inline float gauss_val(double D, const GaussScale& g, float*)
{
    double a = D * g.neg_log2e_inv_eps;
    double ni = rint(a > -200.0 ? a : -200.0);
    float f = (float)fma(D, g.neg_log2e_inv_eps, -ni);  // the fraction from the unrounded product
    float e = ldexpf(__builtin_amdgcn_exp2f(f), (int)ni);
    return a < -160.0 ? 0.0f : e;
}
