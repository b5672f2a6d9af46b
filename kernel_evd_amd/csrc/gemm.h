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

template <typename T, int BM, int BN, int KC, int WM, int WN, int TM, int TN>
struct NTGemm {
    typedef Mfma<T> M;
    typedef typename M::acc_t acc_t;
    typedef typename M::vec_t vec_t;
    static constexpr int MT = M::MT;
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
    static_assert(KC % 8 == 0 && BM % RPP == 0 && BN % RPP == 0, "load shape");

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
        return wm * TM * MT + a * MT + acc_row<T>(v, h);
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
        for (int s = 0; s < KC / 8; ++s) {
            vec_t a[TM], b[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[t] = *reinterpret_cast<const vec_t*>(
                    As + (wm * TM * MT + t * MT + li) * PITCH + s * 8 + h * M::VEC);
#pragma unroll
            for (int t = 0; t < TN; ++t)
                if (t < nact)
                    b[t] = *reinterpret_cast<const vec_t*>(
                        Bs + (wn * TN * MT + t * MT + li) * PITCH + s * 8 + h * M::VEC);
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

    // acc += Ts[0:BM, 0:K] Bg[0:BN, 0:K]^T with Ts in LDS (row pitch tp) and
    // Bg read straight from global memory (a small operand every block
    // shares, served from L2).  K multiple of 8.  The B operand of sub-step
    // s + 1 is loaded while sub-step s multiplies, hiding the L2 latency.
    __device__ inline void lds_global(const T* Ts, int tp, const T* Bg, int64_t ldb, int K)
    {
        int ns = K / 8;
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
                            Bg + (int64_t)(wn * TN * MT + t * MT + li) * ldb + (s + 1) * 8 + h * M::VEC);
            }
            vec_t a[TM];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                a[t] = *reinterpret_cast<const vec_t*>(
                    Ts + (wm * TM * MT + t * MT + li) * tp + s * 8 + h * M::VEC);
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
    typedef NTGemm<float, 128, 128, 32, 4, 1, 1, 4> G;
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
// accuracy (the FP32 rounding of a itself would cost |a| ulps).
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
    if (a < -160.0) return 0.0f;
    double ni = rint(a);
    float f = (float)(a - ni);
    return ldexpf(__builtin_amdgcn_exp2f(f), (int)ni);
}
