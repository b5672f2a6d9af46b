// bf16x6: FP32 products on the bf16 matrix cores.
//
// gfx950 multiplies bfloat16 operands 14x faster than FP32 operands
// (v_mfma_f32_32x32x16_bf16 against v_mfma_f32_32x32x2_f32).  An FP32 value
// splits exactly into three bfloat16 values, x = hi + mid + lo: each is the
// round-to-nearest of what the previous ones leave, and each residual is
// exact in FP32 (8 significant bits per part, 24 in all).  A product
// x y is the sum of the nine part products; every part product is exact in
// the FP32 accumulator.  The six with weight down to 2^-16 of |x y|
// (hi hi; hi mid, mid hi; hi lo, mid mid, lo hi) are kept and the three
// below 2^-24 (mid lo, lo mid, lo lo) dropped, which leaves an error of the
// order of one FP32 rounding per product.  This is the bf16x6 scheme of
// Henry, Tang and Heinecke, "Leveraging the bfloat16 Artificial
// Intelligence Datatype For Higher-Precision Computations", ARITH 2019.
// Six bf16 MFMAs cost about 0.43 of one FP32 MFMA of the same K.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// x = hi + mid + lo, each part rounded to nearest even (v_cvt_pk_bf16_f32)
__device__ inline void bf16_split3(float x, __bf16& hi, __bf16& mid, __bf16& lo)
{
    hi = (__bf16)x;
    float r = x - (float)hi;
    mid = (__bf16)r;
    float r2 = r - (float)mid;
    lo = (__bf16)r2;
}

// The part pairs (A part, B part) of the six kept products.
constexpr int X6_NPROD = 6;
__host__ __device__ constexpr int x6_a(int p) { return p == 0 ? 0 : p == 1 ? 0 : p == 2 ? 1 : p == 3 ? 0 : p == 4 ? 1 : 2; }
__host__ __device__ constexpr int x6_b(int p) { return p == 0 ? 0 : p == 1 ? 1 : p == 2 ? 0 : p == 3 ? 2 : p == 4 ? 1 : 0; }

// v_mfma_f32_32x32x16_bf16: lane l supplies A(l % 32, 8 (l / 32) + e) and
// B(8 (l / 32) + e, l % 32) for e = 0..7; the 32 x 32 FP32 result has the
// layout of v_mfma_f32_32x32x2_f32 (value v of lane l is row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).
__device__ inline void mfma_bf16_32(floatx16& c, const bf16x8& a, const bf16x8& b)
{
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16: lane l supplies A(l % 16, 8 (l / 16) + e) and
// B(8 (l / 16) + e, l % 16); the 16 x 16 FP32 result has the layout of
// v_mfma_f32_16x16x4_f32 (value v of lane l is row 4 (l / 16) + v, column
// l % 16).
__device__ inline void mfma_bf16_16(floatx4& c, const bf16x8& a, const bf16x8& b)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// acc += A[0:BM, 0:K] B[0:BN, 0:K]^T with bf16x6 products on
// v_mfma_f32_16x16x32_bf16, into the accumulators of an FP32 NTGemm G with
// 16 x 16 tiles, one wave column (WN = 1) and KC = 32 (same accumulator
// layout, so the caller's epilogues are unchanged).  NACT live column tiles.
// A is FP32 with K-contiguous rows (as NTGemm::mainloop); B comes already
// split: part q at Bx + q * part_stride, row j at + j * ldb, K-contiguous
// and zero from K up to a multiple of 32 (the downdate's pivot rows, split
// once per round when they are gathered).
//
// One LDS stage [part][row][32] bf16 for the BM + BN rows.  A row's four
// 16-byte slots (8 K values each) are XOR-swizzled by (row >> 2) & 3, which
// makes every access conflict-free: a fragment read (16 rows, one slot
// each), an A split store (4 rows x 8 eight-byte halves per 32 lanes) and a
// B copy (4 rows x 4 slots per 16 lanes).  The next chunk's loads are in
// flight in registers while the current one multiplies; two barriers per
// chunk, the second workgroup on the CU covering them.  A's rows past arows
// read as zero through a buffer resource and its columns past K are zeroed.
// Ends with a barrier.  (B's fragments read straight from L2 by every wave
// instead of through LDS moved four times the bytes and took 60 ms against
// 37.5 at N = 1M, r = 2000.)
template <typename G>
struct X6Stage {
    static constexpr int BM = G::NBM, BN = G::NBN, KC = 32, MT = 16;
    static constexpr int ROWB = KC * 2;  // bytes per row of one part
    static constexpr int PART = (BM + BN) * ROWB;
    static constexpr int BYTES = 3 * PART;
    static constexpr int NVA = BM * (KC / 4) / 256;      // A float4 loads per thread
    static constexpr int NVB = 3 * BN * (KC / 8) / 256;  // B 16-byte part loads per thread
    static_assert(G::NKC == KC && G::M::MT == MT && G::NBM == BM, "16x16 tiles, KC = 32");
    static_assert((BM * (KC / 4)) % 256 == 0 && (3 * BN * (KC / 8)) % 256 == 0, "load shape");
};

template <typename G, int NACT>
__device__ inline void x6_nt_mainloop_n(G& g, unsigned char* lds, const float* A, int64_t lda, int64_t arows,
                                        const __bf16* Bx, int64_t ldb, int64_t part_stride, int64_t K)
{
    typedef X6Stage<G> S;
    constexpr int BM = S::BM, KC = S::KC, NVA = S::NVA, NVB = S::NVB, TM = G::NTM;
    typedef __attribute__((ext_vector_type(4))) float f4;
    typedef __attribute__((ext_vector_type(4))) unsigned u4;
    const int64_t nk = (K + KC - 1) / KC, nfull = K / KC;
    if (nk <= 0) return;
    const __amdgpu_buffer_rsrc_t rsA = G::rows_rsrc(A, lda, arows, BM);
    const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<__bf16*>(Bx), (short)0, (int)(3 * part_stride * 2), 0x00020000);
    f4 rg[NVA];
    u4 rb[NVB];
    auto load = [&](int64_t c) {
        const int64_t k0 = c * KC;
#pragma unroll
        for (int j = 0; j < NVA; ++j) {
            const int lin = threadIdx.x + 256 * j;
            const int row = lin / (KC / 4), cv = lin % (KC / 4);
            const int64_t kk = k0 + 4 * cv;
            rg[j] = __builtin_bit_cast(
                f4, __builtin_amdgcn_raw_buffer_load_b128(rsA, (uint32_t)(((int64_t)row * lda + kk) * 4), 0, 0));
            if (c >= nfull) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (kk + q >= K) rg[j][q] = 0.0f;
            }
        }
        // B: part q, row, 16-byte slot (8 K values); no mask (zero past K)
#pragma unroll
        for (int j = 0; j < NVB; ++j) {
            const int lin = threadIdx.x + 256 * j;
            const int q = lin / (S::BN * 4), rem = lin % (S::BN * 4), row = rem / 4, sl = rem % 4;
            rb[j] = __builtin_amdgcn_raw_buffer_load_b128(
                rsB, (uint32_t)((q * part_stride + (int64_t)row * ldb + k0 + 8 * sl) * 2), 0, 0);
        }
    };
    auto store = [&]() {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int j = 0; j < NVB; ++j) {
            const int lin = threadIdx.x + 256 * j;
            const int q = lin / (S::BN * 4), rem = lin % (S::BN * 4), row = rem / 4, sl = rem % 4;
            const int srow = BM + row;
            *reinterpret_cast<u4*>(lds + q * S::PART + srow * S::ROWB + ((sl ^ ((srow >> 2) & 3)) << 4)) = rb[j];
        }
#pragma unroll
        for (int j = 0; j < NVA; ++j) {
            const int lin = threadIdx.x + 256 * j;
            const int srow = lin / (KC / 4), cv = lin % (KC / 4);  // A rows
            bf16x4 p0, p1, p2;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                __bf16 a0, a1, a2;
                bf16_split3(rg[j][q], a0, a1, a2);
                p0[q] = a0, p1[q] = a1, p2[q] = a2;
            }
            unsigned char* dst = lds + srow * S::ROWB + (((cv >> 1) ^ ((srow >> 2) & 3)) << 4) + (cv & 1) * 8;
            *reinterpret_cast<bf16x4*>(dst) = p0;
            *reinterpret_cast<bf16x4*>(dst + S::PART) = p1;
            *reinterpret_cast<bf16x4*>(dst + 2 * S::PART) = p2;
        }
    };
    // a fragment row is 16 t + li: its slot swizzle depends on li only
    const int slot = (g.h ^ ((g.li >> 2) & 3)) << 4;
    auto compute = [&]() {
        bf16x8 af[TM][3];
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                af[x][q] = *reinterpret_cast<const bf16x8*>(lds + q * S::PART +
                                                            (g.wm * TM * 16 + x * 16 + g.li) * S::ROWB + slot);
        // two column tiles at a time: four independent accumulators between
        // two MFMAs into the same one
#pragma unroll
        for (int y = 0; y < NACT; y += 2) {
            bf16x8 bf[2][3];
#pragma unroll
            for (int yo = 0; yo < 2; ++yo)
                if (y + yo < NACT)
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        bf[yo][q] = *reinterpret_cast<const bf16x8*>(
                            lds + q * S::PART + (BM + (y + yo) * 16 + g.li) * S::ROWB + slot);
#pragma unroll
            for (int p = 0; p < X6_NPROD; ++p)
#pragma unroll
                for (int x = 0; x < TM; ++x)
#pragma unroll
                    for (int yo = 0; yo < 2; ++yo)
                        if (y + yo < NACT) mfma_bf16_16(g.acc[x][y + yo], af[x][x6_a(p)], bf[yo][x6_b(p)]);
            // one column pair's fragments live at a time
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // A second register set (loads two chunks ahead) does not fit beside
    // the accumulators and fragments at two waves per SIMD (it spills).
    load(0);
    for (int64_t c = 0; c < nk; ++c) {
        if (c > 0) __syncthreads();  // every wave is done with chunk c - 1
        store();
        __syncthreads();
        if (c + 1 < nk) load(c + 1);
        compute();
    }
    __syncthreads();
}

template <typename G, int N = G::NTN>
__device__ inline void x6_nt_mainloop_live(G& g, unsigned char* lds, const float* A, int64_t lda, int64_t arows,
                                           const __bf16* Bx, int64_t ldb, int64_t part_stride, int64_t K)
{
    if (g.nact == N)
        x6_nt_mainloop_n<G, N>(g, lds, A, lda, arows, Bx, ldb, part_stride, K);
    else if constexpr (N > 1)
        x6_nt_mainloop_live<G, N - 1>(g, lds, A, lda, arows, Bx, ldb, part_stride, K);
}
