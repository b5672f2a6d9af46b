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
// A is FP32 with K-contiguous rows (the factor's row block); B (BN = 128
// rows) comes already split, chunk-major: the parts of K values
// [32 c, 32 c + 32) are the contiguous block Bx + c X6_CHUNK in fragment
// order (x6_idx), zero from K up to a multiple of 32 (the downdate's pivot
// rows, split once per round when they are gathered).
//
// A row is used by one wave only (the waves split the rows), so A never
// passes through LDS: each lane loads the 8 K values its MFMA fragment
// needs (row 16 x + lane % 16, K group lane / 16; a wave reads whole
// 128-byte row segments) straight from HBM into registers, two chunks
// ahead, and splits them into the three bf16 fragments in registers.  Only
// B, which every wave reads, is staged (its 24 KB chunk is one contiguous
// block, so its loads read whole cache lines): two LDS stages holding the
// chunk as it lies in memory, read lane-linearly, one barrier per chunk, the next chunk's B
// rows (from L2) loaded during the current chunk's products.  A's rows past
// arows read as zero through a buffer resource; its columns past K are
// zeroed at the split.  Ends with a barrier.  (B's fragments read straight
// from L2 by every wave instead of through LDS moved four times the bytes
// and took 60 ms against 37.5 at N = 1M, r = 2000.)
// bf16 elements of one chunk of split B rows: 3 parts x 128 rows x 32.
// Within a chunk the layout is the MFMA fragment order: part, then the
// 16-row tile y, the 8-value K group h, the row i in the tile and the 8
// values, so the fragment of (part, y) is 1 KB read by lane 16 h + i at
// 16 (16 h + i), lane-linear as LDS reads serve fastest, and the copy into
// LDS is linear too.
constexpr int X6_CHUNK = 3 * 128 * 32;
// element index of (row, K value kk) of part 0 in a chunk (kk < 32)
__host__ __device__ constexpr int x6_idx(int row, int kk)
{
    return (((row >> 4) * 4 + (kk >> 3)) * 16 + (row & 15)) * 8 + (kk & 7);
}

template <typename G>
struct X6Stage {
    static constexpr int BM = G::NBM, BN = G::NBN, KC = 32, MT = 16;
    static constexpr int ROWB = KC * 2;  // bytes per row of one part
    static constexpr int PART = BN * ROWB;
    static constexpr int STAGE = 3 * PART;
    static constexpr int BYTES = 2 * STAGE;
    static constexpr int NVB = 3 * BN * (KC / 8) / 256;  // B 16-byte part loads per thread
    static_assert(G::NKC == KC && G::M::MT == MT && G::NBM == BM, "16x16 tiles, KC = 32");
    static_assert((3 * BN * (KC / 8)) % 256 == 0, "load shape");
};

template <typename G, int NACT>
__device__ inline void x6_nt_mainloop_n(G& g, unsigned char* lds, const float* A, int64_t lda, int64_t arows,
                                        const __bf16* Bx, int64_t K)
{
    typedef X6Stage<G> S;
    constexpr int BM = S::BM, KC = S::KC, NVB = S::NVB, TM = G::NTM;
    static_assert(NVB == 6 && S::BN == 128, "B copy: 3 parts x 2 row halves per thread");
    typedef __attribute__((ext_vector_type(4))) float f4;
    typedef __attribute__((ext_vector_type(4))) unsigned u4;
    const int nk = (int)((K + KC - 1) / KC), nfull = (int)(K / KC);
    if (nk <= 0) return;
    const int tid = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rsA = G::rows_rsrc(A, lda, arows, BM);
    // B: the chunks that exist (the loads past the last one read zero)
    const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<__bf16*>(Bx), (short)0, nk * X6_CHUNK * 2, 0x00020000);
    // A: byte offset of the lane's row of tile x and its K group; the chunk's
    // K offset is the uniform soffset
    uint32_t aoff[TM];
#pragma unroll
    for (int x = 0; x < TM; ++x) aoff[x] = (uint32_t)(((g.wm * TM * 16 + x * 16 + g.li) * lda + 8 * g.h) * 4);
    auto loadA = [&](f4 (&r)[TM][2], int c) {
        // past the last chunk: the last chunk again (an L2 hit, not a
        // second read of the row from HBM)
        const int kb = (c < nk ? c : nk - 1) * KC * 4;
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh)
                r[x][hh] = __builtin_bit_cast(
                    f4, __builtin_amdgcn_raw_buffer_load_b128(rsA, aoff[x] + 16 * hh, kb, 0));
    };
    // B: load j of thread tid moves bytes 4096 j + 16 tid of the chunk
    u4 rb[NVB];
    auto loadB = [&](int c) {
#pragma unroll
        for (int j = 0; j < NVB; ++j)
            rb[j] = __builtin_amdgcn_raw_buffer_load_b128(rsB, (uint32_t)(tid * 16), c * X6_CHUNK * 2 + j * 4096, 0);
    };
    auto storeB = [&](int stage) {
        unsigned char* base = lds + stage * S::STAGE + tid * 16;
#pragma unroll
        for (int j = 0; j < NVB; ++j) *reinterpret_cast<u4*>(base + j * 4096) = rb[j];
    };
    bf16x8 af[TM][3];
    auto split = [&](const f4 (&r)[TM][2], int c) {
        const bool tail = c >= nfull;
        const int kl = (int)K - c * KC - 8 * g.h;  // K values of the lane's group still below K
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float v = r[x][e / 4][e % 4];
                v = tail && e >= kl ? 0.0f : v;
                __bf16 a0, a1, a2;
                bf16_split3(v, a0, a1, a2);
                af[x][0][e] = a0, af[x][1][e] = a1, af[x][2][e] = a2;
            }
    };
    auto compute = [&](int stage) {
        const unsigned char* base = lds + stage * S::STAGE + (threadIdx.x & 63) * 16;
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
                        bf[yo][q] = *reinterpret_cast<const bf16x8*>(base + q * S::PART + (y + yo) * 1024);
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
    // Chunk c: B(c) is in stage c & 1, A(c) in registers.  Per chunk the
    // loads go out as B(c + 1), then A(c + 2): the wait for B(c + 1) before
    // its copy to LDS then leaves A(c + 2) in flight (the memory counter
    // completes in order), so each A chunk has two chunks' time to arrive.
    // Every step of the loop issues the same loads, also past the last
    // chunk (the buffer resources keep them in range; those values are never
    // multiplied), and the loop runs whole pairs of steps with an odd last
    // chunk after it: a load or split on only some paths would leave the
    // counter state at the loop head unknown or keep two register sets of
    // one operand live, and the compiler would then wait for every load in
    // flight.
    auto step = [&](f4 (&r)[TM][2], int c) {
        __builtin_amdgcn_sched_barrier(0);  // each step's code stays in its step
        loadB(c + 1);
        split(r, c);
        // the next A loads reuse r's registers only if they stay behind
        // the split (else the loop carries a copy that waits for them)
        __builtin_amdgcn_sched_barrier(0);
        loadA(r, c + 2);
        compute(c & 1);
        storeB((c + 1) & 1);
        __syncthreads();  // stage (c + 1) & 1 written, stage c & 1 free
    };
    f4 ra0[TM][2], ra1[TM][2];
    loadA(ra0, 0);
    loadA(ra1, 1);
    loadB(0);
    storeB(0);
    __syncthreads();
    int c = 0;
    for (; c + 1 < nk; c += 2) {
        step(ra0, c);
        step(ra1, c + 1);
    }
    if (c < nk) {  // odd chunk count: the last chunk, in ra0
        split(ra0, c);
        compute(c & 1);
        __syncthreads();
    }
}

// acc += Ts[0:BM, 0:K] L^T for a lower-triangular L (BN = 128 rows,
// L(n, m) = 0 for m > n; K <= 16 NACT), as bf16x6 on
// v_mfma_f32_16x16x32_bf16 into the accumulators of the FP32 NTGemm G.  Ts is
// FP32 in LDS (row pitch tp), each wave reads only its own rows, and its A
// fragments are split in registers per 32-value chunk.  L comes already
// split, chunk-major in fragment order like the pivot rows (Lx, X6_CHUNK
// elements per chunk of 32 columns), and is read from global memory (a small
// operand every block shares, served from L2), one column-tile pair ahead of
// its products.  Column tile y needs only the chunks c <= y / 2: about half
// the products of the full square.
template <typename G, int NACT>
__device__ inline void x6_tri_n(G& g, const float* Ts, int tp, const __bf16* Lx)
{
    typedef __attribute__((ext_vector_type(4))) float f4;
    constexpr int TM = G::NTM, NC = (NACT + 1) / 2;
    static_assert(G::M::MT == 16 && G::NBN == 128, "16x16 tiles, 128 columns");
    const int lane = threadIdx.x & 63;
    const __bf16* lb = Lx + lane * 8;
    // the three parts of column tiles y0, y0 + 1 of chunk c
    auto loadB = [&](bf16x8 (&b)[2][3], int c, int y0) {
#pragma unroll
        for (int yo = 0; yo < 2; ++yo)
            if (y0 + yo < NACT)
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    b[yo][q] = *reinterpret_cast<const bf16x8*>(lb + c * X6_CHUNK + q * (X6_CHUNK / 3) +
                                                                (y0 + yo) * 512);
    };
    bf16x8 bfr[2][2][3];
    int par = 0;
    loadB(bfr[0], 0, 0);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        bf16x8 af[TM][3];
#pragma unroll
        for (int x = 0; x < TM; ++x) {
            const float* src = Ts + (g.wm * TM * 16 + x * 16 + g.li) * tp + 32 * c + 8 * g.h;
            const f4 v[2] = {*reinterpret_cast<const f4*>(src), *reinterpret_cast<const f4*>(src + 4)};
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                __bf16 a0, a1, a2;
                bf16_split3(v[e / 4][e % 4], a0, a1, a2);
                af[x][0][e] = a0, af[x][1][e] = a1, af[x][2][e] = a2;
            }
        }
#pragma unroll
        for (int y0 = 2 * c; y0 < NACT; y0 += 2) {
            // the next pair: the following tiles of this chunk, else the
            // first pair of the next chunk
            const bool last = y0 + 2 >= NACT;
            const int nc = last ? c + 1 : c, ny = last ? 2 * (c + 1) : y0 + 2;
            if (nc < NC) loadB(bfr[par ^ 1], nc, ny);
#pragma unroll
            for (int p = 0; p < X6_NPROD; ++p)
#pragma unroll
                for (int x = 0; x < TM; ++x)
#pragma unroll
                    for (int yo = 0; yo < 2; ++yo)
                        if (y0 + yo < NACT)
                            mfma_bf16_16(g.acc[x][y0 + yo], af[x][x6_a(p)], bfr[par][yo][x6_b(p)]);
            par ^= 1;
        }
    }
}

template <typename G, int N = G::NTN>
__device__ inline void x6_tri_live(G& g, const float* Ts, int tp, const __bf16* Lx)
{
    if (g.nact == N)
        x6_tri_n<G, N>(g, Ts, tp, Lx);
    else if constexpr (N > 1)
        x6_tri_live<G, N - 1>(g, Ts, tp, Lx);
}

template <typename G, int N = G::NTN>
__device__ inline void x6_nt_mainloop_live(G& g, unsigned char* lds, const float* A, int64_t lda, int64_t arows,
                                           const __bf16* Bx, int64_t K)
{
    if (g.nact == N)
        x6_nt_mainloop_n<G, N>(g, lds, A, lda, arows, Bx, K);
    else if constexpr (N > 1)
        x6_nt_mainloop_live<G, N - 1>(g, lds, A, lda, arows, Bx, K);
}
