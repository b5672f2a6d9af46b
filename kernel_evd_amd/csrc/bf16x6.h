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
