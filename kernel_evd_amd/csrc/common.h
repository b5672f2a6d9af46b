// common.h -- shared device/host helpers of libkevd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "kevd.h"

// ----------------------------------------------------------------- Philox

// Philox4x32-10 (Salmon et al., SC'11).  Counter layout used everywhere in
// libkevd: ctr = {idx, round, stream, 0}, key = {seed lo, seed hi}.  Replaces
// the cuRAND XORWOW states of ks_arpclm.cu:175-180.
enum : uint32_t {
    KEVD_STREAM_PROPOSE = 0,  // proposal draws of a round
    KEVD_STREAM_REJECT = 1,   // acceptance uniforms of a round
    KEVD_STREAM_GAUSS = 2,    // KS initial condition
    KEVD_STREAM_MEDIAN = 3,   // median pair sampling
    KEVD_STREAM_ROWS = 4      // bandwidth row sampling
};

__host__ __device__ inline void kevd_philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                            uint32_t c3, uint32_t k0, uint32_t k1,
                                            uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__host__ __device__ inline uint64_t kevd_draw_u64(uint64_t seed, uint32_t stream,
                                                  uint32_t round, uint32_t idx)
{
    uint32_t o[4];
    kevd_philox(idx, round, stream, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o);
    return ((uint64_t)o[1] << 32) | o[0];
}

// Uniform double in (0,1].
__host__ __device__ inline double kevd_u64_to_unit(uint64_t u)
{
    return (double)((u >> 11) + 1) * 0x1.0p-53;
}

__host__ __device__ inline uint64_t kevd_mulhi64(uint64_t a, uint64_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// ------------------------------------------------------------------- MFMA

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Per-precision MFMA shape.  Both use the "8 K values per sub-step" scheme:
// a lane loads VEC consecutive K values of its operand row with one 16-byte
// LDS read and feeds them to VEC consecutive MFMAs; lane group h = lane / MT
// owns K offsets [VEC*h, VEC*h + VEC) of the sub-step.  The K order inside a
// sub-step is a permutation, identical for A and B, so the sum is unchanged.
template <typename T> struct Mfma;

template <> struct Mfma<float> {
    static constexpr int MT = 32;     // 32x32 output tile (v_mfma_f32_32x32x2_f32)
    static constexpr int NH = 2;      // lane groups along K
    static constexpr int VEC = 4;     // K values per lane per sub-step
    static constexpr int ACC = 16;    // accumulator values per lane
    typedef floatx16 acc_t;
    typedef float4 vec_t;
    __device__ static inline void mma(acc_t& c, float a, float b)
    {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline float get(const vec_t& v, int j)
    {
        return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    }
    // output row of accumulator value v = row_lane(h) + row_off(v)
    __device__ static inline int row_lane(int h) { return 4 * h; }
    __host__ __device__ static constexpr int row_off(int v) { return (v >> 2) * 8 + (v & 3); }
};

template <> struct Mfma<double> {
    static constexpr int MT = 16;     // 16x16 output tile (v_mfma_f64_16x16x4_f64)
    static constexpr int NH = 4;
    static constexpr int VEC = 2;
    static constexpr int ACC = 4;
    typedef doublex4 acc_t;
    typedef double2 vec_t;
    __device__ static inline void mma(acc_t& c, double a, double b)
    {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline double get(const vec_t& v, int j) { return j == 0 ? v.x : v.y; }
    __device__ static inline int row_lane(int h) { return h; }
    __host__ __device__ static constexpr int row_off(int v) { return 4 * v; }
};

// FP32 16x16 tiles (v_mfma_f32_16x16x4_f32): same rate as 32x32x2, finer
// column granularity for GEMMs whose live column count varies (the
// downdate's accepted pivots, the U product's k_out + 1 columns).
struct Mfma16f {
    static constexpr int MT = 16;
    static constexpr int NH = 4;
    static constexpr int VEC = 4;
    static constexpr int ACC = 4;
    typedef floatx4 acc_t;
    typedef float4 vec_t;
    __device__ static inline void mma(acc_t& c, float a, float b)
    {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline float get(const vec_t& v, int j)
    {
        return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    }
    // row 4 h + v (read off the hardware with a single-MFMA probe)
    __device__ static inline int row_lane(int h) { return 4 * h; }
    __host__ __device__ static constexpr int row_off(int v) { return v; }
};

// Row (within the MT x MT tile) of accumulator value v held by lane group h;
// the column is lane % MT.  FP32 32x32: 8 (v / 4) + 4 h + v % 4; FP64 16x16:
// h + 4 v; FP32 16x16: 4 h + v (checked against the kernel-block parity tests).
template <typename T>
__device__ inline int acc_row(int v, int h)
{
    return Mfma<T>::row_lane(h) + Mfma<T>::row_off(v);
}

// --------------------------------------------------------------- reductions

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ------------------------------------------------------------- host side

#define KEVD_ROUND_UP(a, b) ((((a) + (b) - 1) / (b)) * (b))

#ifndef KEVD_LDF_PAD
#define KEVD_LDF_PAD 32
#endif
// leading dimension of a resident rank-r factor (row-major n x ldf)
inline int64_t kevd_factor_ld(int64_t r) { return KEVD_ROUND_UP(r, 64) + KEVD_LDF_PAD; }

struct kevd_ctx;
kevd_status kevd_fail(kevd_ctx* ctx, kevd_status s, const std::string& msg);
kevd_status kevd_hip_fail(kevd_ctx* ctx, hipError_t e, const char* what, const char* file,
                          int line);

#define KEVD_HIP(ctx, call)                                                   \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return kevd_hip_fail((ctx), e_, #call, __FILE__, __LINE__); \
    } while (0)

#define KEVD_TRY(call)                                                        \
    do {                                                                      \
        kevd_status s_ = (call);                                              \
        if (s_ != KEVD_OK) return s_;                                         \
    } while (0)
