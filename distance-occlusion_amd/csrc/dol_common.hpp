// dol_common.hpp -- shared definitions for the gfx950 kernels and the host
// runtime of libdolhip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace dolhip {

constexpr int kMaxDelta2 = 32767;   // DistanceOcclusionLib.h:56
constexpr int kMaxDelta3 = 25000;   // DistanceOcclusionLib.h:57
constexpr int32_t kEdtInf = 0x7fffffff;
constexpr int kRsrcWord3 = 0x00020000;   // buffer descriptor word 3 for raw dword access (gfx9)

__host__ __device__ inline int mask_stride(int w) { return (w + 31) >> 5; }

// ---------------------------------------------------------------- device

// int16 store of an int, modular (the reference stores out-of-range JumpFlood
// candidates into int16, DistanceOcclusionLib.cpp:797-866)
__device__ __forceinline__ int wrap16(int v) { return (int)(int16_t)(uint16_t)(uint32_t)v; }

// squared length as the reference's Closer computes it: int16 squares summed
// in 32 bits, wrapping (DistanceOcclusionLib.cpp:624-630, 2177-2183).  Done
// unsigned so the wrap is defined.
__device__ __forceinline__ int len2(int x, int y)
{
    return (int)((uint32_t)(x * x) + (uint32_t)(y * y));
}
__device__ __forceinline__ int len2(int x, int y, int z)
{
    return (int)((uint32_t)(x * x) + (uint32_t)(y * y) + (uint32_t)(z * z));
}

// Correctly rounded fp32 square root.  The hardware v_sqrt_f32 is within
// 1 ulp; one residual step picks the right neighbour: with r the estimate and
// u its ulp, x <= r*(r-u) (computed exactly by fma) means sqrt(x) < r - u/2, and
// x > r*(r+u) means sqrt(x) > r + u/2.  A float x cannot fall strictly between
// r*(r-u) and (r-u/2)^2 = r*(r-u) + u^2/4, so the test is exact.  Below 2^-60
// the residual can underflow, so tiny inputs use the double-precision root,
// which rounds correctly to float as 53 >= 2*24 + 2.
__device__ __forceinline__ float sqrt_rn(float x)
{
    if (!(x >= 0x1p-60f)) return (float)__builtin_sqrt((double)x);
    float r = __builtin_amdgcn_sqrtf(x);
    const float rm = __uint_as_float(__float_as_uint(r) - 1u);
    const float rp = __uint_as_float(__float_as_uint(r) + 1u);
    const float em = __builtin_fmaf(-rm, r, x);
    const float ep = __builtin_fmaf(-rp, r, x);
    r = em <= 0.0f ? rm : r;
    r = ep > 0.0f ? rp : r;
    return r;
}

// sqrt_rn of a squared distance given as an int32 >= 0, with kEdtInf ("no
// site") giving sqrt_rn(FLT_MAX).  The input is 0 or >= 1, so the residual
// step above needs no tiny-input branch (for 0 the estimate is 0 and neither
// neighbour test fires, the lower neighbour being NaN).
__device__ __forceinline__ float sqrt_d2(int v)
{
    const float x = v == kEdtInf ? 3.40282347e+38f : (float)v;
    float r = __builtin_amdgcn_sqrtf(x);
    const float rm = __uint_as_float(__float_as_uint(r) - 1u);
    const float rp = __uint_as_float(__float_as_uint(r) + 1u);
    const float em = __builtin_fmaf(-rm, r, x);
    const float ep = __builtin_fmaf(-rp, r, x);
    r = em <= 0.0f ? rm : r;
    r = ep > 0.0f ? rp : r;
    return r;
}

// The same for squared distances below 2^20 (or kEdtInf).  On gfx950 the
// hardware root of every integer below 2^20 is either correctly rounded or
// one ulp low, never high (every such integer checked on the device; the
// tests compare every root this path produces with IEEE sqrt), so one residual
// decides: the upper neighbour rp is nearer iff x > (r + ulp/2)^2, i.e.
// x - r * rp > ulp^2 / 4, and x - r * rp is a multiple of ulp^2 whose sign the
// fma keeps.
__device__ __forceinline__ float sqrt_d2_small(int v)
{
    const float x = (float)v;
    const float r = __builtin_amdgcn_sqrtf(x);
    const float rp = __uint_as_float(__float_as_uint(r) + 1u);
    const float root = __builtin_fmaf(-rp, r, x) > 0.0f ? rp : r;
    return v == kEdtInf ? __uint_as_float(0x5f7fffffu) : root;   // sqrt_rn(FLT_MAX)
}

// delta2 packed in a uint32: x in the low half, y in the high half
__device__ __forceinline__ uint32_t pack2(int x, int y)
{
    return ((uint32_t)(uint16_t)x) | ((uint32_t)(uint16_t)y << 16);
}
__device__ __forceinline__ int unpack2x(uint32_t v) { return (int)(int16_t)(v & 0xffffu); }
__device__ __forceinline__ int unpack2y(uint32_t v) { return (int)(int16_t)(v >> 16); }

// one mask bit; callers keep x, row in range
__device__ __forceinline__ int mask_bit(const uint32_t* __restrict__ mask, int ws, size_t row, int x)
{
    return (int)((mask[row * ws + (x >> 5)] >> (x & 31)) & 1u);
}

// ---------------------------------------------------------------- host

// kernel launchers (each file of kernels exposes plain host functions)
struct Launch {
    hipStream_t stream;
};

void set_error(const char* fmt, ...);
int check_launch(const char* what);

// per-kernel event timing (dolhip_profile_*)
struct ProfileScope {
    ProfileScope(const char* name, hipStream_t s);
    ~ProfileScope();
    const char* name_;
    hipStream_t s_;
    void* ev0_;
};

inline unsigned grid1d(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace dolhip
