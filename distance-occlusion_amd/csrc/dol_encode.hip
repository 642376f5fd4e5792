// dol_encode.hip -- distances and image encoders of the reference tool
// (DistanceOcclusionTool.cpp:46-75, 678-698, 1029-1040, 1240-1331, 1367-1469),
// computed on the device so results leave the GPU as float or 8-bit images.
//
// Float semantics follow the reference exactly: the squared length is an
// int32, converted to float and passed to a correctly rounded sqrtf; divides
// are correctly rounded; 128 + v * 127 is a separate multiply and add (no
// FMA); lrintf is round-half-to-even.
#include "dol_cells.hpp"

#include <cfloat>

namespace dolhip {

// encode_u8_signed, encode_u8, dist_of and sdf_u8_cell: dol_cells.hpp

__global__ void __launch_bounds__(256) k_dist_2d(size_t n, const uint32_t* __restrict__ delta,
                                                 float* __restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v = delta[i];
    out[i] = dist_of(len2(unpack2x(v), unpack2y(v)));
}

void launch_dist_2d(size_t n, const int16_t* delta, float* out, hipStream_t s)
{
    ProfileScope p("dist_2d", s);
    hipLaunchKernelGGL(k_dist_2d, dim3(grid1d(n, 256)), dim3(256), 0, s, n, (const uint32_t*)delta,
                       out);
}

// mode 0: signed float distance in cells, (bit ? -0.5 : 0.5) * Dist
// mode 1: EncodeU8Signed((bit ? -Dist : Dist) / max_len)
template <int MODE>
__global__ void __launch_bounds__(256) k_signed_2d(int w, int xb, const uint32_t* __restrict__ mask,
                                                   const uint32_t* __restrict__ delta, float max_len,
                                                   void* __restrict__ out)
{
    const int y = blockIdx.x / xb;
    const int x = (blockIdx.x - y * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    const int ws = mask_stride(w);
    const size_t i = (size_t)y * w + x;
    const uint32_t v = delta[i];
    const int bit = mask_bit(mask, ws, (size_t)y, x);
    const float dd = dist_of(len2(unpack2x(v), unpack2y(v)));
    if (MODE == 0) {
        const float h = __fmul_rn(0.5f, dd);
        ((float*)out)[i] = bit ? -h : h;
    } else {
        ((uint8_t*)out)[i] = sdf_u8_cell(v, bit, max_len);
    }
}

// Four cells a thread for w % 4 == 0: one 16-byte load of deltas, the four
// mask bits from one word, one 16-byte (or 4-byte) store.  Same arithmetic
// as k_signed_2d.
template <int MODE>
__global__ void __launch_bounds__(256) k_signed_2d_x4(int w, size_t cells, const uint32_t* __restrict__ mask,
                                                      const uint4* __restrict__ delta, float max_len,
                                                      void* __restrict__ out)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x, i = 4 * q;
    if (i >= cells) return;
    const size_t row = i / (size_t)w;
    const int x = (int)(i - row * (size_t)w);
    const uint4 v4 = delta[q];
    const uint32_t v[4] = {v4.x, v4.y, v4.z, v4.w};
    const uint32_t bits = mask[row * mask_stride(w) + (x >> 5)] >> (x & 31);
    float f[4];
    uint32_t bytes = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int bit = (bits >> k) & 1;
        if (MODE == 0) {
            const float h = __fmul_rn(0.5f, dist_of(len2(unpack2x(v[k]), unpack2y(v[k]))));
            f[k] = bit ? -h : h;
        } else {
            bytes |= (uint32_t)sdf_u8_cell(v[k], bit, max_len) << (8 * k);
        }
    }
    if (MODE == 0) ((float4*)out)[q] = make_float4(f[0], f[1], f[2], f[3]);
    else ((uint32_t*)out)[q] = bytes;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_signed_3d(int w, int xb, const uint32_t* __restrict__ mask,
                                                   const int16_t* __restrict__ delta, float max_len,
                                                   void* __restrict__ out)
{
    const int row = blockIdx.x / xb;
    const int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    const int ws = mask_stride(w);
    const size_t i = (size_t)row * w + x;
    const int16_t* c = delta + 3 * i;
    const int bit = mask_bit(mask, ws, (size_t)row, x);
    const float dd = dist_of(len2(c[0], c[1], c[2]));
    if (MODE == 0) {
        const float h = __fmul_rn(0.5f, dd);
        ((float*)out)[i] = bit ? -h : h;
    } else {
        const float q = __fdiv_rn(dd, max_len);
        ((uint8_t*)out)[i] = encode_u8_signed(bit ? -q : q);
    }
}

void launch_signed_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                      void* out, int mode, hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p(mode == 0 ? "signed_dist_2d" : "encode_sdf_u8_2d", s);
    if (w % 4 == 0 && (((uintptr_t)delta | (uintptr_t)out) & (mode == 0 ? 15 : 3)) == 0) {
        const size_t cells = (size_t)w * h;
        const dim3 grid(grid1d(cells / 4, 256));
        if (mode == 0)
            hipLaunchKernelGGL(k_signed_2d_x4<0>, grid, dim3(256), 0, s, w, cells, mask,
                               (const uint4*)delta, max_len, out);
        else
            hipLaunchKernelGGL(k_signed_2d_x4<1>, grid, dim3(256), 0, s, w, cells, mask,
                               (const uint4*)delta, max_len, out);
        return;
    }
    if (mode == 0)
        hipLaunchKernelGGL(k_signed_2d<0>, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                           (const uint32_t*)delta, max_len, out);
    else
        hipLaunchKernelGGL(k_signed_2d<1>, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                           (const uint32_t*)delta, max_len, out);
}

// Four cells a thread for w % 4 == 0: their 24 bytes of deltas come in
// three aligned 8-byte loads (cell i starts at byte 6i, 8-aligned for
// i % 4 == 0) and the four results leave in one store.  Same arithmetic as
// k_signed_3d.
template <int MODE>
__global__ void __launch_bounds__(256) k_signed_3d_x4(int w, size_t cells, const uint32_t* __restrict__ mask,
                                                      const int16_t* __restrict__ delta, float max_len,
                                                      void* __restrict__ out)
{
    const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= cells) return;
    const int ws = mask_stride(w);
    const size_t row = i / (size_t)w;
    const int x = (int)(i - row * (size_t)w);
    const uint2* src = (const uint2*)(delta + 3 * i);
    const uint2 a = src[0], b = src[1], c = src[2];
    const uint32_t word[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
    const uint32_t bits = mask[row * ws + (x >> 5)] >> (x & 31);
    float f[4];
    uint32_t bytes = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        // int16 number 3k + j of the four cells
        auto comp = [&](int j) {
            const int n = 3 * k + j;
            const uint32_t v = word[n >> 1];
            return (int)(int16_t)((n & 1) ? (v >> 16) : (v & 0xffffu));
        };
        const float dd = dist_of(len2(comp(0), comp(1), comp(2)));
        const int bit = (bits >> k) & 1;
        if (MODE == 0) {
            const float hh = __fmul_rn(0.5f, dd);
            f[k] = bit ? -hh : hh;
        } else {
            const float q = __fdiv_rn(dd, max_len);
            bytes |= (uint32_t)encode_u8_signed(bit ? -q : q) << (8 * k);
        }
    }
    if (MODE == 0)
        *(float4*)((float*)out + i) = make_float4(f[0], f[1], f[2], f[3]);
    else
        *(uint32_t*)((uint8_t*)out + i) = bytes;
}

void launch_signed_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                      float max_len, void* out, int mode, hipStream_t s)
{
    int xb = (w + 255) / 256;
    unsigned blocks = (unsigned)((size_t)xb * h * d);
    ProfileScope p(mode == 0 ? "signed_dist_3d" : "encode_sdf_u8_3d", s);
    const size_t cells = (size_t)w * h * d;
    if (w % 4 == 0 && ((uintptr_t)delta & 7) == 0 && ((uintptr_t)out & (mode == 0 ? 15 : 3)) == 0) {
        const unsigned b4 = (unsigned)((cells / 4 + 255) / 256);
        if (mode == 0)
            hipLaunchKernelGGL(k_signed_3d_x4<0>, dim3(b4), dim3(256), 0, s, w, cells, mask, delta, max_len, out);
        else
            hipLaunchKernelGGL(k_signed_3d_x4<1>, dim3(b4), dim3(256), 0, s, w, cells, mask, delta, max_len, out);
        return;
    }
    if (mode == 0)
        hipLaunchKernelGGL(k_signed_3d<0>, dim3(blocks), dim3(256), 0, s, w, xb, mask, delta,
                           max_len, out);
    else
        hipLaunchKernelGGL(k_signed_3d<1>, dim3(blocks), dim3(256), 0, s, w, xb, mask, delta,
                           max_len, out);
}

// delta image (DistanceOcclusionTool.cpp:1281-1288): x and y of the delta
// over max_len, signed-encoded, blue 0, alpha 255
__global__ void __launch_bounds__(256) k_sdf_delta_rgba(size_t n, const uint32_t* __restrict__ delta,
                                                        float max_len, uint32_t* __restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v = delta[i];
    uint32_t r = encode_u8_signed(__fdiv_rn((float)unpack2x(v), max_len));
    uint32_t g = encode_u8_signed(__fdiv_rn((float)unpack2y(v), max_len));
    out[i] = r | (g << 8) | (255u << 24);
}

// red/green image (DistanceOcclusionTool.cpp:1289-1307): distance in the
// channel given by the mask bit
__global__ void __launch_bounds__(256) k_sdf_rg_rgba(int w, int xb, const uint32_t* __restrict__ mask,
                                                     const uint32_t* __restrict__ delta, float max_len,
                                                     uint32_t* __restrict__ out)
{
    const int y = blockIdx.x / xb;
    const int x = (blockIdx.x - y * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const uint32_t v = delta[i];
    const int bit = mask_bit(mask, mask_stride(w), (size_t)y, x);
    const uint32_t e = encode_u8(__fdiv_rn(dist_of(len2(unpack2x(v), unpack2y(v))), max_len));
    out[i] = (bit ? (e << 8) : e) | (255u << 24);
}

void launch_sdf_delta_rgba(int w, int h, const int16_t* delta, float max_len, uint8_t* out,
                           hipStream_t s)
{
    size_t n = (size_t)w * h;
    ProfileScope p("encode_sdf_delta", s);
    hipLaunchKernelGGL(k_sdf_delta_rgba, dim3(grid1d(n, 256)), dim3(256), 0, s, n,
                       (const uint32_t*)delta, max_len, (uint32_t*)out);
}

void launch_sdf_rg_rgba(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                        uint8_t* out, hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("encode_sdf_rg", s);
    hipLaunchKernelGGL(k_sdf_rg_rgba, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                       (const uint32_t*)delta, max_len, (uint32_t*)out);
}

// ---- two-field signed images: the tool's -S mode, exterior field d0 (seeds
// at set cells) and interior field d1 (seeds at clear cells), written as
// DistanceOcclusionTool.cpp:1029-1080 (2D) and 1126-1195 (3D).  type 0:
// EncodeU8Signed((Dist(d0) - Dist(d1)) / maxLen), one byte a cell; types 1-3
// are RGBA: 1 the summed deltas over maxLen, 2 EncodeU8 of each distance over
// maxLen, 3 the summed deltas over Dist(d0) - Dist(d1).  Every quotient is a
// single rounded division, as there.
template <int TYPE>
__global__ void __launch_bounds__(256) k_two_field_2d(size_t n, const uint32_t* __restrict__ d0,
                                                      const uint32_t* __restrict__ d1, float max_len,
                                                      void* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = d0[i], b = d1[i];
    const float da = dist_of(len2(unpack2x(a), unpack2y(a)));
    const float db = dist_of(len2(unpack2x(b), unpack2y(b)));
    if (TYPE == 0) {
        ((uint8_t*)out)[i] = encode_u8_signed(__fdiv_rn(__fsub_rn(da, db), max_len));
        return;
    }
    const float sx = (float)(unpack2x(a) + unpack2x(b)), sy = (float)(unpack2y(a) + unpack2y(b));
    uint32_t r, g;
    if (TYPE == 1) {
        r = encode_u8_signed(__fdiv_rn(sx, max_len));
        g = encode_u8_signed(__fdiv_rn(sy, max_len));
    } else if (TYPE == 2) {
        r = encode_u8(__fdiv_rn(da, max_len));
        g = encode_u8(__fdiv_rn(db, max_len));
    } else {
        const float dd = __fsub_rn(da, db);
        r = encode_u8_signed(__fdiv_rn(sx, dd));
        g = encode_u8_signed(__fdiv_rn(sy, dd));
    }
    ((uint32_t*)out)[i] = r | (g << 8) | (255u << 24);
}

// Same, four cells a thread (n % 4 == 0, 16-byte aligned delta images): one
// 16-byte load from each field and one 4-byte (type 0) or 16-byte store.
template <int TYPE>
__device__ __forceinline__ uint32_t two_field_cell(uint32_t a, uint32_t b, float max_len)
{
    const float da = dist_of(len2(unpack2x(a), unpack2y(a)));
    const float db = dist_of(len2(unpack2x(b), unpack2y(b)));
    if (TYPE == 0) return encode_u8_signed(__fdiv_rn(__fsub_rn(da, db), max_len));
    const float sx = (float)(unpack2x(a) + unpack2x(b)), sy = (float)(unpack2y(a) + unpack2y(b));
    uint32_t r, g;
    if (TYPE == 1) {
        r = encode_u8_signed(__fdiv_rn(sx, max_len));
        g = encode_u8_signed(__fdiv_rn(sy, max_len));
    } else if (TYPE == 2) {
        r = encode_u8(__fdiv_rn(da, max_len));
        g = encode_u8(__fdiv_rn(db, max_len));
    } else {
        const float dd = __fsub_rn(da, db);
        r = encode_u8_signed(__fdiv_rn(sx, dd));
        g = encode_u8_signed(__fdiv_rn(sy, dd));
    }
    return r | (g << 8) | (255u << 24);
}

template <int TYPE>
__global__ void __launch_bounds__(256) k_two_field_2d_x4(size_t n4, const uint4* __restrict__ d0,
                                                         const uint4* __restrict__ d1, float max_len,
                                                         void* __restrict__ out)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n4) return;
    const uint4 a = d0[q], b = d1[q];
    const uint32_t c0 = two_field_cell<TYPE>(a.x, b.x, max_len), c1 = two_field_cell<TYPE>(a.y, b.y, max_len);
    const uint32_t c2 = two_field_cell<TYPE>(a.z, b.z, max_len), c3 = two_field_cell<TYPE>(a.w, b.w, max_len);
    if (TYPE == 0)
        ((uint32_t*)out)[q] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
    else
        ((uint4*)out)[q] = make_uint4(c0, c1, c2, c3);
}

// 3D: the same per slice, blue the z channel of type 1 and 3.  Type 3 reads
// the first slice's deltas for every slice, as the tool does
// (DistanceOcclusionTool.cpp:1178-1186 index deltas0/deltas1, not the slice's).
template <int TYPE>
__global__ void __launch_bounds__(256) k_two_field_3d(size_t n, size_t slice, const int16_t* __restrict__ d0,
                                                      const int16_t* __restrict__ d1, float max_len,
                                                      void* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t j = TYPE == 3 ? i % slice : i;
    const int ax = d0[3 * j], ay = d0[3 * j + 1], az = d0[3 * j + 2];
    const int bx = d1[3 * j], by = d1[3 * j + 1], bz = d1[3 * j + 2];
    const float da = dist_of(len2(ax, ay, az)), db = dist_of(len2(bx, by, bz));
    if (TYPE == 0) {
        ((uint8_t*)out)[i] = encode_u8_signed(__fdiv_rn(__fsub_rn(da, db), max_len));
        return;
    }
    uint32_t r, g, bl = 0;
    if (TYPE == 2) {
        r = encode_u8(__fdiv_rn(da, max_len));
        g = encode_u8(__fdiv_rn(db, max_len));
    } else {
        const float q = TYPE == 1 ? max_len : __fsub_rn(da, db);
        r = encode_u8_signed(__fdiv_rn((float)(ax + bx), q));
        g = encode_u8_signed(__fdiv_rn((float)(ay + by), q));
        bl = encode_u8_signed(__fdiv_rn((float)(az + bz), q));
    }
    ((uint32_t*)out)[i] = r | (g << 8) | (bl << 16) | (255u << 24);
}

void launch_two_field_2d(int w, int h, const int16_t* d0, const int16_t* d1, float max_len,
                         int type, uint8_t* out, hipStream_t s)
{
    const size_t n = (size_t)w * h;
    const dim3 grid(grid1d(n, 256)), block(256);
    const uint32_t *a = (const uint32_t*)d0, *b = (const uint32_t*)d1;
    ProfileScope p("encode_two_field_2d", s);
    if (n % 4 == 0 && (((uintptr_t)d0 | (uintptr_t)d1) & 15) == 0 &&
        ((uintptr_t)out & (type == 0 ? 3 : 15)) == 0) {
        const size_t n4 = n / 4;
        const dim3 g4(grid1d(n4, 256));
        const uint4 *a4 = (const uint4*)d0, *b4 = (const uint4*)d1;
        switch (type) {
        case 0: hipLaunchKernelGGL(k_two_field_2d_x4<0>, g4, block, 0, s, n4, a4, b4, max_len, out); break;
        case 1: hipLaunchKernelGGL(k_two_field_2d_x4<1>, g4, block, 0, s, n4, a4, b4, max_len, out); break;
        case 2: hipLaunchKernelGGL(k_two_field_2d_x4<2>, g4, block, 0, s, n4, a4, b4, max_len, out); break;
        default: hipLaunchKernelGGL(k_two_field_2d_x4<3>, g4, block, 0, s, n4, a4, b4, max_len, out); break;
        }
        return;
    }
    switch (type) {
    case 0: hipLaunchKernelGGL(k_two_field_2d<0>, grid, block, 0, s, n, a, b, max_len, out); break;
    case 1: hipLaunchKernelGGL(k_two_field_2d<1>, grid, block, 0, s, n, a, b, max_len, out); break;
    case 2: hipLaunchKernelGGL(k_two_field_2d<2>, grid, block, 0, s, n, a, b, max_len, out); break;
    default: hipLaunchKernelGGL(k_two_field_2d<3>, grid, block, 0, s, n, a, b, max_len, out); break;
    }
}

void launch_two_field_3d(int w, int h, int d, const int16_t* d0, const int16_t* d1, float max_len,
                         int type, uint8_t* out, hipStream_t s)
{
    const size_t slice = (size_t)w * h, n = slice * d;
    const dim3 grid(grid1d(n, 256)), block(256);
    ProfileScope p("encode_two_field_3d", s);
    switch (type) {
    case 0: hipLaunchKernelGGL(k_two_field_3d<0>, grid, block, 0, s, n, slice, d0, d1, max_len, out); break;
    case 1: hipLaunchKernelGGL(k_two_field_3d<1>, grid, block, 0, s, n, slice, d0, d1, max_len, out); break;
    case 2: hipLaunchKernelGGL(k_two_field_3d<2>, grid, block, 0, s, n, slice, d0, d1, max_len, out); break;
    default: hipLaunchKernelGGL(k_two_field_3d<3>, grid, block, 0, s, n, slice, d0, d1, max_len, out); break;
    }
}

// two exact fields -> signed distance sqrt(out) - sqrt(in)
__global__ void __launch_bounds__(256) k_signed_edt(size_t n, const int32_t* __restrict__ dout,
                                                    const int32_t* __restrict__ din,
                                                    float* __restrict__ sdf)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t a = dout[i], b = din[i];
    const float fa = sqrt_d2(a);
    const float fb = sqrt_d2(b);
    sdf[i] = __fsub_rn(fa, fb);
}

void launch_signed_edt(size_t n, const int32_t* dout, const int32_t* din, float* sdf,
                       hipStream_t s)
{
    ProfileScope p("signed_edt_combine", s);
    hipLaunchKernelGGL(k_signed_edt, dim3(grid1d(n, 256)), dim3(256), 0, s, n, dout, din, sdf);
}

// ---- every image of the border method (DistanceOcclusionTool.cpp:1240-1331
// 2D, 1367-1469 3D; DIMS = 2 or 3 int16 components a cell): TYPE 0 sdf, one
// byte a cell, EncodeU8Signed(-/+ Dist / maxLen) negative on set bits; 1
// sdf-delta, the components over maxLen, signed; 2 sdf-rg, EncodeU8(Dist /
// maxLen) in red outside and green inside; 3 sdf-angles, the components over
// Dist, signed.  RGBA alpha 255, blue 0 in 2D.  One row of a slice a
// blockIdx.y, so the mask bit needs no division.
template <int TYPE, int DIMS>
__global__ void __launch_bounds__(256) k_border_image(int w, const uint32_t* __restrict__ mask,
                                                      const int16_t* __restrict__ delta, float max_len,
                                                      uint8_t* __restrict__ out)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= w) return;
    const size_t row = blockIdx.y, i = row * w + x;
    const int16_t* c = delta + (size_t)DIMS * i;
    const int c0 = c[0], c1 = c[1], c2 = DIMS == 3 ? c[2] : 0;
    const float dist = dist_of(DIMS == 3 ? len2(c0, c1, c2) : len2(c0, c1));
    const int b = mask_bit(mask, mask_stride(w), row, x);
    if (TYPE == 0) {
        const float v = __fdiv_rn(dist, max_len);
        out[i] = encode_u8_signed(b ? -v : v);
        return;
    }
    uint32_t r, g, bl;
    if (TYPE == 1) {
        r = encode_u8_signed(__fdiv_rn((float)c0, max_len));
        g = encode_u8_signed(__fdiv_rn((float)c1, max_len));
        bl = DIMS == 3 ? encode_u8_signed(__fdiv_rn((float)c2, max_len)) : 0u;
    } else if (TYPE == 2) {
        const uint32_t e = encode_u8(__fdiv_rn(dist, max_len));
        r = b ? 0u : e;
        g = b ? e : 0u;
        bl = 0u;
    } else {
        r = encode_u8_signed(__fdiv_rn((float)c0, dist));
        g = encode_u8_signed(__fdiv_rn((float)c1, dist));
        bl = DIMS == 3 ? encode_u8_signed(__fdiv_rn((float)c2, dist)) : 0u;
    }
    ((uint32_t*)out)[i] = r | (g << 8) | (bl << 16) | (255u << 24);
}

void launch_border_image(int w, int h, int d, int dims, const uint32_t* mask, const int16_t* delta,
                         float max_len, int type, uint8_t* out, hipStream_t s)
{
    ProfileScope p("encode_border_image", s);
    const dim3 grid((unsigned)((w + 255) / 256), (unsigned)((size_t)h * d)), block(256);
#define DOLHIP_BI(T, D) hipLaunchKernelGGL((k_border_image<T, D>), grid, block, 0, s, w, mask, delta, max_len, out)
    if (dims == 2) {
        switch (type) {
        case 0: DOLHIP_BI(0, 2); break;
        case 1: DOLHIP_BI(1, 2); break;
        case 2: DOLHIP_BI(2, 2); break;
        default: DOLHIP_BI(3, 2); break;
        }
    } else {
        switch (type) {
        case 0: DOLHIP_BI(0, 3); break;
        case 1: DOLHIP_BI(1, 3); break;
        case 2: DOLHIP_BI(2, 3); break;
        default: DOLHIP_BI(3, 3); break;
        }
    }
#undef DOLHIP_BI
}

// ---- unsigned distance images (DistanceOcclusionTool.cpp:678-698, 886-899):
// EncodeU8(Dist / maxLen) of a delta field, EncodeU8(v / maxLen) of scalar
// distances (float, int32, int16; integers converted to float first, as the
// division there does), and sqrtf of float squared distances (:753-754)
template <int DIMS>
__global__ void __launch_bounds__(256) k_distance_u8(size_t n, const int16_t* __restrict__ delta, float max_len,
                                                     uint8_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int16_t* c = delta + (size_t)DIMS * i;
    const int l2 = DIMS == 3 ? len2(c[0], c[1], c[2]) : len2(c[0], c[1]);
    out[i] = encode_u8(__fdiv_rn(dist_of(l2), max_len));
}

template <class T>
__global__ void __launch_bounds__(256) k_scalar_u8(size_t n, const T* __restrict__ v, float max_len,
                                                   uint8_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i] = encode_u8(__fdiv_rn((float)v[i], max_len));
}

__global__ void __launch_bounds__(256) k_sqrt_f32(size_t n, const float* __restrict__ v, float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i] = sqrt_rn(v[i]);
}

void launch_distance_u8(size_t n, int dims, const int16_t* delta, float max_len, uint8_t* out, hipStream_t s)
{
    ProfileScope p("encode_distance_u8", s);
    if (dims == 2)
        hipLaunchKernelGGL(k_distance_u8<2>, dim3(grid1d(n, 256)), dim3(256), 0, s, n, delta, max_len, out);
    else
        hipLaunchKernelGGL(k_distance_u8<3>, dim3(grid1d(n, 256)), dim3(256), 0, s, n, delta, max_len, out);
}

template <class T>
void launch_scalar_u8(size_t n, const T* v, float max_len, uint8_t* out, hipStream_t s)
{
    ProfileScope p("encode_scalar_u8", s);
    hipLaunchKernelGGL(k_scalar_u8<T>, dim3(grid1d(n, 256)), dim3(256), 0, s, n, v, max_len, out);
}
template void launch_scalar_u8<float>(size_t, const float*, float, uint8_t*, hipStream_t);
template void launch_scalar_u8<int32_t>(size_t, const int32_t*, float, uint8_t*, hipStream_t);
template void launch_scalar_u8<int16_t>(size_t, const int16_t*, float, uint8_t*, hipStream_t);

void launch_sqrt_f32(size_t n, const float* v, float* out, hipStream_t s)
{
    ProfileScope p("sqrt_f32", s);
    hipLaunchKernelGGL(k_sqrt_f32, dim3(grid1d(n, 256)), dim3(256), 0, s, n, v, out);
}

}  // namespace dolhip
