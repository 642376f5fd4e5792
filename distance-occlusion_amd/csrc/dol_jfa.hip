// dol_jfa.hip -- JumpFlood passes for gfx950.
//
// Each pass is out of place, as the reference's (DistanceOcclusionLib.cpp:782-873,
// 2398-2447): a cell takes the nearest of its own delta and the deltas at
// +-step, each offset by cw*step and stored to int16 with wrap, comparing
// 32-bit squared lengths with a strict ">" so that the first of equal
// candidates in the reference's tap order wins.  The pass schedule (steps
// FloorPow2(max dim - 1) down to 1, plus one step-1 pass when the count is
// odd so the result ends in the caller's buffer) is the host driver's.
#include "dol_cells.hpp"

namespace dolhip {

// ------------------------------------------------------------------ 2D

// one thread per cell; see jfa2_cell (dol_cells.hpp)
__global__ void __launch_bounds__(256) k_jfa2d(int step, int w, int h, int xb,
                                               const uint32_t* __restrict__ in,
                                               uint32_t* __restrict__ out, int cw)
{
    const int y = blockIdx.x / xb;
    const int x = (blockIdx.x - y * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    out[(size_t)y * w + x] = jfa2_cell(in, w, h, x, y, step, cw);
}

void launch_jfa2d_pass(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("jfa2d_pass", s);
    hipLaunchKernelGGL(k_jfa2d, dim3((unsigned)(xb * h)), dim3(256), 0, s, step, w, h, xb,
                       (const uint32_t*)in, (uint32_t*)out, cw);
}

// ------------------------------------------------------------------ 3D

// 27 taps, z offset outermost, then y, then x, each -s, 0, +s where inside
// the volume; a source whose x equals kMaxDelta3 is skipped
// (DistanceOcclusionLib.cpp:2423-2438).
__global__ void __launch_bounds__(256) k_jfa3d(int step, int w, int h, int d, int xb,
                                               const int16_t* __restrict__ in,
                                               int16_t* __restrict__ out, int cw)
{
    const int row = blockIdx.x / xb;   // z*h + y
    const int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    const int z = row / h, y = row - z * h;
    const size_t slice = (size_t)w * h;
    const size_t idx = (size_t)row * w + x;
    const int16_t* c0 = in + 3 * idx;
    int bx = c0[0], by = c0[1], bz = c0[2];
    int bl = len2(bx, by, bz);
    const int off = cw * step;
    const int kc0 = z - step >= 0 ? -1 : 0, kc1 = z + step < d ? 1 : 0;
    const int jc0 = y - step >= 0 ? -1 : 0, jc1 = y + step < h ? 1 : 0;
    const int ic0 = x - step >= 0 ? -1 : 0, ic1 = x + step < w ? 1 : 0;
    for (int kc = kc0; kc <= kc1; kc++)
        for (int jc = jc0; jc <= jc1; jc++)
            for (int ic = ic0; ic <= ic1; ic++) {
                const ptrdiff_t o = (ptrdiff_t)kc * step * (ptrdiff_t)slice
                                  + (ptrdiff_t)jc * step * w + (ptrdiff_t)ic * step;
                const int16_t* src = in + 3 * ((ptrdiff_t)idx + o);
                const int sx = src[0];
                if (sx == kMaxDelta3) continue;
                const int cx = wrap16(sx + ic * off), cy = wrap16(src[1] + jc * off),
                          cz = wrap16(src[2] + kc * off);
                const int cl = len2(cx, cy, cz);
                if (bl > cl) { bx = cx; by = cy; bz = cz; bl = cl; }
            }
    int16_t* o = out + 3 * idx;
    o[0] = (int16_t)bx; o[1] = (int16_t)by; o[2] = (int16_t)bz;
}

void launch_jfa3d_pass(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("jfa3d_pass", s);
    hipLaunchKernelGGL(k_jfa3d, dim3((unsigned)((size_t)xb * h * d)), dim3(256), 0, s, step, w, h,
                       d, xb, in, out, cw);
}

}  // namespace dolhip
