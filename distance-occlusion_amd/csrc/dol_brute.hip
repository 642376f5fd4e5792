// dol_brute.hip -- BruteForce 2D and 3D for gfx950 (DistanceOcclusionLib.cpp:1087-1110,
// 2475-2505).
//
// The reference visits the cells in raster order; a cell whose delta is a
// boundary delta at its turn (every component below cw in magnitude) relaxes
// every cell of the field towards itself, in place, so a cell relaxed before
// its own turn may become (or stop being) a source.  One workgroup follows
// that order exactly: it looks for the next source from the cell after the
// last one (a chunk of blockDim cells at a time, the lowest index found by
// an LDS atomic), then all of its lanes relax the whole field with that
// source, then it looks again.  A source's own cell never changes while it
// relaxes (its candidate for itself is its own value, and Closer is strict),
// so every lane reads the same source value.  The field stays in global
// memory; the workgroup's lanes share one L1, so a barrier makes each
// relaxation visible to the next search.  It is a checking tool, as in the
// reference: the work is (sources x cells).
#include "dol_cells.hpp"

#include <climits>

namespace dolhip {

// the lowest index in [start, n) whose cell is a source, or INT_MAX
template <class IsSource>
__device__ int next_source(int start, int n, IsSource is_source, int* s_min)
{
    for (int base = start; base < n; base += blockDim.x) {
        if (threadIdx.x == 0) *s_min = INT_MAX;
        __syncthreads();
        const int i = base + threadIdx.x;
        if (i < n && is_source(i)) atomicMin(s_min, i);
        __syncthreads();
        const int found = *s_min;
        __syncthreads();
        if (found != INT_MAX) return found;
    }
    return INT_MAX;
}

// 2D: a source at column sx, row sy with delta v gives cell (x, y) the
// candidate ((sx - x) cw + v.x, (sy - y) cw + v.y), each stored to int16;
// Closer (strict, 32-bit lengths with wrap) decides (Lib:1088-1110)
__global__ void __launch_bounds__(1024) k_brute_force_2d(int w, int h, uint32_t* __restrict__ cells, int cw)
{
    __shared__ int s_min;
    const int n = w * h;
    auto is_source = [&](int i) {
        const uint32_t v = cells[i];
        return abs(unpack2x(v)) < cw && abs(unpack2y(v)) < cw;
    };
    for (int start = 0;;) {
        const int src = next_source(start, n, is_source, &s_min);
        if (src == INT_MAX) break;
        const uint32_t v = cells[src];
        const int sy = src / w, sx = src - sy * w;
        const int vx = unpack2x(v), vy = unpack2y(v);
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int y = i / w, x = i - y * w;
            const uint32_t c = pack2(wrap16((sx - x) * cw + vx), wrap16((sy - y) * cw + vy));
            const uint32_t cur = cells[i];
            if (len2_packed(cur) > len2_packed(c)) cells[i] = c;
        }
        __syncthreads();
        start = src + 1;
    }
}

// 3D: a source at (si, sj, sk) gives cell (x, y, z) the offset
// (si - x, sj - y, sk - z) if its squared length is below the cell's (the
// reference's 3D form takes neither cw nor the source's delta into the
// candidate, Lib:2476-2505)
__global__ void __launch_bounds__(1024) k_brute_force_3d(int w, int h, int d, int16_t* __restrict__ cells,
                                                         int cw)
{
    __shared__ int s_min;
    const int n = w * h * d;
    auto is_source = [&](int i) {
        const int16_t* c = cells + 3 * (size_t)i;
        return abs((int)c[0]) < cw && abs((int)c[1]) < cw && abs((int)c[2]) < cw;
    };
    const int wh = w * h;
    for (int start = 0;;) {
        const int src = next_source(start, n, is_source, &s_min);
        if (src == INT_MAX) break;
        const int sk = src / wh, rem = src - sk * wh, sj = rem / w, si = rem - sj * w;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int z = i / wh, r = i - z * wh, y = r / w, x = r - y * w;
            const int ox = si - x, oy = sj - y, oz = sk - z;
            const int d0 = ox * ox + oy * oy + oz * oz;
            int16_t* c = cells + 3 * (size_t)i;
            if (d0 < len2(c[0], c[1], c[2])) {
                c[0] = (int16_t)ox;
                c[1] = (int16_t)oy;
                c[2] = (int16_t)oz;
            }
        }
        __syncthreads();
        start = src + 1;
    }
}

void launch_brute_force_2d(int w, int h, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("brute_force_2d", s);
    hipLaunchKernelGGL(k_brute_force_2d, dim3(1), dim3(1024), 0, s, w, h, (uint32_t*)delta, cw);
}

void launch_brute_force_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("brute_force_3d", s);
    hipLaunchKernelGGL(k_brute_force_3d, dim3(1), dim3(1024), 0, s, w, h, d, delta, cw);
}

}  // namespace dolhip
