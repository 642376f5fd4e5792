// dol_raster.hip -- the reference's raster-order methods off the data-parallel
// path, for gfx950: the occlusion image init and the shadow sweep (the
// tool's shadows), Chamfer (its -m 4) and Danielsson (its -m 2).
//
// These are recurrences along rows: a line depends on the finished line
// before it.  Each runs in one workgroup that walks the lines in order with
// a barrier between them (the field stays in global memory; the workgroup's
// lanes share one L1, so a barrier makes a finished line visible):
//   shadow sweep  every cell of a line depends only on the line before:
//                 the line is spread over the lanes;
//   Chamfer       a cell also takes its neighbour along the line plus 3,
//                 which is a prefix minimum: the line is a min-scan over the
//                 lanes (exact in integers);
//   Danielsson    the scans along a row keep the reference's running
//                 "closer" choice, which is not associative (a candidate
//                 dropped once never returns), so one lane runs them over a
//                 copy of the row in LDS; the steps between rows use all lanes.
// Float results are the reference's operations in its order (no FMA
// contraction), so they are bit-identical.
#include "dol_cells.hpp"

namespace dolhip {

// ------------------------------------------------------- occlusion init

// InitOccFromBitMask (DistanceOcclusionLib.cpp:2864-2867): scale on set bits, 0
// elsewhere
__global__ void __launch_bounds__(256) k_init_occ(int w, int h, const uint32_t* __restrict__ mask,
                                                  float* __restrict__ occ, float scale)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    const uint32_t word = mask[(size_t)y * mask_stride(w) + (x >> 5)];
    occ[i] = ((word >> (x & 31)) & 1u) ? scale : 0.0f;
}

void launch_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale, hipStream_t s)
{
    ProfileScope p("init_occ", s);
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_init_occ, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, h, mask, occ,
                       scale);
}

// ------------------------------------------------------- shadow sweep

// ShadowSweepRows / ShadowSweepCols (DistanceOcclusionLib.cpp:2773-2862).  Line
// j (a row if `rows`, else a column, in the sweep's order from the lit side)
// and position t along it map to the cell (x, y); the first line copies the
// occlusion, a later cell takes m = prev(t - 1) g1 + prev(t) g0 (at t = 0,
// m = prev(0) g0) and stores m + (1 - m) occ.
__global__ void __launch_bounds__(1024) k_shadow_sweep(float g, int sx, int sy, int rows, int w, int h,
                                                       const float* __restrict__ occ, float* __restrict__ sh)
{
    const float g1 = g;
    const float g0 = __fsub_rn(1.0f, g1);
    const int lines = rows ? h : w, len = rows ? w : h;
    auto cell = [&](int j, int t) -> size_t {
        const int a = rows ? (sy ? h - 1 - j : j) : (sy ? h - 1 - t : t);   // y
        const int b = rows ? (sx ? w - 1 - t : t) : (sx ? w - 1 - j : j);   // x
        return (size_t)a * w + b;
    };
    for (int t = threadIdx.x; t < len; t += blockDim.x) sh[cell(0, t)] = occ[cell(0, t)];
    __syncthreads();
    for (int j = 1; j < lines; j++) {
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            const float m = t == 0 ? __fmul_rn(sh[cell(j - 1, 0)], g0)
                                   : __fadd_rn(__fmul_rn(sh[cell(j - 1, t - 1)], g1),
                                               __fmul_rn(sh[cell(j - 1, t)], g0));
            const size_t c = cell(j, t);
            sh[c] = __fadd_rn(m, __fmul_rn(__fsub_rn(1.0f, m), occ[c]));
        }
        __syncthreads();
    }
}

void launch_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow, hipStream_t s)
{
    ProfileScope p("shadow_sweep", s);
    // DistanceOcclusionLib.cpp:2869-2876
    const int sx = gx > 0.0f ? 0 : 1;
    const int sy = gy > 0.0f ? 0 : 1;
    const bool rows = fabsf(gx) <= fabsf(gy);
    const float g = rows ? fabsf(gx / gy) : fabsf(gy / gx);
    hipLaunchKernelGGL(k_shadow_sweep, dim3(1), dim3(1024), 0, s, g, sx, sy, (int)rows, w, h, occ, shadow);
}

// ------------------------------------------------------- Chamfer

// Chamfer 3-4 (DistanceOcclusionLib.cpp:912-1006) in 64-bit arithmetic.  A row
// after the first first takes, at each cell, the least of itself and its
// neighbours in the finished row (+3 straight, +4 diagonal; a row of one cell
// has no diagonal), a[x]; the reference's running "row[x] = min(row[x],
// row[x-1] + 3)" then makes row[x] = min over j <= x of a[j] + 3 (x - j)
// = 3x + prefix-min(a[j] - 3j), one inclusive min-scan over the row (the
// reverse pass: the same along the reversed row).  Integers, so exact; the
// sums do not overflow (the reference's int32 form does within 4 of
// INT32_MAX).  A lane takes kChamferRun consecutive cells of a chunk and
// scans them itself; the lanes' totals are scanned across each wave with
// shuffles and across the waves through LDS.
constexpr int kChamferRun = 4;
constexpr long long kChamferInf = 0x3fffffffffffffffLL;

template <class T>
__global__ void __launch_bounds__(1024) k_chamfer(int w, int h, T* __restrict__ dist)
{
    __shared__ long long wave_total[16];
    __shared__ long long carry;
    const long long d0 = 3, d1 = 4;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, waves = blockDim.x >> 6;
    const int chunk = blockDim.x * kChamferRun;
    for (int dir = 1; dir >= -1; dir -= 2) {
        for (int k = 0; k < h; k++) {
            const int y = dir > 0 ? k : h - 1 - k;
            T* r1 = dist + (size_t)y * w;
            const T* r0 = k > 0 ? dist + (size_t)(y - dir) * w : nullptr;
            if (t == 0) carry = kChamferInf;
            __syncthreads();
            for (int base = 0; base < w; base += chunk) {
                // this lane's run, in the scan's direction
                long long v[kChamferRun];
                long long run = kChamferInf;
#pragma unroll
                for (int e = 0; e < kChamferRun; e++) {
                    const int q = base + t * kChamferRun + e;
                    const int x = dir > 0 ? q : w - 1 - q;
                    long long a = kChamferInf;
                    if (q < w) {
                        a = r1[x];
                        if (r0) {
                            a = min(a, (long long)r0[x] + d0);
                            if (x > 0) a = min(a, (long long)r0[x - 1] + d1);
                            if (x + 1 < w) a = min(a, (long long)r0[x + 1] + d1);
                        }
                        a -= d0 * q;
                    }
                    run = min(run, a);
                    v[e] = run;
                }
                // exclusive prefix over the lanes before this one
                long long incl = run;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const long long u = __shfl_up(incl, o, 64);
                    if (lane >= o) incl = min(incl, u);
                }
                if (lane == 63) wave_total[wave] = incl;
                __syncthreads();
                long long before = carry;
                for (int i = 0; i < wave; i++) before = min(before, wave_total[i]);
                long long up = __shfl_up(incl, 1, 64);
                if (lane > 0) before = min(before, up);
#pragma unroll
                for (int e = 0; e < kChamferRun; e++) {
                    const int q = base + t * kChamferRun + e;
                    if (q < w) r1[dir > 0 ? q : w - 1 - q] = (T)(min(v[e], before) + d0 * q);
                }
                __syncthreads();
                if (t == 0) {
                    long long c = carry;
                    for (int i = 0; i < waves; i++) c = min(c, wave_total[i]);
                    carry = c;
                }
                __syncthreads();
            }
        }
    }
}

template <class T>
void launch_chamfer(int w, int h, T* dist, hipStream_t s)
{
    ProfileScope p("chamfer", s);
    hipLaunchKernelGGL(k_chamfer<T>, dim3(1), dim3(1024), 0, s, w, h, dist);
}
template void launch_chamfer<int16_t>(int, int, int16_t*, hipStream_t);
template void launch_chamfer<int32_t>(int, int, int32_t*, hipStream_t);

// ------------------------------------------------------- Danielsson

// Closer of DistanceOcclusionLib.cpp:624-630 on packed (x, y) int16 pairs
__device__ __forceinline__ uint32_t closer2p(uint32_t cur, uint32_t cand)
{
    return len2_packed(cur) > len2_packed(cand) ? cand : cur;
}

// Danielsson 2D (DistanceOcclusionLib.cpp:702-778): per row, in a down then an
// up sweep, the step from the finished row (all lanes), then lane 0's two
// scans over the row in LDS.  `row` is the LDS copy (w cells).
__device__ void danielsson_rows2(int w, int h, uint32_t* cells, int cw, uint32_t* row)
{
    const int t = threadIdx.x, T = blockDim.x;
    const uint32_t mcw = (uint32_t)(uint16_t)(-cw), pcw = (uint32_t)(uint16_t)cw;
    auto add = [](uint32_t v, uint32_t o) {
        jfa_u16x2 a = __builtin_bit_cast(jfa_u16x2, v);
        a += __builtin_bit_cast(jfa_u16x2, o);
        return __builtin_bit_cast(uint32_t, a);
    };
    for (int dir = 1; dir >= -1; dir -= 2) {
        for (int k = 1; k < h; k++) {
            const int y = dir > 0 ? k : h - 1 - k;
            uint32_t* r1 = cells + (size_t)y * w;
            const uint32_t* r0 = cells + (size_t)(y - dir) * w;
            const uint32_t oy = (dir > 0 ? mcw : pcw) << 16;
            for (int i = t; i < w; i += T) row[i] = closer2p(r1[i], add(r0[i], oy));
            __syncthreads();
            if (t == 0) {
                uint32_t b = row[0];
                for (int i = 1; i < w; i++) {
                    b = closer2p(row[i], add(b, mcw));
                    row[i] = b;
                }
                for (int i = w - 2; i >= 0; i--) {
                    b = closer2p(row[i], add(b, pcw));
                    row[i] = b;
                }
            }
            __syncthreads();
            for (int i = t; i < w; i += T) r1[i] = row[i];
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(1024) k_danielsson_2d(int w, int h, uint32_t* __restrict__ cells, int cw)
{
    extern __shared__ uint32_t row[];
    danielsson_rows2(w, h, cells, cw, row);
}

// 3D cells as three int16; the LDS row holds (x, y) pairs and z
__device__ __forceinline__ void closer3p(uint32_t& cxy, int& cz, uint32_t nxy, int nz)
{
    const int lc = len2(unpack2x(cxy), unpack2y(cxy), cz);
    const int ln = len2(unpack2x(nxy), unpack2y(nxy), nz);
    if (lc > ln) {
        cxy = nxy;
        cz = nz;
    }
}

// DanielssonSlice (DistanceOcclusionLib.cpp:2264-2346) of slice `s`
__device__ void danielsson_slice3(int w, int h, int16_t* s, int cw, uint32_t* rxy, int16_t* rz)
{
    const int t = threadIdx.x, T = blockDim.x;
    for (int dir = 1; dir >= -1; dir -= 2) {
        for (int k = 1; k < h; k++) {
            const int y = dir > 0 ? k : h - 1 - k;
            int16_t* r1 = s + 3 * (size_t)y * w;
            const int16_t* r0 = s + 3 * (size_t)(y - dir) * w;
            for (int i = t; i < w; i += T) {
                uint32_t cxy = pack2(r1[3 * i], r1[3 * i + 1]);
                int cz = r1[3 * i + 2];
                closer3p(cxy, cz, pack2(r0[3 * i], wrap16(r0[3 * i + 1] - dir * cw)), r0[3 * i + 2]);
                rxy[i] = cxy;
                rz[i] = (int16_t)cz;
            }
            __syncthreads();
            if (t == 0) {
                uint32_t bxy = rxy[0];
                int bz = rz[0];
                for (int i = 1; i < w; i++) {
                    uint32_t cxy = rxy[i];
                    int cz = rz[i];
                    closer3p(cxy, cz, pack2(wrap16(unpack2x(bxy) - cw), unpack2y(bxy)), bz);
                    rxy[i] = bxy = cxy;
                    rz[i] = (int16_t)(bz = cz);
                }
                for (int i = w - 2; i >= 0; i--) {
                    uint32_t cxy = rxy[i];
                    int cz = rz[i];
                    closer3p(cxy, cz, pack2(wrap16(unpack2x(bxy) + cw), unpack2y(bxy)), bz);
                    rxy[i] = bxy = cxy;
                    rz[i] = (int16_t)(bz = cz);
                }
            }
            __syncthreads();
            for (int i = t; i < w; i += T) {
                r1[3 * i] = (int16_t)unpack2x(rxy[i]);
                r1[3 * i + 1] = (int16_t)unpack2y(rxy[i]);
                r1[3 * i + 2] = rz[i];
            }
            __syncthreads();
        }
    }
}

// Danielsson 3D (DistanceOcclusionLib.cpp:2348-2394): each slice from the
// second takes the slice before with (0, 0, -cw) (all lanes) and is swept as
// a slice; then the same back from the last slice with (0, 0, +cw)
__global__ void __launch_bounds__(1024) k_danielsson_3d(int w, int h, int d, int16_t* __restrict__ cells,
                                                        int cw)
{
    extern __shared__ uint32_t lds[];
    uint32_t* rxy = lds;
    int16_t* rz = (int16_t*)(lds + w);
    const size_t n = (size_t)w * h;
    for (int dir = 1; dir >= -1; dir -= 2) {
        for (int k = 1; k < d; k++) {
            const int z = dir > 0 ? k : d - 1 - k;
            int16_t* s1 = cells + 3 * (size_t)z * n;
            const int16_t* s0 = cells + 3 * (size_t)(z - dir) * n;
            for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
                uint32_t cxy = pack2(s1[3 * i], s1[3 * i + 1]);
                int cz = s1[3 * i + 2];
                closer3p(cxy, cz, pack2(s0[3 * i], s0[3 * i + 1]), wrap16(s0[3 * i + 2] - dir * cw));
                s1[3 * i] = (int16_t)unpack2x(cxy);
                s1[3 * i + 1] = (int16_t)unpack2y(cxy);
                s1[3 * i + 2] = (int16_t)cz;
            }
            __syncthreads();
            danielsson_slice3(w, h, s1, cw, rxy, rz);
        }
    }
}

void launch_danielsson_2d(int w, int h, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("danielsson_2d", s);
    const size_t lds = (size_t)w * 4;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_danielsson_2d, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    hipLaunchKernelGGL(k_danielsson_2d, dim3(1), dim3(1024), lds, s, w, h, (uint32_t*)delta, cw);
}

void launch_danielsson_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("danielsson_3d", s);
    const size_t lds = (size_t)w * 6 + 16;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_danielsson_3d, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    hipLaunchKernelGGL(k_danielsson_3d, dim3(1), dim3(1024), lds, s, w, h, d, delta, cw);
}

}  // namespace dolhip
