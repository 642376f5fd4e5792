// dol_seeds.hip -- distance to the nearest of n continuous seed points on a
// w x h grid of cell centres (x, y), and its gradient with respect to the
// seeds.  (An extension: the reference's seeds are cells of a bitmask.)
//
// Forward:
//   1. each seed marks the cell nearest to it (lowest index wins a cell);
//   2. the library's 2D JumpFlood (cw = 1) from the marked cells gives every
//      cell a candidate: the marked cell its delta points to, and that
//      cell's seed (a delta that leads nowhere, as the reference's wrap
//      rules can make one, leaves the cell without a candidate);
//   3. exact fix-up: seeds are bucketed on a coarse grid, and each cell
//      compares, in float64, every seed of the buckets that meet the square
//      of half-side |c - candidate| around it.  Every seed at least as near
//      as the candidate lies in those buckets, so the result is the exact
//      nearest seed under float64 comparison, ties to the lowest index,
//      whatever JumpFlood returned.
// Backward: grad p_k = sum over cells of seed k of -gout(c) (c - p_k) / d(c)
// (zero where d = 0), summed per workgroup in LDS, then one float atomic
// per (workgroup, seed, component).
#include "../../include/dol_hip.h"
#include "dol_common.hpp"

#include <cfloat>
#include <climits>

namespace dolhip {

int check_launch(const char* what);
void set_error(const char* fmt, ...);
void launch_jfa2d_pass(int step, int w, int h, const int16_t* in, int16_t* out, int cw, hipStream_t s);

namespace seeds {

constexpr int kNone = INT_MAX;

struct Grid {
    int bs;      // bucket side in cells
    int nbx, nby;
};

__host__ __device__ inline Grid make_grid(int w, int h)
{
    Grid g;
    int bs = 8;
    while ((w + bs - 1) / bs > 64 || (h + bs - 1) / bs > 64) bs *= 2;
    g.bs = bs;
    g.nbx = (w + bs - 1) / bs;
    g.nby = (h + bs - 1) / bs;
    return g;
}

__global__ void k_fill(int* p, size_t n, int v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_splat(int w, int h, int n, const float* __restrict__ s, int* __restrict__ idx)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int cx = (int)__builtin_rintf(s[2 * k]), cy = (int)__builtin_rintf(s[2 * k + 1]);
    cx = min(max(cx, 0), w - 1);
    cy = min(max(cy, 0), h - 1);
    atomicMin(&idx[(size_t)cy * w + cx], k);
}

__device__ __forceinline__ float d2f(int x, int y, const float* s, int k)
{
    const float dx = (float)x - s[2 * k], dy = (float)y - s[2 * k + 1];
    return dx * dx + dy * dy;
}

// JumpFlood start field: delta 0 at a marked cell, kMaxDelta2 elsewhere
__global__ void k_seed_start(size_t cells, const int* __restrict__ idx0, uint32_t* __restrict__ delta)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) delta[c] = idx0[c] == kNone ? ((uint32_t)kMaxDelta2 * 0x10001u) : 0u;
}

// counting sort of the seeds into buckets, one workgroup (nb <= 4096); the
// list holds each seed as (x, y, index) so the fix-up reads a seed in one go
__global__ void __launch_bounds__(1024) k_bucket(int n, Grid g, const float* __restrict__ s,
                                                 int* __restrict__ start, float4* __restrict__ list)
{
    __shared__ int cnt[4096 + 1];
    const int nb = g.nbx * g.nby;
    for (int i = threadIdx.x; i <= nb; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    auto bucket_of = [&](int k) {
        int bx = (int)floorf(s[2 * k] / (float)g.bs), by = (int)floorf(s[2 * k + 1] / (float)g.bs);
        bx = min(max(bx, 0), g.nbx - 1);
        by = min(max(by, 0), g.nby - 1);
        return by * g.nbx + bx;
    };
    for (int k = threadIdx.x; k < n; k += blockDim.x) atomicAdd(&cnt[bucket_of(k)], 1);
    __syncthreads();
    // exclusive scan of cnt[0 .. nb) (nb <= 4096): each thread sums a run of
    // four, the runs' sums are scanned across the workgroup (wavefront
    // prefix sums, then the wavefronts' totals), then each run is rewritten
    __shared__ int wsum[1024 / 64];
    const int t = threadIdx.x;
    int run[4], tot = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int b = 4 * t + j;
        run[j] = b < nb ? cnt[b] : 0;
        tot += run[j];
    }
    int incl = tot;   // inclusive prefix of the run sums within the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if ((t & 63) >= o) incl += v;
    }
    if ((t & 63) == 63) wsum[t >> 6] = incl;
    __syncthreads();
    int before = incl - tot;
    for (int v = 0; v < (t >> 6); v++) before += wsum[v];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int b = 4 * t + j;
        if (b < nb) {
            cnt[b] = before;
            start[b] = before;
        }
        before += run[j];
    }
    if (t == blockDim.x - 1) start[nb] = before;   // the total (every run ends by the last thread's)
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int pos = atomicAdd(&cnt[bucket_of(k)], 1);
        list[pos] = make_float4(s[2 * k], s[2 * k + 1], __int_as_float(k), 0.0f);
    }
}

// exact nearest seed: float64 comparison, ties to the lowest index
__global__ void __launch_bounds__(256) k_fixup(int w, int h, Grid g, const float* __restrict__ s,
                                               const int* __restrict__ idx0, const uint32_t* __restrict__ delta,
                                               const int* __restrict__ start, const float4* __restrict__ list,
                                               int32_t* __restrict__ idx, float* __restrict__ dist)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t c = (size_t)y * w + x;
    // the candidate: the seed of the marked cell the delta points to
    int best = kNone;
    {
        const uint32_t d = delta[c];
        const int sx = x + (int16_t)(d & 0xffffu), sy = y + (int16_t)(d >> 16);
        if (sx >= 0 && sx < w && sy >= 0 && sy < h) best = idx0[(size_t)sy * w + sx];
    }
    double bd = DBL_MAX;
    int bx0 = 0, bx1 = g.nbx - 1, by0 = 0, by1 = g.nby - 1;
    if (best != kNone) {
        const double dx = (double)x - (double)s[2 * best], dy = (double)y - (double)s[2 * best + 1];
        bd = dx * dx + dy * dy;
        // every seed with distance <= sqrt(bd) lies in this bucket range;
        // the margin covers the rounding of the square root
        const double r = sqrt(bd) * (1.0 + 1e-12) + 1e-9;
        bx0 = max(0, (int)fmax(-1.0, floor(((double)x - r) / g.bs)));
        bx1 = min(g.nbx - 1, (int)fmin((double)g.nbx, floor(((double)x + r) / g.bs)));
        by0 = max(0, (int)fmax(-1.0, floor(((double)y - r) / g.bs)));
        by1 = min(g.nby - 1, (int)fmin((double)g.nby, floor(((double)y + r) / g.bs)));
    }
    for (int by = by0; by <= by1; by++)
        for (int bx = bx0; bx <= bx1; bx++) {
            const int b = by * g.nbx + bx;
            for (int p = start[b]; p < start[b + 1]; p++) {
                const float4 e = list[p];
                const int k = __float_as_int(e.z);
                const double dx = (double)x - (double)e.x, dy = (double)y - (double)e.y;
                const double d = dx * dx + dy * dy;
                if (d < bd || (d == bd && k < best)) { bd = d; best = k; }
            }
        }
    idx[c] = best;
    dist[c] = (float)sqrt(bd);
}

// backward: per-workgroup LDS table of (seed, gx, gy), filled with one sum a
// wavefront and seed, flushed with atomics
constexpr int kSlots = 128;

__global__ void __launch_bounds__(256) k_seed_grad(int w, int h, const float* __restrict__ s,
                                                   const int32_t* __restrict__ idx,
                                                   const float* __restrict__ dist,
                                                   const float* __restrict__ gout,
                                                   float* __restrict__ grad)
{
    __shared__ int key[kSlots];
    __shared__ float gx[kSlots], gy[kSlots];
    for (int i = threadIdx.x; i < kSlots; i += blockDim.x) {
        key[i] = -1;
        gx[i] = 0.0f;
        gy[i] = 0.0f;
    }
    __syncthreads();
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    int mk = -1;   // this cell's seed, or -1 (no term)
    float vx = 0.0f, vy = 0.0f;
    if (x < w && y < h) {
        const size_t c = (size_t)y * w + x;
        const int k = idx[c];
        const float d = dist[c];
        if (k >= 0 && d > 0.0f) {
            const float a = -gout[c] / d;
            mk = k;
            vx = a * ((float)x - s[2 * k]);
            vy = a * ((float)y - s[2 * k + 1]);
        }
    }
    // The cells of a wavefront belong to few seeds: for each seed in turn,
    // the wavefront sums its cells' terms (butterfly over the lanes) and one
    // lane adds the sum to the workgroup's LDS table (open addressing; a full
    // table falls back to global atomics).
    const int lane = threadIdx.x & 63;
    uint64_t left = __ballot(mk >= 0);
    while (left) {
        const int lead = __ffsll((unsigned long long)left) - 1;
        const int kl = __shfl(mk, lead, 64);
        const bool mine = mk == kl;
        left &= ~__ballot(mine);
        float sx = mine ? vx : 0.0f, sy = mine ? vy : 0.0f;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
        }
        if (lane == lead) {
            int slot = (kl * 40503) & (kSlots - 1);
            bool done = false;
            for (int probe = 0; probe < kSlots; probe++) {
                const int prev = atomicCAS(&key[slot], -1, kl);
                if (prev == -1 || prev == kl) {
                    atomicAdd(&gx[slot], sx);
                    atomicAdd(&gy[slot], sy);
                    done = true;
                    break;
                }
                slot = (slot + 1) & (kSlots - 1);
            }
            if (!done) {
                atomicAdd(&grad[2 * kl], sx);
                atomicAdd(&grad[2 * kl + 1], sy);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSlots; i += blockDim.x) {
        const int k = key[i];
        if (k >= 0) {
            atomicAdd(&grad[2 * k], gx[i]);
            atomicAdd(&grad[2 * k + 1], gy[i]);
        }
    }
}

}  // namespace seeds
}  // namespace dolhip

using namespace dolhip;
using namespace dolhip::seeds;

namespace {

inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

int bad(const char* m)
{
    set_error("%s", m);
    return DOLHIP_ERR_INVALID_ARG;
}

}  // namespace

extern "C" {

size_t dolhip_seed_workspace(int w, int h, int n)
{
    if (w < 1 || h < 1 || n < 1) return 0;
    const Grid g = make_grid(w, h);
    const size_t cells = (size_t)w * h;
    return 3 * al(cells * 4) + al(((size_t)g.nbx * g.nby + 1) * 4) + al((size_t)n * 16);
}

int dolhip_seed_nearest_2d_dev(int w, int h, int n, const float* seeds, int32_t* idx, float* dist,
                               void* work, void* stream)
{
    if (w < 1 || h < 1 || n < 1) return bad("need w, h, n >= 1");
    if ((size_t)w * h > ((size_t)1 << 31)) return bad("grid too large");
    if (!seeds || !idx || !dist || !work) return bad("seeds, idx, dist and work must be set");
    hipStream_t s = (hipStream_t)stream;
    const Grid g = make_grid(w, h);
    if (g.nbx * g.nby > 4096) return bad("bucket grid too large");
    const size_t cells = (size_t)w * h;
    char* p = (char*)work;
    int* idx0 = (int*)p;
    uint32_t* da = (uint32_t*)(p + al(cells * 4));
    uint32_t* db = (uint32_t*)(p + 2 * al(cells * 4));
    int* start = (int*)(p + 3 * al(cells * 4));
    float4* list = (float4*)(p + 3 * al(cells * 4) + al(((size_t)g.nbx * g.nby + 1) * 4));
    {
        ProfileScope ps("seed_splat", s);
        hipLaunchKernelGGL(k_fill, dim3(grid1d(cells, 256)), dim3(256), 0, s, idx0, cells, kNone);
        hipLaunchKernelGGL(k_splat, dim3(grid1d((size_t)n, 256)), dim3(256), 0, s, w, h, n, seeds, idx0);
        hipLaunchKernelGGL(k_bucket, dim3(1), dim3(1024), 0, s, n, g, seeds, start, list);
        hipLaunchKernelGGL(k_seed_start, dim3(grid1d(cells, 256)), dim3(256), 0, s, cells, idx0, da);
    }
    {
        // the reference's pass schedule (steps FloorPow2(max - 1) .. 1, and
        // one more step-1 pass when the count is odd), cw = 1
        ProfileScope ps("seed_jfa", s);
        int m = (w >= h ? w : h) - 1, step = 1;
        while (step * 2 <= m) step *= 2;
        uint32_t* a = da;
        uint32_t* b = db;
        for (; step > 0 && m > 0; step /= 2) {
            launch_jfa2d_pass(step, w, h, (const int16_t*)a, (int16_t*)b, 1, s);
            uint32_t* t = a; a = b; b = t;
        }
        if (a != da) launch_jfa2d_pass(1, w, h, (const int16_t*)a, (int16_t*)da, 1, s);
    }
    const dim3 grid((w + 63) / 64, (h + 3) / 4);
    {
        ProfileScope ps("seed_fixup", s);
        hipLaunchKernelGGL(k_fixup, grid, dim3(256), 0, s, w, h, g, seeds, idx0, da, start, list, idx, dist);
    }
    return check_launch("seed_nearest");
}

int dolhip_seed_grad_2d_dev(int w, int h, int n, const float* seeds, const int32_t* idx,
                            const float* dist, const float* gout, float* grad, void* stream)
{
    if (w < 1 || h < 1 || n < 1) return bad("need w, h, n >= 1");
    if (!seeds || !idx || !dist || !gout || !grad) return bad("all pointers must be set");
    hipStream_t s = (hipStream_t)stream;
    ProfileScope ps("seed_grad", s);
    hipLaunchKernelGGL(k_fill, dim3(grid1d((size_t)2 * n, 256)), dim3(256), 0, s, (int*)grad,
                       (size_t)2 * n, 0);
    hipLaunchKernelGGL(k_seed_grad, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, s, w, h, seeds,
                       idx, dist, gout, grad);
    return check_launch("seed_grad");
}

}  // extern "C"
