// dol_seeds.hip -- distance to the nearest of n continuous seed points on a
// w x h grid of cell centres (x, y), and its gradient with respect to the
// seeds.  (An extension: the reference's seeds are cells of a bitmask.)
//
// Forward:
//   1. each seed marks the cell nearest to it (lowest index wins a cell);
//   2. JumpFlood on the seed-index field (steps w/2 ... 1, then one more
//      step-1 pass) gives every cell a candidate seed;
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

// one JumpFlood pass on the index field (float distances; only a candidate)
__global__ void __launch_bounds__(256) k_seed_jfa(int w, int h, int step, const float* __restrict__ s,
                                                  const int* __restrict__ in, int* __restrict__ out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    int best = in[(size_t)y * w + x];
    float bd = best == kNone ? FLT_MAX : d2f(x, y, s, best);
    for (int dy = -1; dy <= 1; dy++) {
        const int yy = y + dy * step;
        if (yy < 0 || yy >= h) continue;
        for (int dx = -1; dx <= 1; dx++) {
            const int xx = x + dx * step;
            if (xx < 0 || xx >= w || (dx == 0 && dy == 0)) continue;
            const int k = in[(size_t)yy * w + xx];
            if (k == kNone) continue;
            const float d = d2f(x, y, s, k);
            if (d < bd || (d == bd && k < best)) { bd = d; best = k; }
        }
    }
    out[(size_t)y * w + x] = best;
}

// counting sort of the seeds into buckets, one workgroup (nb <= 4096)
__global__ void __launch_bounds__(1024) k_bucket(int n, Grid g, const float* __restrict__ s,
                                                 int* __restrict__ start, int* __restrict__ list)
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
    if (threadIdx.x == 0) {   // exclusive scan (nb <= 4096, once per call)
        int acc = 0;
        for (int b = 0; b < nb; b++) {
            const int c = cnt[b];
            cnt[b] = acc;
            start[b] = acc;
            acc += c;
        }
        start[nb] = acc;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int pos = atomicAdd(&cnt[bucket_of(k)], 1);
        list[pos] = k;
    }
}

// exact nearest seed: float64 comparison, ties to the lowest index
__global__ void __launch_bounds__(256) k_fixup(int w, int h, Grid g, const float* __restrict__ s,
                                               const int* __restrict__ cand, const int* __restrict__ start,
                                               const int* __restrict__ list, int32_t* __restrict__ idx,
                                               float* __restrict__ dist)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t c = (size_t)y * w + x;
    int best = cand[c];
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
                const int k = list[p];
                const double dx = (double)x - (double)s[2 * k], dy = (double)y - (double)s[2 * k + 1];
                const double d = dx * dx + dy * dy;
                if (d < bd || (d == bd && k < best)) { bd = d; best = k; }
            }
        }
    idx[c] = best;
    dist[c] = (float)sqrt(bd);
}

// backward: per-workgroup LDS table of (seed, gx, gy), flushed with atomics
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
    if (x < w && y < h) {
        const size_t c = (size_t)y * w + x;
        const int k = idx[c];
        const float d = dist[c];
        if (k >= 0 && d > 0.0f) {
            const float a = -gout[c] / d;
            const float vx = a * ((float)x - s[2 * k]), vy = a * ((float)y - s[2 * k + 1]);
            // open addressing in LDS; a full table falls back to global atomics
            int slot = (k * 40503) & (kSlots - 1);
            bool done = false;
            for (int probe = 0; probe < kSlots; probe++) {
                const int prev = atomicCAS(&key[slot], -1, k);
                if (prev == -1 || prev == k) {
                    atomicAdd(&gx[slot], vx);
                    atomicAdd(&gy[slot], vy);
                    done = true;
                    break;
                }
                slot = (slot + 1) & (kSlots - 1);
            }
            if (!done) {
                atomicAdd(&grad[2 * k], vx);
                atomicAdd(&grad[2 * k + 1], vy);
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
    return 2 * al(cells * 4) + al(((size_t)g.nbx * g.nby + 1) * 4) + al((size_t)n * 4);
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
    int* fa = (int*)p;
    int* fb = (int*)(p + al(cells * 4));
    int* start = (int*)(p + 2 * al(cells * 4));
    int* list = (int*)(p + 2 * al(cells * 4) + al(((size_t)g.nbx * g.nby + 1) * 4));
    {
        ProfileScope ps("seed_splat", s);
        hipLaunchKernelGGL(k_fill, dim3(grid1d(cells, 256)), dim3(256), 0, s, fa, cells, kNone);
        hipLaunchKernelGGL(k_splat, dim3(grid1d((size_t)n, 256)), dim3(256), 0, s, w, h, n, seeds, fa);
        hipLaunchKernelGGL(k_bucket, dim3(1), dim3(1024), 0, s, n, g, seeds, start, list);
    }
    const dim3 grid((w + 63) / 64, (h + 3) / 4);
    int m = w >= h ? w : h;
    int step = 1;
    while (step * 2 < m) step *= 2;
    int* a = fa;
    int* b = fb;
    {
        ProfileScope ps("seed_jfa", s);
        for (; step >= 1; step /= 2) {
            hipLaunchKernelGGL(k_seed_jfa, grid, dim3(256), 0, s, w, h, step, seeds, a, b);
            int* t = a; a = b; b = t;
        }
        hipLaunchKernelGGL(k_seed_jfa, grid, dim3(256), 0, s, w, h, 1, seeds, a, b);
        int* t = a; a = b; b = t;
    }
    {
        ProfileScope ps("seed_fixup", s);
        hipLaunchKernelGGL(k_fixup, grid, dim3(256), 0, s, w, h, g, seeds, a, start, list, idx, dist);
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
