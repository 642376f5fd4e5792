// dol_init.hip -- mask packing and delta/distance initialisation for gfx950.
//
// Border initialisation: the reference builds the signed border seeds by
// scattering writes from 2x2 (2x2x2) mask windows, row pair after row pair,
// word after word, and lets later writes win (SetBorderCellsCombo,
// DistanceOcclusionLib.cpp:504-599 and 1820-2150).  Here every thread owns one
// output cell and replays, in that same global order, only the writes that
// land on its cell.  The windows that can reach cell (x, y[, z]) are those at
// bit b in {x, x+1}, row pair q in {y-1, y} and slice pair p in {z-1, z}, so a
// thread needs just the 3x3(x3) pixels around it; cells whose neighbourhood is
// uniform receive no write and stay at kMaxDelta.  In 2D the same replay is
// evaluated for the 32 cells of a mask word at once, as bit masks.
#include "dol_cells.hpp"

namespace dolhip {

// ------------------------------------------------------------ byte -> bit

// One thread per padded pixel; 32 consecutive threads form one mask word,
// gathered with a wavefront ballot (threshold rule of
// DistanceOcclusionTool.cpp:662).
__global__ void __launch_bounds__(256) k_mask_from_bytes(int w, int ws, size_t total,
                                                         const uint8_t* __restrict__ bytes,
                                                         int threshold, int less_than,
                                                         uint32_t* __restrict__ mask)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int pred = 0;
    if (t < total) {
        size_t row = t / (size_t)(ws * 32);
        int x = (int)(t - row * (size_t)(ws * 32));
        if (x < w) {
            int v = bytes[row * (size_t)w + x];
            pred = ((v >= threshold) != (less_than != 0));
        }
    }
    unsigned long long b = __ballot(pred);
    int lane = threadIdx.x & 63;
    if ((lane & 31) == 0 && t < total) {
        uint32_t word = lane == 0 ? (uint32_t)(b & 0xffffffffull) : (uint32_t)(b >> 32);
        mask[t >> 5] = word;
    }
}

void launch_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                            uint32_t* mask, hipStream_t s)
{
    int ws = mask_stride(w);
    size_t total = (size_t)rows * ws * 32;
    ProfileScope p("mask_from_bytes", s);
    hipLaunchKernelGGL(k_mask_from_bytes, dim3(grid1d(total, 256)), dim3(256), 0, s, w, ws, total,
                       bytes, threshold, less_than, mask);
}

// ------------------------------------------------------------ unsigned init

// rows x ceil(w/256) blocks; block b covers row b / xb, columns (b % xb)*256 + tid
__global__ void __launch_bounds__(256) k_init_deltas_2d(int w, int xb, const uint32_t* __restrict__ mask,
                                                        uint32_t* __restrict__ delta, uint32_t invert)
{
    int row = blockIdx.x / xb;
    int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    int ws = mask_stride(w);
    uint32_t bit = ((mask[(size_t)row * ws + (x >> 5)] >> (x & 31)) & 1u) ^ invert;
    delta[(size_t)row * w + x] = bit ? 0u : pack2(kMaxDelta2, kMaxDelta2);
}

// w % 4 == 0: four cells a thread over the flat image, one 16-byte store
__global__ void __launch_bounds__(256) k_init_deltas_2d_x4(int w, size_t n4, const uint32_t* __restrict__ mask,
                                                           uint4* __restrict__ delta, uint32_t invert)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n4) return;
    const size_t i = 4 * q, row = i / (size_t)w;
    const int x = (int)(i - row * (size_t)w);
    const uint32_t bits = (mask[row * mask_stride(w) + (x >> 5)] >> (x & 31)) ^ (invert ? 15u : 0u);
    const uint32_t far = pack2(kMaxDelta2, kMaxDelta2);
    delta[q] = make_uint4(bits & 1u ? 0u : far, bits & 2u ? 0u : far, bits & 4u ? 0u : far,
                          bits & 8u ? 0u : far);
}

__global__ void __launch_bounds__(256) k_init_deltas_3d(int w, int xb, const uint32_t* __restrict__ mask,
                                                        int16_t* __restrict__ delta, uint32_t invert)
{
    int row = blockIdx.x / xb;
    int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    int ws = mask_stride(w);
    uint32_t bit = ((mask[(size_t)row * ws + (x >> 5)] >> (x & 31)) & 1u) ^ invert;
    int16_t v = bit ? (int16_t)0 : (int16_t)kMaxDelta3;
    int16_t* c = delta + 3 * ((size_t)row * w + x);
    c[0] = v; c[1] = v; c[2] = v;
}

void launch_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert,
                           hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("init_deltas_2d", s);
    if (w % 4 == 0 && ((uintptr_t)delta & 15) == 0) {
        const size_t n4 = (size_t)w * h / 4;
        hipLaunchKernelGGL(k_init_deltas_2d_x4, dim3(grid1d(n4, 256)), dim3(256), 0, s, w, n4, mask,
                           (uint4*)delta, invert ? 1u : 0u);
        return;
    }
    hipLaunchKernelGGL(k_init_deltas_2d, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                       (uint32_t*)delta, invert ? 1u : 0u);
}

void launch_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert,
                           hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("init_deltas_3d", s);
    hipLaunchKernelGGL(k_init_deltas_3d, dim3((unsigned)((size_t)xb * h * d)), dim3(256), 0, s, w,
                       xb, mask, delta, invert ? 1u : 0u);
}

template <class T>
__global__ void __launch_bounds__(256) k_init_dist(int w, int xb, const uint32_t* __restrict__ mask,
                                                   T* __restrict__ dist, T max_d)
{
    int row = blockIdx.x / xb;
    int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    int ws = mask_stride(w);
    uint32_t bit = (mask[(size_t)row * ws + (x >> 5)] >> (x & 31)) & 1u;
    dist[(size_t)row * w + x] = bit ? (T)0 : max_d;
}

template <class T>
void launch_init_dist(int w, int h, const uint32_t* mask, T* dist, T max_d, hipStream_t s)
{
    int xb = (w + 255) / 256;
    ProfileScope p("init_distances", s);
    hipLaunchKernelGGL(k_init_dist<T>, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                       dist, max_d);
}
template void launch_init_dist<float>(int, int, const uint32_t*, float*, float, hipStream_t);
template void launch_init_dist<int32_t>(int, int, const uint32_t*, int32_t*, int32_t, hipStream_t);
template void launch_init_dist<int16_t>(int, int, const uint32_t*, int16_t*, int16_t, hipStream_t);

// ------------------------------------------------------------ border init 2D

// Pixel x-1, x and x+1 of the 32 cells of mask word j of a row, as three
// words (bit i = cell 32j + i): x = -1 replicates x = 0, and the right
// neighbour of the row's last cell is made equal to the cell so that it
// never counts.
struct RowBits {
    uint32_t l, m, r;
};

__device__ __forceinline__ RowBits row_bits(const uint32_t* row, int j, int ws, int w)
{
    RowBits b;
    b.m = row[j];
    const uint32_t prev = j > 0 ? row[j - 1] : 0u, next = j + 1 < ws ? row[j + 1] : 0u;
    b.l = (b.m << 1) | (j > 0 ? prev >> 31 : (b.m & 1u));
    b.r = (b.m >> 1) | (next << 31);
    if (j == ws - 1) {
        const uint32_t last = 1u << ((w - 1) & 31);
        b.r = (b.r & ~last) | (b.m & last);
    }
    return b;
}

// Writes of SetBorderCellsCombo (DistanceOcclusionLib.cpp:504-599) to the 32
// cells of one mask word, for all cells at once: bit i of each word belongs
// to cell 32j + i.  It is the per-cell replay of border2_cell with every
// condition evaluated as a bit mask.  Among the writes that land on a cell the
// last unconditional one wins (order e3, f3, f4, f5; at a word's last bit the
// reference's loop order makes it e3, f3, f5, f4), and with none of those the
// first conditional one (e1, e2, f1, f2).
struct Border2Word {
    uint32_t any, xn, xp, yn, yp;   // written; x = -1; x = +1; y = -1; y = +1
};

__device__ __forceinline__ Border2Word border2_word(const RowBits& up, const RowBits& mid,
                                                    const RowBits& dn, bool has_up, bool has_dn,
                                                    uint32_t right)
{
    uint32_t e1 = 0, e2 = 0, e3 = 0, f1 = 0, f2 = 0, f5 = 0;
    if (has_up) {   // row pair (y-1, y): this row is the lower one
        const uint32_t vert0 = up.m ^ mid.m, horz0 = up.m ^ up.l, vprev0 = up.l ^ mid.l;
        const uint32_t vert1 = up.r ^ mid.r, horz1 = up.r ^ up.m;
        e1 = (vert0 ^ vprev0) & vprev0 & horz0;
        e2 = right & (vert1 ^ vert0) & ~vert0 & horz1;
        e3 = vert0;
    }
    const uint32_t horz0 = mid.m ^ mid.l, horz1 = mid.r ^ mid.m;
    if (has_dn) {   // row pair (y, y+1): this row is the upper one
        const uint32_t vert0 = mid.m ^ dn.m, vprev0 = mid.l ^ dn.l, vert1 = mid.r ^ dn.r;
        f1 = (vert0 ^ vprev0) & vprev0 & ~horz0;
        f2 = right & (vert1 ^ vert0) & ~vert0 & ~horz1;
        f5 = vert0;
    }
    const uint32_t f3 = horz0, f4 = right & horz1;
    const uint32_t last = 0x80000000u;
    const uint32_t s5 = f5 & ~(f4 & last), s4 = f4 & ~(f5 & ~last);
    const uint32_t s3 = f3 & ~f4 & ~f5, se = e3 & ~f3 & ~f4 & ~f5;
    const uint32_t set = e3 | f3 | f4 | f5;
    const uint32_t u1 = ~set & e1, u2 = ~set & ~e1 & e2;
    const uint32_t u3 = ~set & ~e1 & ~e2 & f1, u4 = ~set & ~e1 & ~e2 & ~f1 & f2;
    Border2Word b;
    b.any = set | e1 | e2 | f1 | f2;
    b.xn = s3 | u1 | u3;
    b.xp = s4 | u2 | u4;
    b.yn = se | u1 | u2;
    b.yp = s5 | u3 | u4;
    return b;
}

__device__ __forceinline__ uint32_t border2_value(const Border2Word& b, int i)
{
    if (!((b.any >> i) & 1u)) return pack2(kMaxDelta2, kMaxDelta2);
    const int sx = (int)((b.xp >> i) & 1u) - (int)((b.xn >> i) & 1u);
    const int sy = (int)((b.yp >> i) & 1u) - (int)((b.yn >> i) & 1u);
    return pack2(sx, sy);
}

// Border masks of mask word j of row y.
__device__ __forceinline__ Border2Word border2_masks(const uint32_t* mask, int ws, int w, int h,
                                                     int j, int y)
{
    const RowBits zero{0u, 0u, 0u};
    const bool has_up = y >= 1, has_dn = y + 1 < h;
    const RowBits mid = row_bits(mask + (size_t)y * ws, j, ws, w);
    const RowBits up = has_up ? row_bits(mask + (size_t)(y - 1) * ws, j, ws, w) : zero;
    const RowBits dn = has_dn ? row_bits(mask + (size_t)(y + 1) * ws, j, ws, w) : zero;
    const uint32_t right = j == ws - 1 ? ~(1u << ((w - 1) & 31)) : ~0u;
    return border2_word(up, mid, dn, has_up, has_dn, right);
}

// One thread per 32 cells (one mask word of a row): the word's fill value is
// stored in bulk and the written cells are overwritten.
__global__ void __launch_bounds__(256) k_border_2d_words(int w, int h,
                                                         const uint32_t* __restrict__ mask,
                                                         uint32_t* __restrict__ delta)
{
    const int ws = mask_stride(w);
    const int gid = blockIdx.x * 256 + threadIdx.x;
    if (gid >= ws * h) return;
    const int y = gid / ws, j = gid - y * ws;
    const int x0 = 32 * j, n = min(32, w - x0);
    const Border2Word b = border2_masks(mask, ws, w, h, j, y);
    uint32_t any = n < 32 ? b.any & ((1u << n) - 1u) : b.any;
    uint32_t* out = delta + (size_t)y * w + x0;
    const uint32_t fill = pack2(kMaxDelta2, kMaxDelta2);
    if (n == 32 && ((uintptr_t)out & 15) == 0) {
#pragma unroll
        for (int k = 0; k < 8; k++) ((uint4*)out)[k] = make_uint4(fill, fill, fill, fill);
    } else {
        for (int k = 0; k < n; k++) out[k] = fill;
    }
    while (any) {
        const int i = __ffs(any) - 1;
        any &= any - 1u;
        out[i] = border2_value(b, i);
    }
}

// Same result for w % 32 == 0, where mask word g covers the flat cells
// 32g .. 32g+31: a wave's 64 words are 2048 contiguous cells, which it writes
// as 8 stores of 64 consecutive 16-byte chunks (4 cells a lane).  A chunk
// whose 8 words have no written cell is the fill; otherwise its lanes take the
// masks from the word's owning lane by shuffles.
__global__ void __launch_bounds__(256) k_border_2d_flat(int w, int h,
                                                        const uint32_t* __restrict__ mask,
                                                        uint4* __restrict__ delta)
{
    const int ws = w >> 5;
    const int words = ws * h;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int word0 = gid - lane;
    Border2Word b{0u, 0u, 0u, 0u, 0u};
    if (gid < words) {
        const int y = gid / ws;
        b = border2_masks(mask, ws, w, h, gid - y * ws, y);
    }
    const uint64_t busy = __ballot(b.any != 0u);
    const uint32_t fill = pack2(kMaxDelta2, kMaxDelta2);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int u = k * 64 + lane;
        const int g = word0 + (u >> 3);
        uint4 v = make_uint4(fill, fill, fill, fill);
        if ((busy >> (8 * k)) & 0xffu) {
            const int src = u >> 3;
            Border2Word c;
            c.any = __shfl(b.any, src);
            c.xn = __shfl(b.xn, src);
            c.xp = __shfl(b.xp, src);
            c.yn = __shfl(b.yn, src);
            c.yp = __shfl(b.yp, src);
            const int q = (u & 7) * 4;
            v = make_uint4(border2_value(c, q), border2_value(c, q + 1), border2_value(c, q + 2),
                           border2_value(c, q + 3));
        }
        if (g < words) delta[(size_t)g * 8 + (u & 7)] = v;
    }
}

void launch_border_2d(int w, int h, const uint32_t* mask, int16_t* delta, hipStream_t s)
{
    ProfileScope p("border_2d", s);
    const int words = mask_stride(w) * h;
    const dim3 grid((unsigned)((words + 255) / 256));
    if ((w & 31) == 0 && ((uintptr_t)delta & 15) == 0)
        hipLaunchKernelGGL(k_border_2d_flat, grid, dim3(256), 0, s, w, h, mask, (uint4*)delta);
    else
        hipLaunchKernelGGL(k_border_2d_words, grid, dim3(256), 0, s, w, h, mask,
                           (uint32_t*)delta);
}

// ------------------------------------------------------------ border init 3D

struct Cell3 {
    int x, y, z;
    __device__ void set(int vx, int vy, int vz) { x = vx; y = vy; z = vz; }
};

// Loop kinds of one window, in the reference's order
// (DistanceOcclusionLib.cpp:1899-1973)
enum : int { L_BC = 1, L_FXX = 2, L_FY = 4, L_FZ = 8, L_EX = 16, L_EYX = 32, L_EZX = 64 };

// Pixel cube around the cell: bit (dz*9 + dy*3 + dx) is pixel
// (x-1+dx, y-1+dy, z-1+dz), x = -1 replicating x = 0; outside the volume 0.
__device__ __forceinline__ int cube_bit(uint32_t cube, int dx, int dy, int dz)
{
    return (int)((cube >> (dz * 9 + dy * 3 + dx)) & 1u);
}

struct Win3 {
    int ex, ey, ez, ezx, eyx, fx, fy, fz, fxx, bc;
};

// window whose "x-1" column is cube column kx (0 or 1), rows ky/ky+1,
// slices kz/kz+1 (DistanceOcclusionLib.cpp:1862-1894)
__device__ __forceinline__ Win3 window3(uint32_t cube, int kx, int ky, int kz)
{
    const int p000 = cube_bit(cube, kx, ky, kz), p001 = cube_bit(cube, kx + 1, ky, kz);
    const int p010 = cube_bit(cube, kx, ky + 1, kz), p011 = cube_bit(cube, kx + 1, ky + 1, kz);
    const int p100 = cube_bit(cube, kx, ky, kz + 1), p101 = cube_bit(cube, kx + 1, ky, kz + 1);
    const int p110 = cube_bit(cube, kx, ky + 1, kz + 1), p111 = cube_bit(cube, kx + 1, ky + 1, kz + 1);
    Win3 r;
    r.ex = p000 ^ p001;
    r.ey = p000 ^ p010;
    r.ez = p000 ^ p100;
    r.ezx = p001 ^ p101;
    const int eyz = p100 ^ p110, exy = p010 ^ p011;
    r.fx = r.ey ^ eyz;
    r.fy = r.ez ^ r.ezx;
    r.fz = r.ex ^ exy;
    r.eyx = p001 ^ p011;
    const int eyxz = p101 ^ p111;
    r.fxx = r.eyx ^ eyxz;
    r.bc = r.fz ^ eyz ^ eyxz;
    return r;
}

// Apply loop `kind` of window (b, q, p) to the cell (tx, ty, tz relative
// offsets: the cell sits at (b - bx, q - qy, p - pz) for window origin
// offsets bx in {0, 1} (b = x or x+1), etc.  Coordinates here are relative to
// the window: the cell's position inside the 2x2x2 box is (cx, cy, cz) with
// cx in {0: column b-1, 1: column b}, cy in {0: row q, 1: row q+1},
// cz in {0: slice p, 1: slice p+1}.
__device__ __forceinline__ void apply3(Cell3& s, const Win3& v, int kind, int cx, int cy, int cz)
{
    switch (kind) {
    case L_BC:
        if (v.bc && cx == v.fx && cy == v.fy && cz == v.fz && s.x == kMaxDelta3)
            s.set(1 - 2 * v.fx, 1 - 2 * v.fy, 1 - 2 * v.fz);
        break;
    case L_FXX:
        if (v.fxx && cx == 1 && cy == v.ezx && cz == v.eyx &&
            (s.x == kMaxDelta3 || abs(s.x) + abs(s.y) + abs(s.z) > 1))
            s.set(0, 1 - 2 * v.ezx, 1 - 2 * v.eyx);
        break;
    case L_FY:
        if (v.fy && cx == v.ez && cy == 0 && cz == v.ex &&
            (s.x == kMaxDelta3 || abs(s.x) + abs(s.y) + abs(s.z) > 1))
            s.set(1 - 2 * v.ez, 0, 1 - 2 * v.ex);
        break;
    case L_FZ:
        if (v.fz && cx == v.ey && cy == v.ex && cz == 0 &&
            (s.x == kMaxDelta3 || abs(s.x) + abs(s.y) + abs(s.z) > 1))
            s.set(1 - 2 * v.ey, 1 - 2 * v.ex, 0);
        break;
    case L_EX:
        if (v.ex && cy == 0 && cz == 0) {
            if (cx == 0) s.set(+1, 0, 0);
            else s.set(-1, 0, 0);
        }
        break;
    case L_EYX:
        if (v.eyx && cx == 1 && cz == 0) {
            if (cy == 0) s.set(0, +1, 0);
            else s.set(0, -1, 0);
        }
        break;
    case L_EZX:
        if (v.ezx && cx == 1 && cy == 0) {
            if (cz == 0) s.set(0, 0, +1);
            else s.set(0, 0, -1);
        }
        break;
    }
}

// One row-pair step (fixed q, p) for the windows at b = x and b = x + 1 with
// the loop kinds in `kinds`, ordered as the reference's word loop orders
// them: loop-major when both bits share a word, bit-major across a word end.
__device__ __forceinline__ void row_step3(Cell3& s, uint32_t cube, int kinds, bool right,
                                          bool word_end, int ky, int kz, int cy, int cz)
{
    // window at b = x: columns x-1, x -> kx = 0, cell is column b (cx = 1)
    // window at b = x+1: columns x, x+1 -> kx = 1, cell is column b-1 (cx = 0)
    const Win3 w0 = window3(cube, 0, ky, kz);
    const Win3 w1 = window3(cube, 1, ky, kz);
    if (!word_end) {
#pragma unroll
        for (int k = 1; k <= L_EZX; k <<= 1) {
            if (!(kinds & k)) continue;
            apply3(s, w0, k, 1, cy, cz);
            if (right) apply3(s, w1, k, 0, cy, cz);
        }
    } else {
#pragma unroll
        for (int k = 1; k <= L_EZX; k <<= 1)
            if (kinds & k) apply3(s, w0, k, 1, cy, cz);
        if (right) {
#pragma unroll
            for (int k = 1; k <= L_EZX; k <<= 1)
                if (kinds & k) apply3(s, w1, k, 0, cy, cz);
        }
    }
}

// The cell's value given its pixel cube and the cube bits inside the volume:
// uniform neighbourhoods get no write; otherwise the reference's writes that
// land on the cell are replayed in order (DistanceOcclusionLib.cpp:1899-1973).
__device__ __forceinline__ Cell3 border3_from_cube(uint32_t cube, uint32_t inside, bool right,
                                                   bool word_end, int y, int z, int h, int d)
{
    Cell3 s{kMaxDelta3, kMaxDelta3, kMaxDelta3};
    const uint32_t in_set = cube & inside;
    if (in_set != 0 && in_set != inside) {
        const int all = L_BC | L_FXX | L_FY | L_FZ | L_EX | L_EYX | L_EZX;
        // slice pairs p = z-1 (cell in slice p+1, cz = 1, kz = 0) then p = z
        for (int pp = 0; pp < 2; pp++) {
            const int p = z - 1 + pp;
            if (p < 0 || p + 1 >= d) continue;
            const int kz = pp, cz = 1 - pp;
            for (int qq = 0; qq < 2; qq++) {
                const int q = y - 1 + qq;
                if (q < 0 || q + 1 >= h) continue;
                row_step3(s, cube, all, right, word_end, qq, kz, 1 - qq, cz);
            }
            if (y == h - 1)   // last row of the pair: y faces, x and z edges
                row_step3(s, cube, L_FY | L_EX | L_EZX, right, word_end, 1, kz, 0, cz);
        }
        if (z == d - 1) {     // last slice: z faces, x and y edges
            for (int qq = 0; qq < 2; qq++) {
                const int q = y - 1 + qq;
                if (q < 0 || q + 1 >= h) continue;
                row_step3(s, cube, L_FZ | L_EX | L_EYX, right, word_end, qq, 1, 1 - qq, 0);
            }
            if (y == h - 1) row_step3(s, cube, L_EX, right, word_end, 1, 1, 0, 0);
        }
    }
    return s;
}

// One thread per 32 cells (one mask word of a row).  A cell whose
// neighbourhood inside the volume is uniform gets no write, and such cells are
// found for the whole word with bit operations; their value is stored in bulk
// and only the others are replayed cell by cell.
__global__ void __launch_bounds__(256) k_border_3d_words(int w, int h, int d,
                                                         const uint32_t* __restrict__ mask,
                                                         int16_t* __restrict__ delta)
{
    const int ws = mask_stride(w);
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (size_t)ws * h * d) return;
    const int row = (int)(gid / ws), j = (int)(gid - (size_t)row * ws);
    const int z = row / h, y = row - z * h;
    const int x0 = 32 * j, n = min(32, w - x0);
    RowBits rb[9];
    bool rin[9];
    uint32_t diff = 0;
    const uint32_t centre = mask[(size_t)row * ws + j];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int zz = z - 1 + k / 3, yy = y - 1 + k % 3;
        rin[k] = zz >= 0 && zz < d && yy >= 0 && yy < h;
        rb[k] = RowBits{0u, 0u, 0u};
        if (rin[k]) {
            rb[k] = row_bits(mask + ((size_t)zz * h + yy) * ws, j, ws, w);
            diff |= (centre ^ rb[k].l) | (centre ^ rb[k].m) | (centre ^ rb[k].r);
        }
    }
    if (n < 32) diff &= (1u << n) - 1u;
    int16_t* out = delta + 3 * ((size_t)row * w + x0);
    // every cell first gets the no-write value (kMaxDelta3 in all three
    // components, 16-byte stores when aligned), then the others are replaced
    const uint32_t fill = (uint32_t)(uint16_t)kMaxDelta3 * 0x10001u;
    if (n == 32 && ((uintptr_t)out & 15) == 0) {
#pragma unroll
        for (int k = 0; k < 12; k++) ((uint4*)out)[k] = make_uint4(fill, fill, fill, fill);
    } else {
        for (int k = 0; k < 3 * n; k++) out[k] = (int16_t)kMaxDelta3;
    }
    while (diff) {
        const int i = __ffs(diff) - 1;
        diff &= diff - 1u;
        const int x = x0 + i;
        const bool right = x + 1 < w;
        uint32_t cube = 0, inside = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            if (!rin[k]) continue;
            const uint32_t b = ((rb[k].l >> i) & 1u) | (((rb[k].m >> i) & 1u) << 1) |
                               (right ? ((rb[k].r >> i) & 1u) << 2 : 0u);
            cube |= b << (3 * k);
            inside |= (right ? 7u : 3u) << (3 * k);
        }
        const Cell3 s = border3_from_cube(cube, inside, right, (x & 31) == 31, y, z, h, d);
        out[3 * i] = (int16_t)s.x; out[3 * i + 1] = (int16_t)s.y; out[3 * i + 2] = (int16_t)s.z;
    }
}

void launch_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, hipStream_t s)
{
    ProfileScope p("border_3d", s);
    const size_t words = (size_t)mask_stride(w) * h * d;
    hipLaunchKernelGGL(k_border_3d_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, w,
                       h, d, mask, delta);
}

}  // namespace dolhip
