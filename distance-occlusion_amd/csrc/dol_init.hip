// dol_init.hip -- mask packing and delta/distance initialisation for gfx950.
//
// Border initialisation: the reference builds the signed border seeds by
// scattering writes from 2x2 (2x2x2) mask windows, row pair after row pair,
// word after word, and lets later writes win (SetBorderCellsCombo,
// DistanceOcclusionLib.cpp:504-599 and 1820-2150).  Here each output cell's
// value is found by replaying, in that same global order, only the writes
// that land on it.  The windows that can reach cell (x, y[, z]) are those at
// bit b in {x, x+1}, row pair q in {y-1, y} and slice pair p in {z-1, z}, so
// only the 3x3(x3) pixels around it matter; cells whose neighbourhood is
// uniform receive no write and stay at kMaxDelta.  The replay is evaluated
// for the 32 cells of a mask word at once, as bit masks, one thread a word.
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

// The same, 16 bytes of a row a thread (4 float or int32 cells, 8 int16
// cells, all from one mask word) with one nontemporal store: w a multiple
// of the cells a thread, dist 16-byte aligned (checked by the launcher).
template <class T>
__global__ void __launch_bounds__(256) k_init_dist_vec(int w, int xb, const uint32_t* __restrict__ mask,
                                                       T* __restrict__ dist, T max_d)
{
    constexpr int C = 16 / sizeof(T);
    static_assert(32 % C == 0, "a thread's cells lie in one mask word");
    const int row = blockIdx.x / xb;
    const int x = ((blockIdx.x - row * xb) * 256 + threadIdx.x) * C;
    if (x >= w) return;
    const uint32_t bits = mask[(size_t)row * mask_stride(w) + (x >> 5)] >> (x & 31);
    union {
        T v[C];
        uint32_t u[4];
    } out;
#pragma unroll
    for (int c = 0; c < C; c++) out.v[c] = ((bits >> c) & 1u) ? (T)0 : max_d;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 q = {out.u[0], out.u[1], out.u[2], out.u[3]};
    __builtin_nontemporal_store(q, (u32x4*)(dist + (size_t)row * w + x));
}

template <class T>
void launch_init_dist(int w, int h, const uint32_t* mask, T* dist, T max_d, hipStream_t s)
{
    constexpr int C = 16 / sizeof(T);
    ProfileScope p("init_distances", s);
    if (w % C == 0 && ((uintptr_t)dist & 15) == 0) {
        const int xb = (w / C + 255) / 256;
        hipLaunchKernelGGL(k_init_dist_vec<T>, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                           dist, max_d);
        return;
    }
    int xb = (w + 255) / 256;
    hipLaunchKernelGGL(k_init_dist<T>, dim3((unsigned)(xb * h)), dim3(256), 0, s, w, xb, mask,
                       dist, max_d);
}
template void launch_init_dist<float>(int, int, const uint32_t*, float*, float, hipStream_t);
template void launch_init_dist<int32_t>(int, int, const uint32_t*, int32_t*, int32_t, hipStream_t);
template void launch_init_dist<int16_t>(int, int, const uint32_t*, int16_t*, int16_t, hipStream_t);

// ------------------------------------------------------------ border init 2D

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

// Loop kinds of one window, in the reference's order
// (DistanceOcclusionLib.cpp:1899-1973)
enum : int { L_BC = 1, L_FXX = 2, L_FY = 4, L_FZ = 8, L_EX = 16, L_EYX = 32, L_EZX = 64 };

// The writes that land on cell (x, y, z) come from the 2x2x2 windows at bit
// b in {x, x+1}, row pair q in {y-1, y} and slice pair p in {z-1, z}
// (window features: DistanceOcclusionLib.cpp:1862-1894; loops and their
// order: 1899-1973; the last row's and last slice's extra loops: 1975-2150).
// Every such write has a value fixed by the loop kind and the cell's place
// in the window ((cx, cy, cz) below), so the 32 cells of a mask word are
// replayed together as bit masks: a write's fire mask is its window test
// over the word, and the cells' state is kept as bit planes (the signs of x,
// y and z, "written", and "last write was an edge", which is what the face
// writes' condition abs sum > 1 excludes).  Windows at b = x and b = x + 1
// alternate loop kind by loop kind; at a word's last bit the reference runs
// all of the first window's kinds before the second's, so the second
// window's writes to bit 31 are replayed after the alternating run.  A cell
// whose neighbourhood inside the volume is uniform gets no write and keeps
// kMaxDelta3 in all three components.
struct Planes3 {
    uint32_t xn, xp, yn, yp, zn, zp, set, edge;
};

struct Win3b {
    uint32_t ex, ey, ez, ezx, eyx, fx, fy, fz, fxx, bc;
};

template <int C>
__device__ __forceinline__ uint32_t rb_col(const RowBits& b)
{
    return C == 0 ? b.l : (C == 1 ? b.m : b.r);
}

// rb[dz * 3 + dy]: rows y - 1 + dy of slices z - 1 + dz (zero outside)
template <int KX, int KY, int KZ>
__device__ __forceinline__ Win3b window3b(const RowBits* rb)
{
    const RowBits &r00 = rb[KZ * 3 + KY], &r01 = rb[KZ * 3 + KY + 1];
    const RowBits &r10 = rb[(KZ + 1) * 3 + KY], &r11 = rb[(KZ + 1) * 3 + KY + 1];
    const uint32_t p000 = rb_col<KX>(r00), p001 = rb_col<KX + 1>(r00);
    const uint32_t p010 = rb_col<KX>(r01), p011 = rb_col<KX + 1>(r01);
    const uint32_t p100 = rb_col<KX>(r10), p101 = rb_col<KX + 1>(r10);
    const uint32_t p110 = rb_col<KX>(r11), p111 = rb_col<KX + 1>(r11);
    Win3b r;
    r.ex = p000 ^ p001;
    r.ey = p000 ^ p010;
    r.ez = p000 ^ p100;
    r.ezx = p001 ^ p101;
    const uint32_t eyz = p100 ^ p110, exy = p010 ^ p011;
    r.fx = r.ey ^ eyz;
    r.fy = r.ez ^ r.ezx;
    r.fz = r.ex ^ exy;
    r.eyx = p001 ^ p011;
    const uint32_t eyxz = p101 ^ p111;
    r.fxx = r.eyx ^ eyxz;
    r.bc = r.fz ^ eyz ^ eyxz;
    return r;
}

// bits where the window feature equals the cell's coordinate V (0 or 1)
template <int V>
__device__ __forceinline__ uint32_t eqm(uint32_t f)
{
    return V ? f : ~f;
}

// write (VX, VY, VZ) to the cells of m; EDGE: an edge write
template <int VX, int VY, int VZ, bool EDGE>
__device__ __forceinline__ void write3(Planes3& s, uint32_t m)
{
    s.xn = VX < 0 ? (s.xn | m) : (s.xn & ~m);
    s.xp = VX > 0 ? (s.xp | m) : (s.xp & ~m);
    s.yn = VY < 0 ? (s.yn | m) : (s.yn & ~m);
    s.yp = VY > 0 ? (s.yp | m) : (s.yp & ~m);
    s.zn = VZ < 0 ? (s.zn | m) : (s.zn & ~m);
    s.zp = VZ > 0 ? (s.zp | m) : (s.zp & ~m);
    s.set |= m;
    s.edge = EDGE ? (s.edge | m) : (s.edge & ~m);
}

// loop kind KIND of a window to the cells of `bits`, which sit at (CX, CY,
// CZ) in it: the kind's feature test, the cell's place, and the condition
// on what the cell holds (corner: unwritten; faces: not an edge)
template <int KIND, int CX, int CY, int CZ>
__device__ __forceinline__ void apply3b(Planes3& s, const Win3b& v, uint32_t bits)
{
    if (KIND == L_BC) {
        const uint32_t m = bits & v.bc & eqm<CX>(v.fx) & eqm<CY>(v.fy) & eqm<CZ>(v.fz) & ~s.set;
        write3<1 - 2 * CX, 1 - 2 * CY, 1 - 2 * CZ, false>(s, m);
    } else if (KIND == L_FXX) {
        if (CX != 1) return;
        const uint32_t m = bits & v.fxx & eqm<CY>(v.ezx) & eqm<CZ>(v.eyx) & ~s.edge;
        write3<0, 1 - 2 * CY, 1 - 2 * CZ, false>(s, m);
    } else if (KIND == L_FY) {
        if (CY != 0) return;
        const uint32_t m = bits & v.fy & eqm<CX>(v.ez) & eqm<CZ>(v.ex) & ~s.edge;
        write3<1 - 2 * CX, 0, 1 - 2 * CZ, false>(s, m);
    } else if (KIND == L_FZ) {
        if (CZ != 0) return;
        const uint32_t m = bits & v.fz & eqm<CX>(v.ey) & eqm<CY>(v.ex) & ~s.edge;
        write3<1 - 2 * CX, 1 - 2 * CY, 0, false>(s, m);
    } else if (KIND == L_EX) {
        if (CY != 0 || CZ != 0) return;
        write3<CX == 0 ? 1 : -1, 0, 0, true>(s, bits & v.ex);
    } else if (KIND == L_EYX) {
        if (CX != 1 || CZ != 0) return;
        write3<0, CY == 0 ? 1 : -1, 0, true>(s, bits & v.eyx);
    } else if (KIND == L_EZX) {
        if (CX != 1 || CY != 0) return;
        write3<0, 0, CZ == 0 ? 1 : -1, true>(s, bits & v.ezx);
    }
}

template <int KINDS, int KIND, int CX, int CY, int CZ>
__device__ __forceinline__ void apply3b_if(Planes3& s, const Win3b& v, uint32_t bits)
{
    if (KINDS & KIND) apply3b<KIND, CX, CY, CZ>(s, v, bits);
}

// one row-pair step (fixed q, p) with the loop kinds KINDS: the window at
// b = x (cell column b, cx = 1) and at b = x + 1 (cell column b - 1, cx = 0;
// cells with x + 1 < w only); window rows KY, KY + 1 and slices KZ, KZ + 1
// of the cell's 3x3x3 neighbourhood
template <int KINDS, int KY, int KZ, int CY, int CZ>
__device__ __forceinline__ void row_step3b(Planes3& s, const RowBits* rb, uint32_t right)
{
    const Win3b w0 = window3b<0, KY, KZ>(rb);
    const Win3b w1 = window3b<1, KY, KZ>(rb);
    const uint32_t last = 0x80000000u;
    const uint32_t inner = right & ~last, end = right & last;
#define DOL_STEP3B(K)                                   \
    apply3b_if<KINDS, K, 1, CY, CZ>(s, w0, ~0u);        \
    apply3b_if<KINDS, K, 0, CY, CZ>(s, w1, inner);
    DOL_STEP3B(L_BC)
    DOL_STEP3B(L_FXX)
    DOL_STEP3B(L_FY)
    DOL_STEP3B(L_FZ)
    DOL_STEP3B(L_EX)
    DOL_STEP3B(L_EYX)
    DOL_STEP3B(L_EZX)
#undef DOL_STEP3B
    if (end) {
        apply3b_if<KINDS, L_BC, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_FXX, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_FY, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_FZ, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_EX, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_EYX, 0, CY, CZ>(s, w1, end);
        apply3b_if<KINDS, L_EZX, 0, CY, CZ>(s, w1, end);
    }
}

// the row steps of slice pair z - 1 + PP: row pairs y - 1 and y, and the
// last row's extra step
template <int PP>
__device__ __forceinline__ void slice_pair3b(Planes3& s, const RowBits* rb, uint32_t right, int y, int h)
{
    constexpr int kAll = L_BC | L_FXX | L_FY | L_FZ | L_EX | L_EYX | L_EZX;
    if (y - 1 >= 0) row_step3b<kAll, 0, PP, 1, 1 - PP>(s, rb, right);
    if (y + 1 < h) row_step3b<kAll, 1, PP, 0, 1 - PP>(s, rb, right);
    if (y == h - 1) row_step3b<L_FY | L_EX | L_EZX, 1, PP, 0, 1 - PP>(s, rb, right);
}

// every row step that can reach the word's cells, in the reference's order
// (slice pairs z - 1 then z, then the last slice's extra steps)
__device__ __forceinline__ Planes3 border3_word(const RowBits* rb, uint32_t right, int y, int z, int h, int d)
{
    Planes3 s{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (z - 1 >= 0) slice_pair3b<0>(s, rb, right, y, h);
    if (z + 1 < d) slice_pair3b<1>(s, rb, right, y, h);
    if (z == d - 1) {
        if (y - 1 >= 0) row_step3b<L_FZ | L_EX | L_EYX, 0, 1, 1, 0>(s, rb, right);
        if (y + 1 < h) row_step3b<L_FZ | L_EX | L_EYX, 1, 1, 0, 0>(s, rb, right);
        if (y == h - 1) row_step3b<L_EX, 1, 1, 0, 0>(s, rb, right);
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
    if (!diff) return;
    const uint32_t right = j == ws - 1 ? ~(1u << ((w - 1) & 31)) : ~0u;
    const Planes3 p = border3_word(rb, right, y, z, h, d);
    uint32_t set = p.set & diff;
    while (set) {
        const int i = __ffs(set) - 1;
        set &= set - 1u;
        out[3 * i] = (int16_t)((int)((p.xp >> i) & 1u) - (int)((p.xn >> i) & 1u));
        out[3 * i + 1] = (int16_t)((int)((p.yp >> i) & 1u) - (int)((p.yn >> i) & 1u));
        out[3 * i + 2] = (int16_t)((int)((p.zp >> i) & 1u) - (int)((p.zn >> i) & 1u));
    }
}

// Same result for w % 32 == 0, where mask word g's cells are the flat bytes
// 192g .. 192g + 191 of the delta image: a workgroup's 256 words are one
// contiguous 48 KB run, built in LDS (fill, then the written cells) and
// stored as consecutive 16-byte chunks.
__global__ void __launch_bounds__(256) k_border_3d_flat(int w, int h, int d,
                                                        const uint32_t* __restrict__ mask,
                                                        uint4* __restrict__ delta)
{
    constexpr int kChunks = 12;   // 16-byte chunks of a word's 32 cells
    __shared__ uint4 stage[256 * kChunks];
    const int ws = w >> 5;
    const size_t words = (size_t)ws * h * d;
    const size_t word0 = (size_t)blockIdx.x * 256;
    const size_t gid = word0 + threadIdx.x;
    const uint32_t fill = (uint32_t)(uint16_t)kMaxDelta3 * 0x10001u;
    uint4* mine = stage + threadIdx.x * kChunks;
#pragma unroll
    for (int k = 0; k < kChunks; k++) mine[k] = make_uint4(fill, fill, fill, fill);
    if (gid < words) {
        const int row = (int)(gid / ws), j = (int)(gid - (size_t)row * ws);
        const int z = row / h, y = row - z * h;
        RowBits rb[9];
        uint32_t diff = 0;
        const uint32_t centre = mask[gid];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const int zz = z - 1 + k / 3, yy = y - 1 + k % 3;
            rb[k] = RowBits{0u, 0u, 0u};
            if (zz >= 0 && zz < d && yy >= 0 && yy < h) {
                rb[k] = row_bits(mask + ((size_t)zz * h + yy) * ws, j, ws, w);
                diff |= (centre ^ rb[k].l) | (centre ^ rb[k].m) | (centre ^ rb[k].r);
            }
        }
        if (diff) {
            const uint32_t right = j == ws - 1 ? 0x7fffffffu : ~0u;
            const Planes3 p = border3_word(rb, right, y, z, h, d);
            uint32_t set = p.set & diff;
            int16_t* out = (int16_t*)mine;
            while (set) {
                const int i = __ffs(set) - 1;
                set &= set - 1u;
                out[3 * i] = (int16_t)((int)((p.xp >> i) & 1u) - (int)((p.xn >> i) & 1u));
                out[3 * i + 1] = (int16_t)((int)((p.yp >> i) & 1u) - (int)((p.yn >> i) & 1u));
                out[3 * i + 2] = (int16_t)((int)((p.zp >> i) & 1u) - (int)((p.zn >> i) & 1u));
            }
        }
    }
    __syncthreads();
    const size_t n = min((size_t)256, words - word0) * kChunks;
    uint4* dst = delta + word0 * kChunks;
#pragma unroll
    for (int k = 0; k < kChunks; k++) {
        const size_t c = (size_t)k * 256 + threadIdx.x;
        if (c < n) dst[c] = stage[c];
    }
}

void launch_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, hipStream_t s)
{
    ProfileScope p("border_3d", s);
    if ((w & 31) == 0 && ((uintptr_t)delta & 15) == 0) {
        const size_t words = (size_t)(w >> 5) * h * d;
        hipLaunchKernelGGL(k_border_3d_flat, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                           w, h, d, mask, (uint4*)delta);
        return;
    }
    const size_t words = (size_t)mask_stride(w) * h * d;
    hipLaunchKernelGGL(k_border_3d_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, w,
                       h, d, mask, delta);
}

}  // namespace dolhip
