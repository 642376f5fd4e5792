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

// Four cells a thread, for w % 4 == 0 and 16-byte aligned fields: every tap
// of the four cells comes from one 16-byte load (the cells x..x+3 of a tap
// row), so each lane has 144 bytes in flight.  SM is the step when it is 1
// or 2 (the shifted taps then come from the windows at x - 4, x, x + 4) and
// 4 for every step >= 4 (x +- step stays 4-aligned).  Taps, order and
// arithmetic as jfa2_cell.
template <int SM>
__global__ void __launch_bounds__(256) k_jfa2d_x4(int step, int w, int h, int xb,
                                                  const uint32_t* __restrict__ in,
                                                  uint32_t* __restrict__ out, int cw)
{
    // rows in residue classes mod step (y = r, r + step, r + 2 step, ...)
    // when step divides h: consecutive rows then share two of their three
    // source rows.  Workgroups are dealt to the 8 XCDs round-robin; when the
    // grid splits evenly, XCD c takes the c-th eighth of the rows in that
    // order, so the rows a row's taps read (and, for steps past a block's
    // width, the columns its x taps read) are fetched into the same L2.
    int b = blockIdx.x;
    if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
    const int q = b / xb;
    const int nq = h / step;
    const int y = nq * step == h ? (q % nq) * step + q / nq : q;
    const int x = ((b - q * xb) * 256 + threadIdx.x) * 4;
    if (x >= w) return;
    const int off = cw * step;
    const bool up = y - step >= 0, down = y + step < h;
    const bool rok[3] = {up, true, down};
    const int lsh = SM < 4 ? 4 : step;   // distance of the left and right windows
    uint4 L[3], C[3], R[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        L[r] = C[r] = R[r] = make_uint4(0, 0, 0, 0);
        if (!rok[r]) continue;
        const uint32_t* row = in + (size_t)(y + (r - 1) * step) * w;
        C[r] = *(const uint4*)(row + x);
        if (x - lsh >= 0) L[r] = *(const uint4*)(row + x - lsh);
        if (x + lsh < w) R[r] = *(const uint4*)(row + x + lsh);
    }
    auto el = [](const uint4& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; };
    uint32_t res[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        auto left = [&](int r) { return SM < 4 ? (i >= SM ? el(C[r], i - SM) : el(L[r], 4 + i - SM)) : el(L[r], i); };
        auto right = [&](int r) { return SM < 4 ? (i + SM < 4 ? el(C[r], i + SM) : el(R[r], i + SM - 4)) : el(R[r], i); };
        uint32_t best = el(C[1], i);
        int bl = len2_packed(best);
        if (x + i - step >= 0) {
            jfa2_tap(best, bl, left(1), -off, 0);
            if (up) jfa2_tap(best, bl, left(0), -off, -off);
            if (down) jfa2_tap(best, bl, left(2), -off, +off);
        }
        if (x + i + step < w) {
            jfa2_tap(best, bl, right(1), +off, 0);
            if (up) jfa2_tap(best, bl, right(0), +off, -off);
            if (down) jfa2_tap(best, bl, right(2), +off, +off);
        }
        if (up) jfa2_tap(best, bl, el(C[0], i), 0, -off);
        if (down) jfa2_tap(best, bl, el(C[2], i), 0, +off);
        res[i] = best;
    }
    *(uint4*)(out + (size_t)y * w + x) = make_uint4(res[0], res[1], res[2], res[3]);
}

void launch_jfa2d_pass(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s)
{
    ProfileScope p("jfa2d_pass", s);
    if (w % 4 == 0 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0) {
        const int xb = (w / 4 + 255) / 256;
        const dim3 grid((unsigned)(xb * h));
        const uint32_t* i32 = (const uint32_t*)in;
        uint32_t* o32 = (uint32_t*)out;
        if (step == 1)
            hipLaunchKernelGGL(k_jfa2d_x4<1>, grid, dim3(256), 0, s, step, w, h, xb, i32, o32, cw);
        else if (step == 2)
            hipLaunchKernelGGL(k_jfa2d_x4<2>, grid, dim3(256), 0, s, step, w, h, xb, i32, o32, cw);
        else
            hipLaunchKernelGGL(k_jfa2d_x4<4>, grid, dim3(256), 0, s, step, w, h, xb, i32, o32, cw);
        return;
    }
    int xb = (w + 255) / 256;
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

// Row-staged form, for widths w % 8 == 0 up to kRowMax: one workgroup per
// output row (y, z) copies the (up to) 9 input rows (y + jc*step,
// z + kc*step) whole into LDS with 16-byte loads, then each thread reads its
// taps from LDS.  Same taps, order, skips and arithmetic as k_jfa3d.
//
// Between passes the field is kept in the same 6 bytes a cell, split into
// two planes: packed (x, y) int16 pairs for all cells, then z for all cells.
// A tap then costs one 32-bit and one 16-bit LDS read, x and y are offset
// with one packed 16-bit add (wrapping as the reference's int16 store does).
// The first pass reads the caller's interleaved cCellDelta3 layout (and
// converts the staged rows in LDS) and the last writes it back.
constexpr int kRowMax = 1024;

typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));

// squared length of (x, y) packed as int16 pairs plus z*z, wrapping in 32
// bits as len2 does
__device__ __forceinline__ int len2_packed(uint32_t xy, int z)
{
    const short2v v = __builtin_bit_cast(short2v, xy);
    return __builtin_amdgcn_sdot2(v, v, z * z, false);
}

// W > 0: the row width as a constant (LDS offsets become immediates)
template <bool IN_AOS, bool OUT_AOS, int W>
__global__ void __launch_bounds__(kRowMax) k_jfa3d_rows(int step, int w_arg, int h, int d,
                                                        size_t cells,
                                                        const int16_t* __restrict__ in,
                                                        int16_t* __restrict__ out, int cw)
{
    const int w = W > 0 ? W : w_arg;
    extern __shared__ __align__(16) uint32_t lds[];
    // [9][w] packed (x, y), then [9][w] int16 z (IN_AOS: staged first as
    // [9][3w] int16 and converted in place)
    const int16_t* lds16 = (const int16_t*)lds;
    const int row = blockIdx.x;   // z*h + y
    const int z = row / h, y = row - z * h;
    auto in_row = [&](int r, int& src_row) {
        const int zz = z + (r / 3 - 1) * step, yy = y + (r % 3 - 1) * step;
        src_row = zz * h + yy;
        return zz >= 0 && zz < d && yy >= 0 && yy < h;
    };
    if constexpr (W == 512) {
        // one thread per cell of the row: the 9 x 192 16-byte chunks of the
        // staged rows are split so that a thread's chunks share their row's
        // y offset and column: threads t < 384 take column t % 192 of rows g,
        // g + 3, g + 6 (g = t / 192), threads t >= 384 column t - 384 of rows
        // 2, 5, 8, and threads t < 192 also column 128 + t % 64 of row
        // 2 + 3 (t / 64).  Rows outside the volume read row 0 instead and are
        // not stored; all loads are issued before the first LDS store.
        constexpr int kCh = 3 * W / 8;   // 192 chunks a row in either layout
        static_assert(IN_AOS || W / 4 + W / 8 == kCh, "chunks a row");
        const int t = threadIdx.x;
        const int g = t < 384 ? (t < 192 ? 0 : 1) : 2;
        const int k = t < 384 ? t - 192 * g : t - 384;
        const int ke = 128 + (t & 63), je = t >> 6;   // the extra chunk (t < 192)
        // source offset (16-byte units) and LDS word of chunk kk of staged
        // row r = rg + 3 j, whose source row is (z + (j - 1) step, yy)
        auto chunk = [&](int rg, int j, int kk, int yy, int& dst) -> size_t {
            const int zz = z + (j - 1) * step;
            const int r = rg + 3 * j;
            const bool ok = zz >= 0 && zz < d && yy >= 0 && yy < h;
            const size_t sr = ok ? (size_t)zz * h + yy : 0;
            size_t off;
            if (IN_AOS) {
                off = sr * kCh + kk;
                dst = r * (3 * W / 2) + 4 * kk;
            } else if (kk < W / 4) {
                off = sr * (W / 4) + kk;
                dst = r * W + 4 * kk;
            } else {
                off = cells / 4 + sr * (W / 8) + (kk - W / 4);   // the z plane starts at byte 4 * cells
                dst = 9 * W + r * (W / 2) + 4 * (kk - W / 4);
            }
            if (!ok) dst = -1;
            return off;
        };
        const int yg = y + (g - 1) * step;
        int d0, d1, d2, d3;
        const size_t o0 = chunk(g, 0, k, yg, d0), o1 = chunk(g, 1, k, yg, d1), o2 = chunk(g, 2, k, yg, d2);
        const size_t o3 = chunk(2, je < 3 ? je : 0, ke, y + step, d3);
        if (t >= 192) d3 = -1;
        const uint4* in4 = (const uint4*)in;
        const uint4 v0 = in4[o0], v1 = in4[o1], v2 = in4[o2];
        const uint4 v3 = in4[t < 192 ? o3 : 0];
        if (d0 >= 0) *(uint4*)(lds + d0) = v0;
        if (d1 >= 0) *(uint4*)(lds + d1) = v1;
        if (d2 >= 0) *(uint4*)(lds + d2) = v2;
        if (d3 >= 0) *(uint4*)(lds + d3) = v3;
    } else {
        if (IN_AOS) {
            const int chunks = (3 * w) / 8;   // 16-byte chunks in a row
            for (int e = threadIdx.x; e < 9 * chunks; e += blockDim.x) {
                const int r = e / chunks, k = e - r * chunks;
                int sr;
                if (!in_row(r, sr)) continue;
                ((uint4*)(lds + r * (3 * w / 2)))[k] = ((const uint4*)(in + (size_t)3 * w * sr))[k];
            }
        } else {
            const int cxy = w / 4, cz = w / 8;   // 16-byte chunks in an xy row and a z row
            const uint32_t* pxy = (const uint32_t*)in;
            const int16_t* pz = in + 2 * cells;
            for (int e = threadIdx.x; e < 9 * (cxy + cz); e += blockDim.x) {
                const int r = e / (cxy + cz), k = e - r * (cxy + cz);
                int sr;
                if (!in_row(r, sr)) continue;
                if (k < cxy)
                    ((uint4*)(lds + r * w))[k] = ((const uint4*)(pxy + (size_t)w * sr))[k];
                else
                    ((uint4*)(lds + 9 * w + r * (w / 2)))[k - cxy] = ((const uint4*)(pz + (size_t)w * sr))[k - cxy];
            }
        }
    }
    __syncthreads();
    const int x = threadIdx.x;
    if (IN_AOS) {
        // turn the staged rows into the planar form in place: each thread
        // takes its own column of the 9 rows into registers, then writes it
        // back (rows that were not staged carry garbage and are never read)
        // cell x's 6 bytes start at byte 6x, which is 4-aligned only for
        // even x: read the two aligned words that hold them and take the
        // halves (an unaligned LDS access would be split and serialised)
        uint32_t cxy[9];
        int16_t cz[9];
        if (x < w) {
            const int k0 = (3 * x) >> 1;   // word holding the cell's first byte
            const bool odd = x & 1;
#pragma unroll
            for (int r = 0; r < 9; r++) {
                const uint32_t* wrow = lds + r * (3 * w / 2);
                const uint32_t a = wrow[k0], b = wrow[k0 + 1];
                cxy[r] = odd ? __builtin_amdgcn_alignbit(b, a, 16) : a;
                cz[r] = (int16_t)(odd ? (b >> 16) : (b & 0xffffu));
            }
        }
        __syncthreads();
        if (x < w) {
            int16_t* z16 = (int16_t*)lds + 2 * 9 * w;
#pragma unroll
            for (int r = 0; r < 9; r++) {
                lds[r * w + x] = cxy[r];
                z16[r * w + x] = cz[r];
            }
        }
        __syncthreads();
    }
    if (x >= w) return;
    // tap ic of staged row r.  The x of a tap is clamped into the row so
    // every read is in range (an out-of-row tap is masked off instead of
    // branched around), and the per-ic bases are formed once so that row
    // offsets fold into the LDS instructions' immediates when W is known.
    const int xs[3] = {max(x - step, 0), x, min(x + step, w - 1)};
    const bool okx[3] = {x - step >= 0, true, x + step < w};
    // byte offsets of the three tap columns, held in registers (the empty
    // asm keeps the compiler from re-forming them in every row block)
    typedef __attribute__((address_space(3))) const uint32_t* lds_u32;
    typedef __attribute__((address_space(3))) const int16_t* lds_i16;
    lds_u32 pxy[3];
    lds_i16 pz[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        pxy[k] = (lds_u32)(lds + xs[k]);
        pz[k] = (lds_i16)(lds16 + 2 * 9 * w + xs[k]);
        asm volatile("" : "+v"(pxy[k]), "+v"(pz[k]));
    }
    auto tap = [&](int r, int ic, uint32_t& xy, int& tz) {
        xy = pxy[ic + 1][r * w];
        tz = pz[ic + 1][r * w];
    };
    const int off = cw * step;
    // candidate of tap t = 9 (kc+1) + 3 (jc+1) + (ic+1): the source offset by
    // (ic, jc, kc) * off, each component wrapped to int16
    auto candidate = [&](int t, uint32_t sxy, int sz, uint32_t& cxy, int& cz) {
        const int ic = t % 3 - 1, jc = (t / 3) % 3 - 1, kc = t / 9 - 1;
        ushort2v v = __builtin_bit_cast(ushort2v, sxy);
        v += ushort2v{(unsigned short)(ic * off), (unsigned short)(jc * off)};
        cxy = __builtin_bit_cast(uint32_t, v);
        cz = wrap16(sz + kc * off);
    };
    // the loop keeps only the best length and its tap; the winner's value is
    // formed again at the end
    uint32_t sxy;
    int sz;
    tap(4, 0, sxy, sz);
    int bl = len2_packed(sxy, sz);
    int bt = 13;
#pragma unroll
    for (int kc = -1; kc <= 1; kc++) {
        const int zz = z + kc * step;
        if (zz < 0 || zz >= d) continue;
#pragma unroll
        for (int jc = -1; jc <= 1; jc++) {
            const int yy = y + jc * step;
            if (yy < 0 || yy >= h) continue;
            const int r = (kc + 1) * 3 + (jc + 1);
#pragma unroll
            for (int ic = -1; ic <= 1; ic++) {
                const int t = 3 * r + ic + 1;
                if (t == 13) continue;   // the cell itself never beats itself (strict >)
                tap(r, ic, sxy, sz);
                uint32_t cxy;
                int cz;
                candidate(t, sxy, sz, cxy, cz);
                const int cl = len2_packed(cxy, cz);
                const bool take = okx[ic + 1] && (int16_t)sxy != kMaxDelta3 && bl > cl;
                bl = take ? cl : bl;
                bt = take ? t : bt;
            }
        }
    }
    uint32_t bxy;
    int bz;
    {
        const int r = bt / 3, ic = bt % 3 - 1;
        const int xx = xs[ic + 1];
        sxy = lds[r * w + xx];
        sz = lds16[2 * 9 * w + r * w + xx];
        candidate(bt, sxy, sz, bxy, bz);
    }
    const size_t i = (size_t)row * w + x;
    if (OUT_AOS) {
        int16_t* o = out + 3 * i;
        o[0] = (int16_t)bxy; o[1] = (int16_t)(bxy >> 16); o[2] = (int16_t)bz;
    } else {
        ((uint32_t*)out)[i] = bxy;
        out[2 * cells + i] = (int16_t)bz;
    }
}

void launch_jfa3d_pass(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s)
{
    ProfileScope p("jfa3d_pass", s);
    int xb = (w + 255) / 256;
    hipLaunchKernelGGL(k_jfa3d, dim3((unsigned)((size_t)xb * h * d)), dim3(256), 0, s, step, w, h,
                       d, xb, in, out, cw);
}

static void launch_jfa3d_rows(bool in_aos, bool out_aos, int step, int w, int h, int d,
                              const int16_t* in, int16_t* out, int cw, hipStream_t s)
{
    ProfileScope p("jfa3d_pass", s);
    const int threads = (w + 63) / 64 * 64;
    const size_t lds = (size_t)9 * w * 6;
    const size_t cells = (size_t)w * h * d;
    const dim3 grid((unsigned)((size_t)h * d));
#define DOL_JFA3D_ROWS(W)                                                                          \
    if (in_aos && out_aos)                                                                         \
        hipLaunchKernelGGL((k_jfa3d_rows<true, true, W>), grid, dim3(threads), lds, s, step, w, h, \
                           d, cells, in, out, cw);                                                 \
    else if (in_aos)                                                                               \
        hipLaunchKernelGGL((k_jfa3d_rows<true, false, W>), grid, dim3(threads), lds, s, step, w,   \
                           h, d, cells, in, out, cw);                                              \
    else if (out_aos)                                                                              \
        hipLaunchKernelGGL((k_jfa3d_rows<false, true, W>), grid, dim3(threads), lds, s, step, w,   \
                           h, d, cells, in, out, cw);                                              \
    else                                                                                           \
        hipLaunchKernelGGL((k_jfa3d_rows<false, false, W>), grid, dim3(threads), lds, s, step, w,  \
                           h, d, cells, in, out, cw);
    if (w == 512) {
        DOL_JFA3D_ROWS(512)
    } else {
        DOL_JFA3D_ROWS(0)
    }
#undef DOL_JFA3D_ROWS
}

// FloorPow2 of DistanceOcclusionLib.cpp:875-884
static int floor_pow2_host(int v)
{
    uint32_t u = (uint32_t)v;
    u |= u >> 1; u |= u >> 2; u |= u >> 4; u |= u >> 8; u |= u >> 16;
    return (int)((int32_t)(u + 1u) >> 1);
}

// The pass schedule of DistanceOcclusionLib.cpp:2450-2472: steps
// FloorPow2(max dim - 1) down to 1, ping-ponging between delta and tmp, plus
// one step-1 pass when the count is odd so the result ends in delta.
void jfa3d_run(int w, int h, int d, int16_t* delta, int cw, int16_t* tmp, hipStream_t s)
{
    int m = w >= h ? w : h;
    m = m >= d ? m : d;
    int steps[40], n = 0;
    for (int step = floor_pow2_host(m - 1); step > 0; step /= 2) steps[n++] = step;
    if (n & 1) steps[n++] = 1;
    const bool rows = w % 8 == 0 && w <= kRowMax && ((uintptr_t)delta & 15) == 0 &&
                      ((uintptr_t)tmp & 15) == 0;
    int16_t* a = delta;
    int16_t* b = tmp;
    for (int k = 0; k < n; k++) {
        if (rows)
            launch_jfa3d_rows(k == 0, k == n - 1, steps[k], w, h, d, a, b, cw, s);
        else
            launch_jfa3d_pass(steps[k], w, h, d, a, b, cw, s);
        int16_t* t = a; a = b; b = t;
    }
}

}  // namespace dolhip
