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

#include <type_traits>

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

// 3D passes for widths w % 8 == 0 up to kRowMax (k_jfa3d_ring below): a
// workgroup takes a run of output rows, each lane two columns of a row.
// Same taps, order, skips and results as k_jfa3d.
//
// Between passes the field is kept in the same 6 bytes a cell, split into
// two planes: packed (x, y) int16 pairs for all cells, then z for all cells.
// The first pass reads the caller's interleaved cCellDelta3 layout and the
// last writes it back.
//
// Each pass also writes one flag byte per output row and 64-cell slot:
// kRowUnbounded if a cell that is not kMaxDelta3 has a component beyond
// kFastBound, kRowHasMax if a cell is kMaxDelta3 (k_jfa3d_classify writes
// them for the caller's field before the first pass).  A run whose source
// rows' flags show no unbounded cell, with 54 * cw * step within int16,
// takes the keyed path; otherwise the general path, which follows the
// reference's arithmetic including its int16 and int32 wraps.
//
// Keyed path.  With |components| <= kFastBound and cw * step <= 606 no
// candidate wraps and every candidate length is below 2^24, so a tap's
// outcome is the order of its key 27 * len + rank (the scale 27 stands for
// the key's scale KS below), where rank is the tap's
// place in the reference's order with the cell itself first (rank 0; the
// strict ">" keeps the earliest of equal lengths).  The least key wins.  A
// staged source s on z-row kc holds
//     Bb = 27 * (sx^2 + sy^2 + (sz + kc*off)^2)
// and tap (ic, jc) adds K = 27 * ((ic*off)^2 + (jc*off)^2) + rank and the
// cross term 54 * off * (ic*sx + jc*sy), with one v_dot2_i32_i16 of (sx, sy)
// and the packed constant (54 ic off, 54 jc off):
//     key = 27 * ((sx + ic*off)^2 + (sy + jc*off)^2 + (sz + kc*off)^2) + rank.
// A kMaxDelta3 source (which the reference skips) is staged with Bb =
// poison_b and (sx, sy) = 0, so its key is above every other; so are rows
// outside the volume and the poison column at index w that taps beyond
// either end of the row read.  A kMaxDelta3 cell itself has the key
// self_max_key (above every real candidate, below every skipped one), or 0
// where its own 32-bit length wraps negative (then nothing replaces it, as
// in the reference).  The winner is decoded from the key's rank and rebuilt
// from its source.
constexpr int kRowMax = 1024;
constexpr int kFlagSlots = 16;        // flag bytes a row: one per wave of up to kRowMax cells
constexpr int kFastBound = 2048;
constexpr uint32_t kRowUnbounded = 1, kRowHasMax = 2;
// The key's scale KS is 27 (ranks 0..26 fill the residues) or 32 (the rank
// is the key's low 5 bits, one AND to decode) where the cross-term constant
// 2 * 32 * off still fits 16 bits.
template <int KS> constexpr uint32_t poison_b() { return (uint32_t)KS * 100000000u; }   // above every real key
template <int KS> constexpr uint32_t self_max_key() { return (uint32_t)KS * 67108864u; }  // KS * 2^26
template <int KS> constexpr int fast_max_off() { return 32767 / (2 * KS); }              // 606 for 27, 511 for 32
template <int KS> __device__ __forceinline__ uint32_t key_rank(uint32_t key)
{
    return KS == 32 ? key & 31u : key % (uint32_t)KS;
}
static_assert(poison_b<32>() + 32u * 2u * 511u * 511u + 26u < 0xffffffffu, "poisoned keys stay in 32 bits");

typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));   // 8-byte aligned: one ds_read_b64

// squared length of (x, y) packed as int16 pairs plus z*z, wrapping in 32
// bits as len2 does
__device__ __forceinline__ int len2_packed(uint32_t xy, int z)
{
    const short2v v = __builtin_bit_cast(short2v, xy);
    return __builtin_amdgcn_sdot2(v, v, z * z, false);
}

// rank of tap t = 9 (kc+1) + 3 (jc+1) + (ic+1) in the reference's order with
// the cell itself (t = 13) first
__host__ __device__ constexpr int tap_rank(int t) { return t == 13 ? 0 : (t < 13 ? t + 1 : t); }
// the rank a staged row's B holds: its ic = 0 tap's (on the cell's own row,
// whose ic = 0 tap is the cell, the ic = -1 tap's)
__host__ __device__ constexpr int row_rank(int r) { return r == 4 ? tap_rank(12) : tap_rank(3 * r + 1); }

// the flag bytes of source row srow (its first nw slots), folded into one
// 64-bit word (only bits 0 and 1 of each byte carry anything)
__device__ __forceinline__ uint64_t row_flags(const uint8_t* __restrict__ flags, int srow, int nw)
{
    const uint64_t* fr = (const uint64_t*)(flags + (size_t)srow * kFlagSlots);
    uint64_t a = fr[0], b = nw > 8 ? fr[1] : 0;
    if (nw < 8) a &= ~0ull >> (64 - 8 * nw);
    if (nw > 8 && nw < 16) b &= ~0ull >> (64 - 8 * (nw - 8));
    return a | b;
}

__device__ __forceinline__ uint32_t fold_flags(uint64_t f)
{
    f |= f >> 32;
    f |= f >> 16;
    f |= f >> 8;
    return (uint32_t)f & 0xffu;
}

// One output row (y, z) by the general path, which follows the reference's
// arithmetic including its int16 and int32 wraps: lane t takes columns
// x = t + c T (c < COLS, T lanes; with PAIR x = COLS t + c) from its cells
// cxy[c][r], cz[c][r] of
// staged row r = 3 (kc+1) + (jc+1), source row (z + kc*step, y + jc*step),
// present where rv[r] (uniform).  Stores the result (interleaved if OUT_AOS)
// and the row's flag bytes.  Uses LDS from `lds` on; the caller
// synchronises before using it again.
template <bool OUT_AOS, int W, int COLS, bool PAIR>
__device__ __forceinline__ void jfa3d_row_general(uint32_t* lds, int w, int t, int T, size_t row,
                                                  int step, int cw, const bool (&rv)[9],
                                                  const uint32_t (&cxy)[COLS][9],
                                                  const int (&cz)[COLS][9], size_t cells,
                                                  int16_t* __restrict__ out,
                                                  uint8_t* __restrict__ flags_out)
{
    constexpr bool kAllLive = W > 0 && W % (64 * COLS) == 0;
    const int off = cw * step;
    int xc[COLS];
    bool live[COLS];
#pragma unroll
    for (int c = 0; c < COLS; c++) {
        xc[c] = PAIR ? COLS * t + c : t + c * T;
        live[c] = kAllLive || xc[c] < w;
    }
    uint32_t bxy[COLS];   // the results
    int bz[COLS];
#pragma unroll
    for (int c = 0; c < COLS; c++) {
        bxy[c] = 0;
        bz[c] = 0;
    }
    {
        // [9][w] packed (x, y), then [9][w] int16 z
        const int16_t* lds16 = (const int16_t*)lds;
        int16_t* z16 = (int16_t*)lds + 2 * 9 * w;
#pragma unroll
        for (int c = 0; c < COLS; c++) {
            if (!live[c]) continue;
#pragma unroll
            for (int r = 0; r < 9; r++) {
                lds[r * w + xc[c]] = cxy[c][r];
                z16[r * w + xc[c]] = (int16_t)cz[c][r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < COLS; c++) {
            if (!live[c]) continue;
            const int x = xc[c];
            // tap ic of staged row r.  The x of a tap is clamped into the row
            // so every read is in range (an out-of-row tap is masked off
            // instead of branched around).
            const int xs[3] = {max(x - step, 0), x, min(x + step, w - 1)};
            const bool okx[3] = {x - step >= 0, true, x + step < w};
            // candidate of tap tt = 9 (kc+1) + 3 (jc+1) + (ic+1): the source
            // offset by (ic, jc, kc) * off, each component wrapped to int16
            auto candidate = [&](int tt, uint32_t sxy, int sz, uint32_t& oxy, int& oz) {
                const int ic = tt % 3 - 1, jc = (tt / 3) % 3 - 1, kc = tt / 9 - 1;
                ushort2v v = __builtin_bit_cast(ushort2v, sxy);
                v += ushort2v{(unsigned short)(ic * off), (unsigned short)(jc * off)};
                oxy = __builtin_bit_cast(uint32_t, v);
                oz = wrap16(sz + kc * off);
            };
            // the loop keeps only the best length and its tap; the winner's
            // value is formed again at the end
            int bl = len2_packed(cxy[c][4], cz[c][4]);
            int bt = 13;
#pragma unroll
            for (int r = 0; r < 9; r++) {
                if (!rv[r]) continue;
#pragma unroll
                for (int ic = -1; ic <= 1; ic++) {
                    const int tt = 3 * r + ic + 1;
                    if (tt == 13) continue;   // the cell itself never beats itself (strict >)
                    const uint32_t sxy = lds[r * w + xs[ic + 1]];
                    const int sz = lds16[2 * 9 * w + r * w + xs[ic + 1]];
                    uint32_t oxy;
                    int oz;
                    candidate(tt, sxy, sz, oxy, oz);
                    const int cl = len2_packed(oxy, oz);
                    const bool take = okx[ic + 1] && (int16_t)sxy != kMaxDelta3 && bl > cl;
                    bl = take ? cl : bl;
                    bt = take ? tt : bt;
                }
            }
            if (bt == 13) {
                bxy[c] = cxy[c][4];
                bz[c] = cz[c][4];
            } else {
                const int r = bt / 3, ic = bt % 3 - 1;
                const int xx = xs[ic + 1];
                candidate(bt, lds[r * w + xx], lds16[2 * 9 * w + r * w + xx], bxy[c], bz[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < COLS; c++) {
        if (live[c]) {
            const size_t i = row * w + xc[c];
            if (OUT_AOS) {
                int16_t* o = out + 3 * i;
                o[0] = (int16_t)bxy[c]; o[1] = (int16_t)(bxy[c] >> 16); o[2] = (int16_t)bz[c];
            } else {
                ((uint32_t*)out)[i] = bxy[c];
                out[2 * cells + i] = (int16_t)bz[c];
            }
        }
        if (flags_out) {
            const int ox = (int16_t)bxy[c], oy = (int16_t)(bxy[c] >> 16), oz = (int16_t)bz[c];
            const bool m = ox == kMaxDelta3;
            const bool unb = !m && (abs(ox) > kFastBound || abs(oy) > kFastBound || abs(oz) > kFastBound);
            const uint64_t bu = __ballot(live[c] && unb), bm = __ballot(live[c] && m);
            if ((t & 63) == 0 && xc[c] < w)
                flags_out[row * kFlagSlots + (xc[c] >> 6)] =
                    (uint8_t)((bu ? kRowUnbounded : 0u) | (bm ? kRowHasMax : 0u));
        }
    }
}

constexpr int kRun = 8;   // output rows (at most) a workgroup of k_jfa3d_ring takes

// columns a lane takes: two (half the lanes, so the scalar work of a row is
// shared by twice the cells)
template <int W>
constexpr int jfa3d_cols() { return 2; }
// lanes of the row kernels' workgroups: 256 for 512-wide rows, at most 512
template <int W>
constexpr int jfa3d_lanes() { return W == 512 ? 256 : kRowMax / 2; }

// The 3D passes on runs of rows with the source rows in an LDS ring.
//
// Workgroup b takes a run of up to kRun output rows (y0 + i*step, z),
// i = 0, 1, ..., of one residue y0 mod step.  Row i reads
// source y-rows y0 + (i-1) step, y0 + i step and y0 + (i+1) step (each at
// z - step, z, z + step), so consecutive rows share two of their three
// y-rows.  The ring holds 3 y-rows x 3 z-rows of staged cells: row i + 1's
// new y-row replaces row i - 2's in slot (i + 1) mod 3.  A staged cell
// holds its raw (x, y) and
//     Bb = 27 * (sx^2 + sy^2 + (sz + kc*off)^2)
// (27 standing for the key's scale KS) which does not depend on the y
// offset the cell is read with, so it is formed once per run; a tap adds
//     K = 27 * ((ic*off)^2 + (jc*off)^2) + rank
// and the cross term (54 ic off, 54 jc off) . (sx, sy) with one v_dot2.  A
// kMaxDelta3 cell is staged with (x, y) = 0 and Bb = poison_b.  The whole
// run takes this path when the flags of all its source rows allow it;
// otherwise the workgroup runs its rows through jfa3d_row_general.  With
// IN_AOS the cells are read from the caller's interleaved field (the flags
// then come from k_jfa3d_classify, or are all kRowHasMax after the library's
// own init).  The rows are unrolled, so every LDS offset of the ring's three
// phases is an immediate.
template <bool IN_AOS, bool OUT_AOS, int W, int KS>
__global__ void __launch_bounds__(jfa3d_lanes<W>()) __attribute__((amdgpu_waves_per_eu(W == 512 ? 4 : 2)))
k_jfa3d_ring(int step, int w_arg, int h, int d, size_t cells, const int16_t* __restrict__ in,
             int16_t* __restrict__ out, int cw, const uint8_t* __restrict__ flags_in,
             uint8_t* __restrict__ flags_out, int residues)
{
    constexpr int COLS = jfa3d_cols<W>();
    constexpr bool kAllLive = W > 0 && W % (64 * COLS) == 0;
    const int w = W > 0 ? W : w_arg;
    extern __shared__ __align__(16) uint32_t lds[];
    int q = blockIdx.x;
    if ((gridDim.x & 7) == 0) q = (q & 7) * (gridDim.x >> 3) + (q >> 3);
    const int z = q % d;
    const int qq = q / d;
    const int r0 = qq % residues, ch = qq / residues;
    const int y0 = r0 + ch * kRun * step;
    if (y0 >= h) return;
    const int n = min(kRun, (h - 1 - y0) / step + 1);
    const int t = threadIdx.x, T = blockDim.x;
    const int off = cw * step;
    // The first pass of 512-wide rows reads the caller's interleaved cells
    // two adjacent ones a lane, 12 bytes in one load, so a lane's columns are
    // 2t and 2t + 1 there; otherwise t and t + T.
    constexpr bool kPairAos = IN_AOS && W == 512 && COLS == 2;
    int xc[COLS];
    bool live[COLS];
    unsigned xl[COLS];
#pragma unroll
    for (int c = 0; c < COLS; c++) {
        xc[c] = kPairAos ? COLS * t + c : t + c * T;
        live[c] = kAllLive || xc[c] < w;
        xl[c] = live[c] ? xc[c] : w - 1;
    }
    const bool zv[3] = {z - step >= 0, true, z + step < d};
    const int nw = (w + 63) / 64;
    // this lane's cells of source row sr (rows outside the volume are never
    // read): the row's address in a buffer descriptor, the lane's offset in
    // one register
    // vz gets the raw 16 bits (zero-extended): the sign extension waits for
    // the first use, so that nothing needs a prefetched value early.
    // With ok false the loads are still issued, through a descriptor of no
    // records (they read nothing and return 0), so that every path issues
    // the same loads and waits on them can count exactly.
    auto load = [&](int sr, uint32_t (&vxy)[COLS], int (&vz)[COLS], bool ok = true) {
        sr = __builtin_amdgcn_readfirstlane(ok ? sr : 0);
        const int nb = ok ? 1 : 0;
        if (kPairAos) {
            // raw words d0 d1 d2 = (x0 y0) (z0 x1) (y1 z1) in vxy[0], vz[0],
            // vxy[1]; unpack() sorts them out where they are first used
            const auto rr = __builtin_amdgcn_make_buffer_rsrc((void*)(in + (size_t)sr * w * 3), 0, nb * w * 6, kRsrcWord3);
            const auto v = __builtin_amdgcn_raw_buffer_load_b96(rr, t * 12, 0, 0);
            vxy[0] = v[0];
            vz[0] = (int)v[1];
            vxy[1] = v[2];
            vz[1] = 0;
        } else if (IN_AOS) {
            const auto rr = __builtin_amdgcn_make_buffer_rsrc((void*)(in + (size_t)sr * w * 3), 0, nb * w * 6, kRsrcWord3);
#pragma unroll
            for (int c = 0; c < COLS; c++) {
                // x and y loaded into the two halves of one register (no
                // instruction combines them, so nothing waits on the loads)
                ushort2v v;
                v.x = __builtin_amdgcn_raw_buffer_load_b16(rr, xl[c] * 6, 0, 0);
                v.y = __builtin_amdgcn_raw_buffer_load_b16(rr, xl[c] * 6 + 2, 0, 0);
                vxy[c] = __builtin_bit_cast(uint32_t, v);
                vz[c] = (int)__builtin_amdgcn_raw_buffer_load_b16(rr, xl[c] * 6 + 4, 0, 0);
            }
        } else {
            const size_t base = (size_t)sr * w;
            const auto rxy = __builtin_amdgcn_make_buffer_rsrc((void*)((const uint32_t*)in + base), 0, nb * w * 4, kRsrcWord3);
            const auto rz = __builtin_amdgcn_make_buffer_rsrc((void*)(in + 2 * cells + base), 0, nb * w * 2, kRsrcWord3);
#pragma unroll
            for (int c = 0; c < COLS; c++) {
                vxy[c] = __builtin_amdgcn_raw_buffer_load_b32(rxy, xl[c] * 4, 0, 0);
                vz[c] = (int)__builtin_amdgcn_raw_buffer_load_b16(rz, xl[c] * 2, 0, 0);
            }
        }
    };
    // the loaded cells as (x, y) pairs and 16-bit z, from what load() left
    auto unpack = [&](uint32_t (&vxy)[COLS], int (&vz)[COLS]) __attribute__((always_inline)) {
        if constexpr (kPairAos) {
            const uint32_t d1 = (uint32_t)vz[0], d2 = vxy[1];
            vz[0] = (int)(d1 & 0xffffu);
            vxy[1] = (d1 >> 16) | (d2 << 16);
            vz[1] = (int)(d2 >> 16);
        }
    };
    auto y_ok = [&](int g) { const int yy = y0 + g * step; return yy >= 0 && yy < h; };

    // the path of the whole run, from the flags of every source row
    uint64_t f = 0;
    for (int g = -1; g <= n; g++) {
        if (!y_ok(g)) continue;
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (zv[k]) f |= row_flags(flags_in, (z + (k - 1) * step) * h + y0 + g * step, nw);
    }
    const uint32_t fl = fold_flags(f);
    if ((fl & kRowUnbounded) || off > fast_max_off<KS>()) {
        // general path, one row at a time
        for (int i = 0; i < n; i++) {
            const int y = y0 + i * step;
            bool rv[9];
            uint32_t cxy[COLS][9];
            int cz[COLS][9];
#pragma unroll
            for (int r = 0; r < 9; r++) {
                const int k = r / 3, jc = r % 3 - 1;
                rv[r] = zv[k] && y_ok(i + jc);
                uint32_t vxy[COLS] = {};
                int vz[COLS] = {};
                if (rv[r]) load((z + (k - 1) * step) * h + y + jc * step, vxy, vz);
                unpack(vxy, vz);
#pragma unroll
                for (int c = 0; c < COLS; c++) {
                    cxy[c][r] = vxy[c];
                    cz[c][r] = (int16_t)vz[c];
                }
            }
            jfa3d_row_general<OUT_AOS, W, COLS, kPairAos>(lds, w, t, T, (size_t)z * h + y, step, cw, rv, cxy, cz,
                                                cells, out, flags_out);
            __syncthreads();
        }
        return;
    }
    const bool any_max = (fl & kRowHasMax) != 0;

    // LDS: ring [3 z][3 slots][w + 1] (xy, Bb) pairs, then the decoding
    // tables of the three phases.  The winner's z is read from the source
    // field itself (a cell of the 9 rows just loaded for staging), which keeps
    // the ring at 8 bytes a cell so that 4 workgroups share a CU.
    const int stride = w + 1;
    u32x2* pb = (u32x2*)lds;
    uint4* tbl = (uint4*)(lds + ((9 * stride * 8 + 15) / 16) * 4);
    for (int e = t; e < 81; e += T) {
        // phase p = i mod 3, rank: the winner's cell index relative to
        // (slot 0, column x), the offsets added to its (x, y) and z, and its
        // cell in the source field relative to the output cell
        const int p = e / 27, rk = e % 27;
        const int tt = rk == 0 ? 13 : (rk <= 13 ? rk - 1 : rk);
        const int r = tt / 3, ic = tt % 3 - 1, jc = r % 3 - 1, kc = r / 3 - 1;
        const int slot = (kc + 1) * 3 + (p + jc + 3) % 3;
        const uint32_t dxy = (uint32_t)(uint16_t)(ic * off) | (uint32_t)(uint16_t)(jc * off) << 16;
        const int src = (kc * step * h + jc * step) * w + ic * step;
        tbl[e] = make_uint4((uint32_t)(slot * stride + ic * step), dxy, (uint32_t)(kc * off), (uint32_t)src);
    }
    if (t < 9) pb[t * stride + w] = u32x2{0u, poison_b<KS>()};   // the poison column
    // stage this lane's cells of y-row g of every z-row into ring slot g mod 3;
    // a row outside the volume is staged as poison, so the taps need no test
    auto stage = [&](int g, uint32_t (&vxy)[3][COLS], int (&vz)[3][COLS]) {
        const int slot = (g + 3) % 3;
        // the loads' values are first used here, not where the scheduler
        // would like to convert them (right after the loads)
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int c = 0; c < COLS; c++) asm volatile("" : "+v"(vxy[k][c]), "+v"(vz[k][c]));
#pragma unroll
        for (int k = 0; k < 3; k++) unpack(vxy[k], vz[k]);
        const bool gy = y_ok(g);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (!zv[k] || !gy) {
#pragma unroll
                for (int c = 0; c < COLS; c++)
                    if (live[c]) pb[(k * 3 + slot) * stride + xc[c]] = u32x2{0u, poison_b<KS>()};
                continue;
            }
            const int kc = k - 1;
#pragma unroll
            for (int c = 0; c < COLS; c++) {
                const int czr = (int16_t)vz[k][c] + kc * off;
                const int len = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, vxy[k][c]),
                                                       __builtin_bit_cast(short2v, vxy[k][c]), __mul24(czr, czr), false);
                uint32_t B = __umul24((uint32_t)len, (uint32_t)KS);
                uint32_t X = vxy[k][c];
                if (any_max) {
                    const bool m = (int16_t)X == kMaxDelta3;
                    X = m ? 0u : X;
                    B = m ? poison_b<KS>() : B;
                }
                if (live[c]) pb[(k * 3 + slot) * stride + xc[c]] = u32x2{X, B};
            }
        }
    };
    // raw cells of the middle z-row of the current and the next y-row (the
    // output cell itself and its successor)
    uint32_t self_xy[2][COLS];
    int self_z[2][COLS];
    {
        uint32_t vxy[3][COLS] = {};
        int vz[3][COLS] = {};
#pragma unroll
        for (int g = -1; g <= 1; g++) {
            if (y_ok(g)) {
#pragma unroll
                for (int k = 0; k < 3; k++)
                    if (zv[k]) load((z + (k - 1) * step) * h + y0 + g * step, vxy[k], vz[k]);
            }
            stage(g, vxy, vz);
            if (g >= 0) {
#pragma unroll
                for (int c = 0; c < COLS; c++) {
                    self_xy[g][c] = vxy[1][c];
                    self_z[g][c] = (int16_t)vz[1][c];
                }
            }
        }
    }
    // the uniform per-tap constants
    const uint32_t c54 = (uint32_t)__builtin_amdgcn_readfirstlane(2 * KS * off) & 0xffffu;
    const uint32_t c54n = (0x10000u - c54) & 0xffffu;
    const uint32_t k1 = (uint32_t)__builtin_amdgcn_readfirstlane(KS * off * off);
    // the next y-row's cells, in flight during a row
    uint32_t nxy[3][COLS] = {};
    int nz[3][COLS] = {};
    auto prefetch = [&](int g) {
        const bool gy = g < n + 1 && y_ok(g);
#pragma unroll
        for (int k = 0; k < 3; k++)
            load((z + (k - 1) * step) * h + y0 + g * step, nxy[k], nz[k], gy && zv[k]);
    };
    prefetch(2);
    __syncthreads();
    // A row's results wait in registers until the next row has put its
    // prefetch in flight: the winner's z is loaded from
    // the field, and a wait for that load right after it would also wait for
    // the prefetch issued before it (loads complete in order), stalling every
    // row for a full memory latency.  pxy: the winners' (x, y); pz: the loaded z of each
    // winner's source cell; pdz: the offset added to it.
    uint32_t pxy[COLS];
    int pz[COLS], pdz[COLS];
#pragma unroll
    for (int c = 0; c < COLS; c++) {
        pxy[c] = 0;
        pz[c] = 0;
        pdz[c] = 0;
    }
    auto flush = [&](int i) __attribute__((always_inline)) {
        const size_t row = (size_t)z * h + y0 + i * step;
        // planar output through descriptors of this row (32-bit lane
        // offsets; a column past the row falls outside and stores nothing)
        const auto rxy = __builtin_amdgcn_make_buffer_rsrc((void*)((uint32_t*)out + row * w), 0, w * 4, kRsrcWord3);
        const auto rz = __builtin_amdgcn_make_buffer_rsrc((void*)(out + 2 * cells + row * w), 0, w * 2, kRsrcWord3);
#pragma unroll
        for (int c = 0; c < COLS; c++) {
            const uint32_t bxy = pxy[c];
            const int bz = (int16_t)((uint32_t)pz[c] >> ((pdz[c] & 1) * 16)) + (pdz[c] >> 1);
            if (OUT_AOS) {
                if (live[c]) {
                    int16_t* op = out + 3 * (row * w + xc[c]);
                    op[0] = (int16_t)bxy; op[1] = (int16_t)(bxy >> 16); op[2] = (int16_t)bz;
                }
            } else {
                __builtin_amdgcn_raw_buffer_store_b32(bxy, rxy, xc[c] * 4, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b16((uint16_t)bz, rz, xc[c] * 2, 0, 0);
            }
            if (flags_out) {
                // |component| > kFastBound, each component biased by
                // kFastBound and compared unsigned
                const uint32_t u = __builtin_bit_cast(uint32_t, __builtin_bit_cast(ushort2v, bxy) +
                                                                    ushort2v{kFastBound, kFastBound});
                const uint32_t uz = (uint32_t)(bz + kFastBound);
                const bool m = (bxy & 0xffffu) == (uint32_t)kMaxDelta3;
                const bool unb = !m && max(max(u & 0xffffu, u >> 16), uz) > 2u * kFastBound;
                const uint64_t bu = __ballot(live[c] && unb), bmx = __ballot(live[c] && m);
                if ((t & 63) == 0 && xc[c] < w)
                    flags_out[row * kFlagSlots + (xc[c] >> 6)] =
                        (uint8_t)((bu ? kRowUnbounded : 0u) | (bmx ? kRowHasMax : 0u));
            }
        }
    };
    auto body = [&](int i, auto P) __attribute__((always_inline)) {
        constexpr int p = decltype(P)::value;   // i mod 3
        if (i > 0) {
            // y-row i + 1 into the slot of y-row i - 2, and the row after
            // it into flight; then row i - 1's stores
            if (i + 1 <= n) stage(i + 1, nxy, nz);
#pragma unroll
            for (int c = 0; c < COLS; c++) {
                self_xy[0][c] = self_xy[1][c];
                self_z[0][c] = self_z[1][c];
                self_xy[1][c] = nxy[1][c];
                self_z[1][c] = (int16_t)nz[1][c];
            }
            prefetch(i + 2);
            // keep the prefetch's loads ahead of the stores' wait on the z
            // loads (the scheduler would otherwise hoist their first use)
#pragma unroll
            for (int c = 0; c < COLS; c++) asm volatile("" : "+v"(pz[c])::"memory");
            flush(i - 1);
            __syncthreads();
        }
#pragma unroll
        for (int c = 0; c < COLS; c++) {
            if (!live[c]) continue;
            const int x = xc[c];
            uint32_t best;
            {
                const int sx = (int16_t)self_xy[0][c];
                const int l = len2_packed(self_xy[0][c], self_z[0][c]);
                best = sx == kMaxDelta3 ? (l < 0 ? 0u : self_max_key<KS>()) : __umul24((uint32_t)l, (uint32_t)KS);
            }
            const int colm = x - step >= 0 ? x - step : w;
            const int colp = x + step < w ? x + step : w;
            typedef __attribute__((address_space(3))) const u32x2* lds_u2;
            lds_u2 bm = (lds_u2)(pb + colm), b0 = (lds_u2)(pb + x), bp = (lds_u2)(pb + colp);
            asm volatile("" : "+v"(bm), "+v"(b0), "+v"(bp));
            auto key = [&](int r, int ic, const u32x2& v) {
                const int tt = 3 * r + ic + 1, jc = r % 3 - 1;
                const uint32_t K = (uint32_t)((ic * ic + jc * jc)) * k1 + (uint32_t)tap_rank(tt);
                const uint32_t cx = ic > 0 ? c54 : ic < 0 ? c54n : 0u;
                const uint32_t cy = jc > 0 ? c54 : jc < 0 ? c54n : 0u;
                return (uint32_t)__builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, v.x),
                                                        __builtin_bit_cast(short2v, cx | cy << 16),
                                                        (int)(v.y + K), false);
            };
#pragma unroll
            for (int k = 0; k < 3; k++) {
                u32x2 nb[3][3];
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int r = 3 * k + j;
                    const int slot = k * 3 + (p + j - 1 + 3) % 3;
                    nb[j][0] = bm[slot * stride];
                    nb[j][1] = r != 4 ? b0[slot * stride] : u32x2{0u, poison_b<KS>()};
                    nb[j][2] = bp[slot * stride];
                }
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int r = 3 * k + j;
                    if (r == 4) {
                        best = min(best, min(key(r, -1, nb[j][0]), key(r, 1, nb[j][2])));
                    } else {
                        best = min(best, min(key(r, -1, nb[j][0]), key(r, 0, nb[j][1])));
                        best = min(best, key(r, 1, nb[j][2]));
                    }
                }
            }
            // the winner, rebuilt from its source (rank 0: the cell itself,
            // whose staged copy may have been cleared as a skipped source)
            const uint32_t rank = key_rank<KS>(best);
            const uint4 e = tbl[p * 27 + rank];
            const int idx = x + (int)e.x;
            // (read for every lane, rank 0 included, and then selected: a
            // select rather than a branch around the read)
            ushort2v v = __builtin_bit_cast(ushort2v, pb[idx].x);
            v += __builtin_bit_cast(ushort2v, e.y);
            pxy[c] = rank ? __builtin_bit_cast(uint32_t, v) : self_xy[0][c];
            // the winner's z from its source cell (in range: a winner is never
            // a tap outside the volume); for rank 0 the entry's offsets are 0,
            // so this is the cell's own z.  The 32-bit word holding it is
            // loaded, so that nothing needs the value before the next row's
            // flush; the row's part of the 16-bit index, U, is even (w is),
            // so the word is U / 2 + (vz >> 1) and the half vz & 1.
            const size_t rowoff = ((size_t)z * h + y0 + i * step) * w;
            const int vs = x + (int)e.w;
            const int vz = IN_AOS ? 3 * vs : vs;
            const size_t U = IN_AOS ? 3 * rowoff + 2 : 2 * cells + rowoff;
            pz[c] = (int)((const uint32_t*)in + U / 2)[vz >> 1];
            pdz[c] = (int)e.z * 2 + (vz & 1);   // the offset, and which half
        }
        __syncthreads();
    };
    // the run's rows fully unrolled (n <= kRun): values in flight pass from
    // one row to the next in straight-line code, with no loop to carry them
    // around (whose register copies would wait for the loads)
    auto row = [&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        body(i, std::integral_constant<int, i % 3>{});
        if (i + 1 < n) return false;
        flush(i);
        return true;
    };
    static_assert(kRun == 8, "the run's rows are unrolled for eight");
    if constexpr (IN_AOS && !kPairAos) {
        // the first pass (three 16-bit loads a cell) keeps a loop: unrolled,
        // its loads in flight would need more registers than four workgroups
        // a CU leave
        for (int i = 0; i < n; i += 3) {
            body(i, std::integral_constant<int, 0>{});
            if (i + 1 < n) body(i + 1, std::integral_constant<int, 1>{});
            if (i + 2 < n) body(i + 2, std::integral_constant<int, 2>{});
        }
        flush(n - 1);
        return;
    }
    if (row(std::integral_constant<int, 0>{})) return;
    if (row(std::integral_constant<int, 1>{})) return;
    if (row(std::integral_constant<int, 2>{})) return;
    if (row(std::integral_constant<int, 3>{})) return;
    if (row(std::integral_constant<int, 4>{})) return;
    if (row(std::integral_constant<int, 5>{})) return;
    if (row(std::integral_constant<int, 6>{})) return;
    row(std::integral_constant<int, 7>{});
}

// Flag bytes of the caller's interleaved field, in the passes' format, for
// the first pass.  Widths that are a multiple of 64: a lane reads 8 cells
// (48 bytes) with three 16-byte loads and 8 lanes make one 64-cell slot.
__global__ void __launch_bounds__(256) k_jfa3d_classify64(size_t cells, int w,
                                                         const uint4* __restrict__ in,
                                                         uint8_t* __restrict__ flags)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;   // cells 8g .. 8g + 7
    bool unb = false, m = false;
    if (8 * g < cells) {
        const uint4 a = in[3 * g], b = in[3 * g + 1], c = in[3 * g + 2];
        const uint32_t v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            // cell k's components are the 16-bit halves 3k, 3k + 1, 3k + 2
            int comp[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int hw = 3 * k + j;
                comp[j] = (int16_t)(v[hw >> 1] >> (16 * (hw & 1)));
            }
            const bool mk = comp[0] == kMaxDelta3;
            m |= mk;
            unb |= !mk && (abs(comp[0]) > kFastBound || abs(comp[1]) > kFastBound || abs(comp[2]) > kFastBound);
        }
    }
    const uint64_t bu = __ballot(unb), bm = __ballot(m);
    const int lane = threadIdx.x & 63;
    if ((lane & 7) == 0 && 8 * g < cells) {
        const int sh = lane & ~7;
        const size_t cell = 8 * g, row = cell / (size_t)w;
        const int slot = (int)(cell - row * w) >> 6;
        flags[row * kFlagSlots + slot] = (uint8_t)((((bu >> sh) & 0xff) ? kRowUnbounded : 0u) |
                                                   (((bm >> sh) & 0xff) ? kRowHasMax : 0u));
    }
}

// Flag bytes of the caller's interleaved field, in the passes' format, for
// the first pass: one 64-lane wave per 64-cell slot (any width).
__global__ void __launch_bounds__(256) k_jfa3d_classify(int w, int rows, const int16_t* __restrict__ in,
                                                       uint8_t* __restrict__ flags)
{
    const int nw = (w + 63) / 64;
    const long long slot = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);   // (row, wave of row)
    if (slot >= (long long)rows * nw) return;
    const int row = (int)(slot / nw), wv = (int)(slot - (long long)row * nw);
    const int x = wv * 64 + (threadIdx.x & 63);
    bool unb = false, m = false;
    if (x < w) {
        const int16_t* c = in + 3 * ((size_t)row * w + x);
        const int ox = c[0], oy = c[1], oz = c[2];
        m = ox == kMaxDelta3;
        unb = !m && (abs(ox) > kFastBound || abs(oy) > kFastBound || abs(oz) > kFastBound);
    }
    const uint64_t bu = __ballot(unb), bm = __ballot(m);
    if ((threadIdx.x & 63) == 0)
        flags[(size_t)row * kFlagSlots + wv] = (uint8_t)((bu ? kRowUnbounded : 0u) | (bm ? kRowHasMax : 0u));
}

void launch_jfa3d_pass(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s)
{
    ProfileScope p("jfa3d_pass", s);
    int xb = (w + 255) / 256;
    hipLaunchKernelGGL(k_jfa3d, dim3((unsigned)((size_t)xb * h * d)), dim3(256), 0, s, step, w, h,
                       d, xb, in, out, cw);
}

// dynamic LDS of the run kernel: the larger of the keyed path's ring of
// (xy, Bb) pairs with its decoding tables (38,232 bytes for 512-wide rows:
// four workgroups a CU), and the general path's planes
static size_t jfa3d_rows_lds(int w)
{
    const size_t general = (size_t)9 * w * 6;
    const size_t keyed = ((size_t)9 * (w + 1) * 8 + 15) / 16 * 16 + 81 * 16;
    return general > keyed ? general : keyed;
}


// every flag byte of the first pass set to kRowHasMax, for a start field
// known to hold only kMaxDelta3 cells and cells within kFastBound
__global__ void __launch_bounds__(256) k_jfa3d_flags_max(size_t words, uint4* __restrict__ flags)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < words) flags[i] = make_uint4(kRowHasMax * 0x01010101u, kRowHasMax * 0x01010101u,
                                         kRowHasMax * 0x01010101u, kRowHasMax * 0x01010101u);
}

static void launch_jfa3d_rows(bool first, bool bounded, bool out_aos, int step, int w, int h, int d,
                              const int16_t* in, int16_t* out, int cw, const uint8_t* flags_in,
                              uint8_t* flags_out, hipStream_t s)
{
    ProfileScope p("jfa3d_pass", s);
    const int cols = w == 512 ? jfa3d_cols<512>() : jfa3d_cols<0>();
    const int threads = (w + 64 * cols - 1) / (64 * cols) * 64;
    const size_t lds = jfa3d_rows_lds(w);
    const size_t cells = (size_t)w * h * d;
    // rows of more than 712 cells need over 64 KiB of LDS, which a kernel
    // gets only after asking for it
    auto want_lds = [&](const void* fn) {
        if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    };
    const int residues = step < h ? step : h;
    const int runs = ((h + step - 1) / step + kRun - 1) / kRun;
    const dim3 grid_runs((unsigned)((size_t)d * residues * runs));
    if (first && bounded) {
        // a start field of the library's own init kernels: no cell beyond
        // kFastBound, so flagging every slot "has a kMaxDelta3 cell" (the
        // keyed path with the skip, always exact) replaces reading it
        const size_t words = (size_t)kFlagSlots * h * d / 16;
        hipLaunchKernelGGL(k_jfa3d_flags_max, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, words,
                           (uint4*)flags_in);
    } else if (first) {
        // the caller's field has no flags: classify it first
        if (w % 64 == 0 && ((uintptr_t)in & 15) == 0) {
            hipLaunchKernelGGL(k_jfa3d_classify64, dim3((unsigned)((cells / 8 + 255) / 256)), dim3(256), 0, s,
                               cells, w, (const uint4*)in, (uint8_t*)flags_in);
        } else {
            const int nw = (w + 63) / 64;
            const long long slots = (long long)h * d * nw;
            hipLaunchKernelGGL(k_jfa3d_classify, dim3((unsigned)((slots + 3) / 4)), dim3(256), 0, s, w, h * d,
                               in, (uint8_t*)flags_in);
        }
    }
#define DOL_JFA3D_ROWS(W, KS)                                                                                \
    {                                                                                                        \
        auto fn = first ? (out_aos ? k_jfa3d_ring<true, true, W, KS> : k_jfa3d_ring<true, false, W, KS>)     \
                        : (out_aos ? k_jfa3d_ring<false, true, W, KS> : k_jfa3d_ring<false, false, W, KS>);  \
        want_lds((const void*)fn);                                                                           \
        hipLaunchKernelGGL(fn, grid_runs, dim3(threads), lds, s, step, w, h, d, cells, in, out, cw,          \
                           flags_in, flags_out, residues);                                                   \
    }
    // the key's scale: 32 where the cross term allows it
    const bool k32 = (long long)cw * step <= fast_max_off<32>();
    if (w == 512) {
        if (k32) DOL_JFA3D_ROWS(512, 32) else DOL_JFA3D_ROWS(512, 27)
    } else {
        if (k32) DOL_JFA3D_ROWS(0, 32) else DOL_JFA3D_ROWS(0, 27)
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

// bytes of the JumpFlood 3D workspace: the second field, then the two flag
// arrays of the row kernels (kFlagSlots bytes per row each)
size_t jfa3d_workspace(int w, int h, int d)
{
    const size_t field = ((size_t)w * h * d * 6 + 255) / 256 * 256;
    return field + 2 * (size_t)kFlagSlots * h * d;
}

// The pass schedule of DistanceOcclusionLib.cpp:2450-2472: steps
// FloorPow2(max dim - 1) down to 1, ping-ponging between delta and the
// workspace's field, plus one step-1 pass when the count is odd so the
// result ends in delta.  bounded: the start field comes from the library's
// border or unsigned init (every cell kMaxDelta3 or within kFastBound).
void jfa3d_run(int w, int h, int d, int16_t* delta, int cw, void* work, hipStream_t s, bool bounded)
{
    int m = w >= h ? w : h;
    m = m >= d ? m : d;
    int steps[40], n = 0;
    for (int step = floor_pow2_host(m - 1); step > 0; step /= 2) steps[n++] = step;
    if (n & 1) steps[n++] = 1;
    int16_t* tmp = (int16_t*)work;
    uint8_t* flags[2];
    flags[0] = (uint8_t*)work + ((size_t)w * h * d * 6 + 255) / 256 * 256;
    flags[1] = flags[0] + (size_t)kFlagSlots * h * d;
    const bool rows = w % 8 == 0 && w <= kRowMax && ((uintptr_t)delta & 15) == 0 &&
                      ((uintptr_t)tmp & 15) == 0;
    int16_t* a = delta;
    int16_t* b = tmp;
    for (int k = 0; k < n; k++) {
        if (rows)
            launch_jfa3d_rows(k == 0, bounded, k == n - 1, steps[k], w, h, d, a, b, cw, flags[(k + 1) & 1],
                              k == n - 1 ? nullptr : flags[k & 1], s);
        else
            launch_jfa3d_pass(steps[k], w, h, d, a, b, cw, s);
        int16_t* t = a; a = b; b = t;
    }
}

}  // namespace dolhip
