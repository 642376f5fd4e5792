// dol_cells.hpp -- per-cell device functions shared by the per-image kernels
// and the batched glyph kernel, so both compute identical bits.
#pragma once

#include "dol_common.hpp"

namespace dolhip {

// ------------------------------------------------------------ border init 2D
//
// Value of cell (x, y) after SetBorderCellsCombo (DistanceOcclusionLib.cpp:504-599),
// obtained by replaying, in the reference's order, only the writes that land
// on this cell.  For row pair p the window at bit b has a = P(b, p),
// b' = P(b, p+1) and the same one column left (x = -1 replicating x = 0):
//   vert = a ^ b', horz = a ^ a_left, vprev = a_left ^ b'_left,
//   corner = vert ^ vprev; a corner goes to (b + vprev - 1, p + horz).
// Within a 32-bit word the reference runs the corner loop, then the
// horizontal loop, then the vertical loop over ascending bits, so when bits
// x and x+1 share a word their writes interleave loop by loop; across a word
// end all of bit x's writes come first.  bit(xx, yy) gives the occupancy of
// an in-range pixel (from a bitmask in global memory or LDS, or thresholded
// bytes); the word boundaries are the reference's mask words either way.
template <class Bit>
__device__ __forceinline__ uint32_t border2_cell_by(Bit bit, int w, int h, int x, int y)
{
    const bool right = x + 1 < w;
    const bool word_end = (x & 31) == 31;
    int c[3][3];
    // every read at an in-range pixel (rows and the right column clamped),
    // then masked: no branches around the reads
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int yy = y - 1 + r;
        const bool ok = yy >= 0 && yy < h;
        const int ys = min(max(yy, 0), h - 1);
        const int xl = x > 0 ? x - 1 : 0, xr = right ? x + 1 : x;
        const int b0 = bit(xl, ys), b1 = bit(x, ys), b2 = bit(xr, ys);
        c[r][0] = ok ? b0 : 0;
        c[r][1] = ok ? b1 : 0;
        c[r][2] = (ok && right) ? b2 : 0;
    }
    // rows outside the image never take part in a window; a uniform
    // neighbourhood over the rows inside gets no write
    {
        const int lo = y >= 1 ? 0 : 1, hi = y + 1 < h ? 2 : 1;
        const int ref = c[1][1];
        bool uniform = true;
        for (int r = lo; r <= hi; r++)
            for (int k = 0; k < 3; k++)
                if ((k < 2 || right) && c[r][k] != ref) uniform = false;
        if (uniform) return pack2(kMaxDelta2, kMaxDelta2);
    }
    int sx = kMaxDelta2, sy = kMaxDelta2;
    auto set = [&](int vx, int vy) { sx = vx; sy = vy; };
    auto set_if_unset = [&](int vx, int vy) {
        if (sx == kMaxDelta2) { sx = vx; sy = vy; }
    };
    if (y >= 1) {   // pair y-1: this cell is its lower row
        const int vert0 = c[0][1] ^ c[1][1], horz0 = c[0][1] ^ c[0][0];
        const int vprev0 = c[0][0] ^ c[1][0], corner0 = vert0 ^ vprev0;
        const int vert1 = c[0][2] ^ c[1][2], horz1 = c[0][2] ^ c[0][1];
        const int vprev1 = vert0, corner1 = vert1 ^ vprev1;
        const bool e1 = corner0 && vprev0 && horz0;
        const bool e2 = right && corner1 && !vprev1 && horz1;
        const bool e3 = vert0;
        if (!word_end) {
            if (e1) set_if_unset(-1, -1);
            if (e2) set_if_unset(+1, -1);
            if (e3) set(0, -1);
        } else {
            if (e1) set_if_unset(-1, -1);
            if (e3) set(0, -1);
            if (e2) set_if_unset(+1, -1);
        }
    }
    if (y + 1 < h) {   // pair y: this cell is its upper row
        const int vert0 = c[1][1] ^ c[2][1], horz0 = c[1][1] ^ c[1][0];
        const int vprev0 = c[1][0] ^ c[2][0], corner0 = vert0 ^ vprev0;
        const int vert1 = c[1][2] ^ c[2][2], horz1 = c[1][2] ^ c[1][1];
        const int vprev1 = vert0, corner1 = vert1 ^ vprev1;
        const bool f1 = corner0 && vprev0 && !horz0;
        const bool f2 = right && corner1 && !vprev1 && !horz1;
        const bool f3 = horz0;
        const bool f4 = right && horz1;
        const bool f5 = vert0;
        if (!word_end) {
            if (f1) set_if_unset(-1, +1);
            if (f2) set_if_unset(+1, +1);
            if (f3) set(-1, 0);
            if (f4) set(+1, 0);
            if (f5) set(0, +1);
        } else {
            if (f1) set_if_unset(-1, +1);
            if (f3) set(-1, 0);
            if (f5) set(0, +1);
            if (f2) set_if_unset(+1, +1);
            if (f4) set(+1, 0);
        }
    } else {   // last row: horizontal changes only (DistanceOcclusionLib.cpp:571-598)
        if (c[1][1] ^ c[1][0]) set(-1, 0);
        if (right && (c[1][2] ^ c[1][1])) set(+1, 0);
    }
    return pack2(sx, sy);
}

// the same from a bitmask of row stride ws words (global or LDS)
__device__ __forceinline__ uint32_t border2_cell(const uint32_t* mask, int ws, int w, int h, int x,
                                                 int y)
{
    return border2_cell_by(
        [&](int xx, int yy) { return (int)((mask[(size_t)yy * ws + (xx >> 5)] >> (xx & 31)) & 1u); }, w, h,
        x, y);
}

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

// The same three words for mask word j of a row stored in a bitmask with no
// row padding: the row's pixel x is bit start + x of `flat` (bit b in word
// b >> 5).  Word start + 32 j - 1 >> 5 may be index -1 (the row's first word
// at the bitmask's first bit), so flat[-1] must be readable; its value is
// never used.  Bits past the row's end are left as they are: each bit of
// border2_word depends only on the same bit of its inputs, so they reach only
// cells past the row.
__device__ __forceinline__ RowBits row_bits_flat(const uint32_t* flat, int start, int j, int ws, int w)
{
    const int p = start + 32 * j - 1;     // the bit left of the word's first cell
    const int q = p >> 5, sh = p & 31;
    const uint32_t A = flat[q], B = flat[q + 1], C = flat[q + 2];
    auto funnel = [&](int o) {            // bits p - sh + o .. + 31, o in sh .. sh + 2
        const uint32_t lo = o < 32 ? A : B, hi = o < 32 ? B : C;
        return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(o & 31));
    };
    RowBits b;
    b.m = funnel(sh + 1);
    b.l = j > 0 ? funnel(sh) : (b.m << 1) | (b.m & 1u);
    b.r = funnel(sh + 2);
    if (j == ws - 1) {
        const uint32_t last = 1u << ((w - 1) & 31);
        b.r = (b.r & ~last) | (b.m & last);
    }
    return b;
}

// occupancy of a byte under the tool's threshold (DistanceOcclusionTool.cpp:646-665)
__device__ __forceinline__ int byte_bit(uint8_t v, int threshold, int less_than)
{
    return ((int)v >= threshold) != (less_than != 0);
}

// ------------------------------------------------------------ JumpFlood 2D

// The cell is kept packed: x and y as the int16 halves of one 32-bit word.
// A candidate is the source plus (ox, oy) with one packed 16-bit add, which
// wraps as the reference's int16 store does, and its squared length is one
// dot product of the int16 pair with itself, summed in 32 bits with wrap as
// len2 does.
typedef unsigned short jfa_u16x2 __attribute__((ext_vector_type(2)));
typedef short jfa_i16x2 __attribute__((ext_vector_type(2)));

// (the three-operand v_dot2_i32_i16 with an inline 0: the builtin selects
// the two-operand v_dot2c form, which needs a register zeroed by an extra
// move before every use)
__device__ __forceinline__ int len2_packed(uint32_t xy)
{
    int r;
    asm("v_dot2_i32_i16 %0, %1, %1, 0" : "=v"(r) : "v"(xy));
    return r;
}

__device__ __forceinline__ void jfa2_tap(uint32_t& best, int& bl, uint32_t v, int ox, int oy)
{
    jfa_u16x2 c = __builtin_bit_cast(jfa_u16x2, v);
    c += jfa_u16x2{(unsigned short)ox, (unsigned short)oy};
    const uint32_t cv = __builtin_bit_cast(uint32_t, c);
    const int cl = len2_packed(cv);
    if (bl > cl) { best = cv; bl = cl; }
}

// One pass for cell (x, y) (DistanceOcclusionLib.cpp:782-873).  Tap order:
// self; (-s,0), (-s,-s), (-s,+s) if x-s >= 0; (+s,0), (+s,-s), (+s,+s) if
// x+s < w; (0,-s); (0,+s).  Candidates wrap to int16; strict ">".
// `in` may be global or LDS.
__device__ __forceinline__ uint32_t jfa2_cell(const uint32_t* in, int w, int h, int x, int y,
                                              int step, int cw)
{
    const int off = cw * step;
    const size_t row = (size_t)y * w;
    const bool up = y - step >= 0, down = y + step < h;
    uint32_t best = in[row + x];
    int bl = len2_packed(best);
    if (x - step >= 0) {
        jfa2_tap(best, bl, in[row + x - step], -off, 0);
        if (up) jfa2_tap(best, bl, in[row - (size_t)step * w + x - step], -off, -off);
        if (down) jfa2_tap(best, bl, in[row + (size_t)step * w + x - step], -off, +off);
    }
    if (x + step < w) {
        jfa2_tap(best, bl, in[row + x + step], +off, 0);
        if (up) jfa2_tap(best, bl, in[row - (size_t)step * w + x + step], +off, -off);
        if (down) jfa2_tap(best, bl, in[row + (size_t)step * w + x + step], +off, +off);
    }
    if (up) jfa2_tap(best, bl, in[row - (size_t)step * w + x], 0, -off);
    if (down) jfa2_tap(best, bl, in[row + (size_t)step * w + x], 0, +off);
    return best;
}

// FloorPow2 of DistanceOcclusionLib.cpp:875-884
__host__ __device__ inline int floor_pow2(int v)
{
    uint32_t u = (uint32_t)v;
    u |= u >> 1; u |= u >> 2; u |= u >> 4; u |= u >> 8; u |= u >> 16;
    return (int)((int32_t)(u + 1u) >> 1);
}

// number of passes of the driver (DistanceOcclusionLib.cpp:887-908): the
// halving steps, plus one step-1 pass when their count is odd
__host__ __device__ inline int jfa_pass_count(int maxdim)
{
    int n = 0;
    for (int s = floor_pow2(maxdim - 1); s > 0; s /= 2) n++;
    return n + (n & 1);
}

// step of pass k of that schedule
__host__ __device__ inline int jfa_pass_step(int maxdim, int k)
{
    const int s0 = floor_pow2(maxdim - 1);
    int s = s0;
    for (int i = 0; i < k && s > 0; i++) s /= 2;
    return s > 0 ? s : 1;
}

// ------------------------------------------------------------ encoders

// EncodeU8Signed, DistanceOcclusionTool.cpp:71-75: clamp, 128 + v * 127 as a
// separate multiply and add, round half to even
__device__ __forceinline__ uint8_t encode_u8_signed(float v)
{
    v = (v < -1.0f) ? -1.0f : ((v > 1.0f) ? 1.0f : v);
    return (uint8_t)(int)__builtin_rintf(__fadd_rn(128.0f, __fmul_rn(v, 127.0f)));
}

// EncodeU8, DistanceOcclusionLib.h:137-141
__device__ __forceinline__ uint8_t encode_u8(float v)
{
    v = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
    return (uint8_t)(int)__builtin_rintf(__fmul_rn(v, 255.0f));
}

// Dist of DistanceOcclusionTool.cpp:54-57
__device__ __forceinline__ float dist_of(int l2) { return sqrt_rn((float)l2); }

// standard signed image value (DistanceOcclusionTool.cpp:1259-1263)
__device__ __forceinline__ uint8_t sdf_u8_cell(uint32_t delta, int bit, float max_len)
{
    const float q = __fdiv_rn(dist_of(len2(unpack2x(delta), unpack2y(delta))), max_len);
    return encode_u8_signed(bit ? -q : q);
}

}  // namespace dolhip
