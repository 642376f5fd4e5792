/*
 * dol_oracle.c -- CPU restatement of DistanceOcclusionLib's distance-field
 * path.  TEST INFRASTRUCTURE ONLY (see dol_oracle.h): it exists so the HIP
 * library can be checked, and so bench.py can time a CPU baseline.  Nothing
 * in the product library links, loads or calls it.
 *
 * Every function names the reference lines it restates.  Where the reference
 * relies on two's-complement wrap (int16 stores of out-of-range deltas,
 * DistanceOcclusionLib.cpp:797-866; int32 overflow of the squared length in
 * Closer, DistanceOcclusionLib.cpp:624-630), the wrap is done here in unsigned
 * arithmetic so that it is defined behaviour with the same result.
 */
#include "dol_oracle.h"

#include <float.h>
#include <stddef.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ utils */

static inline int mask_stride(int w) { return (w + 31) >> 5; }

/* bits of word j that are pixels (32j + i < w); ITER_SET_BITS_BEGIN's
 * "break at x >= w" (DistanceOcclusionLib.cpp:56-63) is the same as masking. */
static inline uint32_t valid_bits(int w, int j)
{
    int n = w - 32 * j;
    if (n >= 32) return 0xffffffffu;
    if (n <= 0) return 0u;
    return (1u << n) - 1u;
}

static inline int16_t wrap16(int v) { return (int16_t)(uint16_t)(uint32_t)v; }

int32_t oc_len2_2(const int16_t* c)
{
    uint32_t a = (uint32_t)((int32_t)c[0] * c[0]);
    uint32_t b = (uint32_t)((int32_t)c[1] * c[1]);
    return (int32_t)(a + b);
}

int32_t oc_len2_3(const int16_t* c)
{
    uint32_t a = (uint32_t)((int32_t)c[0] * c[0]);
    uint32_t b = (uint32_t)((int32_t)c[1] * c[1]);
    uint32_t e = (uint32_t)((int32_t)c[2] * c[2]);
    return (int32_t)(a + b + e);
}

static inline int lowest_bit(uint32_t m) { return __builtin_ctz(m); }

/* ------------------------------------------------------------- generators */

/* LCG of DistanceOcclusionLib.cpp:28-32 */
uint32_t oc_rng(uint32_t* seed, int limit)
{
    *seed = (uint32_t)((uint64_t)(*seed) * 1103515245u) + 12345u;
    return (uint32_t)(((uint64_t)(*seed) * (uint64_t)(uint32_t)limit) >> 32);
}

/* DistanceOcclusionLib.cpp:95-113: draws x then y, ORs the bit in */
void oc_mask_add_points_2d(int w, int h, int count, uint32_t seed, uint32_t* mask)
{
    int ws = mask_stride(w);
    for (int n = 0; n < count; n++) {
        int x = (int)oc_rng(&seed, w);
        int y = (int)oc_rng(&seed, h);
        mask[(size_t)y * ws + (x >> 5)] |= 1u << (x & 31);
    }
}

/* DistanceOcclusionLib.cpp:1224-1244: draws x, y, z */
void oc_mask_add_points_3d(int w, int h, int d, int count, uint32_t seed, uint32_t* mask)
{
    int ws = mask_stride(w);
    for (int n = 0; n < count; n++) {
        int x = (int)oc_rng(&seed, w);
        int y = (int)oc_rng(&seed, h);
        int z = (int)oc_rng(&seed, d);
        mask[((size_t)z * h + y) * ws + (x >> 5)] |= 1u << (x & 31);
    }
}

/* inclusive horizontal span [x0, x1] on rows [y0, y1]
 * (FillMaskBlock, DistanceOcclusionLib.cpp:131-156).  The reference is only
 * defined for w >= 2, h >= 3 (d >= 3): below that its block bounds go
 * negative and it writes outside the mask, so negative bounds are clipped. */
static void fill_span(int ws, int x0, int x1, int y0, int y1, uint32_t* mask)
{
    if (x0 < 0 || x1 < 0) return;
    for (int y = y0 < 0 ? 0 : y0; y <= y1; y++) {
        uint32_t* row = mask + (size_t)y * ws;
        for (int x = x0; x <= x1; x++) row[x >> 5] |= 1u << (x & 31);
    }
}

/* DistanceOcclusionLib.cpp:159-179 */
void oc_mask_add_block_2d(int w, int h, int sides, uint32_t* mask)
{
    int ws = mask_stride(w);
    int x0 = w / 4, x1 = x0 + w / 2 - 1;
    int y0 = h / 3, y1 = y0 + h / 3 - 1;
    if (sides > 4) { fill_span(ws, x0, x1, y0, y1, mask); return; }
    if (sides > 0) fill_span(ws, x0, x1, y0, y0, mask);
    if (sides > 1) fill_span(ws, x0, x0, y0, y1, mask);
    if (sides > 2) fill_span(ws, x0, x1, y1, y1, mask);
    if (sides > 3) fill_span(ws, x1, x1, y0, y1, mask);
}

static void fill_box3(int ws, int h, int x0, int x1, int y0, int y1, int z0, int z1,
                      uint32_t* mask)
{
    for (int z = z0 < 0 ? 0 : z0; z <= z1; z++) fill_span(ws, x0, x1, y0, y1, mask + (size_t)z * ws * h);
}

/* DistanceOcclusionLib.cpp:1259-1287 */
void oc_mask_add_block_3d(int w, int h, int d, int sides, uint32_t* mask)
{
    int ws = mask_stride(w);
    int x0 = w / 4, x1 = x0 + w / 2 - 1;
    int y0 = h / 3, y1 = y0 + h / 3 - 1;
    int z0 = d / 3, z1 = z0 + d / 3 - 1;
    if (sides > 6) { fill_box3(ws, h, x0, x1, y0, y1, z0, z1, mask); return; }
    if (sides > 0) fill_box3(ws, h, x0, x1, y0, y1, z0, z0, mask);
    if (sides > 1) fill_box3(ws, h, x0, x0, y0, y1, z0, z1, mask);
    if (sides > 2) fill_box3(ws, h, x0, x1, y0, y0, z0, z1, mask);
    if (sides > 3) fill_box3(ws, h, x1, x1, y0, y1, z0, z1, mask);
    if (sides > 4) fill_box3(ws, h, x0, x1, y1, y1, z0, z1, mask);
    if (sides > 5) fill_box3(ws, h, x0, x1, y0, y1, z1, z1, mask);
}

/* BitMaskFromImageFile's threshold, DistanceOcclusionTool.cpp:646-665:
 * occupied iff (v >= threshold) XOR lessThan; padding bits are zero. */
void oc_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                        uint32_t* mask)
{
    int ws = mask_stride(w);
    for (int y = 0; y < rows; y++)
        for (int j = 0; j < ws; j++) {
            uint32_t m = 0;
            for (int i = 0; i < 32 && 32 * j + i < w; i++) {
                int v = bytes[(size_t)y * w + 32 * j + i];
                if ((v >= threshold) != (less_than != 0)) m |= 1u << i;
            }
            mask[(size_t)y * ws + j] = m;
        }
}

/* ---------------------------------------------------------- initialisation */

/* DistanceOcclusionLib.cpp:224-254 (2D) */
void oc_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int bit = (mask[(size_t)y * ws + (x >> 5)] >> (x & 31)) & 1;
            int16_t v = (bit ^ (invert != 0)) ? 0 : OC_MAX_DELTA2;
            int16_t* c = delta + 2 * ((size_t)y * w + x);
            c[0] = v; c[1] = v;
        }
}

/* DistanceOcclusionLib.cpp:1289-1321 (3D) */
void oc_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert)
{
    int ws = mask_stride(w);
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t r = (size_t)z * h + y;
                int bit = (mask[r * ws + (x >> 5)] >> (x & 31)) & 1;
                int16_t v = (bit ^ (invert != 0)) ? 0 : OC_MAX_DELTA3;
                int16_t* c = delta + 3 * (r * w + x);
                c[0] = v; c[1] = v; c[2] = v;
            }
}

/* InitFromBitMask<T>, DistanceOcclusionLib.cpp:183-222 */
#define INIT_DIST_BODY(T)                                                           \
    int ws = mask_stride(w);                                                        \
    for (int y = 0; y < h; y++)                                                     \
        for (int x = 0; x < w; x++) {                                               \
            int bit = (mask[(size_t)y * ws + (x >> 5)] >> (x & 31)) & 1;            \
            dist[(size_t)y * w + x] = bit ? (T)0 : max_d;                           \
        }

void oc_init_distances_f32(int w, int h, const uint32_t* mask, float* dist, float max_d)
{ INIT_DIST_BODY(float) }
void oc_init_distances_i32(int w, int h, const uint32_t* mask, int32_t* dist, int32_t max_d)
{ INIT_DIST_BODY(int32_t) }
void oc_init_distances_i16(int w, int h, const uint32_t* mask, int16_t* dist, int16_t max_d)
{ INIT_DIST_BODY(int16_t) }

/* ------------------------------------------------------ border init (2D) */

static inline void put2(int16_t* c, int x, int y) { c[0] = (int16_t)x; c[1] = (int16_t)y; }

/* Combo border initialisation, DistanceOcclusionLib.cpp:504-599 and 604-619.
 * Row pairs (y, y+1) in order; per 32-bit word three passes over set bits in
 * ascending order: corners (written only while the cell is still unset, i.e.
 * x == kMaxDelta2), then horizontal-change cells, then vertical-change
 * cells.  Later writes win, which is what decides ties.  The last row then
 * gets its horizontal changes only. */
void oc_border_2d(int w, int h, const uint32_t* mask, int16_t* delta)
{
    const int ws = mask_stride(w);
    for (size_t i = 0; i < (size_t)w * h; i++) put2(delta + 2 * i, OC_MAX_DELTA2, OC_MAX_DELTA2);
    if (w <= 0 || h <= 0) return;

    for (int y = 0; y + 1 < h; y++) {
        const uint32_t* ra = mask + (size_t)y * ws;
        const uint32_t* rb = ra + ws;
        int16_t* da = delta + 2 * (size_t)y * w;
        int16_t* db = da + 2 * (size_t)w;
        uint32_t carry_a = ra[0] & 1u, carry_b = rb[0] & 1u;   /* x = -1 replicates x = 0 */

        for (int j = 0; j < ws; j++) {
            const uint32_t a = ra[j], b = rb[j];
            const uint32_t a_prev = (a << 1) | carry_a;          /* pixel x-1 of row y   */
            const uint32_t b_prev = (b << 1) | carry_b;          /* pixel x-1 of row y+1 */
            carry_a = a >> 31;
            carry_b = b >> 31;
            const uint32_t lim = valid_bits(w, j);
            const uint32_t vert = a ^ b;                 /* change between rows at x   */
            const uint32_t horz = a ^ a_prev;            /* change x-1 -> x on row y   */
            const uint32_t vprev = a_prev ^ b_prev;      /* change between rows at x-1 */
            uint32_t corner = (vert ^ vprev) & lim;
            uint32_t hbits = horz & lim;
            uint32_t vbits = vert & lim;
            const int base = 32 * j;

            while (corner) {
                int i = lowest_bit(corner); corner &= corner - 1;
                int dx = (vprev >> i) & 1, dy = (horz >> i) & 1;
                int16_t* c = (dy ? db : da) + 2 * (base + i + dx - 1);
                if (c[0] == OC_MAX_DELTA2) put2(c, 1 - 2 * dx, 1 - 2 * dy);
            }
            while (hbits) {
                int i = lowest_bit(hbits); hbits &= hbits - 1;
                put2(da + 2 * (base + i - 1), +1, 0);
                put2(da + 2 * (base + i), -1, 0);
            }
            while (vbits) {
                int i = lowest_bit(vbits); vbits &= vbits - 1;
                put2(da + 2 * (base + i), 0, +1);
                put2(db + 2 * (base + i), 0, -1);
            }
        }
    }

    /* last row: horizontal changes only (DistanceOcclusionLib.cpp:571-598) */
    {
        const uint32_t* r = mask + (size_t)(h - 1) * ws;
        int16_t* dr = delta + 2 * (size_t)(h - 1) * w;
        uint32_t carry = r[0] & 1u;
        for (int j = 0; j < ws; j++) {
            uint32_t a = r[j];
            uint32_t hbits = (a ^ ((a << 1) | carry)) & valid_bits(w, j);
            carry = a >> 31;
            while (hbits) {
                int i = lowest_bit(hbits); hbits &= hbits - 1;
                put2(dr + 2 * (32 * j + i - 1), +1, 0);
                put2(dr + 2 * (32 * j + i), -1, 0);
            }
        }
    }
}

/* ------------------------------------------------------ border init (3D) */

static inline void put3(int16_t* c, int x, int y, int z)
{
    c[0] = (int16_t)x; c[1] = (int16_t)y; c[2] = (int16_t)z;
}

/* face and corner writes may replace an unset cell or a diagonal one */
static inline int replaceable3(const int16_t* c)
{
    return c[0] == OC_MAX_DELTA3 || (abs(c[0]) + abs(c[1]) + abs(c[2]) > 1);
}

/* Combo border initialisation, DistanceOcclusionLib.cpp:1820-2150 and
 * 2154-2173.  Slice pairs (z, z+1) in order; within each, row pairs (y, y+1);
 * per 32-bit word, passes in the order: box corners (only while unset),
 * x-face diagonals, y-face diagonals, z-face diagonals (each replacing unset
 * or diagonal cells), then the x, y and z edge cells (unconditional).  The
 * last row of each slice pair gets y-face diagonals and x/z edges; the last
 * slice gets z-face diagonals and x/y edges, and its last row x edges. */
void oc_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta)
{
    const int ws = mask_stride(w);
    const size_t slice_words = (size_t)ws * h;
    const size_t slice_cells = (size_t)w * h;
    for (size_t i = 0; i < slice_cells * d; i++)
        put3(delta + 3 * i, OC_MAX_DELTA3, OC_MAX_DELTA3, OC_MAX_DELTA3);
    if (w <= 0 || h <= 0 || d <= 0) return;

#define CELL3(X, Y, Z) (delta + 3 * ((size_t)(Z) * slice_cells + (size_t)(Y) * w + (X)))

    for (int z = 0; z + 1 < d; z++) {
        const uint32_t* s0 = mask + (size_t)z * slice_words;
        const uint32_t* s1 = s0 + slice_words;

        for (int y = 0; y + 1 < h; y++) {
            const uint32_t* r00 = s0 + (size_t)y * ws;      /* (y,   z)   */
            const uint32_t* r01 = r00 + ws;                 /* (y+1, z)   */
            const uint32_t* r10 = s1 + (size_t)y * ws;      /* (y,   z+1) */
            const uint32_t* r11 = r10 + ws;                 /* (y+1, z+1) */
            uint32_t c00 = r00[0] & 1u, c01 = r01[0] & 1u, c10 = r10[0] & 1u, c11 = r11[0] & 1u;

            for (int j = 0; j < ws; j++) {
                /* suffix 1 = pixel x, suffix 0 = pixel x-1 */
                const uint32_t p001 = r00[j], p011 = r01[j], p101 = r10[j], p111 = r11[j];
                const uint32_t p000 = (p001 << 1) | c00, p010 = (p011 << 1) | c01;
                const uint32_t p100 = (p101 << 1) | c10, p110 = (p111 << 1) | c11;
                c00 = p001 >> 31; c01 = p011 >> 31; c10 = p101 >> 31; c11 = p111 >> 31;

                const uint32_t ex = p000 ^ p001, ey = p000 ^ p010, ez = p000 ^ p100;
                const uint32_t ezx = p001 ^ p101, eyz = p100 ^ p110, exy = p010 ^ p011;
                const uint32_t fx = ey ^ eyz, fy = ez ^ ezx, fz = ex ^ exy;
                const uint32_t eyx = p001 ^ p011, eyxz = p101 ^ p111;
                const uint32_t fxx = eyx ^ eyxz;
                const uint32_t bc = fz ^ (eyz ^ eyxz);
                const uint32_t lim = valid_bits(w, j);
                const int base = 32 * j;
                uint32_t m;

                m = bc & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dx = (fx >> i) & 1, dy = (fy >> i) & 1, dz = (fz >> i) & 1;
                    int16_t* c = CELL3(base + i - 1 + dx, y + dy, z + dz);
                    if (c[0] == OC_MAX_DELTA3) put3(c, 1 - 2 * dx, 1 - 2 * dy, 1 - 2 * dz);
                }
                m = fxx & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dy = (ezx >> i) & 1, dz = (eyx >> i) & 1;
                    int16_t* c = CELL3(base + i, y + dy, z + dz);
                    if (replaceable3(c)) put3(c, 0, 1 - 2 * dy, 1 - 2 * dz);
                }
                m = fy & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dx = (ez >> i) & 1, dz = (ex >> i) & 1;
                    int16_t* c = CELL3(base + i - 1 + dx, y, z + dz);
                    if (replaceable3(c)) put3(c, 1 - 2 * dx, 0, 1 - 2 * dz);
                }
                m = fz & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dx = (ey >> i) & 1, dy = (ex >> i) & 1;
                    int16_t* c = CELL3(base + i - 1 + dx, y + dy, z);
                    if (replaceable3(c)) put3(c, 1 - 2 * dx, 1 - 2 * dy, 0);
                }
                m = ex & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i - 1, y, z), +1, 0, 0);
                    put3(CELL3(base + i, y, z), -1, 0, 0);
                }
                m = eyx & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i, y, z), 0, +1, 0);
                    put3(CELL3(base + i, y + 1, z), 0, -1, 0);
                }
                m = ezx & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i, y, z), 0, 0, +1);
                    put3(CELL3(base + i, y, z + 1), 0, 0, -1);
                }
            }
        }

        /* last row of the slice pair (DistanceOcclusionLib.cpp:1993-2050) */
        {
            const int y = h - 1;
            const uint32_t* r00 = s0 + (size_t)y * ws;
            const uint32_t* r10 = s1 + (size_t)y * ws;
            uint32_t c00 = r00[0] & 1u, c10 = r10[0] & 1u;
            for (int j = 0; j < ws; j++) {
                const uint32_t p001 = r00[j], p101 = r10[j];
                const uint32_t p000 = (p001 << 1) | c00, p100 = (p101 << 1) | c10;
                c00 = p001 >> 31; c10 = p101 >> 31;
                const uint32_t ex = p000 ^ p001, ez = p000 ^ p100, ezx = p001 ^ p101;
                const uint32_t fy = ez ^ ezx;
                const uint32_t lim = valid_bits(w, j);
                const int base = 32 * j;
                uint32_t m;

                m = fy & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dx = (ez >> i) & 1, dz = (ex >> i) & 1;
                    int16_t* c = CELL3(base + i - 1 + dx, y, z + dz);
                    if (replaceable3(c)) put3(c, 1 - 2 * dx, 0, 1 - 2 * dz);
                }
                m = ex & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i - 1, y, z), +1, 0, 0);
                    put3(CELL3(base + i, y, z), -1, 0, 0);
                }
                m = ezx & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i, y, z), 0, 0, +1);
                    put3(CELL3(base + i, y, z + 1), 0, 0, -1);
                }
            }
        }
    }

    /* last slice (DistanceOcclusionLib.cpp:2053-2149) */
    {
        const int z = d - 1;
        const uint32_t* s0 = mask + (size_t)z * slice_words;
        for (int y = 0; y + 1 < h; y++) {
            const uint32_t* r00 = s0 + (size_t)y * ws;
            const uint32_t* r01 = r00 + ws;
            uint32_t c00 = r00[0] & 1u, c01 = r01[0] & 1u;
            for (int j = 0; j < ws; j++) {
                const uint32_t p001 = r00[j], p011 = r01[j];
                const uint32_t p000 = (p001 << 1) | c00, p010 = (p011 << 1) | c01;
                c00 = p001 >> 31; c01 = p011 >> 31;
                const uint32_t ex = p000 ^ p001, ey = p000 ^ p010, eyx = p001 ^ p011;
                const uint32_t fz = ey ^ eyx;
                const uint32_t lim = valid_bits(w, j);
                const int base = 32 * j;
                uint32_t m;

                m = fz & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    int dx = (ey >> i) & 1, dy = (ex >> i) & 1;
                    int16_t* c = CELL3(base + i - 1 + dx, y + dy, z);
                    if (replaceable3(c)) put3(c, 1 - 2 * dx, 1 - 2 * dy, 0);
                }
                m = ex & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i - 1, y, z), +1, 0, 0);
                    put3(CELL3(base + i, y, z), -1, 0, 0);
                }
                m = eyx & lim;
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(base + i, y, z), 0, +1, 0);
                    put3(CELL3(base + i, y + 1, z), 0, -1, 0);
                }
            }
        }
        {
            const int y = h - 1;
            const uint32_t* r00 = s0 + (size_t)y * ws;
            uint32_t c00 = r00[0] & 1u;
            for (int j = 0; j < ws; j++) {
                const uint32_t p001 = r00[j];
                const uint32_t ex = ((p001 << 1) | c00) ^ p001;
                c00 = p001 >> 31;
                uint32_t m = ex & valid_bits(w, j);
                while (m) {
                    int i = lowest_bit(m); m &= m - 1;
                    put3(CELL3(32 * j + i - 1, y, z), +1, 0, 0);
                    put3(CELL3(32 * j + i, y, z), -1, 0, 0);
                }
            }
        }
    }
#undef CELL3
}

/* ------------------------------------------------------------ jump flood */

/* FloorPow2, DistanceOcclusionLib.cpp:875-884 */
int oc_floor_pow2(int v)
{
    uint32_t u = (uint32_t)v;
    u |= u >> 1; u |= u >> 2; u |= u >> 4; u |= u >> 8; u |= u >> 16;
    return (int)((int32_t)(u + 1u) >> 1);
}

/* one 2D pass, DistanceOcclusionLib.cpp:782-873.  Tap order: self; (-s,0),
 * (-s,-s), (-s,+s) when x-s >= 0; (+s,0), (+s,-s), (+s,+s) when x+s < w;
 * (0,-s); (0,+s).  Strict ">" keeps the earliest of equal candidates; no
 * candidate is skipped, and candidates wrap to int16. */
static void jfa2_rows(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                      int y0, int y1)
{
    const int off = cw * step;
    for (int y = y0; y < y1; y++) {
        for (int x = 0; x < w; x++) {
            int16_t best[2];
            const int16_t* self = in + 2 * ((size_t)y * w + x);
            best[0] = self[0]; best[1] = self[1];
            int32_t bl = oc_len2_2(best);
#define JFA2_TAP(DX, DY)                                                          \
            {                                                                     \
                const int16_t* s = in + 2 * ((size_t)(y + (DY) * step) * w + x + (DX) * step); \
                int16_t c[2] = { wrap16(s[0] + (DX) * off), wrap16(s[1] + (DY) * off) }; \
                int32_t cl = oc_len2_2(c);                                        \
                if (bl > cl) { best[0] = c[0]; best[1] = c[1]; bl = cl; }         \
            }
            if (x - step >= 0) {
                JFA2_TAP(-1, 0)
                if (y - step >= 0) JFA2_TAP(-1, -1)
                if (y + step < h) JFA2_TAP(-1, +1)
            }
            if (x + step < w) {
                JFA2_TAP(+1, 0)
                if (y - step >= 0) JFA2_TAP(+1, -1)
                if (y + step < h) JFA2_TAP(+1, +1)
            }
            if (y - step >= 0) JFA2_TAP(0, -1)
            if (y + step < h) JFA2_TAP(0, +1)
#undef JFA2_TAP
            int16_t* o = out + 2 * ((size_t)y * w + x);
            o[0] = best[0]; o[1] = best[1];
        }
    }
}

void oc_jfa_pass_2d(int step, int w, int h, const int16_t* in, int16_t* out, int cw)
{
    jfa2_rows(step, w, h, in, out, cw, 0, h);
}

static void jfa2_pass_mt(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                         int threads)
{
#ifdef _OPENMP
    if (threads > 1) {
        int chunk = (h + threads - 1) / threads;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int t = 0; t < threads; t++) {
            int y0 = t * chunk, y1 = y0 + chunk < h ? y0 + chunk : h;
            if (y0 < y1) jfa2_rows(step, w, h, in, out, cw, y0, y1);
        }
        return;
    }
#endif
    (void)threads;
    jfa2_rows(step, w, h, in, out, cw, 0, h);
}

/* driver, DistanceOcclusionLib.cpp:887-908: steps FloorPow2(max(w,h)-1) down
 * to 1, ping-ponging; an odd pass count ends with one more step-1 pass back
 * into the caller's buffer. */
void oc_jfa_2d_mt(int w, int h, int16_t* delta, int cw, int threads)
{
    size_t n = (size_t)w * h;
    int16_t* tmp = (int16_t*)malloc(n * 2 * sizeof(int16_t) + 4);
    int16_t* a = delta;
    int16_t* b = tmp;
    for (int step = oc_floor_pow2((w >= h ? w : h) - 1); step > 0; step /= 2) {
        jfa2_pass_mt(step, w, h, a, b, cw, threads);
        int16_t* t = a; a = b; b = t;
    }
    if (a != delta) jfa2_pass_mt(1, w, h, a, delta, cw, threads);
    free(tmp);
}

void oc_jfa_2d(int w, int h, int16_t* delta, int cw) { oc_jfa_2d_mt(w, h, delta, cw, 1); }

/* one 3D pass, DistanceOcclusionLib.cpp:2398-2447: 27 taps, loops z offset
 * (outer), y, x (inner), each over {-s, 0, +s} clipped at the volume edge;
 * a source whose x is kMaxDelta3 is skipped; strict ">". */
static void jfa3_slices(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                        int z0, int z1)
{
    const size_t sl = (size_t)w * h;
    for (int z = z0; z < z1; z++) {
        int kc0 = (z - step) >= 0 ? -step : 0, kc1 = (z + step) < d ? step : 0;
        for (int y = 0; y < h; y++) {
            int jc0 = (y - step) >= 0 ? -step : 0, jc1 = (y + step) < h ? step : 0;
            for (int x = 0; x < w; x++) {
                int ic0 = (x - step) >= 0 ? -step : 0, ic1 = (x + step) < w ? step : 0;
                const size_t idx = (size_t)z * sl + (size_t)y * w + x;
                int16_t best[3] = { in[3 * idx], in[3 * idx + 1], in[3 * idx + 2] };
                int32_t bl = oc_len2_3(best);
                for (int kc = kc0; kc <= kc1; kc += step)
                    for (int jc = jc0; jc <= jc1; jc += step)
                        for (int ic = ic0; ic <= ic1; ic += step) {
                            const int16_t* s = in + 3 * (idx + (ptrdiff_t)kc * (ptrdiff_t)sl
                                                         + (ptrdiff_t)jc * w + ic);
                            if (s[0] == OC_MAX_DELTA3) continue;
                            int16_t c[3] = { wrap16(s[0] + cw * ic), wrap16(s[1] + cw * jc),
                                             wrap16(s[2] + cw * kc) };
                            int32_t cl = oc_len2_3(c);
                            if (bl > cl) { best[0] = c[0]; best[1] = c[1]; best[2] = c[2]; bl = cl; }
                        }
                out[3 * idx] = best[0]; out[3 * idx + 1] = best[1]; out[3 * idx + 2] = best[2];
            }
        }
    }
}

void oc_jfa_pass_3d(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw)
{
    jfa3_slices(step, w, h, d, in, out, cw, 0, d);
}

static void jfa3_pass_mt(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                         int threads)
{
#ifdef _OPENMP
    if (threads > 1) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (int z = 0; z < d; z++) jfa3_slices(step, w, h, d, in, out, cw, z, z + 1);
        return;
    }
#endif
    (void)threads;
    jfa3_slices(step, w, h, d, in, out, cw, 0, d);
}

/* driver, DistanceOcclusionLib.cpp:2450-2472 */
void oc_jfa_3d_mt(int w, int h, int d, int16_t* delta, int cw, int threads)
{
    size_t n = (size_t)w * h * d;
    int16_t* tmp = (int16_t*)malloc(n * 3 * sizeof(int16_t) + 8);
    int16_t* a = delta;
    int16_t* b = tmp;
    int m = w >= h ? w : h;
    m = m >= d ? m : d;
    for (int step = oc_floor_pow2(m - 1); step > 0; step /= 2) {
        jfa3_pass_mt(step, w, h, d, a, b, cw, threads);
        int16_t* t = a; a = b; b = t;
    }
    if (a != delta) jfa3_pass_mt(1, w, h, d, a, delta, cw, threads);
    free(tmp);
}

void oc_jfa_3d(int w, int h, int d, int16_t* delta, int cw) { oc_jfa_3d_mt(w, h, d, delta, cw, 1); }

/* ------------------------------------------------------------ Felzenszwalb */

/* 1D lower envelope of parabolas in float, DistanceOcclusionLib.cpp:1026-1062.
 * Same operation order as the reference so the float results are identical:
 * q*q and 2*(q-v) are formed in int and converted, the envelope pops while
 * s <= z[k], and the scan advances while z[k+1] < q. */
static void felz_1d(int n, size_t m, const float* f, float* out, int* v, float* z)
{
    int k = 0;
    v[0] = 0;
    z[0] = -FLT_MAX;
    z[1] = +FLT_MAX;
    for (int q = 1; q < n; q++) {
        float fq = f[m * q] + (float)(q * q);
        float fv = f[m * v[k]] + (float)(v[k] * v[k]);
        float s = (fq - fv) / (float)(2 * (q - v[k]));
        while (s <= z[k]) {
            k--;
            fv = f[m * v[k]] + (float)(v[k] * v[k]);
            s = (fq - fv) / (float)(2 * (q - v[k]));
        }
        k++;
        v[k] = q;
        z[k] = s;
        z[k + 1] = +FLT_MAX;
    }
    k = 0;
    for (int q = 0; q < n; q++) {
        while (z[k + 1] < (float)q) k++;
        out[q] = (float)((q - v[k]) * (q - v[k])) + f[m * v[k]];
    }
}

/* column pass into a transposed temporary, then row pass back
 * (DistanceOcclusionLib.cpp:1065-1085) */
void oc_felzenszwalb_2d(int w, int h, float* dist)
{
    if (w <= 0 || h <= 0) return;
    int n = w >= h ? w : h;
    float* td = (float*)malloc(sizeof(float) * (size_t)w * h);
    int* v = (int*)malloc(sizeof(int) * (size_t)n);
    float* z = (float*)malloc(sizeof(float) * ((size_t)n + 1));
    for (int x = 0; x < w; x++) felz_1d(h, (size_t)w, dist + x, td + (size_t)x * h, v, z);
    for (int y = 0; y < h; y++) felz_1d(w, (size_t)h, td + y, dist + (size_t)y * w, v, z);
    free(td); free(v); free(z);
}

/* ------------------------------------------------------------ brute force */

/* DistanceOcclusionLib.cpp:1088-1110: every boundary cell (|x| < cw and
 * |y| < cw), in raster order, relaxes every cell in place. */
void oc_brute_force_2d(int w, int h, int16_t* delta, int cw)
{
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            int16_t* cell = delta + 2 * ((size_t)i * w + j);
            if (abs(cell[0]) >= cw || abs(cell[1]) >= cw) continue;
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    int16_t cand[2] = { wrap16((j - x) * cw + cell[0]), wrap16((i - y) * cw + cell[1]) };
                    int16_t* c2 = delta + 2 * ((size_t)y * w + x);
                    if (oc_len2_2(c2) > oc_len2_2(cand)) { c2[0] = cand[0]; c2[1] = cand[1]; }
                }
        }
}

/* DistanceOcclusionLib.cpp:2476-2505: as 2D, but the candidate is the plain
 * cell offset -- cw and the boundary cell's own delta are not used. */
void oc_brute_force_3d(int w, int h, int d, int16_t* delta, int cw)
{
    for (int k = 0; k < d; k++)
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                const int16_t* cell = delta + 3 * (((size_t)k * h + j) * w + i);
                if (abs(cell[0]) >= cw || abs(cell[1]) >= cw || abs(cell[2]) >= cw) continue;
                for (int z = 0; z < d; z++)
                    for (int y = 0; y < h; y++)
                        for (int x = 0; x < w; x++) {
                            int16_t* c2 = delta + 3 * (((size_t)z * h + y) * w + x);
                            int32_t d0 = (i - x) * (i - x) + (j - y) * (j - y) + (k - z) * (k - z);
                            if (d0 < oc_len2_3(c2)) put3(c2, i - x, j - y, k - z);
                        }
            }
}

/* ------------------------------------------------------------ fast sweep */

/* Closer(cur, cand) of DistanceOcclusionLib.cpp:624-630 (2177-2183 in 3D):
 * take the candidate when the current cell is strictly longer. */
static inline void closer2(int16_t* cur, int x, int y)
{
    int16_t c[2] = { wrap16(x), wrap16(y) };
    if (oc_len2_2(cur) > oc_len2_2(c)) { cur[0] = c[0]; cur[1] = c[1]; }
}

static inline void closer3(int16_t* cur, int x, int y, int z)
{
    int16_t c[3] = { wrap16(x), wrap16(y), wrap16(z) };
    if (oc_len2_3(cur) > oc_len2_3(c)) { cur[0] = c[0]; cur[1] = c[1]; cur[2] = c[2]; }
}

/* DistanceOcclusionLib.cpp:633-697.  Four sweeps, one per quadrant
 * direction (dx, dy) = (1 - 2 (s & 1), 1 - 2 (s >> 1 & 1)); rows in the
 * sweep's y order from its second row.  Each row first takes the row before
 * it (offset (0, cy)), then, in the sweep's x order from its second cell, the
 * cell before it in the row (cx, 0) and the one before that in the previous
 * row (cx, cy); cx = -cw dx, cy = -cw dy. */
void oc_fast_sweep_2d(int w, int h, int16_t* delta, int cw)
{
    for (int sweep = 0; sweep < 4; sweep++) {
        const int dx = 1 - 2 * (sweep & 1), dy = 1 - 2 * ((sweep >> 1) & 1);
        const int cx = -cw * dx, cy = -cw * dy;
        const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1;
        for (int v = 1; v < h; v++) {
            const int y = y0 + v * dy, yp = y - dy;
            int16_t* row = delta + 2 * (size_t)y * w;
            const int16_t* prev = delta + 2 * (size_t)yp * w;
            for (int u = 0; u < w; u++) {
                const int x = x0 + u * dx;
                closer2(row + 2 * x, prev[2 * x], prev[2 * x + 1] + cy);
            }
            for (int u = 1; u < w; u++) {
                const int x = x0 + u * dx, xp = x - dx;
                closer2(row + 2 * x, row[2 * xp] + cx, row[2 * xp + 1]);
                closer2(row + 2 * x, prev[2 * xp] + cx, prev[2 * xp + 1] + cy);
            }
        }
    }
}

/* DistanceOcclusionLib.cpp:2186-2262.  Eight sweeps, one per octant; slices
 * in the sweep's z order from its second, rows from their second.  A row
 * first takes, cell by cell, the same row of the previous slice (0, 0, cz)
 * and then the previous row of its slice (0, cy, 0); then, in x order from
 * the second cell, the cell before it (cx, 0, 0). */
void oc_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw)
{
    const size_t slice = (size_t)w * h;
    for (int sweep = 0; sweep < 8; sweep++) {
        const int dx = 1 - 2 * (sweep & 1), dy = 1 - 2 * ((sweep >> 1) & 1), dz = 1 - 2 * ((sweep >> 2) & 1);
        const int cx = -cw * dx, cy = -cw * dy, cz = -cw * dz;
        const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1, z0 = dz > 0 ? 0 : d - 1;
        for (int t = 1; t < d; t++) {
            const int z = z0 + t * dz, zp = z - dz;
            for (int v = 1; v < h; v++) {
                const int y = y0 + v * dy, yp = y - dy;
                int16_t* row = delta + 3 * (z * slice + (size_t)y * w);
                const int16_t* back = delta + 3 * (zp * slice + (size_t)y * w);
                const int16_t* prev = delta + 3 * (z * slice + (size_t)yp * w);
                for (int u = 0; u < w; u++) {
                    const int x = x0 + u * dx;
                    closer3(row + 3 * x, back[3 * x], back[3 * x + 1], back[3 * x + 2] + cz);
                    closer3(row + 3 * x, prev[3 * x], prev[3 * x + 1] + cy, prev[3 * x + 2]);
                }
                for (int u = 1; u < w; u++) {
                    const int x = x0 + u * dx, xp = x - dx;
                    closer3(row + 3 * x, row[3 * xp] + cx, row[3 * xp + 1], row[3 * xp + 2]);
                }
            }
        }
    }
}

/* ------------------------------------------------------- occlusion sweeps */

/* DistanceOcclusionLib.cpp:1114-1146 (2D, 4 planes) and 2507-2538 (3D, 8
 * planes): every plane is 1 where the bit is set, else 0. */
static void init_dirw(int w, size_t rows, int planes, const uint32_t* mask, float* dirw)
{
    const size_t cells = (size_t)w * rows;
    const int ws = mask_stride(w);
    for (size_t r = 0; r < rows; r++)
        for (int x = 0; x < w; x++) {
            const float v = ((mask[r * ws + (x >> 5)] >> (x & 31)) & 1u) ? 1.0f : 0.0f;
            for (int p = 0; p < planes; p++) dirw[p * cells + r * w + x] = v;
        }
}
void oc_init_dirw_2d(int w, int h, const uint32_t* mask, float* dirw) { init_dirw(w, (size_t)h, 4, mask, dirw); }
void oc_init_dirw_3d(int w, int h, int d, const uint32_t* mask, float* dirw)
{
    init_dirw(w, (size_t)h * d, 8, mask, dirw);
}

/* QuadrantCombine and the two OctantCombines (Lib:1150-1168, 2543-2567),
 * the same float terms in the same order. */
static float quadrant_combine(float a, float b) { return 0.75f * ((a + b) - 1.0f * a * b); }
static float octant_combine3(float a, float b, float c) { return 0.33333333f * (a + b + c - a * b * c); }
static float octant_combine6(float oa, float ob, float oc, float ia, float ib, float ic)
{
    const float a = ia + (1.0f - ia) * oa;
    const float b = ib + (1.0f - ib) * ob;
    const float c = ic + (1.0f - ic) * oc;
    return 0.3333333f * (a + b + c - a * b * c);
}

/* Lib:1171-1209.  Plane q is swept from its corner (x direction 1 - 2 (q & 1),
 * y 1 - 2 (q >> 1 & 1)), rows in order, each cell in x order: the first row
 * takes the cell before it, the first column the row before it, every other
 * cell both; v += (1 - v) * QuadrantCombine(before, row before). */
void oc_occlusion_sweep_2d(int w, int h, float* dirw)
{
    const size_t cells = (size_t)w * h;
    for (int q = 0; q < 4; q++) {
        float* f = dirw + q * cells;
        const int dx = 1 - 2 * (q & 1), dy = 1 - 2 * ((q >> 1) & 1);
        const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1;
        for (int v = 0; v < h; v++) {
            const int y = y0 + v * dy;
            float* row = f + (size_t)y * w;
            const float* prev = f + (size_t)(y - dy) * w;
            for (int u = 0; u < w; u++) {
                const int x = x0 + u * dx;
                float c;
                if (u == 0 && v == 0) continue;
                if (v == 0) c = quadrant_combine(row[x - dx], 0.0f);
                else if (u == 0) c = quadrant_combine(prev[x], 0.0f);
                else c = quadrant_combine(row[x - dx], prev[x]);
                row[x] += (1.0f - row[x]) * c;
            }
        }
    }
}

/* Lib:2569-2648 (in place) and 2650-2743 (from `in`), plane q over octant
 * (1 - 2 (q & 1), 1 - 2 (q >> 1 & 1), 1 - 2 (q >> 2 & 1)); slices, rows and
 * cells in order; a neighbour before the field's start is 0.  In place:
 * v += (1 - v) * OctantCombine(cell before, row before, slice before); from
 * `in`: out = OctantCombine(their out values, their in values). */
static void occlusion_sweep_3d(int w, int h, int d, const float* in, float* dirw)
{
    const size_t cells = (size_t)w * h * d, slice = (size_t)w * h;
    for (int q = 0; q < 8; q++) {
        float* f = dirw + q * cells;
        const float* g = in ? in + q * cells : NULL;
        const int dx = 1 - 2 * (q & 1), dy = 1 - 2 * ((q >> 1) & 1), dz = 1 - 2 * ((q >> 2) & 1);
        const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1, z0 = dz > 0 ? 0 : d - 1;
        for (int t = 0; t < d; t++)
            for (int v = 0; v < h; v++)
                for (int u = 0; u < w; u++) {
                    if (u == 0 && v == 0 && t == 0) continue;
                    const int x = x0 + u * dx, y = y0 + v * dy, z = z0 + t * dz;
                    const size_t i = z * slice + (size_t)y * w + x;
                    const size_t il = i - dx, iu = i - (ptrdiff_t)dy * w, ib = i - (ptrdiff_t)dz * slice;
                    const float ol = u ? f[il] : 0.0f, ou = v ? f[iu] : 0.0f, ob = t ? f[ib] : 0.0f;
                    if (g) {
                        f[i] = octant_combine6(ol, ou, ob, u ? g[il] : 0.0f, v ? g[iu] : 0.0f, t ? g[ib] : 0.0f);
                    } else {
                        f[i] += (1.0f - f[i]) * octant_combine3(ol, ou, ob);
                    }
                }
    }
}
void oc_occlusion_sweep_3d(int w, int h, int d, float* dirw) { occlusion_sweep_3d(w, h, d, NULL, dirw); }

/* EncodeAO4 / EncodeAO8 (DistanceOcclusionLib.h:143-155): the planes summed
 * in order, EncodeU8(1 - sum / planes) with the reference's 0.25f / 0.125f;
 * one byte a cell.  EncodeDirW4 / EncodeDirW8 (:157-199): the signed sums
 * along x, y (and z) normalised by 1 / sqrtf of their squared length, then
 * EncodeU8 of w * c * 0.5 + 0.5 for each and of w; RGBA a cell (alpha 255 in
 * 2D, w in 3D).  The tool writes these images per cell
 * (DistanceOcclusionTool.cpp:1521-1540, 1595-1615). */
void oc_encode_ao(size_t n, int planes, const float* dirw, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) {
        float s = dirw[i];
        for (int p = 1; p < planes; p++) s += dirw[p * n + i];
        out[i] = oc_encode_u8(1.0f - (planes == 4 ? 0.25f : 0.125f) * s);
    }
}

void oc_encode_dirw(size_t n, int planes, const float* dirw, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) {
        float d[8] = {0};
        for (int p = 0; p < planes; p++) d[p] = dirw[p * n + i];
        uint8_t* o = out + 4 * i;
        if (planes == 4) {
            float fw = d[0] + d[1] + d[2] + d[3];
            float fx = d[0] - d[1] + d[2] - d[3];
            float fy = d[0] + d[1] - d[2] - d[3];
            const float inv = 1.0f / sqrtf(fx * fx + fy * fy);
            fx *= inv;
            fy *= inv;
            fw *= 0.25f;
            fx = fw * fx * 0.5f + 0.5f;
            fy = fw * fy * 0.5f + 0.5f;
            o[0] = oc_encode_u8(fx); o[1] = oc_encode_u8(fy); o[2] = oc_encode_u8(fw); o[3] = 255;
        } else {
            float fw = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7];
            float fx = d[0] - d[1] + d[2] - d[3] + d[4] - d[5] + d[6] - d[7];
            float fy = d[0] + d[1] - d[2] - d[3] + d[4] + d[5] - d[6] - d[7];
            float fz = d[0] + d[1] + d[2] + d[3] - d[4] - d[5] - d[6] - d[7];
            const float inv = 1.0f / sqrtf(fx * fx + fy * fy + fz * fz);
            fx *= inv;
            fy *= inv;
            fz *= inv;
            fw *= 0.125f;
            fx = fw * fx * 0.5f + 0.5f;
            fy = fw * fy * 0.5f + 0.5f;
            fz = fw * fz * 0.5f + 0.5f;
            o[0] = oc_encode_u8(fx); o[1] = oc_encode_u8(fy); o[2] = oc_encode_u8(fz); o[3] = oc_encode_u8(fw);
        }
    }
}
void oc_occlusion_sweep_3d_from(int w, int h, int d, const float* in, float* out)
{
    occlusion_sweep_3d(w, h, d, in, out);
}

/* ------------------------------------------------- exact squared EDT (ext.) */

/* Not in the reference (which only has the float 2D form above).  The
 * textbook separable transform in exact 64-bit integer arithmetic: per-row
 * nearest seed distance, then the lower envelope of parabolas along each
 * further axis with intersections compared by cross-multiplication.
 * Values of INT64_MAX mean "no seed on this line". */
#define EDT_NONE INT64_MAX

static void edt_envelope_1d(int n, const int64_t* f, size_t fs, int64_t* out, size_t os,
                            int* v)
{
    int k = -1;
    for (int q = 0; q < n; q++) {
        int64_t fq = f[fs * q];
        if (fq == EDT_NONE) continue;
        int64_t Fq = fq + (int64_t)q * q;
        while (k >= 1) {
            int a = v[k - 1], b = v[k];
            int64_t Fa = f[fs * a] + (int64_t)a * a, Fb = f[fs * b] + (int64_t)b * b;
            /* pop b when its interval is empty: s(b,q) <= s(a,b) */
            if ((Fq - Fb) * (int64_t)(b - a) <= (Fb - Fa) * (int64_t)(q - b)) k--;
            else break;
        }
        v[++k] = q;
    }
    if (k < 0) {
        for (int q = 0; q < n; q++) out[os * q] = EDT_NONE;
        return;
    }
    int p = 0;
    for (int q = 0; q < n; q++) {
        for (;;) {
            if (p + 1 > k) break;
            int64_t c0 = f[fs * v[p]] + (int64_t)(q - v[p]) * (q - v[p]);
            int64_t c1 = f[fs * v[p + 1]] + (int64_t)(q - v[p + 1]) * (q - v[p + 1]);
            if (c1 <= c0) p++;
            else break;
        }
        out[os * q] = f[fs * v[p]] + (int64_t)(q - v[p]) * (q - v[p]);
    }
}

static void edt_rows(int w, size_t rows, const uint32_t* mask, int invert, int64_t* g)
{
    int ws = mask_stride(w);
    for (size_t r = 0; r < rows; r++) {
        const uint32_t* row = mask + r * ws;
        int64_t* gr = g + r * w;
        int last = -1;
        for (int x = 0; x < w; x++) {
            int bit = ((row[x >> 5] >> (x & 31)) & 1) ^ (invert != 0);
            if (bit) last = x;
            gr[x] = last < 0 ? EDT_NONE : (int64_t)(x - last);
        }
        last = -1;
        for (int x = w - 1; x >= 0; x--) {
            int bit = ((row[x >> 5] >> (x & 31)) & 1) ^ (invert != 0);
            if (bit) last = x;
            if (last >= 0 && (gr[x] == EDT_NONE || gr[x] > last - x)) gr[x] = last - x;
        }
        for (int x = 0; x < w; x++)
            if (gr[x] != EDT_NONE) gr[x] = gr[x] * gr[x];
    }
}

static void edt_store(size_t n, const int64_t* g, int32_t* d2)
{
    for (size_t i = 0; i < n; i++)
        d2[i] = (g[i] == EDT_NONE || g[i] > INT32_MAX) ? OC_EDT_INF : (int32_t)g[i];
}

void oc_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2)
{
    if (w <= 0 || h <= 0) return;
    size_t n = (size_t)w * h;
    int64_t* g = (int64_t*)malloc(n * sizeof(int64_t));
    int64_t* t = (int64_t*)malloc(n * sizeof(int64_t));
    int* v = (int*)malloc(sizeof(int) * (size_t)(w > h ? w : h));
    edt_rows(w, (size_t)h, mask, invert, g);
    for (int x = 0; x < w; x++) edt_envelope_1d(h, g + x, (size_t)w, t + x, (size_t)w, v);
    edt_store(n, t, d2);
    free(g); free(t); free(v);
}

void oc_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2)
{
    if (w <= 0 || h <= 0 || d <= 0) return;
    size_t sl = (size_t)w * h, n = sl * d;
    int64_t* g = (int64_t*)malloc(n * sizeof(int64_t));
    int64_t* t = (int64_t*)malloc(n * sizeof(int64_t));
    int mx = w > h ? w : h;
    mx = mx > d ? mx : d;
    int* v = (int*)malloc(sizeof(int) * (size_t)mx);
    edt_rows(w, (size_t)h * d, mask, invert, g);
    for (int z = 0; z < d; z++)
        for (int x = 0; x < w; x++)
            edt_envelope_1d(h, g + z * sl + x, (size_t)w, t + z * sl + x, (size_t)w, v);
    for (size_t i = 0; i < sl; i++) edt_envelope_1d(d, t + i, sl, g + i, sl, v);
    edt_store(n, g, d2);
    free(g); free(t); free(v);
}

void oc_edt_brute_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t best = EDT_NONE;
            for (int sy = 0; sy < h; sy++)
                for (int sx = 0; sx < w; sx++) {
                    int bit = ((mask[(size_t)sy * ws + (sx >> 5)] >> (sx & 31)) & 1) ^ (invert != 0);
                    if (!bit) continue;
                    int64_t e = (int64_t)(sx - x) * (sx - x) + (int64_t)(sy - y) * (sy - y);
                    if (e < best) best = e;
                }
            d2[(size_t)y * w + x] = best == EDT_NONE ? OC_EDT_INF : (int32_t)best;
        }
}

void oc_edt_brute_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2)
{
    int ws = mask_stride(w);
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t best = EDT_NONE;
                for (int sz = 0; sz < d; sz++)
                    for (int sy = 0; sy < h; sy++)
                        for (int sx = 0; sx < w; sx++) {
                            size_t r = (size_t)sz * h + sy;
                            int bit = ((mask[r * ws + (sx >> 5)] >> (sx & 31)) & 1) ^ (invert != 0);
                            if (!bit) continue;
                            int64_t e = (int64_t)(sx - x) * (sx - x) + (int64_t)(sy - y) * (sy - y)
                                      + (int64_t)(sz - z) * (sz - z);
                            if (e < best) best = e;
                        }
                d2[((size_t)z * h + y) * w + x] = best == EDT_NONE ? OC_EDT_INF : (int32_t)best;
            }
}

/* float distance of an exact squared distance: sqrtf of the int converted to
 * float; "no seed" maps to sqrtf(FLT_MAX), which is what the reference's
 * float path yields for a cell that no seed reaches (DistanceOcclusionTool.cpp:753-754). */
float oc_edt_float(int32_t d2)
{
    if (d2 == OC_EDT_INF) return sqrtf(FLT_MAX);
    return sqrtf((float)d2);
}

/* ------------------------------------------------------------- encoders */

/* EncodeU8, DistanceOcclusionLib.h:137-141 */
uint8_t oc_encode_u8(float v)
{
    v = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
    return (uint8_t)lrintf(v * 255);
}

/* EncodeU8Signed, DistanceOcclusionTool.cpp:71-75 */
uint8_t oc_encode_u8_signed(float v)
{
    v = (v < -1.0f) ? -1.0f : ((v > 1.0f) ? 1.0f : v);
    return (uint8_t)lrintf(128 + v * 127);
}

/* Dist, DistanceOcclusionTool.cpp:54-57 (int32 squared length, as in Dist2) */
void oc_dist_2d(int n, const int16_t* delta, float* out)
{
    for (int i = 0; i < n; i++) out[i] = sqrtf((float)oc_len2_2(delta + 2 * i));
}

static inline int mask_bit(const uint32_t* mask, int ws, size_t row, int x)
{
    return (mask[row * ws + (x >> 5)] >> (x & 31)) & 1;
}

/* signed distance in cells from half-cell deltas of the border method:
 * -/+ 0.5 * sqrtf(|delta|^2), negative inside (set bit) */
void oc_signed_dist_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float* out)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float d = 0.5f * sqrtf((float)oc_len2_2(delta + 2 * i));
            out[i] = mask_bit(mask, ws, (size_t)y, x) ? -d : d;
        }
}

void oc_signed_dist_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta, float* out)
{
    int ws = mask_stride(w);
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t r = (size_t)z * h + y, i = r * w + x;
                float dd = 0.5f * sqrtf((float)oc_len2_3(delta + 3 * i));
                out[i] = mask_bit(mask, ws, r, x) ? -dd : dd;
            }
}

/* standard signed image of the border method, DistanceOcclusionTool.cpp:1242-1271;
 * max_len is the value passed there (twice the tool's maxLen). */
void oc_sdf_u8_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len, uint8_t* out)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float d = sqrtf((float)oc_len2_2(delta + 2 * i)) / max_len;
            out[i] = oc_encode_u8_signed(mask_bit(mask, ws, (size_t)y, x) ? -d : d);
        }
}

/* DistanceOcclusionTool.cpp:1369-1404 */
void oc_sdf_u8_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta, float max_len,
                  uint8_t* out)
{
    int ws = mask_stride(w);
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t r = (size_t)z * h + y, i = r * w + x;
                float dd = sqrtf((float)oc_len2_3(delta + 3 * i)) / max_len;
                out[i] = oc_encode_u8_signed(mask_bit(mask, ws, r, x) ? -dd : dd);
            }
}

/* delta image, DistanceOcclusionTool.cpp:1281-1288 */
void oc_sdf_delta_rgba_2d(int w, int h, const int16_t* delta, float max_len, uint8_t* out)
{
    for (size_t i = 0; i < (size_t)w * h; i++) {
        out[4 * i + 0] = oc_encode_u8_signed(delta[2 * i] / max_len);
        out[4 * i + 1] = oc_encode_u8_signed(delta[2 * i + 1] / max_len);
        out[4 * i + 2] = 0;
        out[4 * i + 3] = 255;
    }
}

/* red/green image, DistanceOcclusionTool.cpp:1289-1307: the distance goes
 * to channel (bit), the other of the first two is zero. */
void oc_sdf_rg_rgba_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                       uint8_t* out)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            int b = mask_bit(mask, ws, (size_t)y, x);
            out[4 * i + b] = oc_encode_u8(sqrtf((float)oc_len2_2(delta + 2 * i)) / max_len);
            out[4 * i + (b ^ 1)] = 0;
            out[4 * i + 2] = 0;
            out[4 * i + 3] = 255;
        }
}

/* Every image of the border method: WriteDistances 2D,
 * DistanceOcclusionTool.cpp:1240-1331, and 3D, 1367-1469 (d = 1 is the 2D
 * writer: the same per-cell formulas, z absent).  type 0 sdf (one byte a
 * cell: EncodeU8Signed of -/+ Dist / maxLen, negative on set bits), 1
 * sdf-delta (the delta's components over maxLen, signed; 2D: blue 0), 2
 * sdf-rg (EncodeU8(Dist / maxLen) in red outside, green inside), 3 sdf-angles
 * (the delta's components over Dist, signed; 2D: blue 0); RGBA alpha 255.
 * max_len is the value passed there (twice the tool's maxLen). */
void oc_border_image(int w, int h, int d, int dims, const uint32_t* mask, const int16_t* delta,
                     float max_len, int type, uint8_t* out)
{
    int ws = mask_stride(w);
    const int nc = dims;   /* int16 components a cell */
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t r = (size_t)z * h + y, i = r * w + x;
                const int16_t* c = delta + (size_t)nc * i;
                int l2 = dims == 2 ? oc_len2_2(c) : oc_len2_3(c);
                float dist = sqrtf((float)l2);
                int b = mask_bit(mask, ws, r, x);
                if (type == 0) {
                    float v = dist / max_len;
                    out[i] = oc_encode_u8_signed(b ? -v : v);
                    continue;
                }
                uint8_t* o = out + 4 * i;
                if (type == 1) {
                    o[0] = oc_encode_u8_signed(c[0] / max_len);
                    o[1] = oc_encode_u8_signed(c[1] / max_len);
                    o[2] = dims == 3 ? oc_encode_u8_signed(c[2] / max_len) : 0;
                } else if (type == 2) {
                    o[b] = oc_encode_u8(dist / max_len);
                    o[b ^ 1] = 0;
                    o[2] = 0;
                } else {
                    o[0] = oc_encode_u8_signed(c[0] / dist);
                    o[1] = oc_encode_u8_signed(c[1] / dist);
                    o[2] = dims == 3 ? oc_encode_u8_signed(c[2] / dist) : 0;
                }
                o[3] = 255;
            }
}

/* Unsigned distance image of the delta methods, WriteDistances
 * DistanceOcclusionTool.cpp:678-686 (2D) and 886-899 (3D, slice by slice):
 * EncodeU8(Dist / maxLen). */
void oc_distance_u8(size_t n, int dims, const int16_t* delta, float max_len, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) {
        int l2 = dims == 2 ? oc_len2_2(delta + 2 * i) : oc_len2_3(delta + 3 * i);
        out[i] = oc_encode_u8(sqrtf((float)l2) / max_len);
    }
}

/* ... and of the scalar methods, DistanceOcclusionTool.cpp:688-697:
 * EncodeU8(distances[i] / maxLen) for float, int32 and int16 distances
 * (Felzenszwalb's float roots, Chamfer's int16 values with maxLen times its
 * multiplier, :753-754, :827). */
void oc_scalar_u8_f32(size_t n, const float* v, float max_len, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) out[i] = oc_encode_u8(v[i] / max_len);
}
void oc_scalar_u8_i32(size_t n, const int32_t* v, float max_len, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) out[i] = oc_encode_u8(v[i] / max_len);
}
void oc_scalar_u8_i16(size_t n, const int16_t* v, float max_len, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) out[i] = oc_encode_u8(v[i] / max_len);
}

/* sqrtf of each value, DistanceOcclusionTool.cpp:753-754 */
void oc_sqrt_f32(size_t n, const float* v, float* out)
{
    for (size_t i = 0; i < n; i++) out[i] = sqrtf(v[i]);
}

/* Two-field signed images of the tool's -S mode: WriteDistances,
 * DistanceOcclusionTool.cpp:1029-1080 (2D).  d0 is the exterior field, d1 the
 * interior one; type 0 sdf (1 byte a cell), 1 sdf-delta, 2 sdf-rg,
 * 3 sdf-angles (RGBA).  Dist is sqrtf of the int32 squared length
 * (DistanceOcclusionTool.cpp:46-57); sums of deltas are int, as there. */
void oc_two_field_image_2d(int w, int h, const int16_t* d0, const int16_t* d1, float max_len,
                           int type, uint8_t* out)
{
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const int16_t* a = d0 + 2 * i;
        const int16_t* b = d1 + 2 * i;
        const float da = sqrtf((float)oc_len2_2(a)), db = sqrtf((float)oc_len2_2(b));
        if (type == 0) {
            out[i] = oc_encode_u8_signed((da - db) / max_len);
            continue;
        }
        uint8_t* o = out + 4 * i;
        if (type == 1) {
            o[0] = oc_encode_u8_signed((a[0] + b[0]) / max_len);
            o[1] = oc_encode_u8_signed((a[1] + b[1]) / max_len);
        } else if (type == 2) {
            o[0] = oc_encode_u8(da / max_len);
            o[1] = oc_encode_u8(db / max_len);
        } else {
            const float d = da - db;
            o[0] = oc_encode_u8_signed((a[0] + b[0]) / d);
            o[1] = oc_encode_u8_signed((a[1] + b[1]) / d);
        }
        o[2] = 0;
        o[3] = 255;
    }
}

/* The same in 3D, slice by slice: DistanceOcclusionTool.cpp:1126-1195.
 * Type 1 and 3 put z in blue.  Type 3 indexes deltas0/deltas1 with the
 * in-slice index (DistanceOcclusionTool.cpp:1178-1186), so every slice shows
 * the first slice's values. */
void oc_two_field_image_3d(int w, int h, int d, const int16_t* d0, const int16_t* d1,
                           float max_len, int type, uint8_t* out)
{
    const size_t n = (size_t)w * h;
    for (int slice = 0; slice < d; slice++)
        for (size_t i = 0; i < n; i++) {
            const size_t c = (size_t)slice * n + i;
            const size_t j = type == 3 ? i : c;
            const int16_t* a = d0 + 3 * j;
            const int16_t* b = d1 + 3 * j;
            const float da = sqrtf((float)oc_len2_3(a)), db = sqrtf((float)oc_len2_3(b));
            if (type == 0) {
                out[c] = oc_encode_u8_signed((da - db) / max_len);
                continue;
            }
            uint8_t* o = out + 4 * c;
            if (type == 2) {
                o[0] = oc_encode_u8(da / max_len);
                o[1] = oc_encode_u8(db / max_len);
                o[2] = 0;
            } else {
                const float q = type == 1 ? max_len : da - db;
                o[0] = oc_encode_u8_signed((a[0] + b[0]) / q);
                o[1] = oc_encode_u8_signed((a[1] + b[1]) / q);
                o[2] = oc_encode_u8_signed((a[2] + b[2]) / q);
            }
            o[3] = 255;
        }
}

/* ------------------------------------------------------------- voxeliser */

/* Surface voxelisation of a triangle list into a bitmask:
 * CreateBitMaskFromTriangles, MeshSupport.cpp:369-475, with the
 * triangle/cell test of cTriCellIntersectDelta, MeshSupport.cpp:158-300
 * (separating axes: the 9 edge x axis products, the triangle's box, its
 * plane).  Every product and sum is a separate float operation in the order
 * written there (this file is built without contraction).  indices == NULL
 * reads consecutive vertices (the reference dereferences indices anyway). */
typedef struct { float x, y, z; } vox_v3;

static vox_v3 v3_sub(vox_v3 a, vox_v3 b) { vox_v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static float v3_dot(vox_v3 a, vox_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float vmin_f(float a, float b) { return a < b ? a : b; }
static float vmax_f(float a, float b) { return b < a ? a : b; }
static void fswap(float* a, float* b) { float t = *a; *a = *b; *b = t; }

typedef struct {
    vox_v3 e1, t1, u1, e2, t2, u2, e3, t3, u3, pmin, pmax, hw, n;
    float ndv_min, ndv_max, nd;
} vox_tri;

/* one edge's three axis tests: t = lower, u = upper bound of the triangle's
 * projection less/plus the box's radius */
static void vox_edge(vox_v3 e, vox_v3 pa, vox_v3 pb, vox_v3 pc, vox_v3 qa, vox_v3 qb, vox_v3 qc,
                     vox_v3 hw, vox_v3* t, vox_v3* u)
{
    /* x from pa, y from pb, z from pc; the other vertex per axis from q* */
    t->x = e.z * pa.y - e.y * pa.z;
    t->y = -e.z * pb.x + e.x * pb.z;
    t->z = e.y * pc.x - e.x * pc.y;
    u->x = e.z * qa.y - e.y * qa.z;
    u->y = -e.z * qb.x + e.x * qb.z;
    u->z = e.y * qc.x - e.x * qc.y;
    if (t->x >= u->x) fswap(&t->x, &u->x);
    if (t->y >= u->y) fswap(&t->y, &u->y);
    if (t->z >= u->z) fswap(&t->z, &u->z);
    const vox_v3 ae = {fabsf(e.x), fabsf(e.y), fabsf(e.z)};
    const vox_v3 r = {ae.z * hw.y + ae.y * hw.z, ae.z * hw.x + ae.x * hw.z, ae.y * hw.x + ae.x * hw.y};
    t->x = t->x - r.x; t->y = t->y - r.y; t->z = t->z - r.z;
    u->x = u->x + r.x; u->y = u->y + r.y; u->z = u->z + r.z;
}

static void vox_setup(vox_v3 p1, vox_v3 p2, vox_v3 p3, vox_v3 hw, vox_tri* T)
{
    T->hw = hw;
    T->e1 = v3_sub(p2, p1);
    T->e2 = v3_sub(p3, p2);
    T->e3 = v3_sub(p1, p3);
    /* TriDoubleAreaNormal, MeshSupport.cpp:141-155 */
    const vox_v3 a = p1, b = p2, c = p3;
    T->n.x = (a.y - b.y) * (a.z + b.z) + (b.y - c.y) * (b.z + c.z) + (c.y - a.y) * (c.z + a.z);
    T->n.y = (a.z - b.z) * (a.x + b.x) + (b.z - c.z) * (b.x + c.x) + (c.z - a.z) * (c.x + a.x);
    T->n.z = (a.x - b.x) * (a.y + b.y) + (b.x - c.x) * (b.y + c.y) + (c.x - a.x) * (c.y + a.y);
    vox_v3 vmn, vmx;
    vmn.x = T->n.x > 0.0f ? -hw.x : hw.x;  vmx.x = T->n.x > 0.0f ? hw.x : -hw.x;
    vmn.y = T->n.y > 0.0f ? -hw.y : hw.y;  vmx.y = T->n.y > 0.0f ? hw.y : -hw.y;
    vmn.z = T->n.z > 0.0f ? -hw.z : hw.z;  vmx.z = T->n.z > 0.0f ? hw.z : -hw.z;
    T->ndv_min = v3_dot(T->n, vmn);
    T->ndv_max = v3_dot(T->n, vmx);
    T->nd = v3_dot(T->n, p1);
    /* per edge, the vertex pairs of MeshSupport.cpp:214-251 */
    vox_edge(T->e1, p1, p1, p2, p3, p3, p3, hw, &T->t1, &T->u1);
    vox_edge(T->e2, p1, p1, p1, p3, p3, p2, hw, &T->t2, &T->u2);
    vox_edge(T->e3, p1, p1, p2, p2, p2, p3, hw, &T->t3, &T->u3);
    T->pmin.x = vmin_f(vmin_f(p1.x, p2.x), p3.x);
    T->pmin.y = vmin_f(vmin_f(p1.y, p2.y), p3.y);
    T->pmin.z = vmin_f(vmin_f(p1.z, p2.z), p3.z);
    T->pmax.x = vmax_f(vmax_f(p1.x, p2.x), p3.x);
    T->pmax.y = vmax_f(vmax_f(p1.y, p2.y), p3.y);
    T->pmax.z = vmax_f(vmax_f(p1.z, p2.z), p3.z);
}

/* IntersectZ / IntersectY / IntersectX, MeshSupport.cpp:279-351 */
static int vox_hit_z(const vox_tri* T, vox_v3 c)
{
    return !(T->pmin.z - c.z > T->hw.z || T->pmax.z - c.z < -T->hw.z);
}

static int vox_hit_y(const vox_tri* T, vox_v3 c)
{
    float dr;
    dr = T->e1.z * c.y - T->e1.y * c.z;
    if (T->t1.x > dr || T->u1.x < dr) return 0;
    dr = T->e2.z * c.y - T->e2.y * c.z;
    if (T->t2.x > dr || T->u2.x < dr) return 0;
    dr = T->e3.z * c.y - T->e3.y * c.z;
    if (T->t3.x > dr || T->u3.x < dr) return 0;
    return !(T->pmin.y - c.y > T->hw.y || T->pmax.y - c.y < -T->hw.y);
}

static int vox_hit_x(const vox_tri* T, vox_v3 c)
{
    float dr;
    dr = -T->e1.z * c.x + T->e1.x * c.z;
    if (T->t1.y > dr || T->u1.y < dr) return 0;
    dr = T->e1.y * c.x - T->e1.x * c.y;
    if (T->t1.z > dr || T->u1.z < dr) return 0;
    dr = -T->e2.z * c.x + T->e2.x * c.z;
    if (T->t2.y > dr || T->u2.y < dr) return 0;
    dr = T->e2.y * c.x - T->e2.x * c.y;
    if (T->t2.z > dr || T->u2.z < dr) return 0;
    dr = -T->e3.z * c.x + T->e3.x * c.z;
    if (T->t3.y > dr || T->u3.y < dr) return 0;
    dr = T->e3.y * c.x - T->e3.x * c.y;
    if (T->t3.z > dr || T->u3.z < dr) return 0;
    if (T->pmin.x - c.x > T->hw.x || T->pmax.x - c.x < -T->hw.x) return 0;
    const float d = T->nd - v3_dot(T->n, c);
    return T->ndv_min <= d && T->ndv_max >= d;
}

void oc_mask_from_triangles(int tri_count, const int* indices, const float* vertices,
                            const float* bbox, int w, int h, int d, uint32_t* mask)
{
    const size_t row = (size_t)mask_stride(w), slice = row * h;
    const vox_v3 whd = {(float)w, (float)h, (float)d};
    const vox_v3 cmin = {bbox[0], bbox[1], bbox[2]};
    const vox_v3 ext = {bbox[3] - bbox[0], bbox[4] - bbox[1], bbox[5] - bbox[2]};
    const vox_v3 cw = {ext.x / whd.x, ext.y / whd.y, ext.z / whd.z};
    const vox_v3 inv = {whd.x / ext.x, whd.y / ext.y, whd.z / ext.z};
    const vox_v3 hw = {cw.x * 0.500001f, cw.y * 0.500001f, cw.z * 0.500001f};
    for (int t = 0; t < tri_count; t++) {
        const int i1 = indices ? indices[3 * t] : 3 * t, i2 = indices ? indices[3 * t + 1] : 3 * t + 1,
                  i3 = indices ? indices[3 * t + 2] : 3 * t + 2;
        const vox_v3 p1 = {vertices[3 * i1], vertices[3 * i1 + 1], vertices[3 * i1 + 2]};
        const vox_v3 p2 = {vertices[3 * i2], vertices[3 * i2 + 1], vertices[3 * i2 + 2]};
        const vox_v3 p3 = {vertices[3 * i3], vertices[3 * i3 + 1], vertices[3 * i3 + 2]};
        /* TriangleBounds, MeshSupport.cpp:353-366 */
        const vox_v3 bmn = {vmin_f(vmin_f(p1.x, p2.x), p3.x), vmin_f(vmin_f(p1.y, p2.y), p3.y),
                            vmin_f(vmin_f(p1.z, p2.z), p3.z)};
        const vox_v3 bmx = {vmax_f(vmax_f(p1.x, p2.x), p3.x), vmax_f(vmax_f(p1.y, p2.y), p3.y),
                            vmax_f(vmax_f(p1.z, p2.z), p3.z)};
        int cx0 = (int)floorf((bmn.x - cmin.x) * inv.x), cy0 = (int)floorf((bmn.y - cmin.y) * inv.y),
            cz0 = (int)floorf((bmn.z - cmin.z) * inv.z);
        int cx1 = (int)ceilf((bmx.x - cmin.x) * inv.x), cy1 = (int)ceilf((bmx.y - cmin.y) * inv.y),
            cz1 = (int)ceilf((bmx.z - cmin.z) * inv.z);
        if (cx0 < 0) cx0 = 0;
        if (cy0 < 0) cy0 = 0;
        if (cz0 < 0) cz0 = 0;
        if (cx1 > w) cx1 = w;
        if (cy1 > h) cy1 = h;
        if (cz1 > d) cz1 = d;
        if (cx0 == cx1) continue;
        if (cx0 + 1 == cx1 && cy0 + 1 == cy1 && cz0 + 1 == cz1) {   /* single cell: no test */
            mask[slice * cz0 + row * cy0 + (cx0 >> 5)] |= 1u << (cx0 & 31);
            continue;
        }
        vox_tri T;
        vox_setup(p1, p2, p3, hw, &T);
        vox_v3 c = {0.0f, 0.0f, 0.0f};
        for (int z = cz0; z < cz1; z++) {
            c.z = cmin.z + cw.z * ((float)z + 0.5f);
            if (!vox_hit_z(&T, c)) continue;
            for (int y = cy0; y < cy1; y++) {
                c.y = cmin.y + cw.y * ((float)y + 0.5f);
                if (!vox_hit_y(&T, c)) continue;
                for (int x = cx0; x < cx1; x++) {
                    c.x = cmin.x + cw.x * ((float)x + 0.5f);
                    if (vox_hit_x(&T, c)) mask[slice * z + row * y + (x >> 5)] |= 1u << (x & 31);
                }
            }
        }
    }
}

/* ------------------------------------------------------------- invariants */

/* CheckDeltas, DistanceOcclusionTool.cpp:371-402: each delta must land in
 * bounds on a boundary cell.  The sum is formed in int16 as there. */
long oc_check_deltas_2d(int w, int h, const int16_t* delta, int cw)
{
    long bad = 0;
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) {
            const int16_t* c = delta + 2 * ((size_t)j * w + i);
            int cx = wrap16(c[0] + i * cw), cy = wrap16(c[1] + j * cw);
            if (cx < 0 || cx >= cw * w || cy < 0 || cy >= cw * h) { bad++; continue; }
            const int16_t* r = delta + 2 * ((size_t)(cy / cw) * w + cx / cw);
            if (!(abs(r[0]) < cw && abs(r[1]) < cw)) bad++;
        }
    return bad;
}

/* DistanceOcclusionTool.cpp:404-439 */
long oc_check_deltas_3d(int w, int h, int d, const int16_t* delta, int cw)
{
    long bad = 0;
    for (int k = 0; k < d; k++)
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                const int16_t* c = delta + 3 * (((size_t)k * h + j) * w + i);
                int cx = wrap16(c[0] + i * cw), cy = wrap16(c[1] + j * cw), cz = wrap16(c[2] + k * cw);
                if (cx < 0 || cx >= cw * w || cy < 0 || cy >= cw * h || cz < 0 || cz >= cw * d) {
                    bad++;
                    continue;
                }
                const int16_t* r = delta + 3 * (((size_t)(cz / cw) * h + cy / cw) * w + cx / cw);
                if (!(abs(r[0]) < cw && abs(r[1]) < cw && abs(r[2]) < cw)) bad++;
            }
    return bad;
}

/* CheckBorderDeltas 2D, DistanceOcclusionTool.cpp:441-500, with the loops
 * over x < w and y < h (the reference runs x to h and y to w, which reads
 * outside the mask and field unless w == h).  Occupancy reads clamp to the
 * image (Occupied, :442-456).  A boundary cell (|x|, |y| < 2) must see the
 * other occupancy at its delta, and when the delta is diagonal the same
 * occupancy at its four axis neighbours (no closer edge); any other cell the
 * same occupancy at all nine.  Returns the number of failed expectations. */
static int occ2_clamped(const uint32_t* mask, int w, int h, int x, int y)
{
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return mask_bit(mask, mask_stride(w), (size_t)y, x);
}

long oc_check_border_deltas_2d(int w, int h, const uint32_t* mask, const int16_t* delta)
{
    long bad = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int16_t* c = delta + 2 * ((size_t)y * w + x);
            const int occ = occ2_clamped(mask, w, h, x, y);
            if (abs(c[0]) < 2 && abs(c[1]) < 2) {
                bad += occ2_clamped(mask, w, h, x + c[0], y + c[1]) != !occ;
                if (c[0] != 0 && c[1] != 0) {
                    bad += occ2_clamped(mask, w, h, x - 1, y) != occ;
                    bad += occ2_clamped(mask, w, h, x + 1, y) != occ;
                    bad += occ2_clamped(mask, w, h, x, y - 1) != occ;
                    bad += occ2_clamped(mask, w, h, x, y + 1) != occ;
                }
            } else {
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) bad += occ2_clamped(mask, w, h, x + dx, y + dy) != occ;
            }
        }
    return bad;
}

/* CheckBorderDeltas 3D, DistanceOcclusionTool.cpp:502-599, loops over x < w and
 * y < h (the reference swaps them as in 2D).  Neighbours outside the volume
 * pass (Expect, :508-521).  A boundary cell (every |component| < 2) must see
 * the other occupancy at its delta; a box diagonal the same occupancy at the
 * twelve face diagonals; each nonzero pair of components (xy, yz, zx) the
 * same occupancy at that plane's four axis neighbours; any other cell the
 * same occupancy at all 27. */
static int expect3(const uint32_t* mask, int w, int h, int d, int x, int y, int z, int value)
{
    if (x < 0 || x >= w || y < 0 || y >= h || z < 0 || z >= d) return 1;
    return mask_bit(mask, mask_stride(w), (size_t)z * h + y, x) == value;
}

long oc_check_border_deltas_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta)
{
    static const int face[12][3] = { { -1, -1, 0 }, { -1, 1, 0 }, { 1, -1, 0 }, { 1, 1, 0 },
                                      { -1, 0, -1 }, { -1, 0, 1 }, { 1, 0, -1 }, { 1, 0, 1 },
                                      { 0, -1, -1 }, { 0, -1, 1 }, { 0, 1, -1 }, { 0, 1, 1 } };
    long bad = 0;
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const int16_t* c = delta + 3 * (((size_t)z * h + y) * w + x);
                const int occ = mask_bit(mask, mask_stride(w), (size_t)z * h + y, x);
                if (abs(c[0]) < 2 && abs(c[1]) < 2 && abs(c[2]) < 2) {
                    bad += !expect3(mask, w, h, d, x + c[0], y + c[1], z + c[2], !occ);
                    if (c[0] && c[1] && c[2])
                        for (int k = 0; k < 12; k++)
                            bad += !expect3(mask, w, h, d, x + face[k][0], y + face[k][1], z + face[k][2], occ);
                    if (c[0] && c[1]) {
                        bad += !expect3(mask, w, h, d, x - 1, y, z, occ) + !expect3(mask, w, h, d, x + 1, y, z, occ);
                        bad += !expect3(mask, w, h, d, x, y - 1, z, occ) + !expect3(mask, w, h, d, x, y + 1, z, occ);
                    }
                    if (c[1] && c[2]) {
                        bad += !expect3(mask, w, h, d, x, y - 1, z, occ) + !expect3(mask, w, h, d, x, y + 1, z, occ);
                        bad += !expect3(mask, w, h, d, x, y, z - 1, occ) + !expect3(mask, w, h, d, x, y, z + 1, occ);
                    }
                    if (c[2] && c[0]) {
                        bad += !expect3(mask, w, h, d, x - 1, y, z, occ) + !expect3(mask, w, h, d, x + 1, y, z, occ);
                        bad += !expect3(mask, w, h, d, x, y, z - 1, occ) + !expect3(mask, w, h, d, x, y, z + 1, occ);
                    }
                } else {
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) bad += !expect3(mask, w, h, d, x + dx, y + dy, z + dz, occ);
                }
            }
    return bad;
}

/* --------------------------------------------- continuous seeds (ext.) */

/* Nearest seed to each cell centre (x, y), in float64; ties go to the lowest
 * seed index.  seeds holds n (x, y) float pairs.  Rows are split over
 * `threads` OpenMP threads (each cell's search is independent). */
void oc_seed_nearest_2d_mt(int w, int h, int n, const float* seeds, int32_t* idx, double* dist,
                           int threads)
{
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(dynamic, 4)
#endif
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double best = INFINITY;
            int bi = -1;
            for (int k = 0; k < n; k++) {
                double dx = (double)x - (double)seeds[2 * k], dy = (double)y - (double)seeds[2 * k + 1];
                double e = dx * dx + dy * dy;
                if (e < best) { best = e; bi = k; }
            }
            idx[(size_t)y * w + x] = bi;
            dist[(size_t)y * w + x] = sqrt(best);
        }
    (void)threads;
}

void oc_seed_nearest_2d(int w, int h, int n, const float* seeds, int32_t* idx, double* dist)
{
    oc_seed_nearest_2d_mt(w, h, n, seeds, idx, dist, 1);
}

/* d(sum w*d)/dp_k = sum over cells of seed k of -w (c - p_k) / d; zero where d == 0 */
void oc_seed_grad_2d(int w, int h, int n, const float* seeds, const int32_t* idx,
                     const float* weight, double* grad)
{
    for (int k = 0; k < 2 * n; k++) grad[k] = 0.0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            int k = idx[i];
            if (k < 0) continue;
            double dx = (double)x - (double)seeds[2 * k], dy = (double)y - (double)seeds[2 * k + 1];
            double dd = sqrt(dx * dx + dy * dy);
            if (dd == 0.0) continue;
            grad[2 * k] -= (double)weight[i] * dx / dd;
            grad[2 * k + 1] -= (double)weight[i] * dy / dd;
        }
}

/* ------------------------------------------------ batched glyphs (ext.) */

/* Signed 8-bit SDF of each of n byte masks, the tool's border path run mask
 * by mask: BitMaskFromImageFile's threshold (DistanceOcclusionTool.cpp:646-665),
 * InitBorderDeltasFromBitMask + JumpFlood(cw = 2)
 * (DistanceOcclusionTool.cpp:1333-1365) and EncodeU8Signed of
 * Dist / (2 maxLen) with maxLen = 0.25 max(w, h)
 * (DistanceOcclusionTool.cpp:1242-1271, 1973-1974).  Mask i (w[i] x h[i]
 * bytes) starts at offset[i] of bytes and of out.  Masks are split over
 * `threads` OpenMP threads. */
void oc_glyph_sdf_batch(int n, const int32_t* w, const int32_t* h, const int64_t* offset,
                        const uint8_t* bytes, int threshold, int less_than, uint8_t* out,
                        int threads)
{
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; i++) {
        int gw = w[i], gh = h[i];
        size_t cells = (size_t)gw * gh;
        uint32_t* mask = (uint32_t*)calloc((size_t)mask_stride(gw) * gh + 1, sizeof(uint32_t));
        int16_t* delta = (int16_t*)malloc(cells * 2 * sizeof(int16_t) + 4);
        oc_mask_from_bytes(gw, gh, bytes + offset[i], threshold, less_than, mask);
        oc_border_2d(gw, gh, mask, delta);
        oc_jfa_2d(gw, gh, delta, 2);
        oc_sdf_u8_2d(gw, gh, mask, delta, 0.5f * (float)(gw > gh ? gw : gh), out + offset[i]);
        free(delta);
        free(mask);
    }
    (void)threads;
}

/* ------------------------------------------- raster-order methods (off the
 * data-parallel path; the tool's -m 2 and -m 4 methods and its shadows) */

/* Danielsson 2D, DistanceOcclusionLib.cpp:702-778: a down sweep (each row from
 * its second takes the row above with (0, -cw), then a left-to-right scan
 * with (-cw, 0) and a right-to-left scan with (+cw, 0)), then an up sweep
 * (from the row below with (0, +cw), then the same two scans). */
static void danielsson_rows(int w, int h, int16_t* delta, int cw)
{
    for (int j = 1; j < h; j++) {
        const int16_t* row0 = delta + 2 * (size_t)(j - 1) * w;
        int16_t* row1 = delta + 2 * (size_t)j * w;
        for (int i = 0; i < w; i++) closer2(row1 + 2 * i, row0[2 * i], row0[2 * i + 1] - cw);
        for (int i = 1; i < w; i++) closer2(row1 + 2 * i, row1[2 * (i - 1)] - cw, row1[2 * (i - 1) + 1]);
        for (int i = w - 2; i >= 0; i--) closer2(row1 + 2 * i, row1[2 * (i + 1)] + cw, row1[2 * (i + 1) + 1]);
    }
    for (int j = h - 1; j > 0; j--) {
        const int16_t* row0 = delta + 2 * (size_t)j * w;
        int16_t* row1 = delta + 2 * (size_t)(j - 1) * w;
        for (int i = 0; i < w; i++) closer2(row1 + 2 * i, row0[2 * i], row0[2 * i + 1] + cw);
        for (int i = 1; i < w; i++) closer2(row1 + 2 * i, row1[2 * (i - 1)] - cw, row1[2 * (i - 1) + 1]);
        for (int i = w - 2; i >= 0; i--) closer2(row1 + 2 * i, row1[2 * (i + 1)] + cw, row1[2 * (i + 1) + 1]);
    }
}

void oc_danielsson_2d(int w, int h, int16_t* delta, int cw) { danielsson_rows(w, h, delta, cw); }

/* DanielssonSlice, DistanceOcclusionLib.cpp:2264-2346: the 2D sweeps of one
 * slice, z carried along */
static void danielsson_slice(int w, int h, int16_t* s, int cw)
{
    for (int j = 1; j < h; j++) {
        const int16_t* row0 = s + 3 * (size_t)(j - 1) * w;
        int16_t* row1 = s + 3 * (size_t)j * w;
        for (int i = 0; i < w; i++) closer3(row1 + 3 * i, row0[3 * i], row0[3 * i + 1] - cw, row0[3 * i + 2]);
        for (int i = 1; i < w; i++)
            closer3(row1 + 3 * i, row1[3 * (i - 1)] - cw, row1[3 * (i - 1) + 1], row1[3 * (i - 1) + 2]);
        for (int i = w - 2; i >= 0; i--)
            closer3(row1 + 3 * i, row1[3 * (i + 1)] + cw, row1[3 * (i + 1) + 1], row1[3 * (i + 1) + 2]);
    }
    for (int j = h - 1; j > 0; j--) {
        const int16_t* row0 = s + 3 * (size_t)j * w;
        int16_t* row1 = s + 3 * (size_t)(j - 1) * w;
        for (int i = 0; i < w; i++) closer3(row1 + 3 * i, row0[3 * i], row0[3 * i + 1] + cw, row0[3 * i + 2]);
        for (int i = 1; i < w; i++)
            closer3(row1 + 3 * i, row1[3 * (i - 1)] - cw, row1[3 * (i - 1) + 1], row1[3 * (i - 1) + 2]);
        for (int i = w - 2; i >= 0; i--)
            closer3(row1 + 3 * i, row1[3 * (i + 1)] + cw, row1[3 * (i + 1) + 1], row1[3 * (i + 1) + 2]);
    }
}

/* Danielsson 3D, DistanceOcclusionLib.cpp:2348-2394: each slice from the
 * second takes the slice before with (0, 0, -cw) and is swept as a 2D slice;
 * then the same from the last slice back with (0, 0, +cw).  The first slice
 * of each direction is not swept on its own. */
void oc_danielsson_3d(int w, int h, int d, int16_t* delta, int cw)
{
    const size_t n = (size_t)w * h;
    for (int k = 1; k < d; k++) {
        const int16_t* s0 = delta + 3 * (k - 1) * n;
        int16_t* s1 = delta + 3 * k * n;
        for (size_t i = 0; i < n; i++) closer3(s1 + 3 * i, s0[3 * i], s0[3 * i + 1], s0[3 * i + 2] - cw);
        danielsson_slice(w, h, s1, cw);
    }
    for (int k = d - 2; k >= 0; k--) {
        const int16_t* s0 = delta + 3 * (k + 1) * n;
        int16_t* s1 = delta + 3 * k * n;
        for (size_t i = 0; i < n; i++) closer3(s1 + 3 * i, s0[3 * i], s0[3 * i + 1], s0[3 * i + 2] + cw);
        danielsson_slice(w, h, s1, cw);
    }
}

/* Chamfer (Borgefors 3-4), DistanceOcclusionLib.cpp:912-1006, on int64
 * copies of the values: d(x) = min(d(x), neighbour + 3 or + 4), a forward
 * raster pass (first row: left neighbour; other rows: the three above, then
 * the left one) and a reverse pass (last row: right neighbour; other rows:
 * the three below, then the right one).  Rows of one cell have no diagonal
 * neighbours.  Returns 3, the divisor of the result.  The sums are taken
 * without overflow (the reference's int32 form overflows on values within
 * 4 of INT32_MAX). */
static int chamfer64(int w, int h, int64_t* v)
{
    const int64_t d0 = 3, d1 = 4;
#define CHM(A, B) do { if ((A) > (B)) (A) = (B); } while (0)
    for (int x = 1; x < w; x++) CHM(v[x], v[x - 1] + d0);
    for (int y = 1; y < h; y++) {
        int64_t* r0 = v + (size_t)(y - 1) * w;
        int64_t* r1 = v + (size_t)y * w;
        CHM(r1[0], r0[0] + d0);
        if (w > 1) CHM(r1[0], r0[1] + d1);
        for (int x = 1; x < w - 1; x++) {
            CHM(r1[x], r0[x - 1] + d1);
            CHM(r1[x], r0[x] + d0);
            CHM(r1[x], r0[x + 1] + d1);
            CHM(r1[x], r1[x - 1] + d0);
        }
        if (w > 1) {
            CHM(r1[w - 1], r0[w - 2] + d1);
            CHM(r1[w - 1], r0[w - 1] + d0);
            CHM(r1[w - 1], r1[w - 2] + d0);
        }
    }
    {
        int64_t* r0 = v + (size_t)(h - 1) * w;
        for (int x = w - 2; x >= 0; x--) CHM(r0[x], r0[x + 1] + d0);
    }
    for (int y = h - 2; y >= 0; y--) {
        int64_t* r0 = v + (size_t)(y + 1) * w;
        int64_t* r1 = v + (size_t)y * w;
        CHM(r1[w - 1], r0[w - 1] + d0);
        if (w > 1) CHM(r1[w - 1], r0[w - 2] + d1);
        for (int x = w - 2; x > 0; x--) {
            CHM(r1[x], r0[x + 1] + d1);
            CHM(r1[x], r0[x] + d0);
            CHM(r1[x], r0[x - 1] + d1);
            CHM(r1[x], r1[x + 1] + d0);
        }
        if (w > 1) {
            CHM(r1[0], r0[1] + d1);
            CHM(r1[0], r0[0] + d0);
            CHM(r1[0], r1[1] + d0);
        }
    }
#undef CHM
    return (int)d0;
}

int oc_chamfer_i16(int w, int h, int16_t* dist)
{
    size_t n = (size_t)w * h;
    int64_t* v = (int64_t*)malloc(n * sizeof(int64_t) + 8);
    for (size_t i = 0; i < n; i++) v[i] = dist[i];
    int r = chamfer64(w, h, v);
    for (size_t i = 0; i < n; i++) dist[i] = (int16_t)v[i];
    free(v);
    return r;
}

int oc_chamfer_i32(int w, int h, int32_t* dist)
{
    size_t n = (size_t)w * h;
    int64_t* v = (int64_t*)malloc(n * sizeof(int64_t) + 8);
    for (size_t i = 0; i < n; i++) v[i] = dist[i];
    int r = chamfer64(w, h, v);
    for (size_t i = 0; i < n; i++) dist[i] = (int32_t)v[i];
    free(v);
    return r;
}

/* InitOccFromBitMask, DistanceOcclusionLib.cpp:2864-2867: scale where a bit is
 * set, 0 elsewhere */
void oc_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale)
{
    int ws = mask_stride(w);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            occ[(size_t)y * w + x] = ((mask[(size_t)y * ws + (x >> 5)] >> (x & 31)) & 1) ? scale : 0.0f;
}

/* ShadowSweep, DistanceOcclusionLib.cpp:2773-2876.  Light along (gx, gy): the
 * sweep runs along y (rows) when |gx| <= |gy|, else along x (columns), from
 * the side the light comes from (s = 0 when the component is positive).  The
 * first line copies the occlusion; every later cell takes
 * m = prev[i - di] g1 + prev[i] g0 (the first cell of a line m = prev[first] g0)
 * and stores m + (1 - m) occ, g1 = |minor / major| and g0 = 1 - g1, in the
 * reference's float operation order. */
static void shadow_lines(float g, int sx, int sy, int w, int h, int rows, const float* occ, float* sh)
{
    const float g1 = g, g0 = 1.0f - g1;
    const int dx = 1 - 2 * sx, dy = 1 - 2 * sy;
    /* a line is a row (rows) or a column; ib, ie, di walk along it; jb, dj
     * step from line to line */
    int ib, ie, di, jb, dj, lines;
    if (rows) {
        di = dx; ib = sx * (w - 1); ie = ib + di * w;
        dj = dy * w; jb = sy * (h - 1) * w; lines = h;
    } else {
        const int wh = w * h;
        ib = (wh - w) * sy; ie = ib + wh * dy; di = dy * w;
        jb = (w - 1) * sx; dj = dx; lines = w;
    }
    const float* i0 = occ + jb;
    float* o0 = sh + jb;
    for (int i = ib; i != ie; i += di) o0[i] = i0[i];
    for (int j = 1; j < lines; j++) {
        const float* i1 = i0 + dj;
        float* o1 = o0 + dj;
        float m = o0[ib] * g0;
        o1[ib] = m + (1.0f - m) * i1[ib];
        for (int i = ib + di; i != ie; i += di) {
            float mm = o0[i - di] * g1 + o0[i] * g0;
            o1[i] = mm + (1.0f - mm) * i1[i];
        }
        i0 = i1;
        o0 = o1;
    }
}

void oc_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow)
{
    const int sx = gx > 0.0f ? 0 : 1;
    const int sy = gy > 0.0f ? 0 : 1;
    if (fabsf(gx) <= fabsf(gy))
        shadow_lines(fabsf(gx / gy), sx, sy, w, h, 1, occ, shadow);
    else
        shadow_lines(fabsf(gy / gx), sx, sy, w, h, 0, occ, shadow);
}
