// dol_sweep.hip -- FastSweep (DistanceOcclusionLib.cpp:633-697, 2186-2262) for
// gfx950 as a wavefront over tiles.
//
// In a sweep's own frame -- u along x in the sweep's x direction, v along y,
// t along z, each from 0 -- the reference updates cell (u, v[, t]) from
// (u, v - 1[, t]), (u - 1, v[, t]), (2D) (u - 1, v - 1) and (3D) (u, v, t - 1),
// each already final when it is read: rows and slices go in order, a row is
// first taken from the row (and slice) before it, cell by cell, then scanned
// in x.  Cells with equal u + v (+ t) are therefore independent, and every
// cell sees the same candidates, in the same order, as in the reference.
//
// One wavefront owns a tile of 64 columns (u) x 64 rows (v) of one slice, one
// lane a row.  Lane l handles column a = s - l of its row at step s, so at
// each step the lane before it has just finished (a, v - 1) and finished
// (a - 1, v - 1) one step earlier: both come across with a one-lane DPP
// shift, and (a - 1, v) is the lane's own previous result.  Lane 0 reads the
// row before the tile, and every lane the column before it, from memory,
// where the tiles before it left them final.  The tiles of one diagonal
// (2D: U + V; 3D: U + V + t) are independent, so each diagonal is one launch,
// in order on the stream; the field is updated in place (a launch reads only
// tiles that earlier launches finished, and its own, which nothing else in it
// touches).
#include "dol_common.hpp"

namespace dolhip {

namespace {

constexpr int kST = 64;               // tile columns and rows (one lane a row)
constexpr int kSStride = kST + 2;     // LDS row stride: lane l's column s - l
                                      // sits at word 65 l + s, a different bank
                                      // for every lane

typedef short sw_i16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short sw_u16x2 __attribute__((ext_vector_type(2)));

// squared length of the int16 pair plus zz, wrapping in 32 bits as the
// reference's Closer does (DistanceOcclusionLib.cpp:624-630, 2177-2183)
__device__ __forceinline__ int sw_len2(uint32_t xy, int zz)
{
    const sw_i16x2 v = __builtin_bit_cast(sw_i16x2, xy);
    return __builtin_amdgcn_sdot2(v, v, zz, false);
}

// (x, y) + (ox, oy), each half wrapping as the reference's int16 store does
__device__ __forceinline__ uint32_t sw_add(uint32_t xy, uint32_t oxy)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(sw_u16x2, xy) + __builtin_bit_cast(sw_u16x2, oxy));
}

struct Frame {
    int w, x0, dx, y0, dy;
    __device__ size_t row(int v) const { return (size_t)(y0 + v * dy) * w; }
    __device__ int col(int u) const { return x0 + u * dx; }
};

__device__ __forceinline__ Frame frame(int w, int h, int sweep)
{
    Frame f;
    f.w = w;
    f.dx = 1 - 2 * (sweep & 1);
    f.dy = 1 - 2 * ((sweep >> 1) & 1);
    f.x0 = f.dx > 0 ? 0 : w - 1;
    f.y0 = f.dy > 0 ? 0 : h - 1;
    return f;
}

}  // namespace

// lane l's row, rotated: register k holds column (k - l) mod 64, so at step s
// every lane needs register s mod 64 -- the same for all lanes, which keeps
// the row in registers through the unrolled steps
__device__ __forceinline__ int rot(int k, int lane) { return (k - lane) & (kST - 1); }

// the value of the lane before (lane 0 gets 0)
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}

// No candidate: a length no int16 pair or triple has (2^31 - 1 is neither a
// sum of two squares, being 3 mod 4, nor of three, being 7 mod 8).
constexpr int kNoLen = 0x7fffffff;

// Closer() applied to the candidates in order keeps the first of least
// length: pick that one of (x0, x1, x2, x3) given their lengths
template <class T>
__device__ __forceinline__ T first_least(T x0, int l0, T x1, int l1, T x2, int l2, T x3, int l3)
{
    const int m = min(min(l0, l1), min(l2, l3));
    T r = l2 == m ? x2 : x3;
    r = l1 == m ? x1 : r;
    return l0 == m ? x0 : r;
}

// Lane l of a tile holds row v0 - 1 + l: lane 0 the row before the tile,
// final already (its values are only passed on), lanes 1..63 the tile's 63
// rows.
constexpr int kRows = kST - 1;

// 2D: tiles (U, V) with U + V = diag; block b is V = max(0, diag - tu + 1) + b
__global__ void __launch_bounds__(kST) k_fast_sweep_2d(int w, int h, int sweep, int cw, int diag,
                                                       int tu, uint32_t* __restrict__ delta)
{
    __shared__ uint32_t tile[kST * kSStride];
    const Frame f = frame(w, h, sweep);
    const uint32_t ox = (uint16_t)(-cw * f.dx), oy = (uint32_t)(uint16_t)(-cw * f.dy) << 16;
    const int V = max(0, diag - (tu - 1)) + (int)blockIdx.x, U = diag - V;
    const int u0 = U * kST, v0 = V * kRows;
    const int nu = min(kST, w - u0);
    const int lane = threadIdx.x;
    const int vr = v0 - 1 + lane;                       // the lane's row
    const int rlo = max(v0 - 1, 0), rhi = min(v0 + kRows - 1, h - 1);
    const int cu = f.col(u0 + min(lane, nu - 1));
    {
        // all rows in flight at once (rows outside the field repeat the nearest)
        uint32_t v[kST];
#pragma unroll
        for (int r = 0; r < kST; r++) v[r] = delta[f.row(min(max(v0 - 1 + r, rlo), rhi)) + cu];
#pragma unroll
        for (int r = 0; r < kST; r++) tile[r * kSStride + lane] = v[r];
    }
    const bool row_ok = vr >= 0 && vr < h, update = row_ok && lane >= 1 && vr >= 1;
    // the lane's last two results; before its first column, the column
    // before the tile
    uint32_t mine = (u0 > 0 && row_ok) ? delta[f.row(vr) + f.col(u0 - 1)] : 0u;
    uint32_t mine_prev = 0u;
    __syncthreads();
    uint32_t R[kST];
#pragma unroll
    for (int k = 0; k < kST; k++) R[k] = tile[lane * kSStride + rot(k, lane)];
#pragma unroll
    for (int s = 0; s < 2 * kST - 1; s++) {
        const int a = s - lane;
        // (a, v - 1) and (a - 1, v - 1), finished by the lane before
        const uint32_t up = from_prev_lane(mine), dg = from_prev_lane(mine_prev);
        const uint32_t cur = R[s & (kST - 1)];
        const uint32_t c0 = sw_add(up, oy), c1 = sw_add(mine, ox), c2 = sw_add(dg, ox | oy);
        const bool left = u0 + a >= 1;
        const uint32_t best = first_least(cur, sw_len2(cur, 0), c0, sw_len2(c0, 0), c1,
                                          left ? sw_len2(c1, 0) : kNoLen, c2, left ? sw_len2(c2, 0) : kNoLen);
        // selects rather than a branch: every lane runs the same instructions
        const bool act = row_ok && a >= 0 && a < nu;
        const uint32_t nxt = act && update ? best : R[s & (kST - 1)];
        R[s & (kST - 1)] = nxt;
        mine_prev = act ? mine : mine_prev;
        mine = act ? nxt : mine;
    }
#pragma unroll
    for (int k = 0; k < kST; k++) tile[lane * kSStride + rot(k, lane)] = R[k];
    __syncthreads();
    {
        uint32_t v[kST];
#pragma unroll
        for (int r = 1; r < kST; r++) v[r] = tile[r * kSStride + lane];
#pragma unroll
        for (int r = 1; r < kST; r++)
            if (v0 - 1 + r <= rhi && lane < nu) delta[f.row(v0 - 1 + r) + cu] = v[r];
    }
}

// 3D: the slices add d / kZC3 to the number of diagonals (u + v + t), so
// short tiles pay: a wavefront holds four tiles, one per 16-lane DPP row,
// each kTW3 = 8 columns by kRows3 = 15 rows (lane 0 of the row carrying the
// row before the tile, as in 2D), which is kTW3 + 15 steps a tile slice.  A lane keeps
// its row and the same row of the slice before in registers, rotated modulo
// kTW3 by its place in the group, so the tile needs no LDS.
constexpr int kTW3 = 8;
constexpr int kGroup3 = 16;                 // lanes per tile (one DPP row)
constexpr int kRows3 = kGroup3 - 1;
constexpr int kTiles3 = kST / kGroup3;      // tiles per wavefront, stacked in v

// the value of the lane before in the lane's DPP row of 16 (row_shr:1); the
// row's first lane gets 0
__device__ __forceinline__ uint32_t from_prev_lane16(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
}

// A tile takes kZC3 consecutive slices in one launch, one after the other:
// everything it reads (the slice before its first, the row before and the
// column before in each of its slices) is final when the launch starts, and
// each slice's result is the next slice's "slice before", kept in registers.
// That halves the launches, which the slices' chain otherwise dominates.
constexpr int kZC3 = 2;

// 3D: the tiles (U, V, T) with U + V + T = diag, tile T holding the slices
// 1 + kZC3 T .. (the sweep's first slice, 0, is never updated); block b holds
// U = b % tu and V = kTiles3 * (b / tu) + g in lane group g.
__global__ void __launch_bounds__(kST) k_fast_sweep_3d(int w, int h, int d, int sweep, int cw, int diag,
                                                       int tu, int16_t* __restrict__ delta)
{
    const int lane = threadIdx.x, j = lane % kGroup3;
    const int U = (int)blockIdx.x % tu, V = kTiles3 * ((int)blockIdx.x / tu) + lane / kGroup3;
    const int T = diag - U - V;
    const int t0 = 1 + T * kZC3;                          // the tile's first slice
    const Frame f = frame(w, h, sweep);
    const int dz = 1 - 2 * ((sweep >> 2) & 1);
    const bool tile_ok = T >= 0 && t0 < d && V * kRows3 < h;
    const size_t slice = (size_t)w * h;
    // slice t of the sweep; a group without a tile, and a slice past the
    // volume's last, read slice 0 (and write nothing)
    auto plane = [&](int t) {
        const int z = tile_ok && t < d ? (dz > 0 ? 0 : d - 1) + t * dz : 0;
        return delta + 3 * (slice * (size_t)z);
    };
    const uint32_t ox = (uint16_t)(-cw * f.dx), oy = (uint32_t)(uint16_t)(-cw * f.dy) << 16;
    const int oz = -cw * dz;
    const int u0 = U * kTW3, v0 = V * kRows3;
    const int nu = min(kTW3, w - u0);
    const int vr = v0 - 1 + j;                            // the lane's row
    const bool row_ok = tile_ok && vr >= 0 && vr < h;
    const bool update = row_ok && j >= 1 && vr >= 1;
    const size_t rbase = f.row(min(max(vr, 0), h - 1));
    auto at = [&](int c) { return 3 * (rbase + f.col(u0 + min(c, nu - 1))); };
    auto pack = [](const int16_t* c) { return (uint32_t)(uint16_t)c[0] | ((uint32_t)(uint16_t)c[1] << 16); };
    // register k holds column (k - j) mod kTW3, so step s reads register s mod kTW3
    uint32_t R[kZC3][kTW3], B[kTW3];
    int RZ[kZC3][kTW3], BZ[kTW3];
    uint32_t first[kZC3];   // per slice, the column before the tile
    int first_z[kZC3];
    {
        const int16_t* back = plane(t0 - 1);
#pragma unroll
        for (int k = 0; k < kTW3; k++) {
            const size_t o = at((k - j) & (kTW3 - 1));
            B[k] = pack(back + o);
            BZ[k] = wrap16(back[o + 2] + oz);
        }
    }
#pragma unroll
    for (int zi = 0; zi < kZC3; zi++) {
        const int16_t* cur = plane(t0 + zi);
#pragma unroll
        for (int k = 0; k < kTW3; k++) {
            const size_t o = at((k - j) & (kTW3 - 1));
            R[zi][k] = pack(cur + o);
            RZ[zi][k] = cur[o + 2];
        }
        first[zi] = 0u;
        first_z[zi] = 0;
        if (u0 > 0 && row_ok) {
            const int16_t* c = cur + 3 * (rbase + f.col(u0 - 1));
            first[zi] = pack(c);
            first_z[zi] = c[2];
        }
    }
#pragma unroll
    for (int zi = 0; zi < kZC3; zi++) {
        const bool live = row_ok && t0 + zi < d;
        const bool upd = update && t0 + zi < d;
        uint32_t mine = first[zi];   // the lane's last result; first, the column before the tile
        int mine_z = first_z[zi];
#pragma unroll
        for (int s = 0; s < kTW3 + kGroup3 - 1; s++) {
            const int a = s - j;
            const int k = s & (kTW3 - 1);
            // (a, v - 1), finished by the lane before
            const uint32_t up = from_prev_lane16(mine);
            const int up_z = (int)(int16_t)from_prev_lane16((uint32_t)mine_z);
            const uint32_t cxy = R[zi][k];
            const int cz = RZ[zi][k];
            const uint32_t uxy = sw_add(up, oy);     // the row before, (0, cy, 0)
            const uint32_t lxy = sw_add(mine, ox);   // the cell before, (cx, 0, 0)
            // candidates in the reference's order: the slice before (0, 0, cz),
            // the row before, the cell before
            const int l0 = sw_len2(cxy, cz * cz), l1 = sw_len2(B[k], BZ[k] * BZ[k]), l2 = sw_len2(uxy, up_z * up_z);
            const int l3 = u0 + a >= 1 ? sw_len2(lxy, mine_z * mine_z) : kNoLen;
            const int m = min(min(l0, l1), min(l2, l3));
            const int pick = l0 == m ? 0 : (l1 == m ? 1 : (l2 == m ? 2 : 3));
            const uint32_t nxy = pick == 0 ? cxy : (pick == 1 ? B[k] : (pick == 2 ? uxy : lxy));
            const int nz = pick == 0 ? cz : (pick == 1 ? BZ[k] : (pick == 2 ? up_z : mine_z));
            if (live && a >= 0 && a < nu) {
                if (upd) {
                    R[zi][k] = nxy;
                    RZ[zi][k] = nz;
                }
                mine = R[zi][k];
                mine_z = RZ[zi][k];
            }
        }
        // this slice's result, offset by one slice, is the next one's slice before
#pragma unroll
        for (int k = 0; k < kTW3; k++) {
            B[k] = R[zi][k];
            BZ[k] = wrap16(RZ[zi][k] + oz);
        }
    }
    if (!update) return;
#pragma unroll
    for (int zi = 0; zi < kZC3; zi++) {
        if (t0 + zi >= d) break;
        int16_t* cur = plane(t0 + zi);
#pragma unroll
        for (int k = 0; k < kTW3; k++) {
            const int c = (k - j) & (kTW3 - 1);
            if (c >= nu) continue;
            int16_t* o = cur + at(c);
            o[0] = (int16_t)(R[zi][k] & 0xffffu);
            o[1] = (int16_t)(R[zi][k] >> 16);
            o[2] = (int16_t)RZ[zi][k];
        }
    }
}

void launch_fast_sweep_2d(int w, int h, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("fast_sweep_2d", s);
    const int tu = (w + kST - 1) / kST, tv = (h + kRows - 1) / kRows;
    for (int sweep = 0; sweep < 4; sweep++)
        for (int diag = 0; diag < tu + tv - 1; diag++) {
            const int lo = diag - (tu - 1) > 0 ? diag - (tu - 1) : 0, hi = diag < tv - 1 ? diag : tv - 1;
            hipLaunchKernelGGL(k_fast_sweep_2d, dim3((unsigned)(hi - lo + 1)), dim3(kST), 0, s, w, h,
                               sweep, cw, diag, tu, (uint32_t*)delta);
        }
}

void launch_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("fast_sweep_3d", s);
    const int tu = (w + kTW3 - 1) / kTW3, tv = (h + kRows3 - 1) / kRows3;
    const int vb = (tv + kTiles3 - 1) / kTiles3;   // wavefronts along v
    const int tt = (d - 1 + kZC3 - 1) / kZC3;      // slice groups after slice 0
    if (tt == 0) return;
    for (int sweep = 0; sweep < 8; sweep++)
        for (int diag = 0; diag < tu + tv + tt - 2; diag++)
            hipLaunchKernelGGL(k_fast_sweep_3d, dim3((unsigned)(tu * vb)), dim3(kST), 0, s, w, h, d,
                               sweep, cw, diag, tu, delta);
}

}  // namespace dolhip
