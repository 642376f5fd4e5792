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

// the value of the lane before (wave_shr:1); lane 0 gets 0
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
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

// 2D: tiles (U, V) with U + V = diag; block b is V = max(0, diag - tu + 1) + b
__global__ void __launch_bounds__(kST) k_fast_sweep_2d(int w, int h, int sweep, int cw, int diag,
                                                       int tu, uint32_t* __restrict__ delta)
{
    __shared__ uint32_t tile[kST * kSStride];
    __shared__ uint32_t top[kST + 1];   // row v0 - 1: [0] column u0 - 1, [1 + a] column u0 + a
    const Frame f = frame(w, h, sweep);
    const uint32_t ox = (uint16_t)(-cw * f.dx), oy = (uint32_t)(uint16_t)(-cw * f.dy) << 16;
    const int V = max(0, diag - (tu - 1)) + (int)blockIdx.x, U = diag - V;
    const int u0 = U * kST, v0 = V * kST;
    const int nu = min(kST, w - u0), nv = min(kST, h - v0);
    const int lane = threadIdx.x;
    const int cu = f.col(u0 + min(lane, nu - 1));
    for (int r = 0; r < nv; r++)
        if (lane < nu) tile[r * kSStride + lane] = delta[f.row(v0 + r) + cu];
    if (v0 > 0) {
        if (lane < nu) top[1 + lane] = delta[f.row(v0 - 1) + cu];
        if (lane == 0 && u0 > 0) top[0] = delta[f.row(v0 - 1) + f.col(u0 - 1)];
    }
    // the lane's last two results; before its first column, the column
    // before the tile
    uint32_t mine = (u0 > 0 && lane < nv) ? delta[f.row(v0 + lane) + f.col(u0 - 1)] : 0u;
    uint32_t mine_prev = 0u;
    __syncthreads();
    const bool row_ok = lane < nv, update = row_ok && v0 + lane >= 1;
    uint32_t* trow = tile + lane * kSStride;
    for (int s = 0; s < kST + kST - 1; s++) {
        const int a = s - lane;
        uint32_t up = from_prev_lane(mine), dg = from_prev_lane(mine_prev);
        if (!row_ok || a < 0 || a >= nu) continue;
        if (lane == 0 && v0 > 0) {
            up = top[1 + a];
            dg = top[a];
        }
        uint32_t cur = trow[a];
        if (update) {
            int lc = sw_len2(cur, 0);
            const uint32_t c0 = sw_add(up, oy);
            const int l0 = sw_len2(c0, 0);
            if (lc > l0) { cur = c0; lc = l0; }
            if (u0 + a >= 1) {
                const uint32_t c1 = sw_add(mine, ox);
                const int l1 = sw_len2(c1, 0);
                if (lc > l1) { cur = c1; lc = l1; }
                const uint32_t c2 = sw_add(dg, ox | oy);
                const int l2 = sw_len2(c2, 0);
                if (lc > l2) cur = c2;
            }
            trow[a] = cur;
        }
        mine_prev = mine;
        mine = cur;
    }
    __syncthreads();
    for (int r = 0; r < nv; r++)
        if (lane < nu) delta[f.row(v0 + r) + cu] = tile[r * kSStride + lane];
}

// 3D: tiles (U, V) of slice t = diag - U - V; block b is (U, V) = (b % tu, b / tu).
// The sweep's first slice (t = 0) is never updated.
__global__ void __launch_bounds__(kST) k_fast_sweep_3d(int w, int h, int d, int sweep, int cw, int diag,
                                                       int tu, int16_t* __restrict__ delta)
{
    __shared__ uint32_t txy[kST * kSStride];
    __shared__ int16_t tz[kST * kSStride];
    __shared__ uint32_t bxy[kST * kSStride];   // the same cells of the slice before
    __shared__ int16_t bz[kST * kSStride];
    __shared__ uint32_t top_xy[kST];            // row v0 - 1
    __shared__ int16_t top_z[kST];
    const int U = (int)blockIdx.x % tu, V = (int)blockIdx.x / tu;
    const int t = diag - U - V;
    if (t < 1 || t >= d) return;   // the sweep's first slice keeps its values
    const Frame f = frame(w, h, sweep);
    const int dz = 1 - 2 * ((sweep >> 2) & 1);
    const int z = (dz > 0 ? 0 : d - 1) + t * dz;
    const size_t slice = (size_t)w * h;
    const int16_t* back = delta + 3 * (slice * (size_t)(z - dz));
    int16_t* cur_slice = delta + 3 * (slice * (size_t)z);
    const uint32_t ox = (uint16_t)(-cw * f.dx), oy = (uint32_t)(uint16_t)(-cw * f.dy) << 16;
    const int oz = -cw * dz;
    const int u0 = U * kST, v0 = V * kST;
    const int nu = min(kST, w - u0), nv = min(kST, h - v0);
    const int lane = threadIdx.x;
    const int cu = f.col(u0 + min(lane, nu - 1));
    auto load = [](const int16_t* c, uint32_t& xy, int16_t& zz) {
        xy = (uint32_t)(uint16_t)c[0] | ((uint32_t)(uint16_t)c[1] << 16);
        zz = c[2];
    };
    for (int r = 0; r < nv; r++) {
        if (lane >= nu) continue;
        const size_t i = 3 * (f.row(v0 + r) + cu);
        load(cur_slice + i, txy[r * kSStride + lane], tz[r * kSStride + lane]);
        load(back + i, bxy[r * kSStride + lane], bz[r * kSStride + lane]);
    }
    if (v0 > 0 && lane < nu) load(cur_slice + 3 * (f.row(v0 - 1) + cu), top_xy[lane], top_z[lane]);
    uint32_t mine = 0u;
    int16_t mine_z = 0;
    if (u0 > 0 && lane < nv) load(cur_slice + 3 * (f.row(v0 + lane) + f.col(u0 - 1)), mine, mine_z);
    __syncthreads();
    const bool row_ok = lane < nv, update = row_ok && v0 + lane >= 1;
    uint32_t* trow = txy + lane * kSStride;
    int16_t* trow_z = tz + lane * kSStride;
    const uint32_t* brow = bxy + lane * kSStride;
    const int16_t* brow_z = bz + lane * kSStride;
    for (int s = 0; s < kST + kST - 1; s++) {
        const int a = s - lane;
        uint32_t up = from_prev_lane(mine);
        int up_z = (int)(int16_t)from_prev_lane((uint32_t)(uint16_t)mine_z);
        if (!row_ok || a < 0 || a >= nu) continue;
        if (lane == 0 && v0 > 0) {
            up = top_xy[a];
            up_z = top_z[a];
        }
        uint32_t cxy = trow[a];
        int cz = trow_z[a];
        if (update) {
            int lc = sw_len2(cxy, cz * cz);
            {   // the slice before, (0, 0, cz)
                const uint32_t c = brow[a];
                const int c_z = wrap16(brow_z[a] + oz);
                const int l = sw_len2(c, c_z * c_z);
                if (lc > l) { cxy = c; cz = c_z; lc = l; }
            }
            {   // the row before, (0, cy, 0)
                const uint32_t c = sw_add(up, oy);
                const int l = sw_len2(c, up_z * up_z);
                if (lc > l) { cxy = c; cz = up_z; lc = l; }
            }
            if (u0 + a >= 1) {   // the cell before, (cx, 0, 0)
                const uint32_t c = sw_add(mine, ox);
                const int l = sw_len2(c, (int)mine_z * mine_z);
                if (lc > l) { cxy = c; cz = mine_z; }
            }
            trow[a] = cxy;
            trow_z[a] = (int16_t)cz;
        }
        mine = cxy;
        mine_z = (int16_t)cz;
    }
    __syncthreads();
    for (int r = 0; r < nv; r++) {
        if (lane >= nu) continue;
        int16_t* c = cur_slice + 3 * (f.row(v0 + r) + cu);
        const uint32_t xy = txy[r * kSStride + lane];
        c[0] = (int16_t)(xy & 0xffffu);
        c[1] = (int16_t)(xy >> 16);
        c[2] = tz[r * kSStride + lane];
    }
}

void launch_fast_sweep_2d(int w, int h, int16_t* delta, int cw, hipStream_t s)
{
    ProfileScope p("fast_sweep_2d", s);
    const int tu = (w + kST - 1) / kST, tv = (h + kST - 1) / kST;
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
    const int tu = (w + kST - 1) / kST, tv = (h + kST - 1) / kST;
    for (int sweep = 0; sweep < 8; sweep++)
        for (int diag = 0; diag < tu + tv + d - 2; diag++)
            hipLaunchKernelGGL(k_fast_sweep_3d, dim3((unsigned)(tu * tv)), dim3(kST), 0, s, w, h, d,
                               sweep, cw, diag, tu, delta);
}

}  // namespace dolhip
