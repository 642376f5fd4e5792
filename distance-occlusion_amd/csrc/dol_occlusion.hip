// dol_occlusion.hip -- direction-field initialisation and occlusion sweeps
// (DistanceOcclusionLib.cpp:1114-1209 2D, 2507-2769 3D) for gfx950.
//
// Each of the 4 (2D) or 8 (3D) planes of the direction field is swept from
// its own corner, and the planes do not interact, so every launch works on
// all of them.  In a plane's sweep frame (u, v, t from its corner) cell
// (u, v, t) combines (u - 1, v, t), (u, v - 1, t) and (u, v, t - 1), each
// already updated, with 0 standing in for a neighbour outside the field;
// only the corner cell is left as it is.  Cells with equal u + v + t are
// independent, so the sweep is a wavefront over tiles, as FastSweep's in
// dol_sweep.hip: in 3D one 16-lane DPP row of a wavefront holds a 16-column
// x 15-row tile of one slice (8 and 32 columns measured slower), in 2D a
// wavefront holds a 64 x 63 tile; lane l
// is on row v0 - 1 + l (lane 0 carrying the row before the tile), lanes
// skewed one column each; a launch is one diagonal U + V (+ t) of tiles.  A
// lane keeps its row (and in 3D the same row of the slice before) in
// registers, rotated by its place in the tile so that every step reads the
// same register.
//
// The float expressions are the reference's, term for term and in its
// order; the library is built without FMA contraction, so each result is
// the reference's to the bit.
#include "dol_cells.hpp"

namespace dolhip {

// ------------------------------------------------------------ init

// one thread per cell: every plane gets 1 where the bit is set, else 0
// (InitDirWFromBitMask, DistanceOcclusionLib.cpp:1114-1146, 2507-2538)
__global__ void __launch_bounds__(256) k_init_dirw(int w, size_t cells, int planes,
                                                   const uint32_t* __restrict__ mask,
                                                   float* __restrict__ dirw)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= cells) return;
    const size_t row = i / (size_t)w;
    const int x = (int)(i - row * (size_t)w);
    const float v = ((mask[row * mask_stride(w) + (x >> 5)] >> (x & 31)) & 1u) ? 1.0f : 0.0f;
    for (int p = 0; p < planes; p++) dirw[(size_t)p * cells + i] = v;
}

// planes: 4 (2D) or 8 (3D)
void launch_init_dirw(int w, int h, int d, int planes, const uint32_t* mask, float* dirw, hipStream_t s)
{
    ProfileScope p("init_dirw", s);
    const size_t cells = (size_t)w * h * d;
    hipLaunchKernelGGL(k_init_dirw, dim3(grid1d(cells, 256)), dim3(256), 0, s, w, cells, planes, mask,
                       dirw);
}

// ------------------------------------------------------------ sweeps

namespace {

constexpr int kOW = 16;                // tile columns
constexpr int kOGroup = 16;            // lanes per tile (one DPP row)
constexpr int kORows = kOGroup - 1;    // tile rows
constexpr int kOTiles = 64 / kOGroup;  // tiles per wavefront, stacked in v

// QuadrantCombine (DistanceOcclusionLib.cpp:1150-1168)
__device__ __forceinline__ float quadrant_combine(float a, float b)
{
    return 0.75f * ((a + b) - 1.0f * a * b);
}

// OctantCombine, three and six arguments (DistanceOcclusionLib.cpp:2543-2567)
__device__ __forceinline__ float octant_combine(float a, float b, float c)
{
    return 0.33333333f * (a + b + c - a * b * c);
}

__device__ __forceinline__ float octant_combine(float oa, float ob, float oc, float ia, float ib,
                                                float ic)
{
    const float a = ia + (1.0f - ia) * oa;
    const float b = ib + (1.0f - ib) * ob;
    const float c = ic + (1.0f - ic) * oc;
    return 0.3333333f * (a + b + c - a * b * c);
}

// the value of the lane before in the lane's DPP row of 16 (row_shr:1)
__device__ __forceinline__ float from_prev_lane16(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf,
                                                                 0xf, false));
}

}  // namespace

// 3D.  MODE 1: in place (OcclusionSweep, :2569-2648); MODE 2: out from in
// (:2650-2743; `in` read only).  Lane group g of block b: plane b / (vb tu),
// r = b % (vb tu), U = r % tu, V = kOTiles * (r / tu) + g, slice group
// T = diag - U - V (slices kOZ T ..).
// A tile takes kOZ consecutive slices in one launch, one after the other:
// what it reads from its neighbours is final when the launch starts, and
// each slice's result (and, MODE 2, input) is the next slice's "slice
// before", kept in registers.  That halves the launches, which the slices'
// chain otherwise dominates.
constexpr int kOZ = 2;

template <int MODE>
__global__ void __launch_bounds__(64) k_occlusion_sweep(int w, int h, int d, int diag, int tu, int vb,
                                                        const float* __restrict__ in,
                                                        float* __restrict__ dirw)
{
    const int lane = threadIdx.x, j = lane % kOGroup;
    const int plane = (int)blockIdx.x / (vb * tu), r = (int)blockIdx.x % (vb * tu);
    const int V = kOTiles * (r / tu) + lane / kOGroup;
    const int Ud = r % tu;
    const int T = diag - Ud - V;
    const int t0 = T * kOZ;                               // the tile's first slice
    const size_t cells = (size_t)w * h * d, slice = (size_t)w * h;
    const int dx = 1 - 2 * (plane & 1), dy = 1 - 2 * ((plane >> 1) & 1), dz = 1 - 2 * ((plane >> 2) & 1);
    const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1, z0 = dz > 0 ? 0 : d - 1;
    const bool tile_ok = Ud >= 0 && Ud < tu && T >= 0 && t0 < d && V * kORows < h;
    const int U = tile_ok ? Ud : 0;   // a group without a tile reads tile 0 (and writes nothing)
    // slice t of the sweep; slices outside the volume (the one before the
    // first, one past the last) and groups without a tile read slice z0
    auto zof = [&](int tt) { return tile_ok && tt >= 0 && tt < d ? z0 + tt * dz : z0; };
    float* out = dirw + (size_t)plane * cells;
    const float* inp = MODE == 2 ? in + (size_t)plane * cells : out;
    static_assert(MODE == 1 || MODE == 2, "3D modes");
    const int u0 = U * kOW, v0 = V * kORows;
    const int nu = min(kOW, w - u0);
    const int vr = v0 - 1 + j;                           // the lane's row
    const bool row_ok = tile_ok && vr >= 0 && vr < h;
    const bool update = row_ok && j >= 1;
    const size_t rb = (size_t)(y0 + min(max(vr, 0), h - 1) * dy) * w;
    auto at = [&](int zz, int c) { return (size_t)zz * slice + rb + (size_t)(x0 + (u0 + min(c, nu - 1)) * dx); };
    // register k holds column (k - j) mod kOW, so step s reads register s mod kOW
    float R[kOZ][kOW], RI[kOZ][kOW], B[kOW], BI[kOW];
    float first[kOZ], first_in[kOZ];   // per slice, the column before the tile
    {
        const int zb = zof(t0 - 1);
#pragma unroll
        for (int k = 0; k < kOW; k++) {
            const int c = (k - j) & (kOW - 1);
            B[k] = out[at(zb, c)];
            BI[k] = MODE == 2 ? inp[at(zb, c)] : 0.0f;
        }
    }
#pragma unroll
    for (int zi = 0; zi < kOZ; zi++) {
        const int z = zof(t0 + zi);
#pragma unroll
        for (int k = 0; k < kOW; k++) {
            const int c = (k - j) & (kOW - 1);
            R[zi][k] = out[at(z, c)];
            RI[zi][k] = MODE == 2 ? inp[at(z, c)] : 0.0f;
        }
        first[zi] = 0.0f;
        first_in[zi] = 0.0f;
        if (u0 > 0 && row_ok) {
            const size_t o = (size_t)z * slice + rb + (size_t)(x0 + (u0 - 1) * dx);
            first[zi] = out[o];
            if (MODE == 2) first_in[zi] = inp[o];
        }
    }
#pragma unroll
    for (int zi = 0; zi < kOZ; zi++) {
        const int t = t0 + zi;
        const bool live = row_ok && t < d;
        // the lane's last results (output and, MODE 2, input); first the
        // column before the tile
        float mine = first[zi], mine_in = first_in[zi];
#pragma unroll
        for (int s = 0; s < kOW + kOGroup - 1; s++) {
            const int a = s - j;
            const int k = s & (kOW - 1);
            const float up = from_prev_lane16(mine);
            const float up_in = MODE == 2 ? from_prev_lane16(mine_in) : 0.0f;
            const bool hl = u0 + a >= 1, hu = vr >= 1, hb = t >= 1;   // neighbours inside
            float v;
            if (MODE == 1) {
                v = R[zi][k] + (1.0f - R[zi][k]) * octant_combine(hl ? mine : 0.0f, hu ? up : 0.0f, hb ? B[k] : 0.0f);
            } else {
                v = octant_combine(hl ? mine : 0.0f, hu ? up : 0.0f, hb ? B[k] : 0.0f, hl ? mine_in : 0.0f,
                                   hu ? up_in : 0.0f, hb ? BI[k] : 0.0f);
            }
            // selects rather than a branch: every lane runs the same instructions
            const bool act = live && a >= 0 && a < nu;
            R[zi][k] = act && update && (hl || hu || hb) ? v : R[zi][k];   // the plane's corner cell stays
            mine = act ? R[zi][k] : mine;
            if (MODE == 2) mine_in = act ? RI[zi][k] : mine_in;
        }
        // this slice's result (and input) is the next one's slice before
#pragma unroll
        for (int k = 0; k < kOW; k++) {
            B[k] = R[zi][k];
            BI[k] = RI[zi][k];
        }
    }
    if (!update) return;
#pragma unroll
    for (int zi = 0; zi < kOZ; zi++) {
        if (t0 + zi >= d) break;
        const int z = zof(t0 + zi);
#pragma unroll
        for (int k = 0; k < kOW; k++) {
            const int c = (k - j) & (kOW - 1);
            if (c < nu) out[at(z, c)] = R[zi][k];
        }
    }
}

// 2D (OcclusionSweep, :1171-1209).  The chain is (tile diagonals) x (steps a
// tile) for any tile shape, but each launch has a cost of its own, so 2D (no
// slices to multiply the diagonals) uses 64-column x 63-row tiles, one a
// wavefront, lane l on row v0 - 1 + l: 258 launches at 8192^2 instead of
// 1570 with 8 x 15 tiles.  The tile passes through LDS to and from HBM
// (coalesced), and each lane's row is held in 64 registers rotated by its
// lane index.  First row: the cell before; first column: the row before;
// elsewhere both (:1191-1205).
constexpr int kOB = 64;                // tile columns = lanes
constexpr int kOBRows = kOB - 1;       // tile rows (lane 0 carries the row before)
constexpr int kOBStride = kOB + 2;     // LDS row stride: lane l's column s - l at word 65 l + s

__device__ __forceinline__ float from_prev_lane64(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf,
                                                                 0xf, false));
}

// block b: plane b / nt, tile V = max(0, diag - tu + 1) + b % nt, U = diag - V
__global__ void __launch_bounds__(kOB) k_occlusion_sweep_2d(int w, int h, int diag, int tu, int nt,
                                                            float* __restrict__ dirw)
{
    __shared__ float tile[kOB * kOBStride];
    const int lane = threadIdx.x;
    const int plane = (int)blockIdx.x / nt;
    const int V = max(0, diag - (tu - 1)) + (int)blockIdx.x % nt, U = diag - V;
    const int dx = 1 - 2 * (plane & 1), dy = 1 - 2 * ((plane >> 1) & 1);
    const int x0 = dx > 0 ? 0 : w - 1, y0 = dy > 0 ? 0 : h - 1;
    float* f = dirw + (size_t)plane * w * h;
    const int u0 = U * kOB, v0 = V * kOBRows;
    const int nu = min(kOB, w - u0);
    const int rlo = max(v0 - 1, 0), rhi = min(v0 + kOBRows - 1, h - 1);
    auto row = [&](int v) { return (size_t)(y0 + v * dy) * w; };
    const int cu = x0 + (u0 + min(lane, nu - 1)) * dx;
    {
        // all rows in flight at once (rows outside the field repeat the nearest)
        float v[kOB];
#pragma unroll
        for (int r = 0; r < kOB; r++) v[r] = f[row(min(max(v0 - 1 + r, rlo), rhi)) + cu];
#pragma unroll
        for (int r = 0; r < kOB; r++) tile[r * kOBStride + lane] = v[r];
    }
    const int vr = v0 - 1 + lane;
    const bool row_ok = vr >= 0 && vr < h, update = row_ok && lane >= 1;
    float mine = (u0 > 0 && row_ok) ? f[row(vr) + x0 + (u0 - 1) * dx] : 0.0f;
    __syncthreads();
    float R[kOB];
#pragma unroll
    for (int k = 0; k < kOB; k++) R[k] = tile[lane * kOBStride + ((k - lane) & (kOB - 1))];
    const bool hu = vr >= 1;
#pragma unroll
    for (int s = 0; s < 2 * kOB - 1; s++) {
        const int a = s - lane;
        const float up = from_prev_lane64(mine);
        const bool hl = u0 + a >= 1;
        const float q = hl ? quadrant_combine(mine, hu ? up : 0.0f) : quadrant_combine(up, 0.0f);
        const float v = R[s & (kOB - 1)] + (1.0f - R[s & (kOB - 1)]) * q;
        // selects rather than a branch: every lane runs the same instructions
        const bool act = row_ok && a >= 0 && a < nu;
        R[s & (kOB - 1)] = act && update && (hl || hu) ? v : R[s & (kOB - 1)];   // the plane's corner cell stays
        mine = act ? R[s & (kOB - 1)] : mine;
    }
#pragma unroll
    for (int k = 0; k < kOB; k++) tile[lane * kOBStride + ((k - lane) & (kOB - 1))] = R[k];
    __syncthreads();
    {
        float v[kOB];
#pragma unroll
        for (int r = 1; r < kOB; r++) v[r] = tile[r * kOBStride + lane];
#pragma unroll
        for (int r = 1; r < kOB; r++)
            if (v0 - 1 + r <= rhi && lane < nu) f[row(v0 - 1 + r) + cu] = v[r];
    }
}

// mode 0: 2D (d == 1, 4 planes); 1: 3D in place; 2: 3D from `in` (8 planes)
void launch_occlusion_sweep(int mode, int w, int h, int d, const float* in, float* dirw, hipStream_t s)
{
    ProfileScope p("occlusion_sweep", s);
    if (mode == 0) {
        const int tu = (w + kOB - 1) / kOB, tv = (h + kOBRows - 1) / kOBRows;
        for (int diag = 0; diag < tu + tv - 1; diag++) {
            const int lo = diag - (tu - 1) > 0 ? diag - (tu - 1) : 0, hi = diag < tv - 1 ? diag : tv - 1;
            const int nt = hi - lo + 1;
            hipLaunchKernelGGL(k_occlusion_sweep_2d, dim3((unsigned)(4 * nt)), dim3(kOB), 0, s, w, h, diag, tu,
                               nt, dirw);
        }
        return;
    }
    const int tu = (w + kOW - 1) / kOW, tv = (h + kORows - 1) / kORows;
    const int vb = (tv + kOTiles - 1) / kOTiles;
    const dim3 grid((unsigned)(8 * vb * tu));
    const int tt = (d + kOZ - 1) / kOZ;   // slice groups
    for (int diag = 0; diag < tu + tv + tt - 2; diag++) {
        if (mode == 1)
            hipLaunchKernelGGL(k_occlusion_sweep<1>, grid, dim3(64), 0, s, w, h, d, diag, tu, vb, nullptr, dirw);
        else
            hipLaunchKernelGGL(k_occlusion_sweep<2>, grid, dim3(64), 0, s, w, h, d, diag, tu, vb, in, dirw);
    }
}

// ------------------------------------------------------------ AO images

// EncodeAO4 / EncodeAO8 (one byte a cell) and EncodeDirW4 / EncodeDirW8 (RGBA
// a cell) of DistanceOcclusionLib.h:143-199 over whole fields of n cells,
// the reference's float terms in its order (the tool's per-cell loops,
// DistanceOcclusionTool.cpp:1521-1540, 1595-1615)
// EncodeDirW4 / EncodeDirW8 of one cell as an RGBA word
template <int PLANES>
__device__ __forceinline__ uint32_t dirw_rgba(const float* d)
{
    uint32_t b0, b1, b2, b3;
    if constexpr (PLANES == 4) {
        float fw = d[0] + d[1] + d[2] + d[3];
        float fx = d[0] - d[1] + d[2] - d[3];
        float fy = d[0] + d[1] - d[2] - d[3];
        const float inv = __fdiv_rn(1.0f, sqrt_rn(fx * fx + fy * fy));
        fx *= inv;
        fy *= inv;
        fw *= 0.25f;
        fx = fw * fx * 0.5f + 0.5f;
        fy = fw * fy * 0.5f + 0.5f;
        b0 = encode_u8(fx); b1 = encode_u8(fy); b2 = encode_u8(fw); b3 = 255;
    } else {
        float fw = d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7];
        float fx = d[0] - d[1] + d[2] - d[3] + d[4] - d[5] + d[6] - d[7];
        float fy = d[0] + d[1] - d[2] - d[3] + d[4] + d[5] - d[6] - d[7];
        float fz = d[0] + d[1] + d[2] + d[3] - d[4] - d[5] - d[6] - d[7];
        const float inv = __fdiv_rn(1.0f, sqrt_rn(fx * fx + fy * fy + fz * fz));
        fx *= inv;
        fy *= inv;
        fz *= inv;
        fw *= 0.125f;
        fx = fw * fx * 0.5f + 0.5f;
        fy = fw * fy * 0.5f + 0.5f;
        fz = fw * fz * 0.5f + 0.5f;
        b0 = encode_u8(fx); b1 = encode_u8(fy); b2 = encode_u8(fz); b3 = encode_u8(fw);
    }
    return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
}

template <int PLANES, bool DIRW>
__global__ void __launch_bounds__(256) k_encode_ao(size_t n, const float* __restrict__ dirw,
                                                   void* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float d[PLANES];
#pragma unroll
    for (int p = 0; p < PLANES; p++) d[p] = dirw[(size_t)p * n + i];
    if (DIRW) {
        ((uint32_t*)out)[i] = dirw_rgba<PLANES>(d);
    } else {
        float s = d[0];
#pragma unroll
        for (int p = 1; p < PLANES; p++) s += d[p];
        ((uint8_t*)out)[i] = encode_u8(1.0f - (PLANES == 4 ? 0.25f : 0.125f) * s);
    }
}

// planes 4 or 8; rgba: EncodeDirW images (4 bytes a cell), else EncodeAO (1)
void launch_encode_ao(size_t n, int planes, bool rgba, const float* dirw, uint8_t* out, hipStream_t s)
{
    ProfileScope p("encode_ao", s);
    const dim3 g(grid1d(n, 256)), b(256);
    if (planes == 4 && !rgba) hipLaunchKernelGGL((k_encode_ao<4, false>), g, b, 0, s, n, dirw, out);
    else if (planes == 4) hipLaunchKernelGGL((k_encode_ao<4, true>), g, b, 0, s, n, dirw, out);
    else if (!rgba) hipLaunchKernelGGL((k_encode_ao<8, false>), g, b, 0, s, n, dirw, out);
    else hipLaunchKernelGGL((k_encode_ao<8, true>), g, b, 0, s, n, dirw, out);
}

}  // namespace dolhip
