// dol_voxel.hip -- surface voxelisation of a triangle list into a bitmask,
// the reference's CreateBitMaskFromTriangles (MeshSupport.cpp:369-475) with
// its triangle/cell test cTriCellIntersectDelta (MeshSupport.cpp:158-351).
//
// A cell (x, y, z) of the triangle's cell range is set when the triangle
// passes, at the cell centre, the slice test (z extent), the row test (the
// three edge x X-axis products and the y extent) and the cell test (the six
// other edge products, the x extent and the plane).  The reference visits
// the range slice by slice and row by row and skips a slice or row whose test
// fails; the set of cells it sets is the cells passing all three, which is
// what is computed here, one cell per lane.  A range of one cell is set
// without a test, as there.
//
// One wavefront takes one triangle at a time (grid-stride): every lane forms
// the triangle's test data (the same values, so no exchange is needed), the
// lanes test the range's (y, z) rows 64 at a time, and for each passing row
// they test its cells 64 at a time and OR the set bits into the mask with
// atomics (any order gives the same mask).  Every float operation
// is a separate rounded operation in the reference's order (no contraction:
// the library is built with -ffp-contract=off); the grid's cell size, its
// inverse and half-width come from the host, computed as there.
#include "dol_common.hpp"

namespace dolhip {

namespace {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float lo_of(float a, float b) { return a < b ? a : b; }   // Min, MeshSupport.cpp:95-98
__device__ __forceinline__ float hi_of(float a, float b) { return b < a ? a : b; }   // Max, MeshSupport.cpp:90-93

struct TriTest {
    V3 e1, t1, u1, e2, t2, u2, e3, t3, u3, pmin, pmax, hw, n;
    float ndv_min, ndv_max, nd;
};

// one edge's three axis intervals: the triangle's projection (the vertex pair
// named per axis) widened by the cell's projected radius
__device__ __forceinline__ void edge_axes(V3 e, V3 pa, V3 pb, V3 pc, V3 qa, V3 qb, V3 qc, V3 hw,
                                          V3& t, V3& u)
{
    t.x = e.z * pa.y - e.y * pa.z;
    t.y = -e.z * pb.x + e.x * pb.z;
    t.z = e.y * pc.x - e.x * pc.y;
    u.x = e.z * qa.y - e.y * qa.z;
    u.y = -e.z * qb.x + e.x * qb.z;
    u.z = e.y * qc.x - e.x * qc.y;
    if (t.x >= u.x) { const float s = t.x; t.x = u.x; u.x = s; }
    if (t.y >= u.y) { const float s = t.y; t.y = u.y; u.y = s; }
    if (t.z >= u.z) { const float s = t.z; t.z = u.z; u.z = s; }
    const V3 ae{fabsf(e.x), fabsf(e.y), fabsf(e.z)};
    const V3 r{ae.z * hw.y + ae.y * hw.z, ae.z * hw.x + ae.x * hw.z, ae.y * hw.x + ae.x * hw.y};
    t = sub3(t, r);
    u = V3{u.x + r.x, u.y + r.y, u.z + r.z};
}

__device__ __forceinline__ void tri_setup(V3 p1, V3 p2, V3 p3, V3 hw, TriTest& T)
{
    T.hw = hw;
    T.e1 = sub3(p2, p1);
    T.e2 = sub3(p3, p2);
    T.e3 = sub3(p1, p3);
    // twice the area-weighted normal, MeshSupport.cpp:141-155
    T.n.x = (p1.y - p2.y) * (p1.z + p2.z) + (p2.y - p3.y) * (p2.z + p3.z) + (p3.y - p1.y) * (p3.z + p1.z);
    T.n.y = (p1.z - p2.z) * (p1.x + p2.x) + (p2.z - p3.z) * (p2.x + p3.x) + (p3.z - p1.z) * (p3.x + p1.x);
    T.n.z = (p1.x - p2.x) * (p1.y + p2.y) + (p2.x - p3.x) * (p2.y + p3.y) + (p3.x - p1.x) * (p3.y + p1.y);
    const V3 vmn{T.n.x > 0.0f ? -hw.x : hw.x, T.n.y > 0.0f ? -hw.y : hw.y, T.n.z > 0.0f ? -hw.z : hw.z};
    const V3 vmx{T.n.x > 0.0f ? hw.x : -hw.x, T.n.y > 0.0f ? hw.y : -hw.y, T.n.z > 0.0f ? hw.z : -hw.z};
    T.ndv_min = dot3(T.n, vmn);
    T.ndv_max = dot3(T.n, vmx);
    T.nd = dot3(T.n, p1);
    edge_axes(T.e1, p1, p1, p2, p3, p3, p3, hw, T.t1, T.u1);
    edge_axes(T.e2, p1, p1, p1, p3, p3, p2, hw, T.t2, T.u2);
    edge_axes(T.e3, p1, p1, p2, p2, p2, p3, hw, T.t3, T.u3);
    T.pmin = V3{lo_of(lo_of(p1.x, p2.x), p3.x), lo_of(lo_of(p1.y, p2.y), p3.y), lo_of(lo_of(p1.z, p2.z), p3.z)};
    T.pmax = V3{hi_of(hi_of(p1.x, p2.x), p3.x), hi_of(hi_of(p1.y, p2.y), p3.y), hi_of(hi_of(p1.z, p2.z), p3.z)};
}

__device__ __forceinline__ bool hit_slice(const TriTest& T, V3 c)
{
    return !(T.pmin.z - c.z > T.hw.z || T.pmax.z - c.z < -T.hw.z);
}

__device__ __forceinline__ bool hit_row(const TriTest& T, V3 c)
{
    float dr = T.e1.z * c.y - T.e1.y * c.z;
    if (T.t1.x > dr || T.u1.x < dr) return false;
    dr = T.e2.z * c.y - T.e2.y * c.z;
    if (T.t2.x > dr || T.u2.x < dr) return false;
    dr = T.e3.z * c.y - T.e3.y * c.z;
    if (T.t3.x > dr || T.u3.x < dr) return false;
    return !(T.pmin.y - c.y > T.hw.y || T.pmax.y - c.y < -T.hw.y);
}

__device__ __forceinline__ bool hit_cell(const TriTest& T, V3 c)
{
    float dr = -T.e1.z * c.x + T.e1.x * c.z;
    if (T.t1.y > dr || T.u1.y < dr) return false;
    dr = T.e1.y * c.x - T.e1.x * c.y;
    if (T.t1.z > dr || T.u1.z < dr) return false;
    dr = -T.e2.z * c.x + T.e2.x * c.z;
    if (T.t2.y > dr || T.u2.y < dr) return false;
    dr = T.e2.y * c.x - T.e2.x * c.y;
    if (T.t2.z > dr || T.u2.z < dr) return false;
    dr = -T.e3.z * c.x + T.e3.x * c.z;
    if (T.t3.y > dr || T.u3.y < dr) return false;
    dr = T.e3.y * c.x - T.e3.x * c.y;
    if (T.t3.z > dr || T.u3.z < dr) return false;
    if (T.pmin.x - c.x > T.hw.x || T.pmax.x - c.x < -T.hw.x) return false;
    const float d = T.nd - dot3(T.n, c);
    return T.ndv_min <= d && T.ndv_max >= d;
}

struct VoxGrid {
    V3 cmin, cw, inv, hw;   // bbox min, cell size, cells per unit, half-width
    int w, h, d;
};

__device__ __forceinline__ V3 vertex(const float* __restrict__ v, int i)
{
    return V3{v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]};
}

}  // namespace

__global__ void __launch_bounds__(256) k_voxelise(int tri_count, const int* __restrict__ indices,
                                                  const float* __restrict__ vertices, VoxGrid g,
                                                  uint32_t* __restrict__ mask)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int waves = (int)((gridDim.x * blockDim.x) >> 6);
    const size_t row = (size_t)mask_stride(g.w), slice = row * g.h;
    for (int t = wave; t < tri_count; t += waves) {
        const int i1 = indices ? indices[3 * t] : 3 * t;
        const int i2 = indices ? indices[3 * t + 1] : 3 * t + 1;
        const int i3 = indices ? indices[3 * t + 2] : 3 * t + 2;
        const V3 p1 = vertex(vertices, i1), p2 = vertex(vertices, i2), p3 = vertex(vertices, i3);
        // the triangle's cell range (TriangleBounds, MeshSupport.cpp:353-366; :409-419)
        const V3 bmn{lo_of(lo_of(p1.x, p2.x), p3.x), lo_of(lo_of(p1.y, p2.y), p3.y), lo_of(lo_of(p1.z, p2.z), p3.z)};
        const V3 bmx{hi_of(hi_of(p1.x, p2.x), p3.x), hi_of(hi_of(p1.y, p2.y), p3.y), hi_of(hi_of(p1.z, p2.z), p3.z)};
        const int cx0 = max((int)floorf((bmn.x - g.cmin.x) * g.inv.x), 0);
        const int cy0 = max((int)floorf((bmn.y - g.cmin.y) * g.inv.y), 0);
        const int cz0 = max((int)floorf((bmn.z - g.cmin.z) * g.inv.z), 0);
        const int cx1 = min((int)ceilf((bmx.x - g.cmin.x) * g.inv.x), g.w);
        const int cy1 = min((int)ceilf((bmx.y - g.cmin.y) * g.inv.y), g.h);
        const int cz1 = min((int)ceilf((bmx.z - g.cmin.z) * g.inv.z), g.d);
        if (cx0 == cx1) continue;
        if (cx0 + 1 == cx1 && cy0 + 1 == cy1 && cz0 + 1 == cz1) {   // one cell: set without a test
            if (lane == 0) atomicOr(mask + slice * cz0 + row * cy0 + (cx0 >> 5), 1u << (cx0 & 31));
            continue;
        }
        const int nx = cx1 - cx0, ny = cy1 - cy0, nz = cz1 - cz0;
        if (nx <= 0 || ny <= 0 || nz <= 0) continue;
        TriTest T;
        tri_setup(p1, p2, p3, g.hw, T);
        // rows first, as the reference skips a failing slice or row: each lane
        // tests one (y, z) row, then the wavefront walks the passing rows'
        // cells 64 at a time
        const int rows = ny * nz;
        if ((long long)rows * nx <= 64) {
            // a small range: one pass over its cells
            const int j = lane;
            if (j < rows * nx) {
                const int x = cx0 + j % nx, r = j / nx;
                const int y = cy0 + r % ny, z = cz0 + r / ny;
                V3 c;
                c.z = g.cmin.z + g.cw.z * ((float)z + 0.5f);
                c.y = g.cmin.y + g.cw.y * ((float)y + 0.5f);
                c.x = g.cmin.x + g.cw.x * ((float)x + 0.5f);
                if (hit_slice(T, c) && hit_row(T, c) && hit_cell(T, c))
                    atomicOr(mask + slice * z + row * y + (x >> 5), 1u << (x & 31));
            }
            continue;
        }
        for (int r0 = 0; r0 < rows; r0 += 64) {
            const int r = r0 + lane;
            bool pass = false;
            if (r < rows) {
                V3 c;
                c.z = g.cmin.z + g.cw.z * ((float)(cz0 + r / ny) + 0.5f);
                c.y = g.cmin.y + g.cw.y * ((float)(cy0 + r % ny) + 0.5f);
                pass = hit_slice(T, c) && hit_row(T, c);
            }
            uint64_t live = __ballot(pass);
            while (live) {
                const int b = __ffsll((unsigned long long)live) - 1;
                live &= live - 1;
                const int rr = r0 + b, y = cy0 + rr % ny, z = cz0 + rr / ny;
                V3 c;
                c.z = g.cmin.z + g.cw.z * ((float)z + 0.5f);
                c.y = g.cmin.y + g.cw.y * ((float)y + 0.5f);
                for (int x = cx0 + lane; x < cx1; x += 64) {
                    c.x = g.cmin.x + g.cw.x * ((float)x + 0.5f);
                    if (hit_cell(T, c)) atomicOr(mask + slice * z + row * y + (x >> 5), 1u << (x & 31));
                }
            }
        }
    }
}

// grid constants of MeshSupport.cpp:382-388, computed on the host in float
// as there: cellW = (max - min) / whd, cellInvW = whd / (max - min),
// hw = cellW * 0.500001
void launch_voxelise(int tri_count, const int* indices, const float* vertices, const float* bbox,
                     int w, int h, int d, uint32_t* mask, hipStream_t s)
{
    VoxGrid g;
    const float whd[3] = {(float)w, (float)h, (float)d};
    float ext[3], cw[3], inv[3], hw[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = bbox[3 + k] - bbox[k];
        cw[k] = ext[k] / whd[k];
        inv[k] = whd[k] / ext[k];
        hw[k] = cw[k] * 0.500001f;
    }
    g.cmin = V3{bbox[0], bbox[1], bbox[2]};
    g.cw = V3{cw[0], cw[1], cw[2]};
    g.inv = V3{inv[0], inv[1], inv[2]};
    g.hw = V3{hw[0], hw[1], hw[2]};
    g.w = w;
    g.h = h;
    g.d = d;
    if (tri_count <= 0) return;
    // a wavefront per triangle, at most 8 per CU in flight per 256 CUs' worth
    const unsigned blocks = (unsigned)min((tri_count + 3) / 4, 256 * 32);
    ProfileScope p("voxelise", s);
    hipLaunchKernelGGL(k_voxelise, dim3(blocks), dim3(256), 0, s, tri_count, indices, vertices, g, mask);
}

}  // namespace dolhip
