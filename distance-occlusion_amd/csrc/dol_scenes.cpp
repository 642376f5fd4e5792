// dol_scenes.cpp -- host-side mask generators.
//
// The reference's own generators (BitMaskAddPoints, BitMaskAddBlock,
// DistanceOcclusionLib.cpp:95-179, 1224-1287) plus seeded scene generators
// for the larger workloads: unions of discs and boxes (2D), of rotated
// ellipsoids and boxes (3D), and ragged sets of glyph-like coverage masks.
// Masks are produced once on the host and then shared by every consumer, so
// the GPU path and any CPU check see identical inputs.  Pixels are cell
// centres at integer coordinates; a shape covers a pixel whose centre is
// inside or on it.
#include "../../include/dol_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline int stride(int w) { return (w + 31) >> 5; }

// LCG of DistanceOcclusionLib.cpp:28-32
inline int lcg(uint32_t& seed, int limit)
{
    seed = (uint32_t)((uint64_t)seed * 1103515245u) + 12345u;
    return (int)(((uint64_t)seed * (uint64_t)(uint32_t)limit) >> 32);
}

// splitmix64 for the scene generators
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int integer(int lo, int hi) { return lo + (int)(uniform() * (double)(hi - lo + 1)); }  // [lo, hi]
};

// set pixels [x0, x1] of one row
void set_span(uint32_t* row, int w, int x0, int x1)
{
    x0 = std::max(x0, 0);
    x1 = std::min(x1, w - 1);
    for (int x = x0; x <= x1;) {
        int j = x >> 5, b = x & 31;
        int n = std::min(32 - b, x1 - x + 1);
        uint32_t bits = (n == 32) ? 0xffffffffu : (((1u << n) - 1u) << b);
        row[j] |= bits;
        x += n;
    }
}

// FillMaskBlock of DistanceOcclusionLib.cpp:131-156 (inclusive bounds);
// negative bounds (the reference's undefined small-size cases) are clipped
void fill_rect(uint32_t* mask, int w, int x0, int x1, int y0, int y1)
{
    if (x0 < 0 || x1 < 0) return;
    for (int y = std::max(y0, 0); y <= y1; y++) set_span(mask + (size_t)y * stride(w), w, x0, x1);
}

// pixel-centre span of {x : a x^2 + 2 b x + c <= 0} around centre cx (a > 0)
bool quad_span(double a, double b, double c, double cx, int& x0, int& x1)
{
    double disc = b * b - a * c;
    if (disc < 0.0) return false;
    double r = std::sqrt(disc);
    double lo = (-b - r) / a, hi = (-b + r) / a;
    x0 = (int)std::ceil(cx + lo);
    x1 = (int)std::floor(cx + hi);
    return x0 <= x1;
}

}  // namespace

extern "C" {

void dolhip_mask_add_points_2d(int w, int h, int count, uint32_t seed, uint32_t* mask)
{
    const int ws = stride(w);
    for (int i = 0; i < count; i++) {
        int x = lcg(seed, w);
        int y = lcg(seed, h);
        mask[(size_t)y * ws + (x >> 5)] |= 1u << (x & 31);
    }
}

void dolhip_mask_add_points_3d(int w, int h, int d, int count, uint32_t seed, uint32_t* mask)
{
    const int ws = stride(w);
    for (int i = 0; i < count; i++) {
        int x = lcg(seed, w);
        int y = lcg(seed, h);
        int z = lcg(seed, d);
        mask[((size_t)z * h + y) * ws + (x >> 5)] |= 1u << (x & 31);
    }
}

void dolhip_mask_add_block_2d(int w, int h, int sides, uint32_t* mask)
{
    int x0 = w / 4, x1 = x0 + w / 2 - 1, y0 = h / 3, y1 = y0 + h / 3 - 1;
    if (sides > 4) { fill_rect(mask, w, x0, x1, y0, y1); return; }
    if (sides > 0) fill_rect(mask, w, x0, x1, y0, y0);
    if (sides > 1) fill_rect(mask, w, x0, x0, y0, y1);
    if (sides > 2) fill_rect(mask, w, x0, x1, y1, y1);
    if (sides > 3) fill_rect(mask, w, x1, x1, y0, y1);
}

void dolhip_mask_add_block_3d(int w, int h, int d, int sides, uint32_t* mask)
{
    const size_t sl = (size_t)stride(w) * h;
    int x0 = w / 4, x1 = x0 + w / 2 - 1, y0 = h / 3, y1 = y0 + h / 3 - 1;
    int z0 = d / 3, z1 = z0 + d / 3 - 1;
    auto box = [&](int a0, int a1, int b0, int b1, int c0, int c1) {
        for (int z = std::max(c0, 0); z <= c1; z++) fill_rect(mask + z * sl, w, a0, a1, b0, b1);
    };
    if (sides > 6) { box(x0, x1, y0, y1, z0, z1); return; }
    if (sides > 0) box(x0, x1, y0, y1, z0, z0);
    if (sides > 1) box(x0, x0, y0, y1, z0, z1);
    if (sides > 2) box(x0, x1, y0, y0, z0, z1);
    if (sides > 3) box(x1, x1, y0, y1, z0, z1);
    if (sides > 4) box(x0, x1, y1, y1, z0, z1);
    if (sides > 5) box(x0, x1, y0, y1, z1, z1);
}

// Union of n shapes, each a disc (centre U[0,w) x U[0,h), radius U[4,200])
// or an axis-aligned box (sides U[8,400] each, corner U[0,w) x U[0,h)),
// with equal odds.
void dolhip_scene_shapes_2d(int w, int h, int n, uint64_t seed, uint32_t* mask)
{
    Rng r(seed);
    const int ws = stride(w);
    for (int s = 0; s < n; s++) {
        if (r.uniform() < 0.5) {
            double cx = r.uniform(0, w), cy = r.uniform(0, h), rad = r.uniform(4, 200);
            int y0 = std::max(0, (int)std::ceil(cy - rad)), y1 = std::min(h - 1, (int)std::floor(cy + rad));
            for (int y = y0; y <= y1; y++) {
                double dy = y - cy;
                int x0, x1;
                if (quad_span(1.0, 0.0, dy * dy - rad * rad, cx, x0, x1))
                    set_span(mask + (size_t)y * ws, w, x0, x1);
            }
        } else {
            int bw = r.integer(8, 400), bh = r.integer(8, 400);
            int x0 = r.integer(0, w - 1), y0 = r.integer(0, h - 1);
            for (int y = y0; y < std::min(h, y0 + bh); y++) set_span(mask + (size_t)y * ws, w, x0, x0 + bw - 1);
        }
    }
}

// Union of n_ell ellipsoids (centre uniform in the volume, semi-axes
// U[4,64], orientation from a uniform random unit quaternion) and n_box
// axis-aligned boxes (sides U[8,128], corner uniform in the volume).
void dolhip_scene_shapes_3d(int w, int h, int d, int n_ell, int n_box, uint64_t seed, uint32_t* mask)
{
    Rng r(seed);
    const int ws = stride(w);
    for (int s = 0; s < n_ell; s++) {
        double c[3] = { r.uniform(0, w), r.uniform(0, h), r.uniform(0, d) };
        double ax[3] = { r.uniform(4, 64), r.uniform(4, 64), r.uniform(4, 64) };
        // uniform quaternion (Shoemake)
        double u1 = r.uniform(), u2 = r.uniform() * 2 * M_PI, u3 = r.uniform() * 2 * M_PI;
        double qa = std::sqrt(1 - u1) * std::sin(u2), qb = std::sqrt(1 - u1) * std::cos(u2);
        double qc = std::sqrt(u1) * std::sin(u3), qd = std::sqrt(u1) * std::cos(u3);
        // rotation matrix R (columns = ellipsoid axes in world space)
        double R[3][3] = {
            { 1 - 2 * (qc * qc + qd * qd), 2 * (qb * qc - qa * qd), 2 * (qb * qd + qa * qc) },
            { 2 * (qb * qc + qa * qd), 1 - 2 * (qb * qb + qd * qd), 2 * (qc * qd - qa * qb) },
            { 2 * (qb * qd - qa * qc), 2 * (qc * qd + qa * qb), 1 - 2 * (qb * qb + qc * qc) },
        };
        // M = R diag(1/a^2) R^T
        double M[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += R[i][k] * R[j][k] / (ax[k] * ax[k]);
                M[i][j] = v;
            }
        double rmax = std::max(ax[0], std::max(ax[1], ax[2]));
        int z0 = std::max(0, (int)std::ceil(c[2] - rmax)), z1 = std::min(d - 1, (int)std::floor(c[2] + rmax));
        int y0 = std::max(0, (int)std::ceil(c[1] - rmax)), y1 = std::min(h - 1, (int)std::floor(c[1] + rmax));
        for (int z = z0; z <= z1; z++)
            for (int y = y0; y <= y1; y++) {
                double qy = y - c[1], qz = z - c[2];
                double a = M[0][0];
                double b = M[0][1] * qy + M[0][2] * qz;
                double cc = M[1][1] * qy * qy + 2 * M[1][2] * qy * qz + M[2][2] * qz * qz - 1.0;
                int x0, x1;
                if (quad_span(a, b, cc, c[0], x0, x1))
                    set_span(mask + ((size_t)z * h + y) * ws, w, x0, x1);
            }
    }
    for (int s = 0; s < n_box; s++) {
        int bw = r.integer(8, 128), bh = r.integer(8, 128), bd = r.integer(8, 128);
        int x0 = r.integer(0, w - 1), y0 = r.integer(0, h - 1), z0 = r.integer(0, d - 1);
        for (int z = z0; z < std::min(d, z0 + bd); z++)
            for (int y = y0; y < std::min(h, y0 + bh); y++)
                set_span(mask + ((size_t)z * h + y) * ws, w, x0, x0 + bw - 1);
    }
}

// Ragged glyph set: mask i has w, h uniform in [min_size, max_size]; its
// bytes (255 inside, 0 outside) start at offset[i] (masks packed back to
// back).  Each is a random closed polygon of 5-40 vertices (star-shaped
// around the centre, filled with the even-odd rule at pixel centres) plus
// 1-4 strokes (segments of width 1-6 px).  With bytes == NULL only the sizes
// and offsets are produced; returns the total byte count.
int64_t dolhip_scene_glyphs(int n, int min_size, int max_size, uint64_t seed, int32_t* w_out,
                            int32_t* h_out, int64_t* offset, uint8_t* bytes)
{
    Rng r(seed);
    int64_t off = 0;
    std::vector<double> px, py;
    for (int g = 0; g < n; g++) {
        const int w = r.integer(min_size, max_size), h = r.integer(min_size, max_size);
        w_out[g] = w;
        h_out[g] = h;
        offset[g] = off;
        // geometry is drawn whether or not bytes are written, so sizes and
        // contents come from one sequence
        const int nv = r.integer(5, 40);
        const double cx = w * r.uniform(0.35, 0.65), cy = h * r.uniform(0.35, 0.65);
        const double rmax = 0.5 * std::min(w, h);
        px.resize(nv);
        py.resize(nv);
        double a0 = r.uniform() * 2 * M_PI;
        for (int k = 0; k < nv; k++) {
            double ang = a0 + 2 * M_PI * (k + r.uniform(0.0, 0.8)) / nv;
            double rad = rmax * r.uniform(0.25, 0.95);
            px[k] = cx + rad * std::cos(ang);
            py[k] = cy + rad * std::sin(ang);
        }
        const int ns = r.integer(1, 4);
        double sx0[4], sy0[4], sx1[4], sy1[4], sw[4];
        for (int k = 0; k < ns; k++) {
            sx0[k] = r.uniform(0, w); sy0[k] = r.uniform(0, h);
            sx1[k] = r.uniform(0, w); sy1[k] = r.uniform(0, h);
            sw[k] = r.uniform(1, 6);
        }
        if (bytes) {
            uint8_t* o = bytes + off;
            std::memset(o, 0, (size_t)w * h);
            // polygon, even-odd rule at pixel centres, one scanline at a time:
            // pixel x is inside when an odd number of edge crossings lie right
            // of it (crossing test (py[k] > y) != (py[j] > y), x < crossing)
            std::vector<double> xs;
            for (int y = 0; y < h; y++) {
                xs.clear();
                for (int k = 0, j = nv - 1; k < nv; j = k++)
                    if ((py[k] > y) != (py[j] > y))
                        xs.push_back((px[j] - px[k]) * (y - py[k]) / (py[j] - py[k]) + px[k]);
                std::sort(xs.begin(), xs.end());
                uint8_t* row = o + (size_t)y * w;
                for (size_t e = 0; e + 1 < xs.size(); e += 2) {
                    // x < xs[e+1] and not x < xs[e]: xs[e] <= x < xs[e+1]
                    int x0 = (int)std::ceil(xs[e]), x1 = (int)std::ceil(xs[e + 1]) - 1;
                    x0 = std::max(x0, 0);
                    x1 = std::min(x1, w - 1);
                    for (int x = x0; x <= x1; x++) row[x] = 255;
                }
            }
            // strokes: capsules of diameter sw around each segment
            for (int k = 0; k < ns; k++) {
                const double rad = 0.5 * sw[k];
                const int xa = std::max(0, (int)std::floor(std::min(sx0[k], sx1[k]) - rad));
                const int xb = std::min(w - 1, (int)std::ceil(std::max(sx0[k], sx1[k]) + rad));
                const int ya = std::max(0, (int)std::floor(std::min(sy0[k], sy1[k]) - rad));
                const int yb = std::min(h - 1, (int)std::ceil(std::max(sy0[k], sy1[k]) + rad));
                const double vx = sx1[k] - sx0[k], vy = sy1[k] - sy0[k];
                const double l2 = vx * vx + vy * vy;
                for (int y = ya; y <= yb; y++)
                    for (int x = xa; x <= xb; x++) {
                        double t = l2 > 0 ? ((x - sx0[k]) * vx + (y - sy0[k]) * vy) / l2 : 0.0;
                        t = t < 0 ? 0 : (t > 1 ? 1 : t);
                        const double ex = x - (sx0[k] + t * vx), ey = y - (sy0[k] + t * vy);
                        if (ex * ex + ey * ey <= rad * rad) o[(size_t)y * w + x] = 255;
                    }
            }
        }
        off += (int64_t)w * h;
    }
    return off;
}

// n seeds uniform in [0, w) x [0, h) as float (x, y) pairs, and (when
// weights is not NULL) one weight per cell uniform in [0, 1)
void dolhip_scene_seeds(int w, int h, int n, uint64_t seed, float* seeds, float* weights)
{
    Rng r(seed);
    for (int k = 0; k < n; k++) {
        float x = (float)r.uniform(0, w), y = (float)r.uniform(0, h);
        seeds[2 * k] = x < (float)w ? x : std::nextafter((float)w, 0.0f);
        seeds[2 * k + 1] = y < (float)h ? y : std::nextafter((float)h, 0.0f);
    }
    if (weights)
        for (size_t i = 0; i < (size_t)w * h; i++) weights[i] = (float)r.uniform();
}

}  // extern "C"
