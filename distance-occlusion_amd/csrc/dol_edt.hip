// dol_edt.hip -- squared Euclidean distance transforms for gfx950.
//
// Two families:
//  * Felzenszwalb float (DistanceOcclusionLib.cpp:1018-1085), bit-identical to
//    the reference: the same float operations in the same order, no FMA
//    contraction, correctly rounded division.  One thread per column (then
//    per row), envelope stacks in a workspace laid out [k][line] so that
//    lanes at equal depth touch consecutive words.
//  * Exact integer EDT (an extension; the reference has only the float 2D
//    form): a per-row nearest-set-bit pass straight from the bitmask, then the
//    lower envelope of parabolas along each further axis with intersections
//    compared by exact cross-multiplication.
#include "dol_common.hpp"

#include <cfloat>

namespace dolhip {

// ------------------------------------------------------ Felzenszwalb float

// 1D envelope of DistanceOcclusionLib.cpp:1026-1062 over n samples of f at
// stride fs, writing n outputs at stride os.  v and z are this line's stacks
// at stride `ls` (number of lines).
__device__ void felz_line(int n, const float* __restrict__ f, size_t fs, float* __restrict__ out,
                          size_t os, int* __restrict__ v, float* __restrict__ z, size_t ls)
{
    int k = 0;
    v[0] = 0;
    z[0] = -FLT_MAX;
    z[ls] = +FLT_MAX;
    int vk = 0;
    float zk = -FLT_MAX;
    for (int q = 1; q < n; q++) {
        const float fq = __fadd_rn(f[fs * q], (float)(q * q));
        float fv = __fadd_rn(f[fs * vk], (float)(vk * vk));
        float s = __fdiv_rn(__fsub_rn(fq, fv), (float)(2 * (q - vk)));
        while (s <= zk && k > 0) {
            k--;
            vk = v[ls * k];
            zk = z[ls * k];
            fv = __fadd_rn(f[fs * vk], (float)(vk * vk));
            s = __fdiv_rn(__fsub_rn(fq, fv), (float)(2 * (q - vk)));
        }
        k++;
        v[ls * k] = q;
        z[ls * k] = s;
        z[ls * (k + 1)] = +FLT_MAX;
        vk = q;
        zk = s;
    }
    k = 0;
    vk = v[0];
    float znext = z[ls];
    for (int q = 0; q < n; q++) {
        while (znext < (float)q) {
            k++;
            vk = v[ls * k];
            znext = z[ls * (k + 1)];
        }
        out[os * q] = __fadd_rn((float)((q - vk) * (q - vk)), f[fs * vk]);
    }
}

// column pass: line x reads dist[q*w + x], writes td[x*h + q]
__global__ void __launch_bounds__(256) k_felz_cols(int w, int h, const float* __restrict__ dist,
                                                   float* __restrict__ td, int* __restrict__ v,
                                                   float* __restrict__ z)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    felz_line(h, dist + x, (size_t)w, td + (size_t)x * h, 1, v + x, z + x, (size_t)w);
}

// row pass: line y reads td[q*h + y], writes dist[y*w + q]
__global__ void __launch_bounds__(256) k_felz_rows(int w, int h, const float* __restrict__ td,
                                                   float* __restrict__ dist, int* __restrict__ v,
                                                   float* __restrict__ z)
{
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    felz_line(w, td + y, (size_t)h, dist + (size_t)y * w, 1, v + y, z + y, (size_t)h);
}

// transposed field td (w*h floats), stack v (w*h ints: a column's h entries
// at stride w, or a row's w entries at stride h) and stack z (one more entry
// per line than v)
static inline size_t felz_part(size_t b) { return (b + 255) & ~(size_t)255; }

size_t felz_workspace(int w, int h)
{
    const size_t cells = (size_t)w * h, lines = (size_t)(w > h ? w : h);
    return felz_part(cells * 4) * 2 + felz_part((cells + lines) * 4);
}

void launch_felzenszwalb_2d(int w, int h, float* dist, void* work, hipStream_t s)
{
    const size_t cells = (size_t)w * h;
    char* p = (char*)work;
    float* td = (float*)p;
    int* v = (int*)(p + felz_part(cells * 4));
    float* z = (float*)(p + 2 * felz_part(cells * 4));
    {
        ProfileScope ps("felzenszwalb_cols", s);
        hipLaunchKernelGGL(k_felz_cols, dim3(grid1d((size_t)w, 256)), dim3(256), 0, s, w, h, dist, td,
                           v, z);
    }
    {
        ProfileScope ps("felzenszwalb_rows", s);
        hipLaunchKernelGGL(k_felz_rows, dim3(grid1d((size_t)h, 256)), dim3(256), 0, s, w, h, td, dist,
                           v, z);
    }
}

// ------------------------------------------------------ exact integer EDT

// Per-row distance to the nearest set bit (clear bit if invert), squared;
// kEdtInf where the row has none.  Bits past w in the last word are ignored.
__global__ void __launch_bounds__(256) k_edt_rows(int w, int xb, const uint32_t* __restrict__ mask,
                                                  uint32_t invert, int32_t* __restrict__ g)
{
    const int row = blockIdx.x / xb;
    const int x = (blockIdx.x - row * xb) * 256 + threadIdx.x;
    if (x >= w) return;
    const int ws = mask_stride(w);
    const uint32_t* r = mask + (size_t)row * ws;
    const int j = x >> 5, i = x & 31;
    const int tail = w - 32 * (ws - 1);
    const uint32_t last_valid = tail >= 32 ? 0xffffffffu : ((1u << tail) - 1u);
    auto word = [&](int jj) -> uint32_t {
        uint32_t m = r[jj] ^ invert;
        return jj == ws - 1 ? (m & last_valid) : m;
    };
    const uint32_t m = word(j);
    int best = 0x7fffffff;
    // nearest at or left of x
    uint32_t lo = m & (i == 31 ? 0xffffffffu : ((2u << i) - 1u));
    if (lo) {
        best = i - (31 - __clz(lo));
    } else {
        for (int jj = j - 1; jj >= 0; jj--) {
            uint32_t t = word(jj);
            if (t) { best = x - (32 * jj + 31 - __clz(t)); break; }
        }
    }
    // nearest at or right of x
    uint32_t hi = m & ~((1u << i) - 1u);
    if (hi) {
        int dx = __ffs(hi) - 1 - i;
        if (dx < best) best = dx;
    } else {
        for (int jj = j + 1; jj < ws; jj++) {
            uint32_t t = word(jj);
            if (t) {
                int dx = 32 * jj + __ffs(t) - 1 - x;
                if (dx < best) best = dx;
                break;
            }
        }
    }
    g[(size_t)row * w + x] = best == 0x7fffffff ? kEdtInf : best * best;
}

// Lower envelope along one axis of length n.  Line c starts at
// (c / inner) * outer + (c % inner) and steps by `stride`.  Input f (kEdtInf
// = no site), output d2 (and optionally the float distance).  Sites are kept
// on a stack in `stk` ([k][line]); the top two are cached in registers.  A
// site b between a and q is popped when its interval is empty:
// (Fq - Fb)(b - a) <= (Fb - Fa)(q - b) with F_i = f_i + i^2, in int64.
__global__ void __launch_bounds__(256) k_edt_lines(int n, long long lines, long long inner,
                                                   long long outer, long long stride,
                                                   const int32_t* __restrict__ f,
                                                   int32_t* __restrict__ d2, float* __restrict__ fl,
                                                   int* __restrict__ stk)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= lines) return;
    const long long base = (c / inner) * outer + (c % inner);
    const int32_t* fp = f + base;
    int* st = stk + c;
    const long long L = lines;
    int k = -1;
    int va = 0, vb = 0;
    long long Fa = 0, Fb = 0;
    for (int q = 0; q < n; q++) {
        const int32_t fq = fp[stride * q];
        if (fq == kEdtInf) continue;
        const long long Fq = (long long)fq + (long long)q * q;
        while (k >= 1) {
            if ((Fq - Fb) * (long long)(vb - va) <= (Fb - Fa) * (long long)(q - vb)) {
                k--;
                vb = va;
                Fb = Fa;
                if (k >= 1) {
                    va = st[L * (k - 1)];
                    Fa = (long long)fp[stride * va] + (long long)va * va;
                }
            } else {
                break;
            }
        }
        k++;
        st[L * k] = q;
        va = vb;
        Fa = Fb;
        vb = q;
        Fb = Fq;
    }
    int32_t* dp = d2 ? d2 + base : nullptr;
    float* lp = fl ? fl + base : nullptr;
    if (k < 0) {
        for (int q = 0; q < n; q++) {
            if (dp) dp[stride * q] = kEdtInf;
            if (lp) lp[stride * q] = sqrt_rn(FLT_MAX);
        }
        return;
    }
    int p = 0;
    int v0 = st[0];
    long long f0 = fp[stride * v0];
    int v1 = k >= 1 ? st[L] : 0;
    long long f1 = k >= 1 ? (long long)fp[stride * v1] : 0;
    for (int q = 0; q < n; q++) {
        long long c0 = f0 + (long long)(q - v0) * (q - v0);
        while (p < k) {
            long long c1 = f1 + (long long)(q - v1) * (q - v1);
            if (c1 > c0) break;
            p++;
            v0 = v1; f0 = f1; c0 = c1;
            if (p < k) {
                v1 = st[L * (p + 1)];
                f1 = fp[stride * v1];
            }
        }
        const int32_t r = c0 > 0x7ffffffeLL ? kEdtInf : (int32_t)c0;
        if (dp) dp[stride * q] = r;
        if (lp) lp[stride * q] = sqrt_d2(r);
    }
}

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

bool edt_tile_fits(int w, int h, int d);
void launch_edt_tile_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                        hipStream_t s);
void launch_edt_tile_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                        float* fl, void* work, hipStream_t s);

size_t edt_tile_workspace_3d(int w, int h, int d);

// Tiled path (every dimension <= 512): 3D needs the y-pass fields only.
// Line path: row field g and the line stack; 3D also the y-pass field.
size_t edt_workspace(int w, int h, int d)
{
    const size_t cells = (size_t)w * h * d;
    if (edt_tile_fits(w, h, d)) return d == 1 ? 256 : edt_tile_workspace_3d(w, h, d);
    return align256(cells * 4) * (d == 1 ? 2 : 3);
}

static void edt_lines(int n, long long lines, long long inner, long long outer, long long stride,
                      const int32_t* f, int32_t* d2, float* fl, int* stk, hipStream_t s,
                      const char* name)
{
    ProfileScope ps(name, s);
    hipLaunchKernelGGL(k_edt_lines, dim3(grid1d((size_t)lines, 256)), dim3(256), 0, s, n, lines,
                       inner, outer, stride, f, d2, fl, stk);
}

static void edt_rows(int w, long long rows, const uint32_t* mask, int invert, int32_t* g,
                     hipStream_t s)
{
    const int xb = (w + 255) / 256;
    ProfileScope ps("edt_rows", s);
    hipLaunchKernelGGL(k_edt_rows, dim3((unsigned)(xb * rows)), dim3(256), 0, s, w, xb, mask,
                       invert ? 0xffffffffu : 0u, g);
}

void launch_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                   void* work, hipStream_t s)
{
    if (edt_tile_fits(w, h, 1)) {
        launch_edt_tile_2d(w, h, mask, invert, d2, fl, s);
        return;
    }
    const size_t part = align256((size_t)w * h * 4);
    int32_t* g = (int32_t*)work;
    int* stk = (int*)((char*)work + part);
    edt_rows(w, h, mask, invert, g, s);
    edt_lines(h, w, w, 0, w, g, d2, fl, stk, s, "edt_cols");
}

void launch_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                   void* work, hipStream_t s)
{
    if (edt_tile_fits(w, h, d)) {
        launch_edt_tile_3d(w, h, d, mask, invert, d2, fl, work, s);
        return;
    }
    const long long sl = (long long)w * h, cells = sl * d;
    const size_t part = align256((size_t)cells * 4);
    int32_t* g = (int32_t*)work;
    int32_t* mid = (int32_t*)((char*)work + part);
    int* stk = (int*)((char*)work + 2 * part);
    edt_rows(w, (long long)h * d, mask, invert, g, s);
    // y lines: (x, z) starts at z*w*h + x, stride w
    edt_lines(h, (long long)w * d, w, sl, w, g, mid, nullptr, stk, s, "edt_cols");
    // z lines: (x, y) starts at y*w + x, stride w*h
    edt_lines(d, sl, sl, 0, sl, mid, d2, fl, stk, s, "edt_slabs");
}

}  // namespace dolhip
