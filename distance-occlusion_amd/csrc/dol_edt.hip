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

// one wavefront a workgroup: the lines (one a thread) spread over as many CUs
// as there are wavefronts, each with an L1 of its own for its dependent reads
constexpr int kFelzThreads = 64;

// column pass: line x reads dist[q*w + x], writes td[x*h + q]
__global__ void __launch_bounds__(kFelzThreads) k_felz_cols(int w, int h, const float* __restrict__ dist,
                                                   float* __restrict__ td, int* __restrict__ v,
                                                   float* __restrict__ z)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    felz_line(h, dist + x, (size_t)w, td + (size_t)x * h, 1, v + x, z + x, (size_t)w);
}

// row pass: line y reads td[q*h + y], writes dist[y*w + q]
__global__ void __launch_bounds__(kFelzThreads) k_felz_rows(int w, int h, const float* __restrict__ td,
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
        hipLaunchKernelGGL(k_felz_cols, dim3(grid1d((size_t)w, kFelzThreads)), dim3(kFelzThreads), 0, s, w, h,
                           dist, td, v, z);
    }
    {
        ProfileScope ps("felzenszwalb_rows", s);
        hipLaunchKernelGGL(k_felz_rows, dim3(grid1d((size_t)h, kFelzThreads)), dim3(kFelzThreads), 0, s, w, h,
                           td, dist, v, z);
    }
}

// ------------------------------------------------------ exact integer EDT

// Per-row distance to the nearest set bit (clear bit if invert), squared;
// kEdtInf where the row has none.  Bits past w in the last word are ignored.
// One workgroup a row (rows of up to 1024 words): the row's words in LDS, a
// bitmap of its non-empty words (one ballot per 64 words), then for each word
// the distances from its bit 0 to the nearest set bit in the words before and
// after it (a search of the bitmap, 32 words a step), and for each cell the
// least of those and of the nearest set bits in its own word (leading and
// trailing zero counts).
constexpr int kRowWords = 1024;
constexpr int kFarBit = 1 << 30;   // no set bit on that side

__global__ void __launch_bounds__(256) k_edt_rows(int w, const uint32_t* __restrict__ mask, uint32_t invert,
                                                  int32_t* __restrict__ g)
{
    __shared__ uint32_t words[kRowWords];
    __shared__ uint32_t nzb[kRowWords / 32];
    __shared__ int dl[kRowWords], dr[kRowWords];
    const size_t row = blockIdx.x;
    const int ws = mask_stride(w), nb = (ws + 31) >> 5;
    const int tail = w - 32 * (ws - 1);
    const uint32_t last_valid = tail >= 32 ? 0xffffffffu : ((1u << tail) - 1u);
    const uint32_t* r = mask + row * ws;
    const int tid = threadIdx.x, lane = tid & 63;
    for (int base = 0; base < ws; base += 256) {
        const int j = base + tid;
        uint32_t m = 0;
        if (j < ws) {
            m = r[j] ^ invert;
            if (j == ws - 1) m &= last_valid;
            words[j] = m;
        }
        const uint64_t bal = __ballot(m != 0u);
        if (j < ws + 63 && (lane == 0 || lane == 32)) {
            const int bw = j >> 5;
            if (bw < nb) nzb[bw] = lane == 0 ? (uint32_t)bal : (uint32_t)(bal >> 32);
        }
    }
    __syncthreads();
    for (int j = tid; j < ws; j += 256) {
        int left = kFarBit, right = kFarBit;
        // the nearest non-empty word below j
        for (int k = (j - 1) >> 5, top = (j - 1) & 31; j > 0 && k >= 0; k--, top = 31) {
            // bits of nzb[k] at positions <= top
            const uint32_t bits = nzb[k] & (top == 31 ? 0xffffffffu : ((2u << top) - 1u));
            if (bits) {
                const int kl = 32 * k + (31 - __clz(bits));
                left = 32 * (j - kl) - (31 - __clz(words[kl]));
                break;
            }
        }
        // and above
        for (int k = (j + 1) >> 5, from = (j + 1) & 31; j + 1 < ws && k < nb; k++, from = 0) {
            const uint32_t bits = nzb[k] & (0xffffffffu << from);
            if (bits) {
                const int kr = 32 * k + __ffs(bits) - 1;
                right = 32 * (kr - j) + __ffs(words[kr]) - 1;
                break;
            }
        }
        dl[j] = left;
        dr[j] = right;
    }
    __syncthreads();
    for (int x = tid; x < w; x += 256) {
        const int j = x >> 5, i = x & 31;
        const uint32_t m = words[j];
        int best = min(dl[j] + i, dr[j] - i);
        const uint32_t lo = m << (31 - i);   // bits at or left of x, x at the top
        if (lo) best = min(best, (int)__clz(lo));
        const uint32_t hi = m >> i;          // bits at or right of x, x at the bottom
        if (hi) best = min(best, __ffs(hi) - 1);
        g[row * w + x] = best >= kFarBit / 2 ? kEdtInf : best * best;
    }
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
                                                   int* __restrict__ stk, const uint8_t* __restrict__ only)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= lines) return;
    if (only && !only[c]) return;   // only the lines the scan left
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

// Outward scan along lines of any length: d(p) = min_q f(q) + (p - q)^2 by
// reading f at p -/+ k for k = 1, 2, ... until k^2 reaches every output's best
// so far (no sample further out can improve it) or the line is exhausted.
// Lane = x (64 consecutive lines of one plane: coalesced reads of every row);
// a wavefront computes kScanRows consecutive outputs of its lines, kScanGroup
// steps at a time from a sliding window of rows.  Values are exact int32
// (kEdtInf = no site; the additions saturate, so a missing site never wins).
// A wavefront that would need more than kScanCap steps (long empty stretches)
// marks its lines in `flags`, and those lines are redone by k_edt_lines.
constexpr int kScanRows = 8;
constexpr int kScanGroup = 8;
constexpr int kScanCap = 1024;

__device__ __forceinline__ uint32_t add_sat_u32(uint32_t a, uint32_t b)
{
    return __builtin_elementwise_add_sat(a, b);
}

__global__ void __launch_bounds__(256) k_edt_scan(int n, int w, long long plane_stride, long long stride,
                                                  const int32_t* __restrict__ f, int32_t* __restrict__ d2,
                                                  float* __restrict__ fl, uint8_t* __restrict__ flags)
{
    constexpr int R = kScanRows, G = kScanGroup, M = G + R - 1;
    const int lane = threadIdx.x & 63;
    const int x = blockIdx.x * 64 + lane;
    const long long plane = blockIdx.y;
    const int p0 = (blockIdx.z * 4 + (threadIdx.x >> 6)) * R;   // the wavefront's first output row
    if (p0 >= n) return;
    const bool active = x < w;
    const int32_t* fp = f + plane * plane_stride + (active ? x : 0);
    auto at = [&](int q) -> uint32_t { return (q >= 0 && q < n) ? (uint32_t)fp[stride * q] : 0xffffffffu; };
    uint32_t best[R];
#pragma unroll
    for (int r = 0; r < R; r++) best[r] = active ? at(p0 + r) : 0u;
    const int last = min(p0 + R, n) - 1;   // the wavefront's last output row
    int k = 1;
    bool capped = false;
    for (;;) {
        uint32_t top = 0;
#pragma unroll
        for (int r = 0; r < R; r++) top = max(top, best[r]);
        // samples left at distance k from some output row: below its last row
        // or above its first
        const bool left = last - k >= 0 || p0 + k < n;
        const bool more = active && left && (uint32_t)k * (uint32_t)k < top;
        if (__builtin_amdgcn_ballot_w64(more) == 0) break;
        if (k > kScanCap) {
            capped = true;
            break;
        }
        // rows p0 + r - k - t (below) are a[G - 1 - t + r], rows p0 + r + k + t (above) b[t + r]
        uint32_t a[M], b[M];
#pragma unroll
        for (int u = 0; u < M; u++) {
            a[u] = at(p0 - k - (G - 1) + u);
            b[u] = at(p0 + k + u);
        }
#pragma unroll
        for (int t = 0; t < G; t++) {
            const uint32_t sq = (uint32_t)((k + t) * (k + t));
#pragma unroll
            for (int r = 0; r < R; r++)
                best[r] = min(best[r], min(add_sat_u32(a[G - 1 - t + r], sq), add_sat_u32(b[t + r], sq)));
        }
        k += G;
    }
    if (!active) return;
    if (capped) {   // the wavefront's lines are redone whole (every writer stores 1)
        flags[plane * w + x] = 1;
        return;
    }
    const long long base = plane * plane_stride + x;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (p0 + r >= n) break;
        const int32_t v = best[r] >= (uint32_t)kEdtInf ? kEdtInf : (int32_t)best[r];
        if (d2) d2[base + stride * (p0 + r)] = v;
        if (fl) fl[base + stride * (p0 + r)] = sqrt_d2(v);
    }
}

__global__ void k_clear_flags(size_t n, uint8_t* __restrict__ flags)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) flags[i] = 0;
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
// Line path: the row field g, the line stack, for 3D the y-pass field, and a
// flag byte a line for the lines the scan leaves to the stack kernel.
size_t edt_workspace(int w, int h, int d)
{
    const size_t cells = (size_t)w * h * d;
    if (edt_tile_fits(w, h, d)) return d == 1 ? 256 : edt_tile_workspace_3d(w, h, d);
    const size_t lines = (size_t)w * (d > h ? d : h);
    return align256(cells * 4) * (d == 1 ? 2 : 3) + align256(lines);
}

// The transform along `planes` x w lines of length n: line (plane, x) starts
// at plane * plane_stride + x and steps by `stride`.  The scan first; the
// lines it flags again with the stack kernel (whose line c = plane * w + x
// has the same start with inner = w, outer = plane_stride).
static void edt_lines(int n, int w, long long planes, long long plane_stride, long long stride,
                      const int32_t* f, int32_t* d2, float* fl, int* stk, uint8_t* flags, hipStream_t s,
                      const char* name)
{
    ProfileScope ps(name, s);
    const size_t lines = (size_t)planes * w;
    hipLaunchKernelGGL(k_clear_flags, dim3(grid1d(lines, 256)), dim3(256), 0, s, lines, flags);
    const dim3 grid((unsigned)((w + 63) / 64), (unsigned)planes, (unsigned)((n + 4 * kScanRows - 1) / (4 * kScanRows)));
    hipLaunchKernelGGL(k_edt_scan, grid, dim3(256), 0, s, n, w, plane_stride, stride, f, d2, fl, flags);
    hipLaunchKernelGGL(k_edt_lines, dim3(grid1d(lines, 256)), dim3(256), 0, s, n, (long long)lines,
                       (long long)w, plane_stride, stride, f, d2, fl, stk, (const uint8_t*)flags);
}

static void edt_rows(int w, long long rows, const uint32_t* mask, int invert, int32_t* g,
                     hipStream_t s)
{
    ProfileScope ps("edt_rows", s);
    hipLaunchKernelGGL(k_edt_rows, dim3((unsigned)rows), dim3(256), 0, s, w, mask, invert ? 0xffffffffu : 0u, g);
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
    uint8_t* flags = (uint8_t*)work + 2 * part;
    edt_rows(w, h, mask, invert, g, s);
    // columns: one plane of w lines, stride w
    edt_lines(h, w, 1, 0, w, g, d2, fl, stk, flags, s, "edt_cols");
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
    uint8_t* flags = (uint8_t*)work + 3 * part;
    edt_rows(w, (long long)h * d, mask, invert, g, s);
    // y lines: plane z, (x, z) starts at z*w*h + x, stride w
    edt_lines(h, w, d, sl, w, g, mid, nullptr, stk, flags, s, "edt_cols");
    // z lines: plane y, (x, y) starts at y*w + x, stride w*h
    edt_lines(d, w, h, w, sl, mid, d2, fl, stk, flags, s, "edt_slabs");
}

}  // namespace dolhip
