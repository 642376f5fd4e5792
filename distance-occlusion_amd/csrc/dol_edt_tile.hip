// dol_edt_tile.hip -- LDS-tiled exact squared EDT for lines of up to 512
// samples (volumes up to 512^3, images up to 512 high).
//
// Separable form (Felzenszwalb): per line, d(p) = min_q f(q) + (p - q)^2,
// with f the squared distance so far (x distance of the row's bits first).
// A workgroup holds a tile of lines in LDS and computes the transform of
// each line in one of two exact ways:
//   scan      read f at p -/+ k for k = 1, 2, ... and stop once k^2 reaches
//             the best value so far; the steps are about the distance
//             computed, so on volumes with sites near every sample this is
//             far cheaper than building an envelope.  Tried first.
//   envelope  the lower envelope of the parabolas F(i) - 2 q i + q^2
//             (F = f + i^2): band hulls in register stacks, a merge tree of
//             bridge walks over an LDS survivor bitmask, then interval-start
//             marks and a running maximum for the owners.  Used for a tile
//             whose scan would need more than a capped number of steps.
// Wide tiles (32 lines, even widths) scan two lines at once as clamped
// 16-bit values packed in one word; a result below 0xffff is exact, and a
// tile with a larger or capped result is redone exactly, half a tile at a
// time, by the 16-line envelope.  Other widths use 16-line tiles throughout.
//
// edt3d_xy: one workgroup per (slice z, tile of x): the x pass straight from
// the bit rows into LDS, then the y transform; writes the y field (for 2D,
// the final int32 and float outputs).  The wide kernel hands the field to
// the z pass as 16-bit values, plus an exact int32 copy for the rare tiles
// whose values reach 0xffff and a per-tile flag naming the field to read.
// edt3d_z: one workgroup per (row y, tile of x): the field's column tile
// along z into LDS, the z transform, int32 and float distances out.
#include "dol_common.hpp"

#include <cfloat>

namespace dolhip {

namespace tile {

constexpr int kCols = 16;      // lines per tile (half a mask word of x)
constexpr int kBand = 16;      // sites per band
constexpr int kMaxN = 512;     // longest line held in LDS
constexpr int kBands = kMaxN / kBand;
constexpr int kThreads = kCols * kBands;   // 512: thread = band * kCols + column
constexpr int kNone = -1;

__device__ __forceinline__ int hi_bit(uint32_t v) { return 31 - __clz(v); }
__device__ __forceinline__ int lo_bit(uint32_t v) { return __ffs(v) - 1; }

// survivor bitmask: word k of column c holds band k in its low kBand bits
__device__ __forceinline__ uint32_t bits_upto(int i) { return i >= 31 ? 0xffffffffu : ((2u << i) - 1u); }

struct Bits {
    uint32_t* w;   // [kBands][kCols]
    __device__ uint32_t get(int k, int c) const { return w[k * kCols + c]; }
    __device__ void put(int k, int c, uint32_t v) { w[k * kCols + c] = v; }

    // highest survivor in [lo, pos), or -1
    __device__ int prev(int c, int pos, int lo) const
    {
        if (pos <= lo) return -1;
        int k = (pos - 1) / kBand;
        uint32_t m = get(k, c) & bits_upto((pos - 1) % kBand);
        for (;;) {
            if (m) {
                const int r = k * kBand + hi_bit(m);
                return r >= lo ? r : -1;
            }
            if (--k < lo / kBand) return -1;
            m = get(k, c);
        }
    }
    // lowest survivor in (pos, hi), or -1
    __device__ int next(int c, int pos, int hi) const
    {
        const int p = pos + 1;
        if (p >= hi) return -1;
        int k = p / kBand;
        const int kend = (hi - 1) / kBand;
        uint32_t m = get(k, c) & ~((1u << (p % kBand)) - 1u);
        for (;;) {
            if (m) {
                const int r = k * kBand + lo_bit(m);
                return r < hi ? r : -1;
            }
            if (++k > kend) return -1;
            m = get(k, c);
        }
    }
    // clear survivors strictly between a and b
    __device__ void clear_between(int c, int a, int b)
    {
        const int p = a + 1, q = b - 1;
        if (p > q) return;
        const int k0 = p / kBand, k1 = q / kBand;
        for (int k = k0; k <= k1; k++) {
            const uint32_t lo = k == k0 ? ~((1u << (p % kBand)) - 1u) : 0xffffffffu;
            const uint32_t hi = k == k1 ? bits_upto(q % kBand) : 0xffffffffu;
            put(k, c, get(k, c) & ~(lo & hi));
        }
    }
};

// ---- LDS words.  Sample q of column c is word cell(q, c) = q * kCols + c
// (rows at a constant stride, so the scan's reads take immediate offsets):
// bits 0..19 hold f (kNoSite: none; every f is below 2^20 on this path),
// bits 20..29 the mark of phase 3 (index + 1 of the site whose interval
// starts at q).

constexpr uint32_t kValueMask = 0xfffffu;
constexpr uint32_t kNoSite = 0xfffffu;
// band maxima of phase 3b: one 16-byte aligned row per column, padded so that
// 16 lanes' rows start in different bank quads
constexpr int kBmStride = kBands + 4;
constexpr int kBmWords = kCols * kBmStride > kBands * kCols ? kCols * kBmStride : kBands * kCols;
constexpr int kWords = kMaxN * kCols;

__device__ __forceinline__ int cell(int q, int c) { return q * kCols + c; }

__device__ __forceinline__ int fval(const uint32_t* fs, int c, int q)
{
    const uint32_t v = fs[cell(q, c)] & kValueMask;
    return v == kNoSite ? kNone : (int)v;
}

// smallest k with k * b >= a, for b in [2, 1022] and |a| < 2^20: the float
// quotient is within 1/8 of a / b, so one integer correction makes it exact
__device__ __forceinline__ int ceil_div(int a, int b)
{
    int k = (int)__builtin_ceilf((float)a * __builtin_amdgcn_rcpf((float)b));
    if (k * b < a) k++;
    else if ((k - 1) * b >= a) k--;
    return k;
}

// v lies on or above the chord u-w (its envelope interval is empty): the
// intersection of u and v is not left of that of v and w.
__device__ __forceinline__ bool removable(int u, int Fu, int v, int Fv, int w, int Fw)
{
    return (Fv - Fu) * (w - v) >= (Fw - Fv) * (v - u);
}

// Envelope of one tile of n-sample lines held in fs.  Out(c, q, value) stores
// a result (kEdtInf where the line has no site).  The survivor bitmask's
// storage is reused for the band maxima of phase 3b.
template <class Out>
__device__ __forceinline__ void envelope(int n, uint32_t* fs, Out& out, Bits bits, int tid)
{
    const int c = tid & (kCols - 1);
    const int b = tid / kCols;
    const int nb = (n + kBand - 1) / kBand;
    const int lo = b * kBand;
    const int hi = min(n, lo + kBand);

    // 1. band hull, register stack.  The band's samples are read up front
    //    (rows past n hold stale words and are masked off), so the loads
    //    overlap, and the loop is unrolled over them.
    if (b < nb) {
        uint32_t raw[kBand];
#pragma unroll
        for (int i = 0; i < kBand; i++) raw[i] = fs[cell(lo + i, c)];
        uint32_t st = 0;
        int t = -1, s = -1, Ft = 0, Fs = 0;
#pragma unroll
        for (int i = 0; i < kBand; i++) {
            const int q = lo + i;
            const uint32_t v = raw[i] & kValueMask;
            const bool site = q < hi && v != kNoSite;
            const int Fq = (int)v + q * q;
            while (site && s >= 0 && removable(s, Fs, t, Ft, q, Fq)) {
                st &= ~(1u << (t - lo));
                t = s;
                Ft = Fs;
                const uint32_t below = st & ((1u << (t - lo)) - 1u);
                if (below) {
                    s = lo + hi_bit(below);
                    Fs = fval(fs, c, s) + s * s;
                } else {
                    s = -1;
                }
            }
            st |= (uint32_t)site << i;
            s = site ? t : s;
            Fs = site ? Ft : Fs;
            t = site ? q : t;
            Ft = site ? Fq : Ft;
        }
        bits.put(b, c, st);
    }
    __syncthreads();

    // 2. tree of bridge merges.  A round's merges are dealt to the first
    //    threads of the workgroup (kCols consecutive threads per group of
    //    bands), so the walks fill whole wavefronts.  Each walk keeps the
    //    bitmask words of its two ends in registers and the sample of each
    //    end's next candidate loaded ahead, so a step waits for one LDS read.
    for (int span = 1; span < nb; span <<= 1) {
        const int groups = (nb - span + 2 * span - 1) / (2 * span);
        if (tid < groups * kCols) {
            const int mc = tid % kCols;
            const int gb = (tid / kCols) * 2 * span;   // first band of the group
            const int glo = gb * kBand;
            const int mid = (gb + span) * kBand;
            const int end = min(n, (gb + 2 * span) * kBand);
            int a = bits.prev(mc, mid, glo);
            int r = bits.next(mc, mid - 1, end);
            if (a >= 0 && r >= 0) {
                const int kfirst = gb, klast = (end - 1) / kBand;
                int ka = a / kBand, kr = r / kBand;
                uint32_t wa = bits.get(ka, mc), wr = bits.get(kr, mc);
                auto right_of = [&](int v) -> int {   // next survivor after v (in word kr)
                    uint32_t m = wr & ~bits_upto(v % kBand);
                    while (!m) {
                        if (++kr > klast) return -1;
                        m = wr = bits.get(kr, mc);
                    }
                    return kr * kBand + lo_bit(m);
                };
                auto left_of = [&](int v) -> int {    // survivor before v (in word ka)
                    uint32_t m = wa & ((1u << (v % kBand)) - 1u);
                    while (!m) {
                        if (--ka < kfirst) return -1;
                        m = wa = bits.get(ka, mc);
                    }
                    return ka * kBand + hi_bit(m);
                };
                auto F = [&](int v) { return v >= 0 ? fval(fs, mc, v) + v * v : 0; };
                int Fa = F(a), Fr = F(r);
                int rn = right_of(r), ap = left_of(a);
                int Frn = F(rn), Fap = F(ap);
                for (;;) {
                    if (rn >= 0 && removable(a, Fa, r, Fr, rn, Frn)) {
                        r = rn;
                        Fr = Frn;
                        rn = right_of(r);
                        Frn = F(rn);
                        continue;
                    }
                    if (ap >= 0 && removable(ap, Fap, a, Fa, r, Fr)) {
                        a = ap;
                        Fa = Fap;
                        ap = left_of(a);
                        Fap = F(ap);
                        continue;
                    }
                    break;
                }
                bits.clear_between(mc, a, r);
            }
        }
        __syncthreads();
    }

    // 3a. interval starts and marks, without branches: a mark that does not
    //    apply ORs zero into a word in range
    if (b < nb) {
        const uint32_t mine = bits.get(b, c);
        int pv = bits.prev(c, lo, 0);                  // hull predecessor of the band
        int pF = pv >= 0 ? fval(fs, c, pv) + pv * pv : 0;
        const int sv = bits.next(c, hi - 1, n);        // hull successor of the band
        const int sF = sv >= 0 ? fval(fs, c, sv) + sv * sv : 0;
        uint32_t raw[kBand];
#pragma unroll
        for (int i = 0; i < kBand; i++) raw[i] = fs[cell(lo + i, c)];
        int cur = -1, cur_l = 0;
        auto mark = [&](bool on, int v, int from, int to) {   // v owns [from, to)
            from = max(from, 0);
            on = on && from < min(to, n);
            atomicOr(&fs[cell(min(from, kMaxN - 1), c)], on ? (uint32_t)(v + 1) << 20 : 0u);
        };
#pragma unroll
        for (int i = 0; i < kBand; i++) {
            const bool sur = (mine >> i) & 1u;
            const int v = lo + i;
            const int Fv = (int)(raw[i] & kValueMask) + v * v;
            const int l = pv < 0 ? -kMaxN : ceil_div(Fv - pF, 2 * (v - pv));
            mark(sur && cur >= 0, cur, cur_l, l);
            cur = sur ? v : cur;
            cur_l = sur ? l : cur_l;
            pv = sur ? v : pv;
            pF = sur ? Fv : pF;
        }
        const int l = sv < 0 ? n : ceil_div(sF - pF, max(2 * (sv - pv), 2));
        mark(cur >= 0, cur, cur_l, l);
    }
    __syncthreads();

    // 3b. owners: running maximum of the marks, carried across bands (the
    //    band maxima are kept column-major, kBmStride words a column, in the
    //    bitmask's storage)
    constexpr uint32_t kMarkMask = ~kValueMask;
    uint32_t* band_max = bits.w;
    uint32_t raw[kBand];
    if (b < nb) {
        uint32_t m = 0;
#pragma unroll
        for (int i = 0; i < kBand; i++) {
            raw[i] = lo + i < hi ? fs[cell(lo + i, c)] & kMarkMask : 0u;
            m = max(m, raw[i]);
        }
        band_max[c * kBmStride + b] = m;
    }
    __syncthreads();
    if (b < nb && out.active(c)) {
        uint32_t run = 0;
        const uint4* bm4 = (const uint4*)(band_max + c * kBmStride);
#pragma unroll
        for (int k4 = 0; k4 < kBands / 4; k4++) {
            const uint4 v = bm4[k4];
            run = max(run, 4 * k4 + 0 < b ? v.x : 0u);
            run = max(run, 4 * k4 + 1 < b ? v.y : 0u);
            run = max(run, 4 * k4 + 2 < b ? v.z : 0u);
            run = max(run, 4 * k4 + 3 < b ? v.w : 0u);
        }
        int own[kBand];
        uint32_t fa[kBand];
#pragma unroll
        for (int i = 0; i < kBand; i++) {
            run = max(run, raw[i]);
            own[i] = (int)(run >> 20) - 1;
            fa[i] = fs[cell(max(own[i], 0), c)] & kValueMask;
        }
#pragma unroll
        for (int i = 0; i < kBand; i++) {
            const int q = lo + i;
            const int a = own[i];
            if (q < hi) out(c, q, a < 0 ? kEdtInf : (int)fa[i] + (q - a) * (q - a));
        }
    }
}

// ---- outputs

struct FieldOut {   // int32 field at (x0 + c, q, z) of a w x h x d volume
    int32_t* f2;
    int w, h, x0, z;
    __device__ bool active(int c) const { return x0 + c < w; }
    __device__ void operator()(int c, int q, int v) const
    {
        f2[((uint32_t)z * h + q) * (uint32_t)w + x0 + c] = v;
    }
    // columns c, c + 1 (c even; w, x0 even and f2 8-byte aligned)
    __device__ void pair(int c, int q, int v0, int v1) const
    {
        int32_t* p = f2 + ((uint32_t)z * h + q) * (uint32_t)w + x0 + c;
        if (x0 + c + 1 < w) *(int2*)p = make_int2(v0, v1);
        else if (x0 + c < w) *p = v0;
    }
};

// the field of the wide tiles as 16-bit values clamped to 0xffff ("0xffff or
// more"; f16) and, written only by tiles whose field reaches that, the exact
// int32 field (f32)
struct Field16Out {   // pairs of exact values below 0xffff, or 0xffff
    uint16_t* f16;
    int w, h, x0, z;
    __device__ bool active(int c) const { return x0 + c < w; }
    // columns c, c + 1 (c even; w, x0 even and f16 4-byte aligned)
    __device__ void pair(int c, int q, int v0, int v1) const
    {
        const uint32_t i = ((uint32_t)z * h + q) * (uint32_t)w + x0 + c;
        *(uint32_t*)(f16 + i) = min((uint32_t)v0, 0xffffu) | min((uint32_t)v1, 0xffffu) << 16;
    }
};

struct FieldBothOut {   // exact value into f32, clamped into f16
    int32_t* f32;
    uint16_t* f16;
    int w, h, x0, z;
    __device__ bool active(int c) const { return x0 + c < w; }
    __device__ void operator()(int c, int q, int v) const
    {
        const uint32_t i = ((uint32_t)z * h + q) * (uint32_t)w + x0 + c;
        f32[i] = v;
        f16[i] = (uint16_t)min((uint32_t)v, 0xffffu);
    }
};

// Every squared distance of this path is at most 3 * 511^2 < 2^20 (or
// kEdtInf), the domain of sqrt_d2_small.
struct FinalOut {   // int32 + float at cell index base(q) + x (< 2^27 on this path)
    int32_t* d2;
    float* fl;
    uint32_t row_stride;   // elements between consecutive q
    uint32_t base;         // element index of (x0, q = 0)
    int valid;             // columns < valid are inside the volume
    __device__ bool active(int c) const { return c < valid; }
    __device__ void operator()(int c, int q, int v) const
    {
        const uint32_t i = base + (uint32_t)q * row_stride + (uint32_t)c;
        if (d2) d2[i] = v;
        if (fl) fl[i] = sqrt_d2_small(v);
    }
    // columns c, c + 1 (c even; base and row_stride even, outputs 8-byte
    // aligned).  Buffer stores take the 32-bit byte offset as it is (no 64-bit
    // address per store); nontemporal: the outputs are not read again here,
    // and these stores leave the caches to the z pass's input.
    __device__ void pair(int c, int q, int v0, int v1) const
    {
        typedef unsigned u32x2 __attribute__((__vector_size__(8)));
        constexpr int kNontemporal = 2;
        const uint32_t i = base + (uint32_t)q * row_stride + (uint32_t)c;
        if (c + 1 < valid) {
            const int off = (int)(i * 4u);
            if (d2)
                __builtin_amdgcn_raw_buffer_store_b64((u32x2){(uint32_t)v0, (uint32_t)v1},
                                                      __builtin_amdgcn_make_buffer_rsrc(d2, 0, -1, kRsrcWord3),
                                                      off, 0, kNontemporal);
            if (fl)
                __builtin_amdgcn_raw_buffer_store_b64(
                    (u32x2){__float_as_uint(sqrt_d2_small(v0)), __float_as_uint(sqrt_d2_small(v1))},
                    __builtin_amdgcn_make_buffer_rsrc(fl, 0, -1, kRsrcWord3), off, 0, kNontemporal);
        } else if (c < valid) {
            (*this)(c, q, v0);
        }
    }
};

// ---- outward scan, tried first on every tile
//
// d(p) = min_q f(q) + (p - q)^2 is found by reading the samples p -/+ k for
// k = 1, 2, ... and stopping once k^2 >= the best value so far (every sample
// further out costs at least k^2) or every sample of the line has been seen.
// The steps are about the distance computed, so on volumes with sites near
// every sample this takes far fewer instructions than the envelope.  A
// wavefront computes kRowsPerWave consecutive samples of all the tile's lines
// at once: k is the same for all its active lanes (the squares are scalars)
// and each read covers 64 consecutive LDS words.  Steps go kStepGroup at a
// time; while the wavefront's rows p -/+ (k + kStepGroup - 1) stay inside
// the line the reads take immediate offsets, and near the line's ends the
// row is clamped (a clamped read repeats an end sample at a larger distance
// than it was already counted at, so it never lowers the minimum).  If some
// output would need more than kScanCap steps the tile is left to the
// envelope (returns false).
constexpr int kScanCap = 160;
constexpr int kStepGroup = 8;
constexpr int kRowsPerWave = 64 / kCols;
constexpr int kRowsPerPass = kThreads / kCols;

template <class Out>
__device__ __forceinline__ bool scan(int n, const uint32_t* fs, Out& out, int* abort_flag)
{
    const int tid = threadIdx.x;
    const int c = tid % kCols;
    const int r0 = tid / kCols;
    const int wr0 = (tid / 64) * kRowsPerWave;   // the wavefront's first row
    const bool active = out.active(c);
    for (int base = 0; base < n; base += kRowsPerPass) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
        const int pw = base + wr0;
        if (pw >= n) break;
        const int p = base + r0;
        const bool live = active && p < n;
        const int reach = max(p, n - 1 - p);   // farthest sample of the line
        int best = live ? (int)(fs[cell(p, c)] & kValueMask) : 0;
        int k = 1;
        bool fail = false;
        for (;;) {
            if (k * k >= best || k > reach) break;
            if (k > kScanCap) {
                fail = true;
                break;
            }
            k = __builtin_amdgcn_readfirstlane(k);   // the same in every active lane
            if (pw - (k + kStepGroup - 1) >= 0 && pw + kRowsPerWave - 1 + k + kStepGroup - 1 < n) {
                const uint32_t* lo = fs + cell(p - k - (kStepGroup - 1), c);
                const uint32_t* hi = fs + cell(p + k, c);
#pragma unroll
                for (int t = 0; t < kStepGroup; t++) {
                    const int a = (int)lo[(kStepGroup - 1 - t) * kCols], b = (int)hi[t * kCols];
                    best = min(best, min(a, b) + (k + t) * (k + t));
                }
            } else {
#pragma unroll
                for (int t = 0; t < kStepGroup; t++) {
                    const int kk = k + t;
                    const int a = (int)fs[cell(max(p - kk, 0), c)];
                    const int b = (int)fs[cell(min(p + kk, n - 1), c)];
                    best = min(best, min(a, b) + kk * kk);
                }
            }
            k += kStepGroup;
        }
        if (fail) __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (live) out(c, p, best >= (int)kNoSite ? kEdtInf : best);
    }
    __syncthreads();
    return *abort_flag == 0;
}

// every output of a tile whose lines have no site
template <class Out>
__device__ void no_sites(int n, Out& out)
{
    const int c = threadIdx.x % kCols;
    if (!out.active(c)) return;
    for (int p = threadIdx.x / kCols; p < n; p += kRowsPerPass) out(c, p, kEdtInf);
}

// ---- the same scan over wide tiles of 2 * kPairs lines held as 16-bit
// values, two adjacent lines in each LDS word (word q * kPairs + i holds
// lines 2i and 2i + 1 at sample q), so one read and three packed operations
// serve two lines.  Values are clamped to kClamp = 2^16 - 1 and the sums
// saturate there.  A result below kClamp is exact: its minimising sample was
// not clamped and its sum did not saturate, and every clamped or saturated
// candidate is at least kClamp.  An output that stays at kClamp (a line with
// no site within reach), or that needs more than kPairScanCap steps, fails
// the tile, which is then redone exactly; the cap keeps every square k^2
// below 2^16.
constexpr int kPairs = 16;
constexpr int kWideCols = 2 * kPairs;                 // lines per wide tile
constexpr int kPairRowsPerWave = 64 / kPairs;
constexpr int kPairRowsPerPass = kThreads / kPairs;
constexpr uint32_t kClamp = 0xffff;
static_assert(kMaxN * kPairs <= kWords, "a wide tile fits the LDS words");

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
// saturating (v_pk_add_u16 with clamp)
__device__ __forceinline__ uint32_t pk_add_sat(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2, a),
                                                                      __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b)   // modular (v_pk_add_u16)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b)   // modular (v_pk_sub_u16)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_mul(uint32_t a, uint32_t b)   // low halves of the products (v_pk_mul_lo_u16)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) * __builtin_bit_cast(u16x2, b));
}
// the packed x pass clamps x distances at 255, so the y scan's results are
// exact below 255^2 (a clamped sample's true value is at least that); the z
// scan's input is exact below kClamp
constexpr uint32_t kXExact = 255u * 255u;

// A lane computes kLaneRows consecutive samples of its two lines: the rows
// its samples read at steps k .. k + kStepGroup - 1 overlap, so each group
// reads kStepGroup + kLaneRows - 1 rows on either side instead of
// kStepGroup * kLaneRows.
#ifndef DOL_LANE_ROWS
#define DOL_LANE_ROWS 4   // a pass of 8 wavefronts x 16 rows: 512-sample lines in 4 passes
#endif
#ifndef DOL_PAIR_GROUP
#define DOL_PAIR_GROUP 8
#endif
constexpr int kLaneRows = DOL_LANE_ROWS;
constexpr int kPairStepGroup = DOL_PAIR_GROUP;
#ifndef DOL_PAIR_CAP
#define DOL_PAIR_CAP (256 - kPairStepGroup)
#endif
constexpr int kPairScanCap = DOL_PAIR_CAP;
static_assert((kPairScanCap + kPairStepGroup - 1) * (kPairScanCap + kPairStepGroup - 1) <= 0xffff,
              "16-bit squares");
constexpr int kPairWaveRows = kPairRowsPerWave * kLaneRows;
constexpr int kPairPassRows = (kThreads / 64) * kPairWaveRows;
static_assert(DOL_PAIR_GROUP + kPairWaveRows - 1 + 7 <= 32, "a side's rows lie in one 32-row window");

// Window minima for skipping: word b * kPairs + i holds the packed minima of
// column pair i over rows 8b .. 8b + 31 (rows past n count as no
// candidate).  A group's rows on one side span at most kPairStepGroup +
// kPairWaveRows - 1 rows from a start within 7 rows of some 8b, so one word
// bounds them all from below.
constexpr int kWinBlocks = kMaxN / 8;
constexpr int kWinWords = kWinBlocks * kPairs;
static_assert(kWinWords % kThreads == 0, "window minima split evenly");
__device__ __forceinline__ void window_minima(int n, const uint32_t* hs, uint32_t* mwin, int tid)
{
    constexpr int kPer = kWinWords / kThreads;
    for (int e = tid; e < kWinWords; e += kThreads) {
        const int b = e / kPairs, i = e % kPairs;
        uint32_t m = 0xffffffffu;
        if (8 * b + 8 <= n) {   // the whole block inside the line (every block of a full line)
#pragma unroll
            for (int r = 0; r < 8; r++) m = pk_min(m, hs[(8 * b + r) * kPairs + i]);
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++)
                if (8 * b + r < n) m = pk_min(m, hs[(8 * b + r) * kPairs + i]);
        }
        mwin[e] = m;
    }
    __syncthreads();
    uint32_t v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = tid + j * kThreads, b = e / kPairs, i = e % kPairs;
        uint32_t m = mwin[e];
        // blocks past the last repeat it (a minimum taken twice is the same)
#pragma unroll
        for (int s = 1; s < 4; s++) m = pk_min(m, mwin[min(b + s, kWinBlocks - 1) * kPairs + i]);
        v[j] = m;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) mwin[tid + j * kThreads] = v[j];
    __syncthreads();
}

// M consecutive rows row0 .. of column pair i; rows outside [0, n) read as
// 0xffff (no candidate).  inside: all of them are inside for every lane of
// the wavefront, so the reads take immediate offsets.
template <int M>
__device__ __forceinline__ void pair_rows(const uint32_t* hs, int row0, int n, int i, bool inside, uint32_t* v)
{
    if (inside) {
        const uint32_t* src = hs + row0 * kPairs + i;
#pragma unroll
        for (int u = 0; u < M; u++) v[u] = src[u * kPairs];
    } else {
#pragma unroll
        for (int u = 0; u < M; u++) {
            const int q = row0 + u;
            const uint32_t x = hs[min(max(q, 0), n - 1) * kPairs + i];
            v[u] = q >= 0 && q < n ? x : 0xffffffffu;
        }
    }
}

template <uint32_t kExact, class Out>
__device__ __forceinline__ bool scan_pairs(int n, const uint32_t* hs, const uint32_t* mwin, Out& out,
                                           int* abort_flag)
{
    constexpr int R = kLaneRows, G = kPairStepGroup;
    const int tid = threadIdx.x;
    const int i = tid % kPairs, c = 2 * i;
    const int wr0 = __builtin_amdgcn_readfirstlane((tid / 64) * kPairWaveRows);   // the wavefront's first row
    const int rr = ((tid % 64) / kPairs) * R;            // the lane's first row in it
    const bool active = out.active(c);
    for (int base = 0; base < n; base += kPairPassRows) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
        const int pw = base + wr0;
        if (pw >= n) break;
        const int p = pw + rr;
        const int reach = max(p + R - 1, n - 1 - p);   // farthest sample from any of the rows
        uint32_t best[R];
#pragma unroll
        for (int r = 0; r < R; r++) best[r] = active && p + r < n ? hs[(p + r) * kPairs + i] : 0u;
        // the loop is the wavefront's: it steps while any lane needs more
        // (further steps never change a finished lane's result), so k and the
        // squares are scalars and the loop needs no per-lane exits
        int k = 1;
        bool fail = false;
        for (;;) {
            uint32_t top = 0;
#pragma unroll
            for (int r = 0; r < R; r++) top = max(top, max(best[r] & 0xffffu, best[r] >> 16));
            const bool more = k * k < (int)top && k <= reach;
            if (__builtin_amdgcn_ballot_w64(more) == 0) break;
            if (k > kPairScanCap) {
                fail = true;
                break;
            }
            // rows p + r - k - t (below) are a[G - 1 - t + r], rows p + r + k + t
            // (above) b[t + r].  Near the ends of the line a side may lie
            // partly or wholly outside it: rows outside give no candidate.
            constexpr int M = G + R - 1;
            const bool lo_all = pw - k - (G - 1) >= 0;
            const bool hi_all = pw + kPairWaveRows - 1 + k + G - 1 < n;
            // a side is skipped when it lies outside the line, or when the
            // least value in its rows plus k^2 reaches every lane's largest
            // best so far (no candidate there can improve an output)
            const uint32_t kk = (uint32_t)(k * k);
            const uint32_t room = top > kk ? top - kk : 0u;
            bool lo_skip = pw + kPairWaveRows - 1 - k < 0;
            bool hi_skip = pw + k >= n;
            if (!lo_skip) {
                const uint32_t m = mwin[(max(pw - k - (G - 1), 0) >> 3) * kPairs + i];
                lo_skip = __builtin_amdgcn_ballot_w64(min(m & 0xffffu, m >> 16) < room) == 0;
            }
            if (!hi_skip) {
                const uint32_t m = mwin[((pw + k) >> 3) * kPairs + i];
                hi_skip = __builtin_amdgcn_ballot_w64(min(m & 0xffffu, m >> 16) < room) == 0;
            }
            uint32_t a[M], b[M];
            if (lo_skip && hi_skip) {
                // nothing in this group
            } else if (lo_skip) {
                pair_rows<M>(hs, p + k, n, i, hi_all, b);
#pragma unroll
                for (int t = 0; t < G; t++) {
                    const uint32_t sq = (uint32_t)((k + t) * (k + t)) * 0x10001u;
#pragma unroll
                    for (int r = 0; r < R; r++) best[r] = pk_min(best[r], pk_add_sat(b[t + r], sq));
                }
            } else if (hi_skip) {
                pair_rows<M>(hs, p - k - (G - 1), n, i, lo_all, a);
#pragma unroll
                for (int t = 0; t < G; t++) {
                    const uint32_t sq = (uint32_t)((k + t) * (k + t)) * 0x10001u;
#pragma unroll
                    for (int r = 0; r < R; r++) best[r] = pk_min(best[r], pk_add_sat(a[G - 1 - t + r], sq));
                }
            } else {
                pair_rows<M>(hs, p - k - (G - 1), n, i, lo_all, a);
                pair_rows<M>(hs, p + k, n, i, hi_all, b);
#pragma unroll
                for (int t = 0; t < G; t++) {
                    const uint32_t sq = (uint32_t)((k + t) * (k + t)) * 0x10001u;
#pragma unroll
                    for (int r = 0; r < R; r++)
                        best[r] = pk_min(best[r], pk_add_sat(pk_min(a[G - 1 - t + r], b[t + r]), sq));
                }
            }
            k += G;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (max(best[r] & 0xffffu, best[r] >> 16) >= kExact) fail = true;
        if (fail) {
            __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (active) {
#pragma unroll
            for (int r = 0; r < R; r++)
                if (p + r < n) out.pair(c, p + r, (int)(best[r] & 0xffffu), (int)(best[r] >> 16));
        }
    }
    __syncthreads();
    return *abort_flag == 0;
}

template <class Out>
__device__ void no_sites_pairs(int n, Out& out)
{
    const int c = 2 * (threadIdx.x % kPairs);
    if (!out.active(c)) return;
    for (int p = threadIdx.x / kPairs; p < n; p += kPairRowsPerPass) out.pair(c, p, kEdtInf, kEdtInf);
}

}  // namespace tile

using namespace tile;

// leading / trailing zero counts that give 0xffffffff for 0: what
// v_ffbh_u32 / v_ffbl_b32 return (written out, so that no select for the
// zero case is added around them)
__device__ __forceinline__ uint32_t clz_or_max(uint32_t x)
{
    uint32_t r;
    asm("v_ffbh_u32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ uint32_t ctz_or_max(uint32_t x)
{
    uint32_t r;
    asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// x distances into LDS of the columns x0 .. of slice z (16 columns, x0 a
// multiple of 16; or, PACKED, the kWideCols columns of mask word x0 / 32 as
// clamped 16-bit values two to a word, 0 past the volume's edge), one thread
// per row.  The row's mask words pass through its own LDS row first (<= 16
// words, the size of its distance row), so the nearest set bits left and
// right of word j are found from a bitmask of the non-empty words and two
// reads.  A column's distance is then the least of: the bits of word j at or
// left of it (leading zeros of the word shifted up), at or right of it
// (trailing zeros shifted down), the set bit before word j, the one after it.
// Without PACKED a column with no site in its row holds kNoSite.  Returns
// whether the slice has a site: a row with a set bit gives every column one.
template <bool PACKED>
__device__ __forceinline__ bool xpass(int w, int h, int z, int x0, const uint32_t* __restrict__ mask,
                      uint32_t invert, uint32_t* fs, int tid)
{
    constexpr int kN = PACKED ? kWideCols : kCols;
    constexpr uint32_t kFar = 0x10000u;                 // no set bit on that side
    constexpr uint32_t kCap = kMaxN;
    static_assert(kPairs == 16 && kCols == 16, "a row of LDS words holds the row's mask words");
    const int ws = mask_stride(w);
    const int j = x0 >> 5, boff = PACKED ? 0 : (x0 & 31);
    const int tail = w - 32 * (ws - 1);
    const uint32_t last_valid = tail >= 32 ? 0xffffffffu : ((1u << tail) - 1u);
    int has_site = 0;
    static_assert(kThreads >= kMaxN, "one row a thread");
    // one row a thread (h <= kMaxN): a plain branch rather than a loop, so no
    // per-word condition is hoisted into registers that live across it
    if (tid < h) {
        const int y = tid;
        const uint32_t* row = mask + ((size_t)z * h + y) * ws;
        uint32_t* srow = fs + y * 16;
        uint32_t words[16];
        if (ws == 16) {
            const uint4* r4 = (const uint4*)row;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint4 v = r4[k];
                words[4 * k] = v.x; words[4 * k + 1] = v.y; words[4 * k + 2] = v.z; words[4 * k + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; k++) words[k] = k < ws ? row[k] : 0u;
        }
        uint32_t nz = 0;   // bit k: word k has a set bit
        if (ws == 16) {    // a full row of words: no per-word range test
#pragma unroll
            for (int k = 0; k < 16; k++) {
                uint32_t m = words[k] ^ invert;
                if (k == 15) m &= last_valid;
                words[k] = m;
                nz |= min(m, 1u) << k;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; k++) {
                uint32_t m = k < ws ? words[k] ^ invert : 0u;
                if (k == ws - 1) m &= last_valid;
                words[k] = m;
                nz |= min(m, 1u) << k;
            }
        }
#pragma unroll
        for (int k = 0; k < 16; k += 4)
            *(uint4*)(srow + k) = make_uint4(words[k], words[k + 1], words[k + 2], words[k + 3]);
        const uint32_t wj = srow[j];
        const uint32_t below = nz & ((1u << j) - 1u), above = nz & ~((2u << j) - 1u);
        uint32_t dL0 = kFar, dR0 = kFar;   // from bit 0 of word j to the set bit before / after it
        if (below) {
            const int kl = hi_bit(below);
            dL0 = (uint32_t)(32 * (j - kl) - hi_bit(srow[kl]));
        }
        if (above) {
            const int kr = lo_bit(above);
            dR0 = (uint32_t)(32 * (kr - j) + lo_bit(srow[kr]));
        }
        has_site |= nz != 0u;
        uint4* out = (uint4*)srow;
        if constexpr (PACKED) {
            // two columns a word from the start: the in-word distances of
            // columns t, t + 1 (leading / trailing zeros, 0xffffffff for none)
            // packed by v_perm into 16-bit halves, the distances through the
            // neighbouring words formed in both halves at once, the least of
            // them clamped to 255 and squared in 16 bits.  255^2 = 65025 then
            // stands for "65025 or more", so the y scan takes results below
            // kXExact as exact.
            const uint32_t DL = min(dL0, 0x3fffu) * 0x10001u, DR = min(dR0, 0x3fffu) * 0x10001u;
            uint32_t v[kPairs];
#pragma unroll
            for (int q = 0; q < kPairs; q++) {
                const int t = 2 * q;
                const uint32_t a = __builtin_amdgcn_perm(clz_or_max(wj << (30 - t)), clz_or_max(wj << (31 - t)),
                                                         0x05040100u);
                const uint32_t b = __builtin_amdgcn_perm(ctz_or_max(wj >> (t + 1)), ctz_or_max(wj >> t), 0x05040100u);
                const uint32_t ii = (uint32_t)t | (uint32_t)(t + 1) << 16;
                const uint32_t dd = pk_min(pk_min(pk_min(a, b), pk_min(pk_add(DL, ii), pk_sub(DR, ii))), 0x00ff00ffu);
                v[q] = pk_mul(dd, dd);
            }
            if (32 * j + kN > w) {   // the last tile of a row: columns past w (w even) hold 0
#pragma unroll
                for (int q = 0; q < kPairs; q++)
                    if (32 * j + 2 * q >= w) v[q] = 0u;
            }
#pragma unroll
            for (int q = 0; q < kPairs; q += 4) out[q >> 2] = make_uint4(v[q], v[q + 1], v[q + 2], v[q + 3]);
        } else {
            uint32_t v[kN];
#pragma unroll
            for (int t = 0; t < kN; t++) {
                const int ii = boff + t;   // bit in word j
                const uint32_t a = clz_or_max(wj << (31 - ii));
                const uint32_t b = ctz_or_max(wj >> ii);
                const uint32_t e = min(min(dL0 + ii, dR0 - ii), kCap);
                const uint32_t dd = min(a, min(b, e));
                v[t] = dd >= kCap ? kNoSite : dd * dd;
            }
#pragma unroll
            for (int t = 0; t < kN; t += 4) out[t >> 2] = make_uint4(v[t], v[t + 1], v[t + 2], v[t + 3]);
        }
    }
    return __syncthreads_or(has_site);
}

// slice z of a tile sequence whose workgroups are dealt to the 8 XCDs
// round-robin: a group of (up to) 8 slices takes 8 * xt consecutive blocks,
// slice-minor, so the tiles of one slice share an XCD and its mask rows are
// fetched into one L2
__device__ __forceinline__ void xy_tile(int d, int xt, int& z, int& jt)
{
    const int bid = blockIdx.x;
    int group = bid / (8 * xt);
    const int gs = min(8, d - group * 8);
    const int inner = bid - group * 8 * xt;
    // groups from both ends of the volume first: the slices near its faces
    // tend to have the most empty space, the longest scans
    if ((d & 7) == 0) {
        const int ng = d >> 3;
        group = (group & 1) ? ng - 1 - (group >> 1) : (group >> 1);
    }
    z = group * 8 + inner % gs;
    jt = inner / gs;
}

// x pass from the bit rows, then the y scan (or envelope) of kCols lines of
// slice z; writes the int32 field f2 (or, for 2D, the final int32 and float
// outputs)
template <bool FINAL>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) k_edt3d_xy(int w, int h, int d, const uint32_t* __restrict__ mask,
                                                       uint32_t invert, int32_t* __restrict__ f2,
                                                       float* __restrict__ fl)
{
    __shared__ __align__(16) uint32_t fs[kWords];
    __shared__ __align__(16) uint32_t bm[kBmWords];
    __shared__ int abort_flag;
    if (threadIdx.x == 0) abort_flag = 0;
    int z, jt;
    xy_tile(d, (w + kCols - 1) / kCols, z, jt);
    const int x0 = jt * kCols;
    const bool any = xpass<false>(w, h, z, x0, mask, invert, fs, threadIdx.x);
    Bits bits{bm};
    if (FINAL) {
        // 2D image (d == 1): final outputs, rows of width w
        FinalOut o{f2, fl, (uint32_t)w, (uint32_t)x0, min(kCols, w - x0)};
        if (!any) no_sites(h, o);
        else if (!scan(h, fs, o, &abort_flag)) envelope(h, fs, o, bits, threadIdx.x);
    } else {
        FieldOut o{f2, w, h, x0, z};
        if (!any) no_sites(h, o);
        else if (!scan(h, fs, o, &abort_flag)) envelope(h, fs, o, bits, threadIdx.x);
    }
}

// the same over wide tiles (kWideCols lines, one mask word): packed scan
// into the 16-bit field.  flags[z * xt + jt] records which field the z pass
// reads for this tile: 0 the 16-bit one (every value below 0xffff), 1 the
// exact one, 2 none (no site: every value is kEdtInf); a tile whose scan fails
// is marked kRedo here and redone exactly by k_edt3d_xy_redo, which sets 1.
// (The redo is a kernel of its own so that this one carries none of the
// envelope's registers.)  w even, f16 4-byte aligned.
constexpr uint8_t kRedo = 3;

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) k_edt3d_xy_wide(int w, int h, int d, const uint32_t* __restrict__ mask,
                                                            uint32_t invert, uint16_t* __restrict__ f16,
                                                            uint8_t* __restrict__ flags)
{
    __shared__ __align__(16) uint32_t fs[kWords];
    __shared__ __align__(16) uint32_t mwin[kWinWords];
    __shared__ int abort_flag;
    if (threadIdx.x == 0) abort_flag = 0;
    const int xt = (w + kWideCols - 1) / kWideCols;
    int z, jt;
    xy_tile(d, xt, z, jt);
    const int x0 = jt * kWideCols;
    const bool any = xpass<true>(w, h, z, x0, mask, invert, fs, threadIdx.x);
    Field16Out o{f16, w, h, x0, z};
    uint8_t flag = 0;
    if (!any) {
        no_sites_pairs(h, o);
        flag = 2;
    } else if (window_minima(h, fs, mwin, threadIdx.x), !scan_pairs<kXExact>(h, fs, mwin, o, &abort_flag)) {
        flag = kRedo;
    }
    if (threadIdx.x == 0) flags[z * xt + jt] = flag;
}

// the wide tiles marked kRedo, each half again exactly (x pass into 32-bit
// words, then the envelope) into both fields.  Workgroup g owns tiles g,
// g + gridDim.x, ...: its threads read those flags at once (one each) and
// list the marked tiles in LDS, so a volume with no such tile costs one read.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) k_edt3d_xy_redo(int w, int h, int d, const uint32_t* __restrict__ mask,
                                                            uint32_t invert, uint16_t* __restrict__ f16,
                                                            int32_t* __restrict__ f32,
                                                            uint8_t* __restrict__ flags)
{
    __shared__ __align__(16) uint32_t fs[kWords];
    __shared__ __align__(16) uint32_t bm[kBmWords];
    __shared__ int list[kThreads];
    __shared__ int count;
    const int xt = (w + kWideCols - 1) / kWideCols;
    const int tiles = xt * d;   // <= kThreads * gridDim.x (checked by the launcher)
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    const int mine = blockIdx.x + threadIdx.x * gridDim.x;
    if (mine < tiles && flags[mine] == kRedo) list[atomicAdd(&count, 1)] = mine;
    __syncthreads();
    const int n = count;
    Bits bits{bm};
    for (int k = 0; k < n; k++) {
        const int t = list[k];
        const int z = t / xt, x0 = (t - z * xt) * kWideCols;
        for (int half = 0; half < 2 && x0 + kCols * half < w; half++) {
            // the thread index is opaque to the compiler here, so nothing of
            // the half's work is hoisted out of the loops (register pressure)
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            __syncthreads();   // the previous half's (or tile's) reads of fs are done
            FieldBothOut oh{f32, f16, w, h, x0 + kCols * half, z};
            xpass<false>(w, h, z, x0 + kCols * half, mask, invert, fs, tid);
            envelope(h, fs, oh, bits, tid);
        }
        if (threadIdx.x == 0) flags[t] = 1;
    }
}

// an f2 sample (< 2^20, or kEdtInf) as an LDS word
__device__ __forceinline__ uint32_t packed(int v) { return v == kEdtInf ? kNoSite : (uint32_t)v; }

// loads the int32 f2 tile of row y, columns x0 .. x0 + valid - 1 of kCols,
// along z into LDS; columns past valid hold kNoSite.  Returns whether the
// tile has a site.
__device__ __forceinline__ bool zload(int w, int h, int d, int y, int x0, int valid,
                                      const int32_t* __restrict__ f2, uint32_t* fs, int tid)
{
    int has_site = 0;
    if (valid == kCols && d == kMaxN && (w & 3) == 0) {
        // full tile: 16-byte loads, all in flight before the first LDS store
        constexpr int kQuads = kCols / 4;
        constexpr int kIter = kMaxN * kQuads / kThreads;
        int4 v[kIter];
#pragma unroll
        for (int k = 0; k < kIter; k++) {
            const int e = tid + k * kThreads;
            v[k] = *(const int4*)(f2 + ((size_t)(e / kQuads) * h + y) * w + x0 + (e % kQuads) * 4);
        }
#pragma unroll
        for (int k = 0; k < kIter; k++) {
            const int e = tid + k * kThreads;
            has_site |= (v[k].x & v[k].y & v[k].z & v[k].w) != kEdtInf;
            const uint4 q = make_uint4(packed(v[k].x), packed(v[k].y), packed(v[k].z), packed(v[k].w));
            *(uint4*)(fs + cell(e / kQuads, (e % kQuads) * 4)) = q;
        }
    } else {
        for (int e = tid; e < d * kCols; e += kThreads) {
            const int z = e / kCols, c = e % kCols;
            const int v = c < valid ? f2[((size_t)z * h + y) * w + x0 + c] : kEdtInf;
            has_site |= v != kEdtInf;
            fs[cell(z, c)] = packed(v);
        }
    }
    return __syncthreads_or(has_site);
}

// loads the 16-bit field of row y, columns x0 .. x0 + valid - 1 of a wide
// tile, along z into LDS, two columns a word (0 past valid).  Returns whether
// some value is below 0xffff.
__device__ __forceinline__ bool zload16(int w, int h, int d, int y, int x0, int valid,
                                        const uint16_t* __restrict__ f16, uint32_t* fs, int tid)
{
    int known = 0;
    if (valid == kWideCols && d == kMaxN && (w & 7) == 0) {
        // full tile: 16-byte loads (8 columns), all in flight before the first LDS store
        constexpr int kOcts = kWideCols / 8;
        constexpr int kIter = kMaxN * kOcts / kThreads;
        uint4 v[kIter];
#pragma unroll
        for (int k = 0; k < kIter; k++) {
            const int e = tid + k * kThreads;
            v[k] = *(const uint4*)(f16 + ((size_t)(e / kOcts) * h + y) * w + x0 + (e % kOcts) * 8);
        }
#pragma unroll
        for (int k = 0; k < kIter; k++) {
            const int e = tid + k * kThreads;
            known |= (v[k].x & v[k].y & v[k].z & v[k].w) != 0xffffffffu;
            *(uint4*)(fs + (e / kOcts) * kPairs + (e % kOcts) * 4) = v[k];
        }
    } else {
        for (int e = tid; e < d * kWideCols; e += kThreads) {
            const int z = e / kWideCols, c = e % kWideCols;
            const uint16_t v = c < valid ? f16[((size_t)z * h + y) * w + x0 + c] : (uint16_t)0;
            known |= c < valid && v != 0xffff;
            ((uint16_t*)fs)[z * kWideCols + c] = v;
        }
    }
    return __syncthreads_or(known);
}

// the exact field of row y, columns x0 .. x0 + valid - 1 of kCols, along z
// into LDS, each slice's values from the field its flag names (x tile jt)
__device__ __forceinline__ void zload_exact(int w, int h, int d, int y, int x0, int valid, int jt,
                                            const uint16_t* __restrict__ f16,
                                            const int32_t* __restrict__ f32,
                                            const uint8_t* __restrict__ flags, uint32_t* fs, int tid)
{
    const int xt = (w + kWideCols - 1) / kWideCols;
    for (int e = tid; e < d * kCols; e += kThreads) {
        const int z = e / kCols, c = e % kCols;
        int v = kEdtInf;
        if (c < valid) {
            const size_t i = ((size_t)z * h + y) * w + x0 + c;
            const uint8_t flag = flags[z * xt + jt];
            if (flag == 0) v = f16[i];
            else if (flag == 1) v = f32[i];
        }
        fs[cell(z, c)] = packed(v);
    }
    __syncthreads();
}

// tile (y, x0) of the z pass; blocks congruent mod 8 share an XCD, and each
// XCD gets a run of consecutive tiles, so the partial rows that neighbouring
// tiles write meet in one L2
__device__ __forceinline__ void z_tile(int xt, int cols, int& y, int& x0)
{
    int bid = blockIdx.x;
    if ((gridDim.x & 7) == 0) bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
    y = bid / xt;
    x0 = (bid - y * xt) * cols;
}

// z scan (or envelope) over the f2 tile of row y, columns x0 .. x0 + kCols - 1;
// writes int32 and float distances
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) k_edt3d_z(int w, int h, int d, const int32_t* __restrict__ f2,
                                                      int32_t* __restrict__ d2, float* __restrict__ fl)
{
    __shared__ __align__(16) uint32_t fs[kWords];
    __shared__ __align__(16) uint32_t bm[kBmWords];
    __shared__ int abort_flag;
    if (threadIdx.x == 0) abort_flag = 0;
    int y, x0;
    z_tile((w + kCols - 1) / kCols, kCols, y, x0);
    const int valid = min(kCols, w - x0);
    const bool any = zload(w, h, d, y, x0, valid, f2, fs, threadIdx.x);
    Bits bits{bm};
    FinalOut o{d2, fl, (uint32_t)(w * h), (uint32_t)(y * w + x0), valid};
    if (!any) no_sites(d, o);
    else if (!scan(d, fs, o, &abort_flag)) envelope(d, fs, o, bits, threadIdx.x);
}

// the same over wide tiles: packed scan of the 16-bit field, and if that
// fails (or no value of the tile is known below 0xffff), each half again
// exactly from the fields the flags name.  w even, outputs 8-byte aligned.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) k_edt3d_z_wide(int w, int h, int d, const uint16_t* __restrict__ f16,
                                                           const int32_t* __restrict__ f32,
                                                           const uint8_t* __restrict__ flags,
                                                           int32_t* __restrict__ d2, float* __restrict__ fl)
{
    __shared__ __align__(16) uint32_t fs[kWords];
    __shared__ __align__(16) uint32_t bm[kBmWords > kWinWords ? kBmWords : kWinWords];   // window minima, then the envelope's bitmask
    __shared__ int abort_flag;
    if (threadIdx.x == 0) abort_flag = 0;
    int y, x0;
    z_tile((w + kWideCols - 1) / kWideCols, kWideCols, y, x0);
    const int valid = min(kWideCols, w - x0);
    const bool known = zload16(w, h, d, y, x0, valid, f16, fs, threadIdx.x);
    FinalOut o{d2, fl, (uint32_t)(w * h), (uint32_t)(y * w + x0), valid};
    if (!known) {
        // every value reads 0xffff: if no slice of the tile has a site, no
        // line has one
        const int xt = (w + kWideCols - 1) / kWideCols, jt = x0 / kWideCols;
        int exact = 0;
        for (int z = threadIdx.x; z < d; z += kThreads) exact |= flags[z * xt + jt] != 2;
        if (!__syncthreads_or(exact)) {
            no_sites_pairs(d, o);
            return;
        }
    }
    if (!known || (window_minima(d, fs, bm, threadIdx.x), !scan_pairs<kClamp>(d, fs, bm, o, &abort_flag))) {
        Bits bits{bm};
        for (int half = 0; half < 2 && kCols * half < valid; half++) {
            // the thread index is opaque to the compiler here, so nothing of
            // the half's work is hoisted out of the loop (register pressure)
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            __syncthreads();
            const int xh = x0 + kCols * half, vh = min(kCols, valid - kCols * half);
            FinalOut oh{d2, fl, (uint32_t)(w * h), (uint32_t)(y * w + xh), vh};
            zload_exact(w, h, d, y, xh, vh, x0 / kWideCols, f16, f32, flags, fs, tid);
            envelope(d, fs, oh, bits, tid);
        }
    }
}


// every dimension <= 512 keeps each F below 2^20 and each product in int32
bool edt_tile_fits(int w, int h, int d) { return w <= kMaxN && h <= kMaxN && d <= kMaxN; }

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// workgroups of the redo pass (each checks its share of the tiles' flags)
constexpr int kRedoGroups = 1024;

// 3D: the exact int32 field, the 16-bit field and the wide tiles' flags
size_t edt_tile_workspace_3d(int w, int h, int d)
{
    const size_t cells = (size_t)w * h * d;
    return align256(cells * 4) + align256(cells * 2) + align256((size_t)d * ((w + kWideCols - 1) / kWideCols));
}

void launch_edt_tile_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                        hipStream_t s)
{
    const int xt = (w + kCols - 1) / kCols;
    ProfileScope p("edt2d_xy", s);
    hipLaunchKernelGGL(k_edt3d_xy<true>, dim3((unsigned)xt), dim3(kThreads), 0, s, w, h, 1, mask,
                       invert ? 0xffffffffu : 0u, d2, fl);
}

void launch_edt_tile_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                        float* fl, void* work, hipStream_t s)
{
    int32_t* f2 = (int32_t*)work;
    // wide tiles store column pairs with 8-byte stores and read the 16-bit
    // field with 16-byte loads
    if ((w & 1) == 0 && (((uintptr_t)d2 | (uintptr_t)fl) & 7) == 0 && ((uintptr_t)work & 15) == 0) {
        const size_t cells = (size_t)w * h * d;
        const int xt = (w + kWideCols - 1) / kWideCols;
        uint16_t* f16 = (uint16_t*)((char*)work + align256(cells * 4));
        uint8_t* flags = (uint8_t*)((char*)f16 + align256(cells * 2));
        {
            ProfileScope p("edt3d_xy", s);
            hipLaunchKernelGGL(k_edt3d_xy_wide, dim3((unsigned)(xt * d)), dim3(kThreads), 0, s, w,
                               h, d, mask, invert ? 0xffffffffu : 0u, f16, flags);
        }
        static_assert((kMaxN / kWideCols) * kMaxN <= kThreads * kRedoGroups, "one flag a redo thread");
        {
            ProfileScope p("edt3d_redo", s);
            hipLaunchKernelGGL(k_edt3d_xy_redo, dim3((unsigned)(xt * d < kRedoGroups ? xt * d : kRedoGroups)), dim3(kThreads), 0, s,
                               w, h, d, mask, invert ? 0xffffffffu : 0u, f16, f2, flags);
        }
        {
            ProfileScope p("edt3d_z", s);
            hipLaunchKernelGGL(k_edt3d_z_wide, dim3((unsigned)(xt * h)), dim3(kThreads), 0, s, w, h,
                               d, f16, f2, flags, d2, fl);
        }
        return;
    }
    const int xt = (w + kCols - 1) / kCols;
    {
        ProfileScope p("edt3d_xy", s);
        hipLaunchKernelGGL(k_edt3d_xy<false>, dim3((unsigned)(xt * d)), dim3(kThreads), 0, s, w, h,
                           d, mask, invert ? 0xffffffffu : 0u, f2, nullptr);
    }
    {
        ProfileScope p("edt3d_z", s);
        hipLaunchKernelGGL(k_edt3d_z, dim3((unsigned)(xt * h)), dim3(kThreads), 0, s, w,
                           h, d, f2, d2, fl);
    }
}

}  // namespace dolhip
