// dol_edt.hip -- squared Euclidean distance transforms for gfx950.
//
// Two families:
//  * Felzenszwalb float (DistanceOcclusionLib.cpp:1018-1085), bit-identical to
//    the reference: the same float operations in the same order, no FMA
//    contraction, correctly rounded division.  Each pass is an envelope
//    kernel (one lane a line, the stack's top in LDS) and an output kernel
//    (one workgroup a line, one output a thread).
//  * Exact integer EDT (an extension; the reference has only the float 2D
//    form): a per-row nearest-set-bit pass straight from the bitmask, then the
//    lower envelope of parabolas along each further axis with intersections
//    compared by exact cross-multiplication.
#include "dol_common.hpp"

#include <cfloat>

namespace dolhip {

// ------------------------------------------------------ Felzenszwalb float
//
// DistanceOcclusionLib.cpp:1026-1062 runs each line in two loops: the lower
// envelope of the parabolas (a stack v, z whose pops depend on float
// comparisons, so a line is one sequential chain), then the outputs, each
// taking the first envelope entry k with z[k + 1] >= q.  Here the two loops
// are two kernels:
//  * k_felz_envelope: one lane a line, 64 lines a wavefront, the same float
//    operations in the same order.  The stack's top entries live in an LDS
//    ring ((v, z) and the key F = f[v] + v*v the reference recomputes on
//    each pop, the same float), the top two also in registers, so that the
//    first pop of a step is a select beside the test; the stack in device
//    memory, read by pops deeper than the ring, is brought up to date from
//    the ring once a block of steps.  A block after one in which no line
//    popped first takes its tests as if none pops, and if so pushes with
//    no branch.  The line's samples are loaded kFelzAhead steps ahead.
//  * k_felz_output: one workgroup a line, one output a thread.  z[1..D] rises
//    strictly (a push follows a failed "s <= z[k]"), so the entry the
//    reference's running scan reaches for q is the number of z[1..D] below
//    q, found by binary search over the line's z in LDS.  A NaN in z breaks
//    the order; such a line (marked by the envelope kernel) runs the
//    reference's scan in one thread.
// Layouts: element q of line l at f2[l + q * m]; outputs at out[l * n + q];
// the stacks at V[l * n + k], Z[l * (n + 1) + k]; D[l] the final depth, or
// -1 - depth for a line with a NaN in z.

constexpr int kRing = 64;        // stack entries a line keeps in LDS (half with 64 lines)
constexpr int kFelzAhead = 16;   // samples a line has in flight

// A line's stacks in device memory: V[l * felz_vs(n) + k], Z[l * felz_zs(n) + k],
// each line's rows 16-byte aligned, with 4 entries past the deepest one a
// copy of the ring can write
__host__ __device__ inline int felz_vs(int n) { return ((n + 3) & ~3) + 4; }
__host__ __device__ inline int felz_zs(int n) { return ((n + 4) & ~3) + 4; }

template <int LPW>
__global__ void __launch_bounds__(LPW) k_felz_envelope(int n, int lines, const float* __restrict__ f2, size_t m,
                                                       int* __restrict__ V, float* __restrict__ Z,
                                                       int* __restrict__ D)
{
    // the ring: entry (v, z, F) of depth k at slot k mod R, one 16-byte
    // record a lane (one LDS write or read a push or pop); it holds depths
    // lo .. k (a push below lo lowers lo).  Pushes go to the ring only;
    // after every block of kFelzAhead steps each line copies the entries it
    // pushed in the block that are still on the stack (depths head .. k) to
    // its stack in device memory, mostly with 16-byte stores.  Depths below
    // head did not change in the block, so after the copy device memory
    // holds the whole stack.  An entry a pop reads from device memory (a
    // depth below lo) was pushed in an earlier block: overwriting its slot
    // within its own block would take R > kFelzAhead more pushes.
    constexpr int R = LPW >= 64 ? kRing / 2 : kRing;
    static_assert(R >= 2 * kFelzAhead, "a block's pushes stay in the ring until it is copied");
    __shared__ uint4 ring[R][LPW];
    const int t = threadIdx.x;
    const int l = blockIdx.x * LPW + t;
    if (l >= lines) return;
    const float* f = f2 + l;
    // the stacks through descriptors based at the workgroup's first line,
    // with 32-bit lane offsets (below 2^32 for lines of up to 2^24 samples)
    const size_t l0 = (size_t)blockIdx.x * LPW;
    const int vs = felz_vs(n), zs = felz_zs(n);
    const auto rV = __builtin_amdgcn_make_buffer_rsrc((void*)(V + l0 * vs), 0, 0xffffffffu, kRsrcWord3);
    const auto rZ = __builtin_amdgcn_make_buffer_rsrc((void*)(Z + l0 * zs), 0, 0xffffffffu, kRsrcWord3);
    const uint32_t oV = (uint32_t)t * (uint32_t)vs * 4u, oZ = (uint32_t)t * (uint32_t)zs * 4u;
    int k = 0, lo = 0;   // stack depth; the lowest depth the ring still holds
    // the top entry (depth k) and the one below it (depth k - 1, when k > 0)
    // in registers: the test after popping the top is computed beside the
    // test against it, and the first pop is a select, not a branch
    int vk = 0, v1 = 0;
    float zk = -FLT_MAX, z1 = -FLT_MAX;
    float Fk = __fadd_rn(f[0], 0.0f), F1 = Fk;
    auto put = [&](int d, int v, float z, float F) __attribute__((always_inline)) {
        ring[d & (R - 1)][t] = make_uint4((uint32_t)v, __float_as_uint(z), __float_as_uint(F), 0u);
    };
    // the ring's entries of depths head .. top (head >= top - kFelzAhead + 1,
    // all at or above lo) to the stacks in device memory: the first quad's
    // entries one at a time, then whole 16-byte aligned quads (at most 4),
    // whose depths above the top carry stale values (depths above the top
    // are not on the stack; the highest, top + 3, is below felz_vs(n)).
    // With last, depth top + 1 of Z gets the terminator +FLT_MAX.
    auto copy = [&](int top, int head, bool last) __attribute__((always_inline)) {
        const int hq = head & ~3;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int d = hq + c;
            if (d >= head && d <= top) {
                const uint4 e = ring[d & (R - 1)][t];
                __builtin_amdgcn_raw_buffer_store_b32(e.x, rV, oV + 4u * (uint32_t)d, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b32(e.y, rZ, oZ + 4u * (uint32_t)d, 0, 0);
            }
        }
#pragma unroll
        for (int i = 1; i <= 4; i++) {
            const int d = hq + 4 * i;
            if (d <= top) {
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 vv, zz;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const uint4 e = ring[(d + c) & (R - 1)][t];
                    vv[c] = e.x;
                    zz[c] = last && d + c == top + 1 ? __float_as_uint(+FLT_MAX) : e.y;
                }
                __builtin_amdgcn_raw_buffer_store_b128(vv, rV, oV + 4u * (uint32_t)d, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(zz, rZ, oZ + 4u * (uint32_t)d, 0, 0);
            }
        }
        // the terminator, unless a quad above wrote it
        if (last && !(top + 1 >= hq + 4 && ((top + 1) & 3) != 0))
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(+FLT_MAX), rZ, oZ + 4u * (uint32_t)(top + 1), 0, 0);
    };
    put(0, 0, -FLT_MAX, Fk);
    __builtin_amdgcn_raw_buffer_store_b32(0u, rV, oV, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(-FLT_MAX), rZ, oZ, 0, 0);
    int head = 1;   // the lowest depth pushed since the last copy
    bool nan = false;
    bool popped = false;   // this line popped in the current block
    auto step = [&](int q, float fq) __attribute__((always_inline)) {
        const float fqt = __fadd_rn(fq, (float)(q * q));
        float s = __fdiv_rn(__fsub_rn(fqt, Fk), (float)(2 * (q - vk)));
        // the same test against the entry below (v1 < q: no zero divisor;
        // used only if the top is popped), computed beside the first: the
        // first pop is a select, not a branch
        const float s1 = __fdiv_rn(__fsub_rn(fqt, F1), (float)(2 * (q - v1)));
        const bool pop = s <= zk && k > 0;
        popped |= pop;
        k -= pop;
        s = pop ? s1 : s;
        vk = pop ? v1 : vk;
        zk = pop ? z1 : zk;
        Fk = pop ? F1 : Fk;
        // further pops read the stack's next entry from the ring, or, for a
        // depth the ring no longer holds, from the stack in device memory
        while (s <= zk && k > 0) {
            k--;
            if (k >= lo) {
                const uint4 e = ring[k & (R - 1)][t];
                vk = (int)e.x;
                zk = __uint_as_float(e.y);
                Fk = __uint_as_float(e.z);
            } else {
                vk = (int)__builtin_amdgcn_raw_buffer_load_b32(rV, oV + 4u * (uint32_t)k, 0, 0);
                zk = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rZ, oZ + 4u * (uint32_t)k, 0, 0));
                Fk = __fadd_rn(f[m * (size_t)vk], (float)(vk * vk));
                // wait for these loads here, in the rare branch: left to
                // the loop, the wait would cover the samples in flight too
                asm volatile("" : "+v"(vk), "+v"(zk), "+v"(Fk));
            }
            s = __fdiv_rn(__fsub_rn(fqt, Fk), (float)(2 * (q - vk)));
        }
        k++;
        put(k, q, s, fqt);
        head = min(head, k);
        lo = max(min(lo, k), k - R + 1);
        nan |= s != s;
        v1 = vk;
        z1 = zk;
        F1 = Fk;
        vk = q;
        zk = s;
        Fk = fqt;
    };
    // samples kFelzAhead steps ahead; every load is issued (past the line's
    // end it reads the last sample), so that the waits on them count exactly
    auto sample = [&](int q) { return f[m * (size_t)min(q, n - 1)]; };
    float cur[kFelzAhead], nxt[kFelzAhead];
#pragma unroll
    for (int j = 0; j < kFelzAhead; j++) cur[j] = sample(1 + j);
    int q0 = 1;
    // A block of kFelzAhead steps after a block in which no line popped
    // first takes the tests as if none pops: the top is then always sample
    // q - 1 (a step that does not pop pushes q), so each test is the
    // reference's with the divisor 2 * (q - (q - 1)) = 2.  If no line of
    // the wavefront pops in the block, those are its steps, and the pushes
    // follow with no branch; otherwise the block runs step by step.
    bool hot = true;
    for (; q0 + kFelzAhead <= n; q0 += kFelzAhead) {
#pragma unroll
        for (int j = 0; j < kFelzAhead; j++) nxt[j] = sample(q0 + kFelzAhead + j);
        bool done = false;
        if (!hot) {
            float ft[kFelzAhead], sj[kFelzAhead];
            bool any = false;
#pragma unroll
            for (int j = 0; j < kFelzAhead; j++) {
                const int q = q0 + j;
                ft[j] = __fadd_rn(cur[j], (float)(q * q));
                sj[j] = __fdiv_rn(__fsub_rn(ft[j], j ? ft[j - 1] : Fk), 2.0f);
                any |= sj[j] <= (j ? sj[j - 1] : zk) && (j || k > 0);
            }
            if (!__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
                for (int j = 0; j < kFelzAhead; j++) {
                    put(k + 1 + j, q0 + j, sj[j], ft[j]);
                    nan |= sj[j] != sj[j];
                }
                head = min(head, k + 1);
                lo = max(min(lo, k + 1), k + kFelzAhead - R + 1);
                k += kFelzAhead;
                v1 = q0 + kFelzAhead - 2;
                z1 = sj[kFelzAhead - 2];
                F1 = ft[kFelzAhead - 2];
                vk = q0 + kFelzAhead - 1;
                zk = sj[kFelzAhead - 1];
                Fk = ft[kFelzAhead - 1];
                done = true;
            }
        }
        if (!done) {
            popped = false;
#pragma unroll
            for (int j = 0; j < kFelzAhead; j++) step(q0 + j, cur[j]);
            hot = __builtin_amdgcn_ballot_w64(popped) != 0;
        }
        copy(k, head, false);
        head = k + 1;
#pragma unroll
        for (int j = 0; j < kFelzAhead; j++) cur[j] = nxt[j];
    }
#pragma unroll
    for (int j = 0; j < kFelzAhead; j++)
        if (q0 + j < n) step(q0 + j, cur[j]);
    copy(k, head, true);
    D[l] = nan ? -1 - k : k;
}

constexpr int kFelzOutThreads = 512;   // 32 waves a CU beside the line's z in LDS (256: 8 % slower on 8192^2; 1024: within 1.5 %)

// outputs of line blockIdx.x; zs holds z[1..D] (n floats at most)
__global__ void __launch_bounds__(kFelzOutThreads) k_felz_output(int n, const float* __restrict__ f2, size_t m,
                                                                 const int* __restrict__ V,
                                                                 const float* __restrict__ Z,
                                                                 const int* __restrict__ D,
                                                                 float* __restrict__ out)
{
    extern __shared__ float zs[];
    // workgroups are dealt to the 8 XCDs in turn: each XCD takes a run of
    // consecutive lines, whose sample gathers f[m * v] share cache lines
    int l = blockIdx.x;
    if ((gridDim.x & 7) == 0) l = (l & 7) * (gridDim.x >> 3) + (l >> 3);
    const float* f = f2 + l;
    const int* Vl = V + (size_t)l * felz_vs(n);
    const float* Zl = Z + (size_t)l * felz_zs(n);
    float* o = out + (size_t)l * n;
    const int dd = D[l];
    if (dd < 0) {
        // a NaN in z: the reference's scan as it is
        if (threadIdx.x == 0) {
            int k = 0;
            for (int q = 0; q < n; q++) {
                while (Zl[k + 1] < (float)q) k++;
                const int vk = Vl[k];
                o[q] = __fadd_rn((float)((q - vk) * (q - vk)), f[m * (size_t)vk]);
            }
        }
        return;
    }
    for (int j = threadIdx.x; j < dd; j += kFelzOutThreads) zs[j] = Zl[1 + j];
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += kFelzOutThreads) {
        // k = #{j in 1..dd : z[j] < q}
        const float fq = (float)q;
        int lo = 0, hi = dd;   // the count lies in [lo, hi]
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (zs[mid] < fq)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int vk = Vl[lo];
        o[q] = __fadd_rn((float)((q - vk) * (q - vk)), f[m * (size_t)vk]);
    }
}

// the transposed field td (w*h floats), the stacks V and Z (felz_vs and
// felz_zs entries a line, for the larger of the two passes), the depths D
// (max(w, h) ints)
static inline size_t felz_part(size_t b) { return (b + 255) & ~(size_t)255; }

static inline size_t felz_stack(int w, int h, int (*stride)(int))
{
    const size_t cols = (size_t)w * stride(h), rows = (size_t)h * stride(w);
    return felz_part((cols > rows ? cols : rows) * 4);
}

size_t felz_workspace(int w, int h)
{
    const size_t cells = (size_t)w * h, lines = (size_t)(w > h ? w : h);
    return felz_part(cells * 4) + felz_stack(w, h, felz_vs) + felz_stack(w, h, felz_zs) + felz_part(lines * 4);
}

// one 1D pass over `lines` lines of n samples
static void felz_pass(int n, int lines, const float* f2, size_t m, float* out, int* V, float* Z, int* D,
                      hipStream_t s)
{
    // 64 lines a wavefront: the lines' samples at each step are one
    // contiguous read, and each line is one chain of dependent steps whose
    // length does not shrink with fewer lines a wavefront (measured 2, 4, 8,
    // 16 and 64 lines on 512^2 to 16384 x 2048 fields: 64 fastest or within
    // 2 % everywhere)
    hipLaunchKernelGGL(k_felz_envelope<64>, dim3(grid1d((size_t)lines, 64)), dim3(64), 0, s, n, lines, f2, m,
                       V, Z, D);
    // a line's z in LDS: at most 32767 samples (the entry points' bound), 128 KiB
    const size_t lds = (size_t)n * 4;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_felz_output, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_felz_output, dim3((unsigned)lines), dim3(kFelzOutThreads), lds, s, n, f2, m, V, Z, D, out);
}

void launch_felzenszwalb_2d(int w, int h, float* dist, void* work, hipStream_t s)
{
    const size_t cells = (size_t)w * h;
    char* p = (char*)work;
    float* td = (float*)p;
    int* V = (int*)(p + felz_part(cells * 4));
    float* Z = (float*)((char*)V + felz_stack(w, h, felz_vs));
    int* D = (int*)((char*)Z + felz_stack(w, h, felz_zs));
    {
        // columns: line x reads dist[q*w + x], writes td[x*h + q]
        ProfileScope ps("felzenszwalb_cols", s);
        felz_pass(h, w, dist, (size_t)w, td, V, Z, D, s);
    }
    {
        // rows: line y reads td[q*h + y], writes dist[y*w + q]
        ProfileScope ps("felzenszwalb_rows", s);
        felz_pass(w, h, td, (size_t)h, dist, V, Z, D, s);
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
