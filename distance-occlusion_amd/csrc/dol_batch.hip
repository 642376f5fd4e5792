// dol_batch.hip -- signed 8-bit SDFs of many small masks in one launch.
//
// Per mask the pipeline of the reference tool's signed border mode
// (DistanceOcclusionTool.cpp:634-670 threshold, :1333-1365 border init +
// JumpFlood(cw = 2), :1240-1271 EncodeU8Signed with 2 * maxLen where
// maxLen = 0.25 max(w, h), :1973-1974).
//
// Batched form, two tiers by mask size:
//  * Bins (every mask whose field fits a workgroup's LDS, about three in
//    four of the cell passes of a 16..256-pixel mix).  The host packs masks
//    into bins of at most kGroups 64-cell groups (best fit, largest first;
//    each mask's cells padded to whole groups).  One persistent kernel: a
//    workgroup takes a bin from an atomic queue and runs the whole pipeline
//    for all of its masks in LDS: the masks as one bitmask in LDS (a ballot
//    a 64-cell group), border init from it (each cell evaluates the
//    reference's writes for its mask word as bit masks), every JumpFlood pass (each pass computes into
//    registers, barrier, writes back -- the reference's out-of-place passes
//    with one buffer), encode.  Group g of the bin belongs to wave g % W, so
//    a wave's groups are static register slots, and every group lies in one
//    mask, so a mask's size, base and pass steps are wavefront-uniform
//    (scalar loads of a per-bin group table).  A cell's (x, y) is one
//    high multiply by ceil(2^31 / w) (exact for cell indices below 2^16).
//  * Large masks (fields above a bin): every JumpFlood step over all of them
//    in one launch each (tiles of 2048 cells, eight a thread), out of place
//    through two device fields; the first launch seeds the border from the
//    bytes, the last one runs the final step-1 pass and encodes.  They all
//    have 8 passes, steps 128 .. 1 (their larger side is above 128).  The
//    tiles are dealt to the 8 XCDs in contiguous runs, so the rows a step
//    reads step rows above and below were fetched into the same L2 shortly
//    before.
// The queue counter is reset by a one-thread kernel (not a memset) so the
// sequence can be captured in a hipGraph and replayed.  A tuner times the
// (threads, groups per wave) variants on the actual batch and keeps the
// fastest; each variant has its own bins (their capacity is the variant's).
//
// Per-image form: the same per-cell functions (dol_cells.hpp) through the
// ordinary one-kernel-per-step launches, for comparison.
#include "../../include/dol_hip.h"
#include "dol_cells.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <vector>

namespace dolhip {

void launch_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                            uint32_t* mask, hipStream_t s);
void launch_border_2d(int w, int h, const uint32_t* mask, int16_t* delta, hipStream_t s);
void launch_jfa2d_pass(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s);
void launch_signed_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                      void* out, int mode, hipStream_t s);
int check_launch(const char* what);
void set_error(const char* fmt, ...);

constexpr int kMaxGlyph = 256;

// a mask inside its bin
struct BinSlot {
    long long off;    // byte offset of the mask in the input and the output
    int cell0;        // its first cell in the bin's LDS field (a multiple of 64)
    int w, h;
    uint32_t magic;   // div_magic(w)
    int passes, step0;
};
struct Bin {
    int slot0;   // first BinSlot
    int groups;  // 64-cell groups in use
    int passes;  // most passes of its masks
    int pad;
};
// a large mask, and a tile of one
struct BigMask {
    long long off;    // byte offset in the input and the output
    long long f0;     // first cell of its fields
    int w, h;
    uint32_t magic;
    int pad;
};
struct BigTile {
    int mask, cell0;
};
constexpr int kBigThreads = 256;
constexpr int kBigCells = 8;                       // cells a thread
constexpr int kBigTile = kBigThreads * kBigCells;  // cells a workgroup

// ceil(2^31 / w): for a cell index i < 2^16 of a w-wide mask, i / w is
// mulhi(2 i, magic) (the rounding error is below 2^-15, under 1 / w)
__host__ __device__ inline uint32_t div_magic(int w) { return (uint32_t)((0x80000000ull + w - 1) / w); }
__device__ __forceinline__ int div_by(int i, uint32_t magic) { return (int)__umulhi((uint32_t)i << 1, magic); }

__global__ void k_reset_counter(int* c) { *c = 0; }

// ------------------------------------------------------------ bins

// per-group word of a bin: slot | (w - 1) << 8 | (h - 1) << 16 | passes << 24 |
// log2(first step) << 28
__host__ __device__ inline uint32_t group_word(int slot, int w, int h, int passes, int step0)
{
    int l = 0;
    while ((1 << (l + 1)) <= step0) l++;
    return (uint32_t)slot | (uint32_t)(w - 1) << 8 | (uint32_t)(h - 1) << 16 | (uint32_t)passes << 24 |
           (uint32_t)l << 28;
}

// The packed squared length as len2_packed; with BUILTIN through the
// compiler's dot product (v_dot2c with a zeroed accumulator, which the
// compiler schedules freely) rather than the inline-asm three-operand form:
// measured faster in the large-mask kernels (taps from device memory),
// slower in the bins (taps from LDS) and in the streaming 2D pass.
template <bool BUILTIN>
__device__ __forceinline__ int len2_packed_b(uint32_t xy)
{
    if constexpr (BUILTIN) {
        const jfa_i16x2 v = __builtin_bit_cast(jfa_i16x2, xy);
        return __builtin_amdgcn_sdot2(v, v, 0, false);
    } else {
        return len2_packed(xy);
    }
}

// One JumpFlood pass for the cell at c, (x, y) of a w x h mask, in half
// cells (DistanceOcclusionLib.cpp:782-873, the taps of jfa2_cell in its
// order), without branches: a tap outside the mask reads the cell itself and
// its length is replaced by INT_MAX, which never passes the strict test.
// best takes a candidate iff bl > cl, and then bl = cl, so bl = min(bl, cl).
template <bool BUILTIN>
__device__ __forceinline__ void jfa2_tap_if(uint32_t& best, int& bl, uint32_t v, int ox, int oy, bool ok)
{
    jfa_u16x2 c = __builtin_bit_cast(jfa_u16x2, v);
    c += jfa_u16x2{(unsigned short)ox, (unsigned short)oy};
    const uint32_t cv = __builtin_bit_cast(uint32_t, c);
    const int d = len2_packed_b<BUILTIN>(cv);
    const int cl = ok ? d : 0x7fffffff;
    best = bl > cl ? cv : best;
    bl = min(bl, cl);
}

// ld(o) reads the cell o cells after this one (o = 0: the cell itself)
template <bool BUILTIN, class Load>
__device__ __forceinline__ uint32_t jfa2_taps(Load ld, int w, int h, int x, int y, int step)
{
    const int off = 2 * step;
    const int sw = step * w;
    const bool l = x - step >= 0, r = x + step < w, u = y - step >= 0, d = y + step < h;
    const int dl = l ? -step : 0, dr = r ? step : 0, du = u ? -sw : 0, dd = d ? sw : 0;
    uint32_t best = ld(0);
    int bl = len2_packed_b<BUILTIN>(best);
    jfa2_tap_if<BUILTIN>(best, bl, ld(dl), -off, 0, l);
    jfa2_tap_if<BUILTIN>(best, bl, ld(du + dl), -off, -off, l && u);
    jfa2_tap_if<BUILTIN>(best, bl, ld(dd + dl), -off, +off, l && d);
    jfa2_tap_if<BUILTIN>(best, bl, ld(dr), +off, 0, r);
    jfa2_tap_if<BUILTIN>(best, bl, ld(du + dr), +off, -off, r && u);
    jfa2_tap_if<BUILTIN>(best, bl, ld(dd + dr), +off, +off, r && d);
    jfa2_tap_if<BUILTIN>(best, bl, ld(du), 0, -off, u);
    jfa2_tap_if<BUILTIN>(best, bl, ld(dd), 0, +off, d);
    return best;
}

__device__ __forceinline__ uint32_t jfa2_cell_at(const uint32_t* c, int w, int h, int x, int y, int step)
{
    return jfa2_taps<false>([&](int o) { return c[o]; }, w, h, x, y, step);
}

// v, made opaque to the compiler: values derived from it are recomputed
// where they are used (a few scalar operations a group) instead of being
// hoisted out of the loops and held, one set per unrolled group, in scalar
// registers that the groups would exhaust
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+s"(v));
    return v;
}

template <int T, int G>
__global__ void __launch_bounds__(T) k_glyph_bins(int nbins, const Bin* __restrict__ bins,
                                                  const BinSlot* __restrict__ slots,
                                                  const uint32_t* __restrict__ gwords,
                                                  const uint8_t* __restrict__ bytes,
                                                  uint8_t* __restrict__ out, int threshold, int less_than,
                                                  int* __restrict__ counter)
{
    constexpr int W = T / 64;        // wavefronts
    constexpr int kGroups = W * G;   // groups a bin holds
    extern __shared__ uint32_t field[];   // kGroups * 64 cells; the bin's bitmask before border init
    __shared__ uint32_t gw[kGroups];      // the bin's group words
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    for (;;) {
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_next);
        if (b >= nbins) break;
        const Bin bin = bins[b];
        for (int g = tid; g < bin.groups; g += T) gw[g] = gwords[(size_t)b * kGroups + g];
        __syncthreads();

        // per group of this wave: the lane's cell (x, y), or ~0 past the mask
        uint32_t xy[G];
        uint32_t r[G];
        uint64_t inside = 0;

        // 1. the masks as one flat bitmask in LDS (bin cell c is bit c, one
        // ballot a group; word 0 is left free so that row_bits_flat may read
        // the word before the first); x, y and the inside bits
        uint32_t* flat = field + 1;
        int wv = opaque(wave);
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int g = wv + W * j;
            xy[j] = ~0u;
            if (g < bin.groups) {
                const uint32_t m = __builtin_amdgcn_readfirstlane(gw[g]);
                const BinSlot& sl = slots[bin.slot0 + (int)(m & 0xffu)];
                const int i = g * 64 - sl.cell0 + lane;
                int bit = 0;
                if (i < sl.w * sl.h) {
                    const int y = div_by(i, sl.magic);
                    xy[j] = (uint32_t)(i - y * sl.w) | (uint32_t)y << 16;
                    bit = byte_bit(bytes[sl.off + i], threshold, less_than);
                }
                inside |= (uint64_t)bit << j;
                const uint64_t bal = __ballot(bit);
                if (lane == 0) flat[2 * g] = (uint32_t)bal;
                if (lane == 32) flat[2 * g + 1] = (uint32_t)(bal >> 32);
            }
        }
        __syncthreads();

        // 2. border init: each cell evaluates the reference's writes for its
        // mask word (border2_word over the rows above, at and below), then
        // takes its own bit
        wv = opaque(wave);
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int g = wv + W * j;
            if (g < bin.groups && xy[j] != ~0u) {
                const uint32_t m = __builtin_amdgcn_readfirstlane(gw[g]);
                const int w = (int)((m >> 8) & 0xffu) + 1, h = (int)((m >> 16) & 0xffu) + 1;
                const int x = (int)(xy[j] & 0xffffu), y = (int)(xy[j] >> 16);
                const int row = (g * 64 + lane) - x;   // flat bit of the row's first pixel
                const int ws = (w + 31) >> 5, jw = x >> 5;
                const RowBits zero{0u, 0u, 0u};
                const bool has_up = y >= 1, has_dn = y + 1 < h;
                const RowBits mid = row_bits_flat(flat, row, jw, ws, w);
                const RowBits up = has_up ? row_bits_flat(flat, row - w, jw, ws, w) : zero;
                const RowBits dn = has_dn ? row_bits_flat(flat, row + w, jw, ws, w) : zero;
                const uint32_t right = jw == ws - 1 ? ~(1u << ((w - 1) & 31)) : ~0u;
                r[j] = border2_value(border2_word(up, mid, dn, has_up, has_dn, right), x & 31);
            }
        }
        __syncthreads();
        wv = opaque(wave);
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int g = wv + W * j;
            if (g < bin.groups && xy[j] != ~0u) field[g * 64 + lane] = r[j];
        }
        __syncthreads();

        // 3. JumpFlood in half cells, every mask on its own schedule
        for (int k = 0; k < bin.passes; k++) {
            // the per-group tests are recomputed each pass (two scalar and one
            // vector compare) rather than held as lane masks in scalar
            // registers, which the unrolled groups would exhaust
            const int groups = opaque(bin.groups);
#pragma unroll
            for (int j = 0; j < G; j++) asm volatile("" : "+v"(xy[j]));
            wv = opaque(wave);
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int g = wv + W * j;
                if (g < groups) {
                    const uint32_t m = __builtin_amdgcn_readfirstlane(gw[g]);
                    if (k < (int)((m >> 24) & 0xfu) && xy[j] != ~0u) {
                        const int w = (int)((m >> 8) & 0xffu) + 1, h = (int)((m >> 16) & 0xffu) + 1;
                        const int step = max((1 << (m >> 28)) >> k, 1);
                        r[j] = jfa2_cell_at(field + g * 64 + lane, w, h, (int)(xy[j] & 0xffffu), (int)(xy[j] >> 16),
                                            step);
                    }
                }
            }
            __syncthreads();
            wv = opaque(wave);
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int g = wv + W * j;
                if (g < groups) {
                    const uint32_t m = __builtin_amdgcn_readfirstlane(gw[g]);
                    if (k < (int)((m >> 24) & 0xfu) && xy[j] != ~0u) field[g * 64 + lane] = r[j];
                }
            }
            __syncthreads();
        }

        // 4. encode
        wv = opaque(wave);
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int g = wv + W * j;
            if (g < bin.groups && xy[j] != ~0u) {
                const uint32_t m = __builtin_amdgcn_readfirstlane(gw[g]);
                const BinSlot& sl = slots[bin.slot0 + (int)(m & 0xffu)];
                const int i = (int)(xy[j] >> 16) * sl.w + (int)(xy[j] & 0xffffu);
                const float max_len = 0.5f * (float)max(sl.w, sl.h);
                out[sl.off + i] = sdf_u8_cell(field[g * 64 + lane], (int)((inside >> j) & 1u), max_len);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ large masks

// the tile of this workgroup: the tiles are dealt to the XCDs (which take
// workgroups round-robin) in contiguous runs
__device__ __forceinline__ BigTile big_tile(const BigTile* tiles)
{
    const int n = gridDim.x, b = blockIdx.x;
    const int per = (n + 7) >> 3, xcd = b & 7, k = b >> 3;
    int t = xcd * per + k;
    // the last XCDs' runs are shorter when n is not a multiple of 8: the
    // blocks past them take the remaining tiles in order
    const int full = n - 8 * (per - 1);   // XCDs whose run has `per` tiles
    if (xcd >= full) t = full * per + (xcd - full) * (per - 1) + k;
    return tiles[t];
}

// Thread t of a tile takes its cells t, t + kBigThreads, ... (kBigCells a
// thread, all loads in flight together)
__device__ __forceinline__ bool big_cell(const BigMask& m, int cell0, int c, int& i, int& x, int& y)
{
    i = cell0 + (int)threadIdx.x + c * kBigThreads;
    if (i >= m.w * m.h) return false;
    y = div_by(i, m.magic);
    x = i - y * m.w;
    return true;
}

__global__ void __launch_bounds__(kBigThreads) k_big_border(const BigTile* __restrict__ tiles,
                                                            const BigMask* __restrict__ masks,
                                                            const uint8_t* __restrict__ bytes, int threshold,
                                                            int less_than, uint32_t* __restrict__ fa)
{
    const BigTile t = big_tile(tiles);
    const BigMask m = masks[t.mask];
    const int w = m.w, area = w * m.h;
    const auto rb = __builtin_amdgcn_make_buffer_rsrc((void*)(bytes + m.off), 0, area, kRsrcWord3);
    uint32_t v[kBigCells];
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        const int i = min(t.cell0 + (int)threadIdx.x + c * kBigThreads, area - 1);
        const int y = div_by(i, m.magic), x = i - y * w;
        v[c] = border2_cell_by(
            [&](int xx, int yy) {
                return byte_bit((uint8_t)__builtin_amdgcn_raw_buffer_load_b8(rb, yy * w + xx, 0, 0), threshold,
                                less_than);
            },
            w, m.h, x, y);
    }
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        const int i = t.cell0 + (int)threadIdx.x + c * kBigThreads;
        if (i < area) fa[m.f0 + i] = v[c];
    }
}

__global__ void __launch_bounds__(kBigThreads) k_big_pass(int step, const BigTile* __restrict__ tiles,
                                                          const BigMask* __restrict__ masks,
                                                          const uint32_t* __restrict__ fin, uint32_t* __restrict__ fout)
{
    const BigTile t = big_tile(tiles);
    const BigMask m = masks[t.mask];
    const int area = m.w * m.h;
    // the mask's field as a buffer: 32-bit offsets, one add a tap
    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)(fin + m.f0), 0, area * 4, kRsrcWord3);
    uint32_t v[kBigCells];
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        // past the mask's end a lane computes its last cell again, unstored
        const int i = min(t.cell0 + (int)threadIdx.x + c * kBigThreads, area - 1);
        const int y = div_by(i, m.magic), x = i - y * m.w;
        v[c] = jfa2_taps<true>([&](int o) { return __builtin_amdgcn_raw_buffer_load_b32(rs, (i + o) * 4, 0, 0); }, m.w, m.h,
                         x, y, step);
    }
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        const int i = t.cell0 + (int)threadIdx.x + c * kBigThreads;
        if (i < area) fout[m.f0 + i] = v[c];
    }
}

// the last (step 1) pass, encoded straight away
__global__ void __launch_bounds__(kBigThreads) k_big_last(const BigTile* __restrict__ tiles,
                                                          const BigMask* __restrict__ masks,
                                                          const uint32_t* __restrict__ fin,
                                                          const uint8_t* __restrict__ bytes, int threshold,
                                                          int less_than, uint8_t* __restrict__ out)
{
    const BigTile t = big_tile(tiles);
    const BigMask m = masks[t.mask];
    const int area = m.w * m.h;
    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)(fin + m.f0), 0, area * 4, kRsrcWord3);
    const float max_len = 0.5f * (float)max(m.w, m.h);
    uint32_t v[kBigCells];
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        const int i = min(t.cell0 + (int)threadIdx.x + c * kBigThreads, area - 1);
        const int y = div_by(i, m.magic), x = i - y * m.w;
        v[c] = jfa2_taps<true>([&](int o) { return __builtin_amdgcn_raw_buffer_load_b32(rs, (i + o) * 4, 0, 0); }, m.w, m.h,
                         x, y, 1);
    }
#pragma unroll
    for (int c = 0; c < kBigCells; c++) {
        const int i = t.cell0 + (int)threadIdx.x + c * kBigThreads;
        if (i < area) out[m.off + i] = sdf_u8_cell(v[c], byte_bit(bytes[m.off + i], threshold, less_than), max_len);
    }
}

// ------------------------------------------------------------ variants

struct Variant {
    int threads, groups;   // threads a workgroup, groups (of 64 cells) a wavefront holds
    const void* fn;
};

#define DOLHIP_VARIANT(T, G) { T, G, (const void*)&k_glyph_bins<T, G> }
static const Variant kVariants[] = {
    DOLHIP_VARIANT(1024, 39),   // 39,936 cells, one workgroup a CU
    DOLHIP_VARIANT(1024, 19),   // 19,456 cells, two a CU
    DOLHIP_VARIANT(768, 26),    // 19,968 cells, two a CU
    DOLHIP_VARIANT(512, 39),    // 19,968 cells, two a CU
};
#undef DOLHIP_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static int variant_cells(const Variant& v) { return v.threads * v.groups; }
static size_t variant_lds(const Variant& v) { return (size_t)variant_cells(v) * 4; }   // plus the static group words

// one variant's split of the batch: bins for the masks that fit, tiles of
// the others, and their device copies
struct Plan {
    std::vector<Bin> bins;
    std::vector<BinSlot> slots;
    std::vector<uint32_t> gwords;
    std::vector<BigMask> big;
    std::vector<BigTile> tiles;
    long long big_cells = 0;
    BigMask* d_big = nullptr;
    BigTile* d_tiles = nullptr;
    Bin* d_bins = nullptr;
    BinSlot* d_slots = nullptr;
    uint32_t* d_gwords = nullptr;
    int grid = 0;
};

struct GlyphDesc {
    int w, h;
    long long off;
};

// best fit, largest first: each mask into the bin whose free groups are the
// fewest that still hold it; bins then ordered by work (cells x passes)
static void make_plan(const std::vector<GlyphDesc>& glyphs, const Variant& v, Plan& p)
{
    const int kGroups = v.threads / 64 * v.groups;
    const int cap = kGroups * 64;
    std::vector<int> order(glyphs.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return glyphs[a].w * glyphs[a].h > glyphs[b].w * glyphs[b].h;
    });
    struct Open {
        std::vector<int> members;
        int used = 0;
    };
    std::vector<int> large;
    std::vector<Open> open;
    std::multimap<int, int> by_free;   // free groups -> bin
    for (int gi : order) {
        const GlyphDesc& g = glyphs[gi];
        const int area = g.w * g.h;
        if (area > cap) {
            large.push_back(gi);
            continue;
        }
        const int need = (area + 63) / 64;
        auto it = by_free.lower_bound(need);
        while (it != by_free.end() && open[it->second].members.size() >= 255) ++it;
        int bi;
        if (it == by_free.end()) {
            bi = (int)open.size();
            open.push_back({});
        } else {
            bi = it->second;
            by_free.erase(it);
        }
        open[bi].members.push_back(gi);
        open[bi].used += need;
        if (kGroups - open[bi].used > 0) by_free.insert({kGroups - open[bi].used, bi});
    }
    // large masks in the order given, their tiles in cell order
    std::sort(large.begin(), large.end());
    for (int gi : large) {
        const GlyphDesc& g = glyphs[gi];
        const int area = g.w * g.h;
        for (int c = 0; c < area; c += kBigTile) p.tiles.push_back({(int)p.big.size(), c});
        p.big.push_back(BigMask{g.off, p.big_cells, g.w, g.h, div_magic(g.w), 0});
        p.big_cells += ((long long)area + 63) / 64 * 64;
    }
    // bins ordered by work (cells x (passes + 2)), largest first
    std::vector<long long> work(open.size(), 0);
    for (size_t b = 0; b < open.size(); b++)
        for (int gi : open[b].members) {
            const int md = std::max(glyphs[gi].w, glyphs[gi].h);
            work[b] += (long long)glyphs[gi].w * glyphs[gi].h * (jfa_pass_count(md) + 2);
        }
    std::vector<int> bo(open.size());
    std::iota(bo.begin(), bo.end(), 0);
    std::stable_sort(bo.begin(), bo.end(), [&](int a, int b) { return work[a] > work[b]; });
    p.gwords.assign(bo.size() * (size_t)kGroups, 0u);
    for (size_t k = 0; k < bo.size(); k++) {
        const Open& o = open[bo[k]];
        Bin bin{(int)p.slots.size(), 0, 0, 0};
        for (size_t s = 0; s < o.members.size(); s++) {
            const GlyphDesc& g = glyphs[o.members[s]];
            const int md = std::max(g.w, g.h);
            BinSlot sl{g.off, bin.groups * 64, g.w, g.h, div_magic(g.w), jfa_pass_count(md), floor_pow2(md - 1)};
            const int need = (g.w * g.h + 63) / 64;
            for (int q = 0; q < need; q++)
                p.gwords[k * kGroups + bin.groups + q] = group_word((int)s, g.w, g.h, sl.passes, sl.step0);
            bin.groups += need;
            bin.passes = std::max(bin.passes, sl.passes);
            p.slots.push_back(sl);
        }
        p.bins.push_back(bin);
    }
}

}  // namespace dolhip

using namespace dolhip;

struct dolhip_batch {
    int n = 0;
    int threshold = 128, less_than = 1;
    int64_t bytes = 0;
    int device = 0;
    int variant = 0;
    bool tuned = false;
    std::vector<GlyphDesc> host;   // as given
    Plan plans[kNumVariants];
    int* d_counter = nullptr;
    uint32_t* d_big_a = nullptr;   // the large masks' two fields
    uint32_t* d_big_b = nullptr;
    // graph state
    hipStream_t gstream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* g_bytes = nullptr;
    void* g_out = nullptr;
    int g_variant = -1;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // per-image workspace
    void* d_work = nullptr;
};

namespace {

int fail_b(int code, const char* msg)
{
    set_error("%s", msg);
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? DOLHIP_ERR_OUT_OF_MEMORY : DOLHIP_ERR_DEVICE;
}

#define BTRY(x)                                          \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #x);   \
    } while (0)

template <class T>
hipError_t upload_vec(const std::vector<T>& v, T** d)
{
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc(d, v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

int launch_variant(dolhip_batch* b, int v, const uint8_t* bytes, uint8_t* out, hipStream_t s)
{
    const Variant& var = kVariants[v];
    Plan& p = b->plans[v];
    if (!p.tiles.empty()) {
        const dim3 grid((unsigned)p.tiles.size()), blk(kBigThreads);
        hipLaunchKernelGGL(k_big_border, grid, blk, 0, s, p.d_tiles, p.d_big, bytes, b->threshold, b->less_than,
                           b->d_big_a);
        uint32_t* fi = b->d_big_a;
        uint32_t* fo = b->d_big_b;
        for (int step = 128; step > 1; step /= 2) {
            hipLaunchKernelGGL(k_big_pass, grid, blk, 0, s, step, p.d_tiles, p.d_big, fi, fo);
            std::swap(fi, fo);
        }
        hipLaunchKernelGGL(k_big_last, grid, blk, 0, s, p.d_tiles, p.d_big, fi, bytes, b->threshold,
                           b->less_than, out);
    }
    if (!p.bins.empty()) {
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, b->d_counter);
        int nb = (int)p.bins.size();
        void* args[] = { &nb, &p.d_bins, &p.d_slots, &p.d_gwords, (void*)&bytes, &out, &b->threshold,
                         &b->less_than, &b->d_counter };
        BTRY(hipLaunchKernel(var.fn, dim3(p.grid), dim3(var.threads), args, variant_lds(var), s));
    }
    return check_launch("glyph batch");
}

}  // namespace

extern "C" {

int dolhip_batch_create(int n, const int32_t* w, const int32_t* h, const int64_t* offset,
                        int threshold, int less_than, dolhip_batch** out)
{
    if (!out) return fail_b(DOLHIP_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (n < 1 || !w || !h || !offset) return fail_b(DOLHIP_ERR_INVALID_ARG, "need n >= 1 and w, h, offset arrays");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail_b(DOLHIP_ERR_NO_DEVICE, "no gfx950 (MI355X) device is available");
    }
    auto* b = new dolhip_batch();
    b->n = n;
    b->threshold = threshold;
    b->less_than = less_than != 0;
    b->host.resize(n);
    for (int i = 0; i < n; i++) {
        if (w[i] < 1 || h[i] < 1 || w[i] > kMaxGlyph || h[i] > kMaxGlyph || offset[i] < 0) {
            delete b;
            return fail_b(DOLHIP_ERR_INVALID_ARG, "every mask needs 1 <= w, h <= 256 and offset >= 0");
        }
        b->host[i] = { w[i], h[i], (long long)offset[i] };
        b->bytes = std::max<int64_t>(b->bytes, offset[i] + (int64_t)w[i] * h[i]);
    }
    (void)hipGetDevice(&b->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        delete b;
        return fail_b(DOLHIP_ERR_DEVICE, "hipGetDeviceProperties failed");
    }
    long long big_cells = 0;
    hipError_t e = hipSuccess;
    for (int v = 0; v < kNumVariants && e == hipSuccess; v++) {
        Plan& p = b->plans[v];
        make_plan(b->host, kVariants[v], p);
        (void)hipFuncSetAttribute(kVariants[v].fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)variant_lds(kVariants[v]));
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kVariants[v].fn, kVariants[v].threads,
                                                         variant_lds(kVariants[v])) != hipSuccess || per_cu < 1)
            per_cu = 1;
        p.grid = (int)std::min<long long>((long long)p.bins.size(), (long long)per_cu * prop.multiProcessorCount);
        big_cells = std::max(big_cells, p.big_cells);
        if ((e = upload_vec(p.bins, &p.d_bins)) != hipSuccess || (e = upload_vec(p.slots, &p.d_slots)) != hipSuccess ||
            (e = upload_vec(p.gwords, &p.d_gwords)) != hipSuccess || (e = upload_vec(p.big, &p.d_big)) != hipSuccess ||
            (e = upload_vec(p.tiles, &p.d_tiles)) != hipSuccess)
            break;
    }
    (void)hipGetLastError();
    if (e == hipSuccess) e = hipMalloc(&b->d_counter, sizeof(int));

    if (e == hipSuccess && big_cells > 0) e = hipMalloc(&b->d_big_a, (size_t)big_cells * 4);
    if (e == hipSuccess && big_cells > 0) e = hipMalloc(&b->d_big_b, (size_t)big_cells * 4);
    if (e != hipSuccess) {
        int r = hip_fail(e, "batch allocation");
        dolhip_batch_destroy(b);
        return r;
    }
    *out = b;
    return DOLHIP_OK;
}

int dolhip_batch_destroy(dolhip_batch* b)
{
    if (!b) return DOLHIP_OK;
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->gstream) (void)hipStreamDestroy(b->gstream);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);

    for (Plan& p : b->plans) {
        if (p.d_bins) (void)hipFree(p.d_bins);
        if (p.d_slots) (void)hipFree(p.d_slots);
        if (p.d_gwords) (void)hipFree(p.d_gwords);
        if (p.d_big) (void)hipFree(p.d_big);
        if (p.d_tiles) (void)hipFree(p.d_tiles);
    }
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->d_big_a) (void)hipFree(b->d_big_a);
    if (b->d_big_b) (void)hipFree(b->d_big_b);
    if (b->d_work) (void)hipFree(b->d_work);
    delete b;
    return DOLHIP_OK;
}

int64_t dolhip_batch_bytes(const dolhip_batch* b) { return b ? b->bytes : 0; }

int dolhip_batch_config(const dolhip_batch* b, int* tile_cells, int* waves_per_group)
{
    if (!b) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch is null");
    const Variant& v = kVariants[b->variant];
    if (tile_cells) *tile_cells = variant_cells(v);
    if (waves_per_group) *waves_per_group = v.threads / 64;
    return DOLHIP_OK;
}

int dolhip_batch_set_variant(dolhip_batch* b, int variant)
{
    if (!b) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch is null");
    if (variant < 0) return kNumVariants;
    if (variant >= kNumVariants) return fail_b(DOLHIP_ERR_INVALID_ARG, "no such variant");
    b->variant = variant;
    return DOLHIP_OK;
}

int dolhip_batch_sdf_u8_dev(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, int graph,
                            void* stream)
{
    if (!b || !bytes || !out) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch, bytes and out must be set");
    hipStream_t s = (hipStream_t)stream;
    if (!graph) return launch_variant(b, b->variant, bytes, out, s);

    // graph replay: the captured launch pair runs on the batch's own stream,
    // ordered after the caller's stream by events
    if (!b->gstream) {
        BTRY(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking));
        BTRY(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
        BTRY(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    }
    if (!b->exec || b->g_bytes != bytes || b->g_out != out || b->g_variant != b->variant) {
        if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        BTRY(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
        int r = launch_variant(b, b->variant, bytes, out, b->gstream);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(b->gstream, &gr);
        if (r != DOLHIP_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return r;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = gr;
        BTRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->g_bytes = bytes;
        b->g_out = out;
        b->g_variant = b->variant;
    }
    BTRY(hipEventRecord(b->ev_in, s));
    BTRY(hipStreamWaitEvent(b->gstream, b->ev_in, 0));
    BTRY(hipGraphLaunch(b->exec, b->gstream));
    BTRY(hipEventRecord(b->ev_out, b->gstream));
    BTRY(hipStreamWaitEvent(s, b->ev_out, 0));
    return DOLHIP_OK;
}

int dolhip_batch_tune(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, void* stream)
{
    if (!b || !bytes || !out) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch, bytes and out must be set");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    BTRY(hipEventCreate(&e0));
    BTRY(hipEventCreate(&e1));
    float best = 1e30f;
    int best_v = b->variant;
    for (int v = 0; v < kNumVariants; v++) {
        int r = launch_variant(b, v, bytes, out, s);   // warm
        if (r != DOLHIP_OK) continue;
        BTRY(hipEventRecord(e0, s));
        for (int k = 0; k < 3; k++) launch_variant(b, v, bytes, out, s);
        BTRY(hipEventRecord(e1, s));
        BTRY(hipEventSynchronize(e1));
        float ms = 0;
        BTRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; best_v = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->variant = best_v;
    b->tuned = true;
    return DOLHIP_OK;
}

// One mask through the per-image launches.  work: mask words + two delta
// fields (NULL: the batch-independent library buffer is not used; callers
// pass memory of glyph_workspace bytes).
static size_t glyph_workspace(int w, int h)
{
    const size_t mw = (size_t)((w + 31) >> 5) * h;
    return (mw + 2 * (size_t)w * h) * 4 + 512;
}

int dolhip_glyph_sdf_u8_dev(int w, int h, const uint8_t* bytes, int threshold, int less_than,
                            uint8_t* out, void* work, void* stream)
{
    if (w < 1 || h < 1 || w > 32767 || h > 32767) return fail_b(DOLHIP_ERR_INVALID_ARG, "bad size");
    if (!bytes || !out || !work) return fail_b(DOLHIP_ERR_INVALID_ARG, "bytes, out and work must be set");
    hipStream_t s = (hipStream_t)stream;
    const size_t mw = (size_t)((w + 31) >> 5) * h;
    uint32_t* mask = (uint32_t*)work;
    int16_t* da = (int16_t*)((char*)work + ((mw * 4 + 255) & ~(size_t)255));
    int16_t* db = da + 2 * (size_t)w * h;
    launch_mask_from_bytes(w, h, bytes, threshold, less_than, mask, s);
    launch_border_2d(w, h, mask, da, s);
    int16_t* a = da;
    int16_t* t = db;
    for (int step = floor_pow2((w >= h ? w : h) - 1); step > 0; step /= 2) {
        launch_jfa2d_pass(step, w, h, a, t, 2, s);
        std::swap(a, t);
    }
    if (a != da) launch_jfa2d_pass(1, w, h, a, da, 2, s);
    launch_signed_2d(w, h, mask, da, 0.5f * (float)(w >= h ? w : h), out, 1, s);
    return check_launch("glyph per-image");
}

// every mask of the batch through the per-image launches, in batch order
int dolhip_batch_per_image_dev(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, void* stream)
{
    if (!b || !bytes || !out) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch, bytes and out must be set");
    if (!b->d_work) BTRY(hipMalloc(&b->d_work, glyph_workspace(kMaxGlyph, kMaxGlyph)));
    for (const GlyphDesc& g : b->host) {
        int r = dolhip_glyph_sdf_u8_dev(g.w, g.h, bytes + g.off, b->threshold, b->less_than,
                                        out + g.off, b->d_work, stream);
        if (r != DOLHIP_OK) return r;
    }
    return DOLHIP_OK;
}

}  // extern "C"
