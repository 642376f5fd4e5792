// dol_batch.hip -- signed 8-bit SDFs of many small masks in one launch.
//
// Per mask the pipeline of the reference tool's signed border mode
// (DistanceOcclusionTool.cpp:634-670 threshold, :1333-1365 border init +
// JumpFlood(cw = 2), :1240-1271 EncodeU8Signed with 2 * maxLen where
// maxLen = 0.25 max(w, h), :1973-1974).
//
// Batched form: one persistent kernel.  Workgroups take masks from an atomic
// queue, largest first (the host sorts them), and run the whole pipeline for
// a mask without leaving the workgroup: the bitmask is built in LDS with
// wavefront ballots; the delta field lives in LDS when it fits the tuned
// budget (each JumpFlood pass computes into registers, barrier, writes back,
// which keeps the reference's out-of-place semantics with one buffer), or
// else in a per-workgroup ping-pong area in global memory that stays in L2.
// The queue counter is reset by a one-thread kernel (not a memset) so the
// pair can be captured in a hipGraph and replayed.  A tuner times the
// (threads, cells-per-thread) variants on the actual batch and keeps the
// fastest.
//
// Per-image form: the same per-cell functions (dol_cells.hpp) through the
// ordinary one-kernel-per-step launches, for comparison.
#include "../../include/dol_hip.h"
#include "dol_cells.hpp"

#include <algorithm>
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

struct GlyphDesc {
    int w, h;
    long long off;
};

constexpr int kMaxGlyph = 256;
constexpr int kMaskWords = (kMaxGlyph / 32) * kMaxGlyph;   // 2048 words for 256 x 256

__global__ void k_reset_counter(int* c) { *c = 0; }

template <int T, int CPT>
__global__ void __launch_bounds__(T) k_glyph_batch(int n, const GlyphDesc* __restrict__ desc,
                                                   const uint8_t* __restrict__ bytes,
                                                   uint8_t* __restrict__ out, int threshold,
                                                   int less_than, uint32_t* __restrict__ scratch,
                                                   int scratch_cells, int* __restrict__ counter)
{
    constexpr int kLdsCells = T * CPT;
    extern __shared__ uint32_t smem[];
    uint32_t* sdelta = smem;                 // kLdsCells
    uint32_t* smask = smem + kLdsCells;      // kMaskWords
    __shared__ int s_next;
    uint32_t* ga = scratch + (size_t)blockIdx.x * 2 * scratch_cells;
    uint32_t* gb = ga + scratch_cells;
    const int tid = threadIdx.x;
    const int lane = tid & 63;

    for (;;) {
        if (tid == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int g = s_next;
        __syncthreads();
        if (g >= n) break;
        const GlyphDesc gd = desc[g];
        const int w = gd.w, h = gd.h, area = w * h;
        const int ws = (w + 31) >> 5;
        const uint8_t* src = bytes + gd.off;
        uint8_t* dst = out + gd.off;
        // thread strides over cells in row-major order: (x, y) += (rT, qT)
        const int qT = T / w, rT = T - qT * w;

        // 1. threshold into the LDS bitmask, one ballot per 64 padded pixels
        const int padded = ws * 32 * h;
        for (int base = 0; base < padded; base += T) {
            const int p = base + tid;
            int pred = 0;
            if (p < padded) {
                const int row = p / (ws * 32), x = p - row * (ws * 32);
                if (x < w) pred = ((int)src[row * w + x] >= threshold) != (less_than != 0);
            }
            const unsigned long long b = __ballot(pred);
            if ((lane & 31) == 0 && p < padded)
                smask[p >> 5] = lane == 0 ? (uint32_t)b : (uint32_t)(b >> 32);
        }
        __syncthreads();

        const bool in_lds = area <= kLdsCells;
        uint32_t* cur = in_lds ? sdelta : ga;
        uint32_t* nxt = gb;

        // 2. border init
        {
            int x = tid % w, y = tid / w;
            for (int i = tid; i < area; i += T) {
                cur[i] = border2_cell(smask, ws, w, h, x, y);
                x += rT; y += qT;
                if (x >= w) { x -= w; y++; }
            }
        }
        __syncthreads();

        // 3. JumpFlood in half cells
        const int maxdim = w >= h ? w : h;
        const int passes = jfa_pass_count(maxdim);
        for (int k = 0; k < passes; k++) {
            const int step = jfa_pass_step(maxdim, k);
            if (in_lds) {
                uint32_t r[CPT];
                int x = tid % w, y = tid / w;
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int i = tid + c * T;
                    if (i < area) r[c] = jfa2_cell(sdelta, w, h, x, y, step, 2);
                    x += rT; y += qT;
                    if (x >= w) { x -= w; y++; }
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int i = tid + c * T;
                    if (i < area) sdelta[i] = r[c];
                }
            } else {
                int x = tid % w, y = tid / w;
                for (int i = tid; i < area; i += T) {
                    nxt[i] = jfa2_cell(cur, w, h, x, y, step, 2);
                    x += rT; y += qT;
                    if (x >= w) { x -= w; y++; }
                }
                uint32_t* t = cur; cur = nxt; nxt = t;
            }
            __syncthreads();
        }

        // 4. encode
        {
            const float max_len = 0.5f * (float)maxdim;
            int x = tid % w, y = tid / w;
            for (int i = tid; i < area; i += T) {
                const int bit = (int)((smask[y * ws + (x >> 5)] >> (x & 31)) & 1u);
                dst[i] = sdf_u8_cell(cur[i], bit, max_len);
                x += rT; y += qT;
                if (x >= w) { x -= w; y++; }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ variants

struct Variant {
    int threads, cpt;
    const void* fn;
};

#define DOLHIP_VARIANT(T, C) { T, C, (const void*)&k_glyph_batch<T, C> }
static const Variant kVariants[] = {
    DOLHIP_VARIANT(256, 16), DOLHIP_VARIANT(256, 32), DOLHIP_VARIANT(256, 64),
    DOLHIP_VARIANT(512, 8),  DOLHIP_VARIANT(512, 16), DOLHIP_VARIANT(512, 32),
    DOLHIP_VARIANT(1024, 8), DOLHIP_VARIANT(1024, 16), DOLHIP_VARIANT(1024, 32),
};
#undef DOLHIP_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static size_t variant_lds(const Variant& v)
{
    return ((size_t)v.threads * v.cpt + kMaskWords) * 4;
}

}  // namespace dolhip

using namespace dolhip;

struct dolhip_batch {
    int n = 0;
    int threshold = 128, less_than = 1;
    int64_t bytes = 0;
    int max_area = 1;
    int device = 0;
    int variant = 5;      // chosen by the tuner, default 512 x 32
    bool tuned = false;
    GlyphDesc* d_desc = nullptr;
    int* d_counter = nullptr;
    uint32_t* d_scratch = nullptr;
    int grid[kNumVariants] = {};
    int scratch_grid = 0;
    std::vector<GlyphDesc> host;   // sorted, largest first
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

int launch_variant(dolhip_batch* b, int v, const uint8_t* bytes, uint8_t* out, hipStream_t s)
{
    const Variant& var = kVariants[v];
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, b->d_counter);
    void* args[] = { &b->n, &b->d_desc, (void*)&bytes, &out, &b->threshold, &b->less_than,
                     &b->d_scratch, &b->max_area, &b->d_counter };
    BTRY(hipLaunchKernel(var.fn, dim3(b->grid[v]), dim3(var.threads), args, variant_lds(var), s));
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
        b->max_area = std::max(b->max_area, w[i] * h[i]);
    }
    std::stable_sort(b->host.begin(), b->host.end(), [](const GlyphDesc& a, const GlyphDesc& c) {
        return a.w * a.h > c.w * c.h;
    });
    (void)hipGetDevice(&b->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        delete b;
        return fail_b(DOLHIP_ERR_DEVICE, "hipGetDeviceProperties failed");
    }
    int maxgrid = 1;
    for (int v = 0; v < kNumVariants; v++) {
        (void)hipFuncSetAttribute(kVariants[v].fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)variant_lds(kVariants[v]));
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kVariants[v].fn, kVariants[v].threads,
                                                         variant_lds(kVariants[v])) != hipSuccess || per_cu < 1)
            per_cu = 1;
        b->grid[v] = std::min(n, per_cu * prop.multiProcessorCount);
        maxgrid = std::max(maxgrid, b->grid[v]);
    }
    (void)hipGetLastError();
    b->scratch_grid = maxgrid;
    hipError_t e;
    if ((e = hipMalloc(&b->d_desc, sizeof(GlyphDesc) * n)) != hipSuccess ||
        (e = hipMalloc(&b->d_counter, sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&b->d_scratch, (size_t)maxgrid * 2 * b->max_area * 4)) != hipSuccess ||
        (e = hipMemcpy(b->d_desc, b->host.data(), sizeof(GlyphDesc) * n, hipMemcpyHostToDevice)) != hipSuccess) {
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
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->d_scratch) (void)hipFree(b->d_scratch);
    if (b->d_work) (void)hipFree(b->d_work);
    delete b;
    return DOLHIP_OK;
}

int64_t dolhip_batch_bytes(const dolhip_batch* b) { return b ? b->bytes : 0; }

int dolhip_batch_config(const dolhip_batch* b, int* tile_cells, int* waves_per_group)
{
    if (!b) return fail_b(DOLHIP_ERR_INVALID_ARG, "batch is null");
    const Variant& v = kVariants[b->variant];
    if (tile_cells) *tile_cells = v.threads * v.cpt;
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
