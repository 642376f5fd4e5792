// dol_capi.cpp -- host runtime and C ABI of libdolhip (include/dol_hip.h).
//
// Host-pointer entry points keep the reference's synchronous, in-place
// contract (DistanceOcclusionLib.h:61-98): copy in, run the kernels, copy
// back, return.  Device-pointer entry points (_dev) enqueue on the caller's
// stream and return.  Failures are reported as status codes with a message
// in dolhip_last_error(); nothing falls back to the CPU.
#include "../../include/dol_hip.h"
#include "dol_common.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace dolhip {

// kernel launchers (dol_init.hip, dol_jfa.hip, dol_edt.hip, dol_encode.hip)
void launch_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                            uint32_t* mask, hipStream_t s);
void launch_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert,
                           hipStream_t s);
void launch_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert,
                           hipStream_t s);
template <class T>
void launch_init_dist(int w, int h, const uint32_t* mask, T* dist, T max_d, hipStream_t s);
void launch_border_2d(int w, int h, const uint32_t* mask, int16_t* delta, hipStream_t s);
void launch_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, hipStream_t s);
void launch_fast_sweep_2d(int w, int h, int16_t* delta, int cw, hipStream_t s);
void launch_init_dirw(int w, int h, int d, int planes, const uint32_t* mask, float* dirw, hipStream_t s);
void launch_occlusion_sweep(int mode, int w, int h, int d, const float* in, float* dirw, hipStream_t s);
void launch_encode_ao(size_t n, int planes, bool rgba, const float* dirw, uint8_t* out, hipStream_t s);
void launch_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s);
void launch_brute_force_2d(int w, int h, int16_t* delta, int cw, hipStream_t s);
void launch_brute_force_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s);
void launch_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale, hipStream_t s);
void launch_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow, hipStream_t s);
template <class T>
void launch_chamfer(int w, int h, T* dist, hipStream_t s);
void launch_danielsson_2d(int w, int h, int16_t* delta, int cw, hipStream_t s);
void launch_danielsson_3d(int w, int h, int d, int16_t* delta, int cw, hipStream_t s);
void launch_border_image(int w, int h, int d, int dims, const uint32_t* mask, const int16_t* delta,
                         float max_len, int type, uint8_t* out, hipStream_t s);
void launch_distance_u8(size_t n, int dims, const int16_t* delta, float max_len, uint8_t* out, hipStream_t s);
template <class T>
void launch_scalar_u8(size_t n, const T* v, float max_len, uint8_t* out, hipStream_t s);
void launch_sqrt_f32(size_t n, const float* v, float* out, hipStream_t s);
void launch_jfa2d_pass(int step, int w, int h, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s);
size_t jfa3d_workspace(int w, int h, int d);
void jfa3d_run(int w, int h, int d, int16_t* delta, int cw, void* work, hipStream_t s, bool bounded = false);
void launch_jfa3d_pass(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw,
                       hipStream_t s);
size_t felz_workspace(int w, int h);
void launch_felzenszwalb_2d(int w, int h, float* dist, void* work, hipStream_t s);
size_t edt_workspace(int w, int h, int d);
void launch_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                   void* work, hipStream_t s);
void launch_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2, float* fl,
                   void* work, hipStream_t s);
void launch_dist_2d(size_t n, const int16_t* delta, float* out, hipStream_t s);
void launch_signed_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                      void* out, int mode, hipStream_t s);
void launch_signed_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                      float max_len, void* out, int mode, hipStream_t s);
void launch_sdf_delta_rgba(int w, int h, const int16_t* delta, float max_len, uint8_t* out,
                           hipStream_t s);
void launch_sdf_rg_rgba(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                        uint8_t* out, hipStream_t s);
void launch_signed_edt(size_t n, const int32_t* dout, const int32_t* din, float* sdf,
                       hipStream_t s);
void launch_voxelise(int tri_count, const int* indices, const float* vertices, const float* bbox,
                     int w, int h, int d, uint32_t* mask, hipStream_t s);
void launch_two_field_2d(int w, int h, const int16_t* d0, const int16_t* d1, float max_len,
                         int type, uint8_t* out, hipStream_t s);
void launch_two_field_3d(int w, int h, int d, const int16_t* d0, const int16_t* d1, float max_len,
                         int type, uint8_t* out, hipStream_t s);

// ------------------------------------------------------------------ errors

static thread_local std::string g_error;

void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

static int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

static int fail_hip(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory)
        return fail(DOLHIP_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
    return fail(DOLHIP_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(e, what);
    return DOLHIP_OK;
}

#define TRY_HIP(x)                                   \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return fail_hip(e_, #x); \
    } while (0)

#define TRY(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_ != DOLHIP_OK) return r_; \
    } while (0)

// ------------------------------------------------------------------ device

static int g_device_ok = -1;   // -1 unknown, 0 none, 1 usable
static std::mutex g_mutex;

static int require_device()
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_device_ok < 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            g_device_ok = 0;
        } else {
            hipDeviceProp_t prop;
            int dev = 0;
            (void)hipGetDevice(&dev);
            g_device_ok = (hipGetDeviceProperties(&prop, dev) == hipSuccess &&
                           std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
                              ? 1
                              : 0;
        }
    }
    if (!g_device_ok)
        return fail(DOLHIP_ERR_NO_DEVICE, "no gfx950 (MI355X) device is available");
    return DOLHIP_OK;
}

// Library-owned work buffers for _dev calls made with work == NULL: one per
// (device, stream), since work on one stream runs in order and can share a
// buffer while work on two streams may overlap.  A buffer that is too small
// is replaced by a larger one, and the old one is kept until the process
// ends rather than freed: work already enqueued with it, or about to be
// enqueued by a thread that obtained it, may still use it.
struct WorkBuf {
    int dev;
    void* stream;
    void* ptr;
    size_t bytes;
};
static std::vector<WorkBuf> g_work;
static std::vector<void*> g_retired_work;

static int library_work(size_t bytes, void* stream, void** out)
{
    int dev = 0;
    TRY_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mutex);
    WorkBuf* wb = nullptr;
    for (auto& b : g_work)
        if (b.dev == dev && b.stream == stream) {
            wb = &b;
            break;
        }
    if (!wb) {
        g_work.push_back({dev, stream, nullptr, 0});
        wb = &g_work.back();
    }
    if (wb->bytes < bytes) {
        const size_t want = bytes > 2 * wb->bytes ? bytes : 2 * wb->bytes;
        void* p = nullptr;
        TRY_HIP(hipMalloc(&p, want));
        if (wb->ptr) g_retired_work.push_back(wb->ptr);
        wb->ptr = p;
        wb->bytes = want;
    }
    *out = wb->ptr;
    return DOLHIP_OK;
}

static int resolve_work(void* given, size_t bytes, void* stream, void** out)
{
    if (given) {
        *out = given;
        return DOLHIP_OK;
    }
    return library_work(bytes, stream, out);
}

// RAII device buffer for the host-pointer entry points
struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes)
    {
        hipError_t e = hipMalloc(&p, bytes ? bytes : 4);
        if (e != hipSuccess) {
            p = nullptr;
            return fail_hip(e, "hipMalloc");
        }
        return DOLHIP_OK;
    }
    int upload(const void* src, size_t bytes)
    {
        TRY(alloc(bytes));
        if (bytes) TRY_HIP(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        return DOLHIP_OK;
    }
    int download(void* dst, size_t bytes)
    {
        if (bytes) TRY_HIP(hipMemcpy(dst, p, bytes, hipMemcpyDeviceToHost));
        return DOLHIP_OK;
    }
    template <class T>
    T* as() { return (T*)p; }
};

// ------------------------------------------------------------------ profiling

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;

ProfileScope::ProfileScope(const char* name, hipStream_t s) : name_(name), s_(s), ev0_(nullptr)
{
    if (!g_prof) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev0_ = e;
}

ProfileScope::~ProfileScope()
{
    if (!ev0_) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s_);
    std::lock_guard<std::mutex> lk(g_mutex);
    g_prof_recs.push_back({name_, (hipEvent_t)ev0_, e});
}

static void prof_clear()
{
    for (auto& r : g_prof_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
}

// ------------------------------------------------------------------ helpers

static inline hipStream_t as_stream(void* s) { return (hipStream_t)s; }

static int check_dims(int w, int h, int d = 1)
{
    if (w < 1 || h < 1 || d < 1)
        return fail(DOLHIP_ERR_INVALID_ARG, "dimensions must be positive (w=%d h=%d d=%d)", w, h, d);
    if ((long long)w * h * d > (1LL << 34))
        return fail(DOLHIP_ERR_INVALID_ARG, "volume too large (%d x %d x %d)", w, h, d);
    if (w > 32767 || h > 32767 || d > 32767)
        return fail(DOLHIP_ERR_INVALID_ARG, "each dimension must be below 32768 (int16 deltas)");
    return DOLHIP_OK;
}

static int check_ptr(const void* p, const char* name)
{
    if (!p) return fail(DOLHIP_ERR_INVALID_ARG, "%s is null", name);
    return DOLHIP_OK;
}

// a cCellDelta2 field on the device: the kernels access its cells as 32-bit
// words, so the pointer must be 4-byte aligned (the struct itself only needs 2)
static int check_cells2(const void* p, const char* name)
{
    TRY(check_ptr(p, name));
    if ((uintptr_t)p & 3) return fail(DOLHIP_ERR_INVALID_ARG, "%s must be 4-byte aligned", name);
    return DOLHIP_OK;
}

// cell width of the delta units: 1 for whole cells, 2 for the half cells of
// the border methods; any positive value the int16 deltas can hold
static int check_cw(int cw)
{
    if (cw < 1 || cw > 16384)
        return fail(DOLHIP_ERR_INVALID_ARG, "cw must be in 1 .. 16384, got %d", cw);
    return DOLHIP_OK;
}

static inline size_t mask_bytes(int w, long long rows) { return (size_t)mask_stride(w) * rows * 4; }

static int finish(hipStream_t s, const char* what)
{
    TRY(check_launch(what));
    (void)s;
    return DOLHIP_OK;
}

static int finish_sync(const char* what)
{
    TRY(check_launch(what));
    TRY_HIP(hipDeviceSynchronize());
    return DOLHIP_OK;
}

// FloorPow2 of DistanceOcclusionLib.cpp:875-884
static int floor_pow2(int v)
{
    uint32_t u = (uint32_t)v;
    u |= u >> 1; u |= u >> 2; u |= u >> 4; u |= u >> 8; u |= u >> 16;
    return (int)((int32_t)(u + 1u) >> 1);
}

// pass schedule of DistanceOcclusionLib.cpp:887-908 / 2450-2472
static void jfa2d_schedule(int w, int h, int16_t* delta, int cw, int16_t* tmp, hipStream_t s)
{
    int16_t* a = delta;
    int16_t* b = tmp;
    for (int step = floor_pow2((w >= h ? w : h) - 1); step > 0; step /= 2) {
        launch_jfa2d_pass(step, w, h, a, b, cw, s);
        int16_t* t = a; a = b; b = t;
    }
    if (a != delta) launch_jfa2d_pass(1, w, h, a, delta, cw, s);
}

template <class T>
int init_dist_host(int w, int h, const uint32_t* mask, T* dist, T max_d)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dist, "dist"));
    TRY(require_device());
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(o.alloc((size_t)w * h * sizeof(T)));
    launch_init_dist<T>(w, h, m.as<uint32_t>(), o.as<T>(), max_d, 0);
    TRY(finish_sync("init_distances"));
    return o.download(dist, (size_t)w * h * sizeof(T));
}

}  // namespace dolhip

using namespace dolhip;

extern "C" {

const char* dolhip_last_error(void) { return g_error.c_str(); }
const char* dolhip_version(void) { return "dolhip 0.1 gfx950"; }

int dolhip_device_count(void)
{
    if (require_device() != DOLHIP_OK) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int dolhip_profile_enable(int enable)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    prof_clear();
    g_prof = enable != 0;
    return DOLHIP_OK;
}

int dolhip_profile_query(const char* prefix, double* avg_ms, int* count)
{
    TRY(check_ptr(prefix, "prefix"));
    TRY_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(g_mutex);
    double total = 0.0;
    int n = 0;
    size_t pl = std::strlen(prefix);
    for (auto& r : g_prof_recs) {
        if (r.name.compare(0, pl, prefix) != 0) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            total += ms;
            n++;
        }
    }
    if (avg_ms) *avg_ms = n ? total / n : 0.0;
    if (count) *count = n;
    return DOLHIP_OK;
}

// ---- masks

int dolhip_mask_from_bytes_dev(int w, int rows, const uint8_t* bytes, int threshold,
                               int less_than, uint32_t* mask, void* stream)
{
    TRY(check_dims(w, rows));
    TRY(check_ptr(bytes, "bytes"));
    TRY(check_ptr(mask, "mask"));
    TRY(require_device());
    launch_mask_from_bytes(w, rows, bytes, threshold, less_than, mask, as_stream(stream));
    return finish(as_stream(stream), "mask_from_bytes");
}

int dolhip_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                           uint32_t* mask)
{
    TRY(check_dims(w, rows));
    TRY(check_ptr(bytes, "bytes"));
    TRY(check_ptr(mask, "mask"));
    TRY(require_device());
    DevBuf b, m;
    TRY(b.upload(bytes, (size_t)w * rows));
    TRY(m.alloc(mask_bytes(w, rows)));
    launch_mask_from_bytes(w, rows, b.as<uint8_t>(), threshold, less_than, m.as<uint32_t>(), 0);
    TRY(finish_sync("mask_from_bytes"));
    return m.download(mask, mask_bytes(w, rows));
}

// ---- initialisation

int dolhip_init_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta, int invert,
                              void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_cells2(delta, "delta"));
    TRY(require_device());
    launch_init_deltas_2d(w, h, mask, delta, invert, as_stream(stream));
    return finish(as_stream(stream), "init_deltas_2d");
}

int dolhip_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(o.alloc((size_t)w * h * 4));
    launch_init_deltas_2d(w, h, m.as<uint32_t>(), o.as<int16_t>(), invert, 0);
    TRY(finish_sync("init_deltas_2d"));
    return o.download(delta, (size_t)w * h * 4);
}

int dolhip_init_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                              int invert, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    launch_init_deltas_3d(w, h, d, mask, delta, invert, as_stream(stream));
    return finish(as_stream(stream), "init_deltas_3d");
}

int dolhip_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    const size_t cells = (size_t)w * h * d;
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(o.alloc(cells * 6));
    launch_init_deltas_3d(w, h, d, m.as<uint32_t>(), o.as<int16_t>(), invert, 0);
    TRY(finish_sync("init_deltas_3d"));
    return o.download(delta, cells * 6);
}

int dolhip_init_border_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta,
                                     void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_cells2(delta, "delta"));
    TRY(require_device());
    launch_border_2d(w, h, mask, delta, as_stream(stream));
    return finish(as_stream(stream), "border_2d");
}

int dolhip_init_border_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(o.alloc((size_t)w * h * 4));
    launch_border_2d(w, h, m.as<uint32_t>(), o.as<int16_t>(), 0);
    TRY(finish_sync("border_2d"));
    return o.download(delta, (size_t)w * h * 4);
}

int dolhip_init_border_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                                     void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    launch_border_3d(w, h, d, mask, delta, as_stream(stream));
    return finish(as_stream(stream), "border_3d");
}

int dolhip_init_border_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    const size_t cells = (size_t)w * h * d;
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(o.alloc(cells * 6));
    launch_border_3d(w, h, d, m.as<uint32_t>(), o.as<int16_t>(), 0);
    TRY(finish_sync("border_3d"));
    return o.download(delta, cells * 6);
}

int dolhip_init_distances_f32(int w, int h, const uint32_t* mask, float* dist, float max_d)
{ return init_dist_host<float>(w, h, mask, dist, max_d); }
int dolhip_init_distances_i32(int w, int h, const uint32_t* mask, int32_t* dist, int32_t max_d)
{ return init_dist_host<int32_t>(w, h, mask, dist, max_d); }
int dolhip_init_distances_i16(int w, int h, const uint32_t* mask, int16_t* dist, int16_t max_d)
{ return init_dist_host<int16_t>(w, h, mask, dist, max_d); }

int dolhip_init_distances_f32_dev(int w, int h, const uint32_t* mask, float* dist, float max_d,
                                  void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dist, "dist"));
    TRY(require_device());
    launch_init_dist<float>(w, h, mask, dist, max_d, as_stream(stream));
    return finish(as_stream(stream), "init_distances");
}

// ---- propagation

size_t dolhip_jump_flood_2d_workspace(int w, int h)
{
    return (w < 1 || h < 1) ? 0 : (size_t)w * h * 4;
}

int dolhip_jump_flood_2d_dev(int w, int h, int16_t* delta, int cw, void* work, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_jump_flood_2d_workspace(w, h), stream, &tmp));
    jfa2d_schedule(w, h, delta, cw, (int16_t*)tmp, as_stream(stream));
    return finish(as_stream(stream), "jump_flood_2d");
}

int dolhip_jump_flood_2d(int w, int h, int16_t* delta, int cw)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf a, t;
    TRY(a.upload(delta, bytes));
    TRY(t.alloc(bytes));
    jfa2d_schedule(w, h, a.as<int16_t>(), cw, t.as<int16_t>(), 0);
    TRY(finish_sync("jump_flood_2d"));
    return a.download(delta, bytes);
}

// ---- JumpFlood 2D without allocation
//
// The second field of the passes lives in `scratch`, a static device array
// of DOLHIP_NOALLOC_SCRATCH_BYTES that every device holds from module load
// until the process ends.  Host code reaches each device's copy through the
// address hipGetSymbolAddress gives with that device current (the array's
// name is only a host-side stand-in), looked up on the first call on a
// device and kept.

static __device__ __align__(256) uint8_t scratch[DOLHIP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_address(void** out)
{
    int dev = 0;
    TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(DOLHIP_ERR_DEVICE, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DOLHIP_OK;
}

int dolhip_jump_flood_2d_noalloc(int w, int h, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta, "delta"));
    TRY(check_cw(cw));
    if ((size_t)w * h * 4 > (size_t)DOLHIP_NOALLOC_SCRATCH_BYTES)
        return fail(DOLHIP_ERR_INVALID_ARG,
                    "dolhip_jump_flood_2d_noalloc: %d x %d cells need %zu bytes of scratch, more than its %d",
                    w, h, (size_t)w * h * 4, DOLHIP_NOALLOC_SCRATCH_BYTES);
    TRY(require_device());
    void* tmp = nullptr;
    TRY(scratch_address(&tmp));
    jfa2d_schedule(w, h, delta, cw, (int16_t*)tmp, as_stream(stream));
    return finish(as_stream(stream), "jump_flood_2d_noalloc");
}

size_t dolhip_jump_flood_3d_workspace(int w, int h, int d)
{
    return (w < 1 || h < 1 || d < 1) ? 0 : jfa3d_workspace(w, h, d);
}

int dolhip_jump_flood_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* work,
                             void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_jump_flood_3d_workspace(w, h, d), stream, &tmp));
    jfa3d_run(w, h, d, delta, cw, tmp, as_stream(stream));
    return finish(as_stream(stream), "jump_flood_3d");
}

int dolhip_jump_flood_3d(int w, int h, int d, int16_t* delta, int cw)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 6;
    DevBuf a, t;
    TRY(a.upload(delta, bytes));
    TRY(t.alloc(dolhip_jump_flood_3d_workspace(w, h, d)));
    jfa3d_run(w, h, d, a.as<int16_t>(), cw, t.p, 0);
    TRY(finish_sync("jump_flood_3d"));
    return a.download(delta, bytes);
}

int dolhip_fast_sweep_2d_dev(int w, int h, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    launch_fast_sweep_2d(w, h, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "fast_sweep_2d");
}

int dolhip_fast_sweep_2d(int w, int h, int16_t* delta, int cw)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_fast_sweep_2d(w, h, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("fast_sweep_2d"));
    return a.download(delta, bytes);
}

int dolhip_fast_sweep_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    launch_fast_sweep_3d(w, h, d, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "fast_sweep_3d");
}

int dolhip_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 6;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_fast_sweep_3d(w, h, d, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("fast_sweep_3d"));
    return a.download(delta, bytes);
}

// ---- BruteForce

static int check_brute(long long cells)
{
    if (cells >= (1LL << 31))
        return fail(DOLHIP_ERR_INVALID_ARG, "BruteForce takes at most 2^31 - 1 cells");
    return DOLHIP_OK;
}

int dolhip_brute_force_2d_dev(int w, int h, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_brute((long long)w * h));
    TRY(require_device());
    launch_brute_force_2d(w, h, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "brute_force_2d");
}

int dolhip_brute_force_2d(int w, int h, int16_t* delta, int cw)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_brute((long long)w * h));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_brute_force_2d(w, h, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("brute_force_2d"));
    return a.download(delta, bytes);
}

int dolhip_brute_force_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_brute((long long)w * h * d));
    TRY(require_device());
    launch_brute_force_3d(w, h, d, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "brute_force_3d");
}

int dolhip_brute_force_3d(int w, int h, int d, int16_t* delta, int cw)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_brute((long long)w * h * d));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 6;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_brute_force_3d(w, h, d, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("brute_force_3d"));
    return a.download(delta, bytes);
}

// ---- raster-order methods: one workgroup walks the lines in order

// the LDS row copy of Danielsson holds a row of cells (4 B in 2D, 6 B in 3D)
static int check_row_lds(int w, int bytes_per_cell, const char* what)
{
    if ((long long)w * bytes_per_cell + 16 > 160 * 1024)
        return fail(DOLHIP_ERR_INVALID_ARG, "%s takes rows of at most %d cells", what,
                    (160 * 1024 - 16) / bytes_per_cell);
    return DOLHIP_OK;
}

int dolhip_danielsson_2d_dev(int w, int h, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_row_lds(w, 4, "Danielsson 2D"));
    TRY(require_device());
    launch_danielsson_2d(w, h, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "danielsson_2d");
}

int dolhip_danielsson_2d(int w, int h, int16_t* delta, int cw)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_row_lds(w, 4, "Danielsson 2D"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_danielsson_2d(w, h, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("danielsson_2d"));
    return a.download(delta, bytes);
}

int dolhip_danielsson_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_row_lds(w, 6, "Danielsson 3D"));
    TRY(require_device());
    launch_danielsson_3d(w, h, d, delta, cw, as_stream(stream));
    return finish(as_stream(stream), "danielsson_3d");
}

int dolhip_danielsson_3d(int w, int h, int d, int16_t* delta, int cw)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta, "delta"));
    TRY(check_cw(cw));
    TRY(check_row_lds(w, 6, "Danielsson 3D"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 6;
    DevBuf a;
    TRY(a.upload(delta, bytes));
    launch_danielsson_3d(w, h, d, a.as<int16_t>(), cw, 0);
    TRY(finish_sync("danielsson_3d"));
    return a.download(delta, bytes);
}

extern "C++" {
// Chamfer's multiplier (DistanceOcclusionLib.cpp:912-1006 return d0): distances
// come out in units of 1/3 cell
static const int kChamferMultiplier = 3;

template <class T>
static int chamfer_dev(int w, int h, T* dist, int* multiplier, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dist, "distances"));
    TRY(require_device());
    launch_chamfer<T>(w, h, dist, as_stream(stream));
    TRY(finish(as_stream(stream), "chamfer"));
    if (multiplier) *multiplier = kChamferMultiplier;
    return DOLHIP_OK;
}

template <class T>
static int chamfer_host(int w, int h, T* dist, int* multiplier)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dist, "distances"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * sizeof(T);
    DevBuf a;
    TRY(a.upload(dist, bytes));
    launch_chamfer<T>(w, h, a.as<T>(), 0);
    TRY(finish_sync("chamfer"));
    TRY(a.download(dist, bytes));
    if (multiplier) *multiplier = kChamferMultiplier;
    return DOLHIP_OK;
}
}  // extern "C++"

int dolhip_chamfer_i16(int w, int h, int16_t* dist, int* multiplier)
{
    return chamfer_host(w, h, dist, multiplier);
}
int dolhip_chamfer_i32(int w, int h, int32_t* dist, int* multiplier)
{
    return chamfer_host(w, h, dist, multiplier);
}
int dolhip_chamfer_i16_dev(int w, int h, int16_t* dist, int* multiplier, void* stream)
{
    return chamfer_dev(w, h, dist, multiplier, stream);
}
int dolhip_chamfer_i32_dev(int w, int h, int32_t* dist, int* multiplier, void* stream)
{
    return chamfer_dev(w, h, dist, multiplier, stream);
}

int dolhip_init_occ_dev(int w, int h, const uint32_t* mask, float* occ, float scale, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(occ, "occlusion"));
    TRY(require_device());
    launch_init_occ(w, h, mask, occ, scale, as_stream(stream));
    return finish(as_stream(stream), "init_occ");
}

int dolhip_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(occ, "occlusion"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * sizeof(float);
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(o.alloc(bytes));
    launch_init_occ(w, h, m.as<uint32_t>(), o.as<float>(), scale, 0);
    TRY(finish_sync("init_occ"));
    return o.download(occ, bytes);
}

int dolhip_shadow_sweep_dev(float gx, float gy, int w, int h, const float* occ, float* shadow, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(occ, "occlusion"));
    TRY(check_ptr(shadow, "shadow"));
    TRY(require_device());
    launch_shadow_sweep(gx, gy, w, h, occ, shadow, as_stream(stream));
    return finish(as_stream(stream), "shadow_sweep");
}

int dolhip_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(occ, "occlusion"));
    TRY(check_ptr(shadow, "shadow"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * sizeof(float);
    DevBuf i, o;
    TRY(i.upload(occ, bytes));
    TRY(o.alloc(bytes));
    launch_shadow_sweep(gx, gy, w, h, i.as<float>(), o.as<float>(), 0);
    TRY(finish_sync("shadow_sweep"));
    return o.download(shadow, bytes);
}

// ---- the tool's image writers

static int check_image_type(int type)
{
    if (type < 0 || type > 3)
        return fail(DOLHIP_ERR_INVALID_ARG, "image type must be 0 (sdf), 1 (sdf-delta), 2 (sdf-rg) or 3 (sdf-angles)");
    return DOLHIP_OK;
}

static int check_rgba(const void* p, int type)
{
    if (type != 0 && ((uintptr_t)p & 3)) return fail(DOLHIP_ERR_INVALID_ARG, "an RGBA image must be 4-byte aligned");
    return DOLHIP_OK;
}

// d = 0: a 2D field
static int border_image(int w, int h, int d, const uint32_t* mask, const int16_t* delta, float max_len, int type,
                        uint8_t* out, void* stream, bool dev)
{
    const int dd = d > 0 ? d : 1, dims = d > 0 ? 3 : 2;
    TRY(check_dims(w, h, dd));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    TRY(check_image_type(type));
    TRY(require_device());
    const size_t cells = (size_t)w * h * dd, obytes = cells * (type == 0 ? 1 : 4);
    if (dev) {
        TRY(check_rgba(out, type));
        launch_border_image(w, h, dd, dims, mask, delta, max_len, type, out, as_stream(stream));
        return finish(as_stream(stream), "encode_border_image");
    }
    DevBuf m, a, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * dd)));
    TRY(a.upload(delta, cells * dims * 2));
    TRY(o.alloc(obytes));
    launch_border_image(w, h, dd, dims, m.as<uint32_t>(), a.as<int16_t>(), max_len, type, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_border_image"));
    return o.download(out, obytes);
}

int dolhip_encode_border_image_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                                  int type, uint8_t* out)
{
    return border_image(w, h, 0, mask, delta, max_len, type, out, nullptr, false);
}
int dolhip_encode_border_image_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                                      int type, uint8_t* out, void* stream)
{
    return border_image(w, h, 0, mask, delta, max_len, type, out, stream, true);
}
int dolhip_encode_border_image_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                                  float max_len, int type, uint8_t* out)
{
    if (d < 1) return fail(DOLHIP_ERR_INVALID_ARG, "d must be positive");
    return border_image(w, h, d, mask, delta, max_len, type, out, nullptr, false);
}
int dolhip_encode_border_image_3d_dev(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                                      float max_len, int type, uint8_t* out, void* stream)
{
    if (d < 1) return fail(DOLHIP_ERR_INVALID_ARG, "d must be positive");
    return border_image(w, h, d, mask, delta, max_len, type, out, stream, true);
}

static int distance_u8(int w, int h, int d, const int16_t* delta, float max_len, uint8_t* out, void* stream,
                       bool dev)
{
    const int dd = d > 0 ? d : 1, dims = d > 0 ? 3 : 2;
    TRY(check_dims(w, h, dd));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    const size_t cells = (size_t)w * h * dd;
    if (dev) {
        launch_distance_u8(cells, dims, delta, max_len, out, as_stream(stream));
        return finish(as_stream(stream), "encode_distance_u8");
    }
    DevBuf a, o;
    TRY(a.upload(delta, cells * dims * 2));
    TRY(o.alloc(cells));
    launch_distance_u8(cells, dims, a.as<int16_t>(), max_len, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_distance_u8"));
    return o.download(out, cells);
}

int dolhip_encode_distance_u8_2d(int w, int h, const int16_t* delta, float max_len, uint8_t* out)
{
    return distance_u8(w, h, 0, delta, max_len, out, nullptr, false);
}
int dolhip_encode_distance_u8_2d_dev(int w, int h, const int16_t* delta, float max_len, uint8_t* out,
                                     void* stream)
{
    return distance_u8(w, h, 0, delta, max_len, out, stream, true);
}
int dolhip_encode_distance_u8_3d(int w, int h, int d, const int16_t* delta, float max_len, uint8_t* out)
{
    if (d < 1) return fail(DOLHIP_ERR_INVALID_ARG, "d must be positive");
    return distance_u8(w, h, d, delta, max_len, out, nullptr, false);
}
int dolhip_encode_distance_u8_3d_dev(int w, int h, int d, const int16_t* delta, float max_len, uint8_t* out,
                                     void* stream)
{
    if (d < 1) return fail(DOLHIP_ERR_INVALID_ARG, "d must be positive");
    return distance_u8(w, h, d, delta, max_len, out, stream, true);
}

extern "C++" {
template <class T>
static int scalar_u8(int w, int h, const T* v, float max_len, uint8_t* out, void* stream, bool dev)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(v, "distances"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    const size_t n = (size_t)w * h;
    if (dev) {
        launch_scalar_u8<T>(n, v, max_len, out, as_stream(stream));
        return finish(as_stream(stream), "encode_scalar_u8");
    }
    DevBuf a, o;
    TRY(a.upload(v, n * sizeof(T)));
    TRY(o.alloc(n));
    launch_scalar_u8<T>(n, a.as<T>(), max_len, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_scalar_u8"));
    return o.download(out, n);
}
}  // extern "C++"

int dolhip_encode_scalar_u8_f32(int w, int h, const float* v, float max_len, uint8_t* out)
{
    return scalar_u8(w, h, v, max_len, out, nullptr, false);
}
int dolhip_encode_scalar_u8_f32_dev(int w, int h, const float* v, float max_len, uint8_t* out, void* stream)
{
    return scalar_u8(w, h, v, max_len, out, stream, true);
}
int dolhip_encode_scalar_u8_i32(int w, int h, const int32_t* v, float max_len, uint8_t* out)
{
    return scalar_u8(w, h, v, max_len, out, nullptr, false);
}
int dolhip_encode_scalar_u8_i32_dev(int w, int h, const int32_t* v, float max_len, uint8_t* out, void* stream)
{
    return scalar_u8(w, h, v, max_len, out, stream, true);
}
int dolhip_encode_scalar_u8_i16(int w, int h, const int16_t* v, float max_len, uint8_t* out)
{
    return scalar_u8(w, h, v, max_len, out, nullptr, false);
}
int dolhip_encode_scalar_u8_i16_dev(int w, int h, const int16_t* v, float max_len, uint8_t* out, void* stream)
{
    return scalar_u8(w, h, v, max_len, out, stream, true);
}

int dolhip_sqrt_f32_dev(int w, int h, const float* v, float* out, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(v, "in"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    launch_sqrt_f32((size_t)w * h, v, out, as_stream(stream));
    return finish(as_stream(stream), "sqrt_f32");
}

int dolhip_sqrt_f32(int w, int h, const float* v, float* out)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(v, "in"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    const size_t n = (size_t)w * h;
    DevBuf a, o;
    TRY(a.upload(v, n * 4));
    TRY(o.alloc(n * 4));
    launch_sqrt_f32(n, a.as<float>(), o.as<float>(), 0);
    TRY(finish_sync("sqrt_f32"));
    return o.download(out, n * 4);
}

int dolhip_init_dirw_2d_dev(int w, int h, const uint32_t* mask, float* dirw4, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dirw4, "dirw4"));
    TRY(require_device());
    launch_init_dirw(w, h, 1, 4, mask, dirw4, as_stream(stream));
    return finish(as_stream(stream), "init_dirw_2d");
}

int dolhip_init_dirw_2d(int w, int h, const uint32_t* mask, float* dirw4)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dirw4, "dirw4"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4 * sizeof(float);
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(o.alloc(bytes));
    launch_init_dirw(w, h, 1, 4, m.as<uint32_t>(), o.as<float>(), 0);
    TRY(finish_sync("init_dirw_2d"));
    return o.download(dirw4, bytes);
}

int dolhip_init_dirw_3d_dev(int w, int h, int d, const uint32_t* mask, float* dirw8, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dirw8, "dirw8"));
    TRY(require_device());
    launch_init_dirw(w, h, d, 8, mask, dirw8, as_stream(stream));
    return finish(as_stream(stream), "init_dirw_3d");
}

int dolhip_init_dirw_3d(int w, int h, int d, const uint32_t* mask, float* dirw8)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(dirw8, "dirw8"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 8 * sizeof(float);
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(o.alloc(bytes));
    launch_init_dirw(w, h, d, 8, m.as<uint32_t>(), o.as<float>(), 0);
    TRY(finish_sync("init_dirw_3d"));
    return o.download(dirw8, bytes);
}

int dolhip_occlusion_sweep_2d_dev(int w, int h, float* dirw4, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dirw4, "dirw4"));
    TRY(require_device());
    launch_occlusion_sweep(0, w, h, 1, nullptr, dirw4, as_stream(stream));
    return finish(as_stream(stream), "occlusion_sweep_2d");
}

int dolhip_occlusion_sweep_2d(int w, int h, float* dirw4)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dirw4, "dirw4"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4 * sizeof(float);
    DevBuf a;
    TRY(a.upload(dirw4, bytes));
    launch_occlusion_sweep(0, w, h, 1, nullptr, a.as<float>(), 0);
    TRY(finish_sync("occlusion_sweep_2d"));
    return a.download(dirw4, bytes);
}

int dolhip_occlusion_sweep_3d_dev(int w, int h, int d, float* dirw8, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(dirw8, "dirw8"));
    TRY(require_device());
    launch_occlusion_sweep(1, w, h, d, nullptr, dirw8, as_stream(stream));
    return finish(as_stream(stream), "occlusion_sweep_3d");
}

int dolhip_occlusion_sweep_3d(int w, int h, int d, float* dirw8)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(dirw8, "dirw8"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 8 * sizeof(float);
    DevBuf a;
    TRY(a.upload(dirw8, bytes));
    launch_occlusion_sweep(1, w, h, d, nullptr, a.as<float>(), 0);
    TRY(finish_sync("occlusion_sweep_3d"));
    return a.download(dirw8, bytes);
}

int dolhip_occlusion_sweep_3d_from_dev(int w, int h, int d, const float* in8, float* out8, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(in8, "in8"));
    TRY(check_ptr(out8, "out8"));
    if (in8 == out8) return fail(DOLHIP_ERR_INVALID_ARG, "in8 and out8 must be different fields");
    TRY(require_device());
    launch_occlusion_sweep(2, w, h, d, in8, out8, as_stream(stream));
    return finish(as_stream(stream), "occlusion_sweep_3d_from");
}

int dolhip_occlusion_sweep_3d_from(int w, int h, int d, const float* in8, float* out8)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(in8, "in8"));
    TRY(check_ptr(out8, "out8"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 8 * sizeof(float);
    DevBuf a, b;
    TRY(a.upload(in8, bytes));
    TRY(b.upload(out8, bytes));
    launch_occlusion_sweep(2, w, h, d, a.as<float>(), b.as<float>(), 0);
    TRY(finish_sync("occlusion_sweep_3d_from"));
    return b.download(out8, bytes);
}

// AO images: n cells of `planes` planes into 1 (ao) or 4 (rgba) bytes a cell
static int encode_ao_dev(size_t n, int planes, bool rgba, const float* dirw, uint8_t* out, void* stream,
                         const char* what)
{
    TRY(check_ptr(dirw, "dirw"));
    TRY(check_ptr(out, "out"));
    if (rgba && ((uintptr_t)out & 3)) return fail(DOLHIP_ERR_INVALID_ARG, "rgba must be 4-byte aligned");
    TRY(require_device());
    launch_encode_ao(n, planes, rgba, dirw, out, as_stream(stream));
    return finish(as_stream(stream), what);
}

static int encode_ao_host(size_t n, int planes, bool rgba, const float* dirw, uint8_t* out, const char* what)
{
    TRY(check_ptr(dirw, "dirw"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    const size_t ob = n * (rgba ? 4 : 1);
    DevBuf a, o;
    TRY(a.upload(dirw, n * planes * sizeof(float)));
    TRY(o.alloc(ob));
    launch_encode_ao(n, planes, rgba, a.as<float>(), o.as<uint8_t>(), 0);
    TRY(finish_sync(what));
    return o.download(out, ob);
}

int dolhip_encode_ao_2d(int w, int h, const float* dirw4, uint8_t* ao)
{
    TRY(check_dims(w, h));
    return encode_ao_host((size_t)w * h, 4, false, dirw4, ao, "encode_ao_2d");
}
int dolhip_encode_ao_2d_dev(int w, int h, const float* dirw4, uint8_t* ao, void* stream)
{
    TRY(check_dims(w, h));
    return encode_ao_dev((size_t)w * h, 4, false, dirw4, ao, stream, "encode_ao_2d");
}
int dolhip_encode_ao_3d(int w, int h, int d, const float* dirw8, uint8_t* ao)
{
    TRY(check_dims(w, h, d));
    return encode_ao_host((size_t)w * h * d, 8, false, dirw8, ao, "encode_ao_3d");
}
int dolhip_encode_ao_3d_dev(int w, int h, int d, const float* dirw8, uint8_t* ao, void* stream)
{
    TRY(check_dims(w, h, d));
    return encode_ao_dev((size_t)w * h * d, 8, false, dirw8, ao, stream, "encode_ao_3d");
}
int dolhip_encode_dirw_2d(int w, int h, const float* dirw4, uint8_t* rgba)
{
    TRY(check_dims(w, h));
    return encode_ao_host((size_t)w * h, 4, true, dirw4, rgba, "encode_dirw_2d");
}
int dolhip_encode_dirw_2d_dev(int w, int h, const float* dirw4, uint8_t* rgba, void* stream)
{
    TRY(check_dims(w, h));
    return encode_ao_dev((size_t)w * h, 4, true, dirw4, rgba, stream, "encode_dirw_2d");
}
int dolhip_encode_dirw_3d(int w, int h, int d, const float* dirw8, uint8_t* rgba)
{
    TRY(check_dims(w, h, d));
    return encode_ao_host((size_t)w * h * d, 8, true, dirw8, rgba, "encode_dirw_3d");
}
int dolhip_encode_dirw_3d_dev(int w, int h, int d, const float* dirw8, uint8_t* rgba, void* stream)
{
    TRY(check_dims(w, h, d));
    return encode_ao_dev((size_t)w * h * d, 8, true, dirw8, rgba, stream, "encode_dirw_3d");
}

size_t dolhip_felzenszwalb_2d_workspace(int w, int h)
{
    return (w < 1 || h < 1) ? 0 : felz_workspace(w, h);
}

int dolhip_felzenszwalb_2d_dev(int w, int h, float* dist, void* work, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dist, "dist"));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, felz_workspace(w, h), stream, &tmp));
    launch_felzenszwalb_2d(w, h, dist, tmp, as_stream(stream));
    return finish(as_stream(stream), "felzenszwalb_2d");
}

int dolhip_felzenszwalb_2d(int w, int h, float* dist)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(dist, "dist"));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf a, t;
    TRY(a.upload(dist, bytes));
    TRY(t.alloc(felz_workspace(w, h)));
    launch_felzenszwalb_2d(w, h, a.as<float>(), t.p, 0);
    TRY(finish_sync("felzenszwalb_2d"));
    return a.download(dist, bytes);
}

// ---- exact EDT

size_t dolhip_edt_2d_workspace(int w, int h)
{
    return (w < 1 || h < 1) ? 0 : edt_workspace(w, h, 1);
}

size_t dolhip_edt_3d_workspace(int w, int h, int d)
{
    return (w < 1 || h < 1 || d < 1) ? 0 : edt_workspace(w, h, d);
}

int dolhip_edt_2d_dev(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* dist,
                      void* work, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    if (!d2 && !dist) return fail(DOLHIP_ERR_INVALID_ARG, "d2 and dist are both null");
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_edt_2d_workspace(w, h), stream, &tmp));
    launch_edt_2d(w, h, mask, invert, d2, dist, tmp, as_stream(stream));
    return finish(as_stream(stream), "edt_2d");
}

int dolhip_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* dist)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    if (!d2 && !dist) return fail(DOLHIP_ERR_INVALID_ARG, "d2 and dist are both null");
    TRY(require_device());
    const size_t cells = (size_t)w * h;
    DevBuf m, o2, of, t;
    TRY(m.upload(mask, mask_bytes(w, h)));
    if (d2) TRY(o2.alloc(cells * 4));
    if (dist) TRY(of.alloc(cells * 4));
    TRY(t.alloc(dolhip_edt_2d_workspace(w, h)));
    launch_edt_2d(w, h, m.as<uint32_t>(), invert, o2.as<int32_t>(), of.as<float>(), t.p, 0);
    TRY(finish_sync("edt_2d"));
    if (d2) TRY(o2.download(d2, cells * 4));
    if (dist) TRY(of.download(dist, cells * 4));
    return DOLHIP_OK;
}

int dolhip_edt_3d_dev(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                      float* dist, void* work, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    if (!d2 && !dist) return fail(DOLHIP_ERR_INVALID_ARG, "d2 and dist are both null");
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_edt_3d_workspace(w, h, d), stream, &tmp));
    launch_edt_3d(w, h, d, mask, invert, d2, dist, tmp, as_stream(stream));
    return finish(as_stream(stream), "edt_3d");
}

int dolhip_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                  float* dist)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    if (!d2 && !dist) return fail(DOLHIP_ERR_INVALID_ARG, "d2 and dist are both null");
    TRY(require_device());
    const size_t cells = (size_t)w * h * d;
    DevBuf m, o2, of, t;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    if (d2) TRY(o2.alloc(cells * 4));
    if (dist) TRY(of.alloc(cells * 4));
    TRY(t.alloc(dolhip_edt_3d_workspace(w, h, d)));
    launch_edt_3d(w, h, d, m.as<uint32_t>(), invert, o2.as<int32_t>(), of.as<float>(), t.p, 0);
    TRY(finish_sync("edt_3d"));
    if (d2) TRY(o2.download(d2, cells * 4));
    if (dist) TRY(of.download(dist, cells * 4));
    return DOLHIP_OK;
}

// two fields (to set bits, to clear bits) in the work area after the EDT's own
static int signed_edt_dev(int w, int h, int d, const uint32_t* mask, float* sdf, void* work,
                          hipStream_t s)
{
    const long long cells = (long long)w * h * d;
    const size_t field = ((size_t)cells * 4 + 255) & ~(size_t)255;
    const size_t inner = d == 1 ? dolhip_edt_2d_workspace(w, h) : dolhip_edt_3d_workspace(w, h, d);
    void* tmp = nullptr;
    TRY(resolve_work(work, inner + 2 * field, (void*)s, &tmp));
    int32_t* fout = (int32_t*)((char*)tmp + inner);
    int32_t* fin = (int32_t*)((char*)tmp + inner + field);
    if (d == 1) {
        launch_edt_2d(w, h, mask, 0, fout, nullptr, tmp, s);
        launch_edt_2d(w, h, mask, 1, fin, nullptr, tmp, s);
    } else {
        launch_edt_3d(w, h, d, mask, 0, fout, nullptr, tmp, s);
        launch_edt_3d(w, h, d, mask, 1, fin, nullptr, tmp, s);
    }
    launch_signed_edt((size_t)cells, fout, fin, sdf, s);
    return check_launch("signed_edt");
}

int dolhip_signed_edt_2d_dev(int w, int h, const uint32_t* mask, float* sdf, void* work,
                             void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(sdf, "sdf"));
    TRY(require_device());
    return signed_edt_dev(w, h, 1, mask, sdf, work, as_stream(stream));
}

int dolhip_signed_edt_3d_dev(int w, int h, int d, const uint32_t* mask, float* sdf, void* work,
                             void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(sdf, "sdf"));
    TRY(require_device());
    return signed_edt_dev(w, h, d, mask, sdf, work, as_stream(stream));
}

static int signed_edt_host(int w, int h, int d, const uint32_t* mask, float* sdf)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(sdf, "sdf"));
    TRY(require_device());
    const size_t cells = (size_t)w * h * d;
    DevBuf m, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(o.alloc(cells * 4));
    TRY(signed_edt_dev(w, h, d, m.as<uint32_t>(), o.as<float>(), nullptr, 0));
    TRY(finish_sync("signed_edt"));
    return o.download(sdf, cells * 4);
}

int dolhip_signed_edt_2d(int w, int h, const uint32_t* mask, float* sdf)
{ return signed_edt_host(w, h, 1, mask, sdf); }
int dolhip_signed_edt_3d(int w, int h, int d, const uint32_t* mask, float* sdf)
{ return signed_edt_host(w, h, d, mask, sdf); }

// ---- distances and encoders

int dolhip_dist_2d_dev(int n, const int16_t* delta, float* out, void* stream)
{
    if (n < 0) return fail(DOLHIP_ERR_INVALID_ARG, "n must be >= 0");
    TRY(check_cells2(delta, "delta"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    if (n) launch_dist_2d((size_t)n, delta, out, as_stream(stream));
    return finish(as_stream(stream), "dist_2d");
}

static int signed_dev(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                      float max_len, void* out, int mode, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(d == 1 ? check_cells2(delta, "delta") : check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    if (mode == 1 && !(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    TRY(require_device());
    if (d == 1 && mode >= 0)
        launch_signed_2d(w, h, mask, delta, max_len, out, mode, as_stream(stream));
    else
        launch_signed_3d(w, h, d, mask, delta, max_len, out, mode, as_stream(stream));
    return finish(as_stream(stream), "signed");
}

static int signed_host(int w, int h, int d, bool three, const uint32_t* mask,
                       const int16_t* delta, float max_len, void* out, int mode)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    if (mode == 1 && !(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    TRY(require_device());
    const size_t cells = (size_t)w * h * d;
    const size_t obytes = cells * (mode == 0 ? 4 : 1);
    DevBuf m, dl, o;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(dl.upload(delta, cells * (three ? 6 : 4)));
    TRY(o.alloc(obytes));
    if (!three)
        launch_signed_2d(w, h, m.as<uint32_t>(), dl.as<int16_t>(), max_len, o.p, mode, 0);
    else
        launch_signed_3d(w, h, d, m.as<uint32_t>(), dl.as<int16_t>(), max_len, o.p, mode, 0);
    TRY(finish_sync("signed"));
    return o.download(out, obytes);
}

int dolhip_signed_distance_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta,
                                  float* out, void* stream)
{ return signed_dev(w, h, 1, mask, delta, 1.0f, out, 0, stream); }

int dolhip_signed_distance_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                              float* out)
{ return signed_host(w, h, 1, false, mask, delta, 1.0f, out, 0); }

int dolhip_signed_distance_3d_dev(int w, int h, int d, const uint32_t* mask,
                                  const int16_t* delta, float* out, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    TRY(require_device());
    launch_signed_3d(w, h, d, mask, delta, 1.0f, out, 0, as_stream(stream));
    return finish(as_stream(stream), "signed_dist_3d");
}

int dolhip_signed_distance_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                              float* out)
{ return signed_host(w, h, d, true, mask, delta, 1.0f, out, 0); }

int dolhip_encode_sdf_u8_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta,
                                float max_len, uint8_t* out, void* stream)
{ return signed_dev(w, h, 1, mask, delta, max_len, out, 1, stream); }

int dolhip_encode_sdf_u8_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                            float max_len, uint8_t* out)
{ return signed_host(w, h, 1, false, mask, delta, max_len, out, 1); }

int dolhip_encode_sdf_u8_3d_dev(int w, int h, int d, const uint32_t* mask,
                                const int16_t* delta, float max_len, uint8_t* out, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    if (!(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    TRY(require_device());
    launch_signed_3d(w, h, d, mask, delta, max_len, out, 1, as_stream(stream));
    return finish(as_stream(stream), "encode_sdf_u8_3d");
}

int dolhip_encode_sdf_u8_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                            float max_len, uint8_t* out)
{ return signed_host(w, h, d, true, mask, delta, max_len, out, 1); }

int dolhip_encode_sdf_delta_rgba_2d(int w, int h, const int16_t* delta, float max_len,
                                    uint8_t* out)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    if (!(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    TRY(require_device());
    const size_t cells = (size_t)w * h;
    DevBuf dl, o;
    TRY(dl.upload(delta, cells * 4));
    TRY(o.alloc(cells * 4));
    launch_sdf_delta_rgba(w, h, dl.as<int16_t>(), max_len, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_sdf_delta"));
    return o.download(out, cells * 4);
}

int dolhip_encode_sdf_rg_rgba_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                                 float max_len, uint8_t* out)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(check_ptr(out, "out"));
    if (!(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    TRY(require_device());
    const size_t cells = (size_t)w * h;
    DevBuf m, dl, o;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(dl.upload(delta, cells * 4));
    TRY(o.alloc(cells * 4));
    launch_sdf_rg_rgba(w, h, m.as<uint32_t>(), dl.as<int16_t>(), max_len, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_sdf_rg"));
    return o.download(out, cells * 4);
}

// ---- signed pipelines

int dolhip_signed_border_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta, void* work,
                                void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_cells2(delta, "delta"));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_jump_flood_2d_workspace(w, h), stream, &tmp));
    launch_border_2d(w, h, mask, delta, as_stream(stream));
    jfa2d_schedule(w, h, delta, 2, (int16_t*)tmp, as_stream(stream));
    return finish(as_stream(stream), "signed_border_2d");
}

int dolhip_signed_border_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                                void* work, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta, "delta"));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_jump_flood_3d_workspace(w, h, d), stream, &tmp));
    launch_border_3d(w, h, d, mask, delta, as_stream(stream));
    jfa3d_run(w, h, d, delta, 2, tmp, as_stream(stream), true);
    return finish(as_stream(stream), "signed_border_3d");
}


// ---- two-field signed images (the tool's -S mode,
// DistanceOcclusionTool.cpp:1082-1124, 1197-1232)

static int check_image_type(int type, float max_len)
{
    if (type < 0 || type > 3)
        return fail(DOLHIP_ERR_INVALID_ARG, "image type must be 0 (sdf), 1 (sdf-delta), 2 (sdf-rg) or 3 (sdf-angles), got %d", type);
    if (!(max_len > 0.0f)) return fail(DOLHIP_ERR_INVALID_ARG, "max_len must be > 0");
    return DOLHIP_OK;
}

size_t dolhip_two_field_2d_workspace(int w, int h) { return dolhip_jump_flood_2d_workspace(w, h); }

size_t dolhip_two_field_3d_workspace(int w, int h, int d) { return dolhip_jump_flood_3d_workspace(w, h, d); }

int dolhip_two_field_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta0,
                                   int16_t* delta1, int cw, void* work, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_cells2(delta0, "delta0"));
    TRY(check_cells2(delta1, "delta1"));
    TRY(check_cw(cw));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_two_field_2d_workspace(w, h), stream, &tmp));
    const hipStream_t s = as_stream(stream);
    launch_init_deltas_2d(w, h, mask, delta0, 0, s);
    launch_init_deltas_2d(w, h, mask, delta1, 1, s);
    jfa2d_schedule(w, h, delta0, cw, (int16_t*)tmp, s);
    jfa2d_schedule(w, h, delta1, cw, (int16_t*)tmp, s);
    return finish(s, "two_field_deltas_2d");
}

int dolhip_two_field_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta0,
                               int16_t* delta1, int cw)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * 4;
    DevBuf m, a, b, t;
    TRY(m.upload(mask, mask_bytes(w, h)));
    TRY(a.alloc(bytes));
    TRY(b.alloc(bytes));
    TRY(t.alloc(dolhip_two_field_2d_workspace(w, h)));
    TRY(dolhip_two_field_deltas_2d_dev(w, h, m.as<uint32_t>(), a.as<int16_t>(), b.as<int16_t>(), cw,
                                       t.p, nullptr));
    TRY(finish_sync("two_field_deltas_2d"));
    TRY(a.download(delta0, bytes));
    return b.download(delta1, bytes);
}

int dolhip_two_field_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta0,
                                   int16_t* delta1, int cw, void* work, void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_cw(cw));
    TRY(require_device());
    void* tmp = nullptr;
    TRY(resolve_work(work, dolhip_two_field_3d_workspace(w, h, d), stream, &tmp));
    const hipStream_t s = as_stream(stream);
    launch_init_deltas_3d(w, h, d, mask, delta0, 0, s);
    launch_init_deltas_3d(w, h, d, mask, delta1, 1, s);
    jfa3d_run(w, h, d, delta0, cw, tmp, s, true);
    jfa3d_run(w, h, d, delta1, cw, tmp, s, true);
    return finish(s, "two_field_deltas_3d");
}

int dolhip_two_field_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta0,
                               int16_t* delta1, int cw)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(mask, "mask"));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_cw(cw));
    TRY(require_device());
    const size_t bytes = (size_t)w * h * d * 6;
    DevBuf m, a, b, t;
    TRY(m.upload(mask, mask_bytes(w, (long long)h * d)));
    TRY(a.alloc(bytes));
    TRY(b.alloc(bytes));
    TRY(t.alloc(dolhip_two_field_3d_workspace(w, h, d)));
    TRY(dolhip_two_field_deltas_3d_dev(w, h, d, m.as<uint32_t>(), a.as<int16_t>(), b.as<int16_t>(),
                                       cw, t.p, nullptr));
    TRY(finish_sync("two_field_deltas_3d"));
    TRY(a.download(delta0, bytes));
    return b.download(delta1, bytes);
}

int dolhip_encode_two_field_2d_dev(int w, int h, const int16_t* delta0, const int16_t* delta1,
                                   float max_len, int type, uint8_t* out, void* stream)
{
    TRY(check_dims(w, h));
    TRY(check_cells2(delta0, "delta0"));
    TRY(check_cells2(delta1, "delta1"));
    TRY(check_ptr(out, "out"));
    TRY(check_image_type(type, max_len));
    TRY(require_device());
    launch_two_field_2d(w, h, delta0, delta1, max_len, type, out, as_stream(stream));
    return finish(as_stream(stream), "encode_two_field_2d");
}

int dolhip_encode_two_field_2d(int w, int h, const int16_t* delta0, const int16_t* delta1,
                               float max_len, int type, uint8_t* out)
{
    TRY(check_dims(w, h));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_ptr(out, "out"));
    TRY(check_image_type(type, max_len));
    TRY(require_device());
    const size_t cells = (size_t)w * h, out_bytes = cells * (type == 0 ? 1 : 4);
    DevBuf a, b, o;
    TRY(a.upload(delta0, cells * 4));
    TRY(b.upload(delta1, cells * 4));
    TRY(o.alloc(out_bytes));
    launch_two_field_2d(w, h, a.as<int16_t>(), b.as<int16_t>(), max_len, type, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_two_field_2d"));
    return o.download(out, out_bytes);
}

int dolhip_encode_two_field_3d_dev(int w, int h, int d, const int16_t* delta0,
                                   const int16_t* delta1, float max_len, int type, uint8_t* out,
                                   void* stream)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_ptr(out, "out"));
    TRY(check_image_type(type, max_len));
    TRY(require_device());
    launch_two_field_3d(w, h, d, delta0, delta1, max_len, type, out, as_stream(stream));
    return finish(as_stream(stream), "encode_two_field_3d");
}

int dolhip_encode_two_field_3d(int w, int h, int d, const int16_t* delta0, const int16_t* delta1,
                               float max_len, int type, uint8_t* out)
{
    TRY(check_dims(w, h, d));
    TRY(check_ptr(delta0, "delta0"));
    TRY(check_ptr(delta1, "delta1"));
    TRY(check_ptr(out, "out"));
    TRY(check_image_type(type, max_len));
    TRY(require_device());
    const size_t cells = (size_t)w * h * d, out_bytes = cells * (type == 0 ? 1 : 4);
    DevBuf a, b, o;
    TRY(a.upload(delta0, cells * 6));
    TRY(b.upload(delta1, cells * 6));
    TRY(o.alloc(out_bytes));
    launch_two_field_3d(w, h, d, a.as<int16_t>(), b.as<int16_t>(), max_len, type, o.as<uint8_t>(), 0);
    TRY(finish_sync("encode_two_field_3d"));
    return o.download(out, out_bytes);
}


// ---- surface voxeliser (MeshSupport.cpp:369-475)

static int check_voxel_args(int tri_count, const float* vertices, const float* bbox, int w, int h,
                            int d, const uint32_t* mask)
{
    TRY(check_dims(w, h, d));
    if (tri_count < 0) return fail(DOLHIP_ERR_INVALID_ARG, "tri_count must be >= 0, got %d", tri_count);
    TRY(check_ptr(bbox, "bbox"));
    TRY(check_ptr(mask, "mask"));
    if (tri_count > 0) TRY(check_ptr(vertices, "vertices"));
    for (int k = 0; k < 3; k++)
        if (!(bbox[3 + k] > bbox[k]))
            return fail(DOLHIP_ERR_INVALID_ARG, "bbox max must exceed min on every axis");
    return DOLHIP_OK;
}

int dolhip_mask_from_triangles_dev(int tri_count, const int* indices, const float* vertices,
                                   const float* bbox, int w, int h, int d, uint32_t* mask,
                                   void* stream)
{
    TRY(check_voxel_args(tri_count, vertices, bbox, w, h, d, mask));
    TRY(require_device());
    launch_voxelise(tri_count, indices, vertices, bbox, w, h, d, mask, as_stream(stream));
    return finish(as_stream(stream), "mask_from_triangles");
}

int dolhip_mask_from_triangles(int tri_count, const int* indices, const float* vertices,
                               const float* bbox, int w, int h, int d, uint32_t* mask)
{
    TRY(check_voxel_args(tri_count, vertices, bbox, w, h, d, mask));
    TRY(require_device());
    // vertices referenced: up to the largest index (or 3 per triangle)
    long long nv = 3LL * tri_count;
    if (indices) {
        nv = 0;
        for (long long i = 0; i < 3LL * tri_count; i++) {
            if (indices[i] < 0) return fail(DOLHIP_ERR_INVALID_ARG, "negative vertex index %d", indices[i]);
            nv = indices[i] + 1 > nv ? indices[i] + 1 : nv;
        }
    }
    const size_t mbytes = mask_bytes(w, (long long)h * d);
    DevBuf m, v, ix;
    TRY(m.upload(mask, mbytes));
    if (nv > 0) TRY(v.upload(vertices, (size_t)nv * 12));
    if (indices && tri_count > 0) TRY(ix.upload(indices, (size_t)tri_count * 12));
    launch_voxelise(tri_count, indices ? ix.as<int>() : nullptr, v.as<float>(), bbox, w, h, d,
                    m.as<uint32_t>(), 0);
    TRY(finish_sync("mask_from_triangles"));
    return m.download(mask, mbytes);
}

}  // extern "C"
