/*
 * dol_hip.h -- C ABI of the MI355X (gfx950) distance-field library.
 *
 * It replaces the data-parallel distance-field path of DistanceOcclusionLib
 * (DistanceOcclusionLib.h:61-98): initialisation from 1-bit masks, signed
 * border initialisation, JumpFlood in 2D and 3D, Felzenszwalb, plus an exact
 * integer squared EDT in 2D and 3D, the output encoders of the reference
 * tool, batched ragged glyph SDFs and a continuous-seed distance with its
 * gradient; and, so that the whole reference header can be served, its
 * sweep and raster-order methods (FastSweep, Danielsson, Chamfer,
 * BruteForce, the direction and occlusion sweeps, the occlusion image and
 * shadow sweep, DistanceOcclusionLib.h:65-117).
 *
 * Layouts are the reference's (DistanceOcclusionLib.h:16-57):
 *   mask    uint32 words, rows padded to ceil(w/32) words, LSB first,
 *           row r of slice z at word (z*h + r) * ceil(w/32)
 *   delta2  int16 {x, y} per cell (4 bytes), index y*w + x
 *   delta3  int16 {x, y, z} per cell (6 bytes, packed), index (z*h + y)*w + x
 *   scalars row-major, same indexing as the deltas
 *
 * Two forms of every entry point:
 *   dolhip_<name>(...)      host pointers; synchronous, like the reference's
 *                           functions: copies in, runs, copies back, returns.
 *   dolhip_<name>_dev(...)  device pointers and a hipStream_t (NULL = the
 *                           default stream); asynchronous on that stream.
 *                           Where a pass needs a temporary, `work` may point
 *                           to device memory of at least the size the matching
 *                           *_workspace() call returns, or be NULL, in which
 *                           case the library uses a buffer of its own, one per
 *                           (device, stream): calls on one stream share it in
 *                           stream order, calls on different streams use
 *                           different ones.  It grows as needed (hipMalloc,
 *                           which may synchronise); a grown-out buffer is kept
 *                           until the process ends, since work enqueued with it
 *                           may still run.  Pass your own `work` under graph
 *                           capture, or when a stream handle may be reused for
 *                           a new stream while the old one still has work.
 *
 * Every call returns a dolhip_status; dolhip_last_error() gives the message
 * of the last failure on the calling thread.  The reference's functions
 * return void and check nothing (DistanceOcclusionLib.cpp:14-18); here bad
 * sizes and null pointers are rejected with DOLHIP_ERR_INVALID_ARG and device
 * failures reported as DOLHIP_ERR_DEVICE.  There is no CPU fallback: without a
 * usable gfx950 device every compute call fails with DOLHIP_ERR_NO_DEVICE.
 */
#ifndef DOL_HIP_H
#define DOL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dolhip_status {
    DOLHIP_OK = 0,
    DOLHIP_ERR_INVALID_ARG = 1,
    DOLHIP_ERR_OUT_OF_MEMORY = 2,
    DOLHIP_ERR_DEVICE = 3,
    DOLHIP_ERR_NO_DEVICE = 4,
} dolhip_status;

#define DOLHIP_MAX_DELTA2 32767   /* kMaxDelta2, DistanceOcclusionLib.h:56 */
#define DOLHIP_MAX_DELTA3 25000   /* kMaxDelta3, DistanceOcclusionLib.h:57 */
#define DOLHIP_EDT_INF    0x7fffffff  /* squared distance of a cell no seed reaches */

/* ---- runtime ------------------------------------------------------------ */
const char* dolhip_last_error(void);
const char* dolhip_version(void);
int dolhip_device_count(void);            /* usable devices, 0 if none */
/* Per-kernel timing with HIP events recorded on the launch stream, for
 * measurement harnesses.  enable != 0 starts recording (clearing old records). */
int dolhip_profile_enable(int enable);
/* Average duration in ms and launch count of kernels whose name starts with
 * `prefix`, over the records collected so far (synchronises the device). */
int dolhip_profile_query(const char* prefix, double* avg_ms, int* count);

/* ---- masks ---------------------------------------------------------------- */
/* BitMaskFromImageFile's thresholding (DistanceOcclusionTool.cpp:646-665):
 * bit set iff (byte >= threshold) XOR less_than; `rows` = h (2D) or h*d (3D).
 * One wavefront ballot per 64 pixels. */
int dolhip_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                           uint32_t* mask);
int dolhip_mask_from_bytes_dev(int w, int rows, const uint8_t* bytes, int threshold,
                               int less_than, uint32_t* mask, void* stream);

/* Host-side generators.  The reference's BitMaskAddPoints / BitMaskAddBlock
 * (DistanceOcclusionLib.h:30-35, .cpp:95-179, 1224-1287; block sizes need
 * w >= 2, h >= 3, d >= 3 there), and seeded scenes for larger workloads:
 * unions of discs and boxes (2D), of rotated ellipsoids and boxes (3D), ragged
 * glyph coverage masks (255 inside) and continuous seed points.  All OR into /
 * fill host memory and touch no device. */
void dolhip_mask_add_points_2d(int w, int h, int count, uint32_t seed, uint32_t* mask);
void dolhip_mask_add_points_3d(int w, int h, int d, int count, uint32_t seed, uint32_t* mask);
void dolhip_mask_add_block_2d(int w, int h, int sides, uint32_t* mask);
void dolhip_mask_add_block_3d(int w, int h, int d, int sides, uint32_t* mask);
void dolhip_scene_shapes_2d(int w, int h, int n, uint64_t seed, uint32_t* mask);
void dolhip_scene_shapes_3d(int w, int h, int d, int n_ellipsoids, int n_boxes, uint64_t seed,
                            uint32_t* mask);
int64_t dolhip_scene_glyphs(int n, int min_size, int max_size, uint64_t seed, int32_t* w,
                            int32_t* h, int64_t* offset, uint8_t* bytes);
void dolhip_scene_seeds(int w, int h, int n, uint64_t seed, float* seeds, float* weights);

/* ---- initialisation ------------------------------------------------------- */
/* InitDeltasFromBitMask 2D (DistanceOcclusionLib.h:75, .cpp:224-254) */
int dolhip_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert);
int dolhip_init_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta, int invert,
                              void* stream);
/* InitDeltasFromBitMask 3D (DistanceOcclusionLib.h:76, .cpp:1289-1321) */
int dolhip_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert);
int dolhip_init_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                              int invert, void* stream);
/* InitBorderDeltasFromBitMask 2D (DistanceOcclusionLib.h:79, .cpp:504-619) */
int dolhip_init_border_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta);
int dolhip_init_border_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta,
                                     void* stream);
/* InitBorderDeltasFromBitMask 3D (DistanceOcclusionLib.h:80, .cpp:1820-2173) */
int dolhip_init_border_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta);
int dolhip_init_border_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                                     void* stream);
/* InitDistancesFromBitMask (DistanceOcclusionLib.h:61-63, .cpp:181-222) */
int dolhip_init_distances_f32(int w, int h, const uint32_t* mask, float* dist, float max_d);
int dolhip_init_distances_f32_dev(int w, int h, const uint32_t* mask, float* dist, float max_d,
                                  void* stream);
int dolhip_init_distances_i32(int w, int h, const uint32_t* mask, int32_t* dist, int32_t max_d);
int dolhip_init_distances_i16(int w, int h, const uint32_t* mask, int16_t* dist, int16_t max_d);

/* ---- propagation ---------------------------------------------------------- */
/* JumpFlood 2D, in place (DistanceOcclusionLib.h:92, .cpp:780-908) */
int dolhip_jump_flood_2d(int w, int h, int16_t* delta, int cw);
size_t dolhip_jump_flood_2d_workspace(int w, int h);
int dolhip_jump_flood_2d_dev(int w, int h, int16_t* delta, int cw, void* work, void* stream);
/* JumpFlood 2D on a device field, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The second field of the passes is a static device array of
 * DOLHIP_NOALLOC_SCRATCH_BYTES (16 MiB) that each device holds until the
 * process ends, so w*h may be at most 4,194,304 cells (2048 x 2048); larger
 * fields are rejected with DOLHIP_ERR_INVALID_ARG.  Every call on a device
 * uses that device's one array: calls that overlap on one device (several
 * streams or threads) race. */
#define DOLHIP_NOALLOC_SCRATCH_BYTES (16 << 20)
int dolhip_jump_flood_2d_noalloc(int w, int h, int16_t* delta, int cw, void* stream);
/* JumpFlood 3D, in place (DistanceOcclusionLib.h:93, .cpp:2396-2472) */
int dolhip_jump_flood_3d(int w, int h, int d, int16_t* delta, int cw);
size_t dolhip_jump_flood_3d_workspace(int w, int h, int d);
int dolhip_jump_flood_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* work,
                             void* stream);

/* FastSweep, in place (DistanceOcclusionLib.h:84-85; .cpp:633-697 2D,
 * 2186-2262 3D): the four quadrant (eight octant) sweeps, bit-identical.
 * No workspace; the _dev forms need a 4-byte aligned 2D field. */
int dolhip_fast_sweep_2d(int w, int h, int16_t* delta, int cw);
int dolhip_fast_sweep_2d_dev(int w, int h, int16_t* delta, int cw, void* stream);
int dolhip_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw);
int dolhip_fast_sweep_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream);

/* BruteForce, in place (DistanceOcclusionLib.h:96-97; .cpp:1087-1110 2D,
 * 2475-2505 3D): the reference's raster order of sources kept exactly (a
 * cell that is a boundary cell at its turn relaxes the whole field), run by
 * one workgroup; (sources x cells) work, for checking.  At most 2^31 - 1
 * cells. */
int dolhip_brute_force_2d(int w, int h, int16_t* delta, int cw);
int dolhip_brute_force_2d_dev(int w, int h, int16_t* delta, int cw, void* stream);
int dolhip_brute_force_3d(int w, int h, int d, int16_t* delta, int cw);
int dolhip_brute_force_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream);

/* Danielsson, in place (DistanceOcclusionLib.h:88-89; .cpp:702-778 2D,
 * 2264-2394 3D), bit-identical.  Its scans along a row keep a running choice
 * that is not associative, so one workgroup walks the rows (slices) in
 * order; rows of at most 40,956 cells (2D) or 27,304 cells (3D). */
int dolhip_danielsson_2d(int w, int h, int16_t* delta, int cw);
int dolhip_danielsson_2d_dev(int w, int h, int16_t* delta, int cw, void* stream);
int dolhip_danielsson_3d(int w, int h, int d, int16_t* delta, int cw);
int dolhip_danielsson_3d_dev(int w, int h, int d, int16_t* delta, int cw, void* stream);

/* Chamfer 3-4, in place (DistanceOcclusionLib.h:65-66; .cpp:912-1006):
 * *multiplier (if not null) gets 3, the divisor to cells.  Exact (each row a
 * min-scan); sums taken in 64 bits. */
int dolhip_chamfer_i16(int w, int h, int16_t* dist, int* multiplier);
int dolhip_chamfer_i16_dev(int w, int h, int16_t* dist, int* multiplier, void* stream);
int dolhip_chamfer_i32(int w, int h, int32_t* dist, int* multiplier);
int dolhip_chamfer_i32_dev(int w, int h, int32_t* dist, int* multiplier, void* stream);

/* Occlusion image and shadow sweep (DistanceOcclusionLib.h:101-103;
 * .cpp:2773-2876): occ = scale on set bits, else 0; shadow from occ for a
 * light along (gx, gy), float results bit-identical. */
int dolhip_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale);
int dolhip_init_occ_dev(int w, int h, const uint32_t* mask, float* occ, float scale, void* stream);
int dolhip_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow);
int dolhip_shadow_sweep_dev(float gx, float gy, int w, int h, const float* occ, float* shadow,
                            void* stream);

/* Direction fields and occlusion sweeps (DistanceOcclusionLib.h:104-112;
 * .cpp:1114-1209, 2507-2769), float results bit-identical.  dirw4 is 4 planes
 * of w*h floats, dirw8 8 planes of w*h*d; the sweeps update in place, or
 * (_from) write out8 from in8, out8's first cell of each plane left as it
 * was.  No workspace. */
int dolhip_init_dirw_2d(int w, int h, const uint32_t* mask, float* dirw4);
int dolhip_init_dirw_2d_dev(int w, int h, const uint32_t* mask, float* dirw4, void* stream);
int dolhip_init_dirw_3d(int w, int h, int d, const uint32_t* mask, float* dirw8);
int dolhip_init_dirw_3d_dev(int w, int h, int d, const uint32_t* mask, float* dirw8, void* stream);
int dolhip_occlusion_sweep_2d(int w, int h, float* dirw4);
int dolhip_occlusion_sweep_2d_dev(int w, int h, float* dirw4, void* stream);
int dolhip_occlusion_sweep_3d(int w, int h, int d, float* dirw8);
int dolhip_occlusion_sweep_3d_dev(int w, int h, int d, float* dirw8, void* stream);
int dolhip_occlusion_sweep_3d_from(int w, int h, int d, const float* in8, float* out8);
int dolhip_occlusion_sweep_3d_from_dev(int w, int h, int d, const float* in8, float* out8,
                                       void* stream);

/* AO images of a swept field (the tool's per-cell EncodeAO4/8 and
 * EncodeDirW4/8 loops, DistanceOcclusionLib.h:143-199): ao one byte a cell,
 * rgba four. */
int dolhip_encode_ao_2d(int w, int h, const float* dirw4, uint8_t* ao);
int dolhip_encode_ao_2d_dev(int w, int h, const float* dirw4, uint8_t* ao, void* stream);
int dolhip_encode_ao_3d(int w, int h, int d, const float* dirw8, uint8_t* ao);
int dolhip_encode_ao_3d_dev(int w, int h, int d, const float* dirw8, uint8_t* ao, void* stream);
int dolhip_encode_dirw_2d(int w, int h, const float* dirw4, uint8_t* rgba);
int dolhip_encode_dirw_2d_dev(int w, int h, const float* dirw4, uint8_t* rgba, void* stream);
int dolhip_encode_dirw_3d(int w, int h, int d, const float* dirw8, uint8_t* rgba);
int dolhip_encode_dirw_3d_dev(int w, int h, int d, const float* dirw8, uint8_t* rgba, void* stream);

/* Felzenszwalb float squared distances in place, bit-identical to the
 * reference's float arithmetic (DistanceOcclusionLib.h:71, .cpp:1018-1085) */
int dolhip_felzenszwalb_2d(int w, int h, float* dist);
size_t dolhip_felzenszwalb_2d_workspace(int w, int h);
int dolhip_felzenszwalb_2d_dev(int w, int h, float* dist, void* work, void* stream);

/* ---- exact squared EDT (extension of DistanceOcclusionLib.cpp:1018-1085) --- */
/* Squared Euclidean distance from each cell to the nearest set bit (clear
 * bit if invert), exact in integers; DOLHIP_EDT_INF where there is none.
 * Either output may be NULL: d2 int32, dist = sqrtf((float)d2) (sqrtf(FLT_MAX)
 * for DOLHIP_EDT_INF, as the reference's float path gives). */
int dolhip_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* dist);
size_t dolhip_edt_2d_workspace(int w, int h);
int dolhip_edt_2d_dev(int w, int h, const uint32_t* mask, int invert, int32_t* d2, float* dist,
                      void* work, void* stream);
int dolhip_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                  float* dist);
size_t dolhip_edt_3d_workspace(int w, int h, int d);
int dolhip_edt_3d_dev(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2,
                      float* dist, void* work, void* stream);
/* Two-field signed distance (the -S semantics of DistanceOcclusionTool.cpp:1029-1036
 * with exact fields): sdf = sqrt(EDT to set bits) - sqrt(EDT to clear bits),
 * negative inside. */
int dolhip_signed_edt_2d(int w, int h, const uint32_t* mask, float* sdf);
int dolhip_signed_edt_2d_dev(int w, int h, const uint32_t* mask, float* sdf, void* work,
                             void* stream);
int dolhip_signed_edt_3d(int w, int h, int d, const uint32_t* mask, float* sdf);
int dolhip_signed_edt_3d_dev(int w, int h, int d, const uint32_t* mask, float* sdf, void* work,
                             void* stream);

/* ---- distances and encoders (DistanceOcclusionTool.cpp:46-75) -------------- */
/* Dist of each delta: sqrtf((float)(x*x + y*y)) (DistanceOcclusionTool.cpp:54-57) */
int dolhip_dist_2d_dev(int n, const int16_t* delta, float* out, void* stream);
/* Signed distance in cells from half-cell border deltas:
 * (bit ? -0.5f : 0.5f) * sqrtf(|delta|^2) */
int dolhip_signed_distance_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                              float* out);
int dolhip_signed_distance_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta,
                                  float* out, void* stream);
int dolhip_signed_distance_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                              float* out);
int dolhip_signed_distance_3d_dev(int w, int h, int d, const uint32_t* mask,
                                  const int16_t* delta, float* out, void* stream);
/* uint8 signed image, zero crossing at 128 (DistanceOcclusionTool.cpp:1242-1271,
 * 1369-1404); max_len is the divisor passed there (twice the tool's maxLen) */
int dolhip_encode_sdf_u8_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                            float max_len, uint8_t* out);
int dolhip_encode_sdf_u8_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta,
                                float max_len, uint8_t* out, void* stream);
int dolhip_encode_sdf_u8_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                            float max_len, uint8_t* out);
int dolhip_encode_sdf_u8_3d_dev(int w, int h, int d, const uint32_t* mask,
                                const int16_t* delta, float max_len, uint8_t* out, void* stream);
/* RGBA8 delta image (DistanceOcclusionTool.cpp:1281-1288) and red/green image
 * (DistanceOcclusionTool.cpp:1289-1307) */
int dolhip_encode_sdf_delta_rgba_2d(int w, int h, const int16_t* delta, float max_len,
                                    uint8_t* out);
int dolhip_encode_sdf_rg_rgba_2d(int w, int h, const uint32_t* mask, const int16_t* delta,
                                 float max_len, uint8_t* out);

/* Every image of the border method: WriteDistances, DistanceOcclusionTool.cpp:1240-1331
 * (2D) and 1367-1469 (3D, the slices one after another).  type 0 sdf (one byte a
 * cell), 1 sdf-delta, 2 sdf-rg, 3 sdf-angles (RGBA8, 4-byte aligned); max_len is
 * the divisor passed there (twice the tool's maxLen).  Bit-identical. */
int dolhip_encode_border_image_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                                  int type, uint8_t* out);
int dolhip_encode_border_image_2d_dev(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                                      int type, uint8_t* out, void* stream);
int dolhip_encode_border_image_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                                  float max_len, int type, uint8_t* out);
int dolhip_encode_border_image_3d_dev(int w, int h, int d, const uint32_t* mask, const int16_t* delta,
                                      float max_len, int type, uint8_t* out, void* stream);
/* The unsigned distance images (WriteDistances, DistanceOcclusionTool.cpp:678-698
 * 2D, 886-899 3D): EncodeU8(Dist(delta) / max_len) of a delta field, and
 * EncodeU8(v / max_len) of scalar distances (Felzenszwalb's float roots; Chamfer's
 * int16 values with max_len times its multiplier); sqrtf of float squared
 * distances (:753-754). */
int dolhip_encode_distance_u8_2d(int w, int h, const int16_t* delta, float max_len, uint8_t* out);
int dolhip_encode_distance_u8_2d_dev(int w, int h, const int16_t* delta, float max_len, uint8_t* out,
                                     void* stream);
int dolhip_encode_distance_u8_3d(int w, int h, int d, const int16_t* delta, float max_len, uint8_t* out);
int dolhip_encode_distance_u8_3d_dev(int w, int h, int d, const int16_t* delta, float max_len, uint8_t* out,
                                     void* stream);
int dolhip_encode_scalar_u8_f32(int w, int h, const float* v, float max_len, uint8_t* out);
int dolhip_encode_scalar_u8_f32_dev(int w, int h, const float* v, float max_len, uint8_t* out, void* stream);
int dolhip_encode_scalar_u8_i32(int w, int h, const int32_t* v, float max_len, uint8_t* out);
int dolhip_encode_scalar_u8_i32_dev(int w, int h, const int32_t* v, float max_len, uint8_t* out, void* stream);
int dolhip_encode_scalar_u8_i16(int w, int h, const int16_t* v, float max_len, uint8_t* out);
int dolhip_encode_scalar_u8_i16_dev(int w, int h, const int16_t* v, float max_len, uint8_t* out, void* stream);
int dolhip_sqrt_f32(int w, int h, const float* v, float* out);
int dolhip_sqrt_f32_dev(int w, int h, const float* v, float* out, void* stream);

/* ---- signed SDF pipelines (DistanceOcclusionTool.cpp:1333-1365, 1471-1498) --- */
/* border init + JumpFlood(cw = 2), delta left in `delta` */
int dolhip_signed_border_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta, void* work,
                                void* stream);
int dolhip_signed_border_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta,
                                void* work, void* stream);

/* ---- two-field signed images: the tool's -S mode (GenerateSignedDistance,
 * DistanceOcclusionTool.cpp:1082-1124 (2D), 1197-1232 (3D)) ---------------- */
/* exterior field delta0 = InitDeltasFromBitMask(mask, invert = false) and
 * interior field delta1 = (invert = true), each then JumpFlood(cw) (the tool
 * uses cw = 1).  work: dolhip_two_field_*_workspace bytes, or NULL. */
size_t dolhip_two_field_2d_workspace(int w, int h);
size_t dolhip_two_field_3d_workspace(int w, int h, int d);
int dolhip_two_field_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta0,
                               int16_t* delta1, int cw);
int dolhip_two_field_deltas_2d_dev(int w, int h, const uint32_t* mask, int16_t* delta0,
                                   int16_t* delta1, int cw, void* work, void* stream);
int dolhip_two_field_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta0,
                               int16_t* delta1, int cw);
int dolhip_two_field_deltas_3d_dev(int w, int h, int d, const uint32_t* mask, int16_t* delta0,
                                   int16_t* delta1, int cw, void* work, void* stream);
/* the images WriteDistances writes from the two fields
 * (DistanceOcclusionTool.cpp:1029-1080 (2D), 1126-1195 (3D)); type:
 *   0 sdf:        EncodeU8Signed((Dist(d0) - Dist(d1)) / max_len), 1 byte a cell
 *   1 sdf-delta:  RGBA, channels EncodeU8Signed((d0.c + d1.c) / max_len) for
 *                 x, y (and z in 3D; blue 0 in 2D), alpha 255
 *   2 sdf-rg:     RGBA, EncodeU8(Dist(d0) / max_len), EncodeU8(Dist(d1) / max_len), 0, 255
 *   3 sdf-angles: as 1 over Dist(d0) - Dist(d1) instead of max_len; in 3D every
 *                 slice shows the first slice's values, as the tool writes them */
int dolhip_encode_two_field_2d(int w, int h, const int16_t* delta0, const int16_t* delta1,
                               float max_len, int type, uint8_t* out);
int dolhip_encode_two_field_2d_dev(int w, int h, const int16_t* delta0, const int16_t* delta1,
                                   float max_len, int type, uint8_t* out, void* stream);
int dolhip_encode_two_field_3d(int w, int h, int d, const int16_t* delta0, const int16_t* delta1,
                               float max_len, int type, uint8_t* out);
int dolhip_encode_two_field_3d_dev(int w, int h, int d, const int16_t* delta0,
                                   const int16_t* delta1, float max_len, int type, uint8_t* out,
                                   void* stream);

/* ---- surface voxeliser: CreateBitMaskFromTriangles (MeshSupport.cpp:369-475,
 * declared MeshSupport.h:44-53) ----------------------------------------------
 * ORs into mask (w x h x d bitmask layout) every cell the triangles touch:
 * the reference's triangle/cell test at each cell of a triangle's cell range
 * (a one-cell range is set without the test).  vertices: float x, y, z
 * triples; indices: 3 per triangle, or NULL for consecutive vertices (the
 * reference reads indices even when NULL); bbox: {min x, y, z, max x, y, z},
 * the volume's extent (the tool passes FindBounds of the mesh,
 * MeshSupport.cpp:833-860).  The _dev form takes device vertices, indices
 * and mask (bbox stays a host pointer) and trusts indices to be in range. */
int dolhip_mask_from_triangles(int tri_count, const int* indices, const float* vertices,
                               const float* bbox, int w, int h, int d, uint32_t* mask);
int dolhip_mask_from_triangles_dev(int tri_count, const int* indices, const float* vertices,
                                   const float* bbox, int w, int h, int d, uint32_t* mask,
                                   void* stream);

/* ---- batched ragged glyphs (extension) ------------------------------------- */
/* n masks of uint8 coverage, mask i of size w[i] x h[i] starting at byte
 * offset[i] of `bytes` (and of `out`).  Per mask: threshold (>= threshold XOR
 * less_than), border init, JumpFlood(cw = 2), EncodeU8Signed with
 * max_len = 0.5 * max(w, h) (the tool's 2 * maxLen with maxLen = 0.25 * max(w, h),
 * DistanceOcclusionTool.cpp:1973-1974).  Every mask needs w, h in [1, 256]. */
typedef struct dolhip_batch dolhip_batch;
int dolhip_batch_create(int n, const int32_t* w, const int32_t* h, const int64_t* offset,
                        int threshold, int less_than, dolhip_batch** out);
int dolhip_batch_destroy(dolhip_batch* b);
int64_t dolhip_batch_bytes(const dolhip_batch* b);
/* bytes / out are device pointers; with graph != 0 the launch is captured in
 * a hipGraph the first time and replayed afterwards for the same pointers. */
int dolhip_batch_sdf_u8_dev(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, int graph,
                            void* stream);
/* the same per-mask pipeline with one kernel launch per step (the per-image
 * form); work: at least (ceil(w/32)*h + 2*w*h)*4 + 512 bytes of device memory */
int dolhip_glyph_sdf_u8_dev(int w, int h, const uint8_t* bytes, int threshold, int less_than,
                            uint8_t* out, void* work, void* stream);
/* every mask of the batch through dolhip_glyph_sdf_u8_dev, one after another
 * on `stream` (the per-image form of the same work) */
int dolhip_batch_per_image_dev(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, void* stream);
/* the chosen variant's LDS capacity in cells (masks above it run as per-step
 * launches) and wavefronts a workgroup (for reports) */
int dolhip_batch_config(const dolhip_batch* b, int* tile_cells, int* waves_per_group);
/* select kernel variant `variant` (0 <= variant < count) for later calls;
 * with variant < 0 returns the number of variants instead */
int dolhip_batch_set_variant(dolhip_batch* b, int variant);
/* times every (threads, groups of 64 cells a wavefront) variant on this batch
 * and keeps the fastest for later dolhip_batch_sdf_u8_dev calls; overwrites
 * `out` */
int dolhip_batch_tune(dolhip_batch* b, const uint8_t* bytes, uint8_t* out, void* stream);

/* ---- continuous seeds (extension) ---------------------------------------- */
/* Nearest of n seeds (float x, y pairs) to each cell centre (x, y) of a w x h
 * grid, exact under float64 comparison with ties to the lowest index; idx int32
 * and dist float32 = |c - p|.  work: dolhip_seed_workspace(w, h, n) bytes. */
size_t dolhip_seed_workspace(int w, int h, int n);
int dolhip_seed_nearest_2d_dev(int w, int h, int n, const float* seeds, int32_t* idx,
                               float* dist, void* work, void* stream);
/* grad[k] = sum over cells c of seed k of -gout(c) (c - p_k) / d(c), zero where
 * d = 0; grad is 2n floats and is overwritten (float atomics: the summation
 * order, hence the last bits, vary between runs) */
int dolhip_seed_grad_2d_dev(int w, int h, int n, const float* seeds, const int32_t* idx,
                            const float* dist, const float* gout, float* grad, void* stream);

#ifdef __cplusplus
}
#endif
#endif
