/*
 * dol_oracle.h -- CPU restatement of DistanceOcclusionLib's distance-field
 * path.  TEST INFRASTRUCTURE ONLY: this is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never links or calls it.
 *
 * Layouts follow DistanceOcclusionLib.h:16-57:
 *   bitmask : rows padded to ceil(w/32) uint32 words, bit i of word j is
 *             pixel x = 32j + i (LSB first); slices of ceil(w/32)*h words.
 *   delta2  : int16 {x, y} per cell, row-major, cell index y*w + x.
 *   delta3  : int16 {x, y, z} per cell (6 bytes, packed), index z*w*h + y*w + x.
 *   kMaxDelta2 = 32767, kMaxDelta3 = 25000 (DistanceOcclusionLib.h:56-57).
 */
#ifndef DOL_ORACLE_H
#define DOL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAX_DELTA2 32767
#define OC_MAX_DELTA3 25000
#define OC_EDT_INF    0x7fffffff   /* squared distance where no seed exists */

/* generators and byte -> bit (Lib:28-32, 95-179, 1224-1287; Tool:634-670) */
uint32_t oc_rng(uint32_t* seed, int limit);
void oc_mask_add_points_2d(int w, int h, int count, uint32_t seed, uint32_t* mask);
void oc_mask_add_points_3d(int w, int h, int d, int count, uint32_t seed, uint32_t* mask);
void oc_mask_add_block_2d(int w, int h, int sides, uint32_t* mask);
void oc_mask_add_block_3d(int w, int h, int d, int sides, uint32_t* mask);
void oc_mask_from_bytes(int w, int rows, const uint8_t* bytes, int threshold, int less_than,
                        uint32_t* mask);

/* initialisation (Lib:181-254, 1289-1321) */
void oc_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert);
void oc_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert);
void oc_init_distances_f32(int w, int h, const uint32_t* mask, float* dist, float max_d);
void oc_init_distances_i32(int w, int h, const uint32_t* mask, int32_t* dist, int32_t max_d);
void oc_init_distances_i16(int w, int h, const uint32_t* mask, int16_t* dist, int16_t max_d);

/* signed border initialisation, Combo write order (Lib:504-619, 1820-2173) */
void oc_border_2d(int w, int h, const uint32_t* mask, int16_t* delta);
void oc_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta);

/* jump flooding (Lib:780-908, 2396-2472) */
int  oc_floor_pow2(int v);
void oc_jfa_pass_2d(int step, int w, int h, const int16_t* in, int16_t* out, int cw);
void oc_jfa_pass_3d(int step, int w, int h, int d, const int16_t* in, int16_t* out, int cw);
void oc_jfa_2d(int w, int h, int16_t* delta, int cw);
void oc_jfa_3d(int w, int h, int d, int16_t* delta, int cw);
void oc_jfa_2d_mt(int w, int h, int16_t* delta, int cw, int threads);
void oc_jfa_3d_mt(int w, int h, int d, int16_t* delta, int cw, int threads);

/* quadrant / octant sweeps, in place (Lib:633-697, 2186-2262) */
void oc_fast_sweep_2d(int w, int h, int16_t* delta, int cw);
void oc_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw);

/* direction fields and occlusion sweeps (Lib:1114-1209, 2507-2769) */
void oc_init_dirw_2d(int w, int h, const uint32_t* mask, float* dirw4);
void oc_init_dirw_3d(int w, int h, int d, const uint32_t* mask, float* dirw8);
void oc_occlusion_sweep_2d(int w, int h, float* dirw4);
void oc_occlusion_sweep_3d(int w, int h, int d, float* dirw8);
void oc_occlusion_sweep_3d_from(int w, int h, int d, const float* in8, float* out8);
/* AO images from n cells of 4 or 8 planes: one byte a cell, or RGBA
 * (Lib.h:143-199) */
void oc_encode_ao(size_t n, int planes, const float* dirw, uint8_t* out);
void oc_encode_dirw(size_t n, int planes, const float* dirw, uint8_t* out);

/* Felzenszwalb float squared distances, in place (Lib:1018-1085) */
void oc_felzenszwalb_2d(int w, int h, float* dist);

/* brute force (Lib:1087-1110, 2475-2505) */
void oc_brute_force_2d(int w, int h, int16_t* delta, int cw);
void oc_brute_force_3d(int w, int h, int d, int16_t* delta, int cw);

/* exact squared EDT to set bits (invert: to clear bits); OC_EDT_INF if none */
void oc_edt_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2);
void oc_edt_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2);
void oc_edt_brute_2d(int w, int h, const uint32_t* mask, int invert, int32_t* d2);
void oc_edt_brute_3d(int w, int h, int d, const uint32_t* mask, int invert, int32_t* d2);
float oc_edt_float(int32_t d2);

/* distances and encoders (Tool:46-75, 678-698, 1029-1040, 1240-1331, 1367-1469) */
int32_t oc_len2_2(const int16_t* c);
int32_t oc_len2_3(const int16_t* c);
uint8_t oc_encode_u8(float v);
uint8_t oc_encode_u8_signed(float v);
void oc_dist_2d(int n, const int16_t* delta, float* out);
void oc_signed_dist_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float* out);
void oc_signed_dist_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta, float* out);
void oc_sdf_u8_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len, uint8_t* out);
void oc_sdf_u8_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta, float max_len,
                  uint8_t* out);
void oc_sdf_delta_rgba_2d(int w, int h, const int16_t* delta, float max_len, uint8_t* out);
void oc_sdf_rg_rgba_2d(int w, int h, const uint32_t* mask, const int16_t* delta, float max_len,
                       uint8_t* out);
void oc_two_field_image_2d(int w, int h, const int16_t* d0, const int16_t* d1, float max_len,
                           int type, uint8_t* out);
void oc_two_field_image_3d(int w, int h, int d, const int16_t* d0, const int16_t* d1,
                           float max_len, int type, uint8_t* out);

/* surface voxeliser of MeshSupport.cpp:369-475 (ORs into mask); bbox =
 * {min xyz, max xyz}; indices NULL reads consecutive vertices */
void oc_mask_from_triangles(int tri_count, const int* indices, const float* vertices,
                            const float* bbox, int w, int h, int d, uint32_t* mask);

/* invariants of Tool:371-439 (CheckDeltas); returns number of failing cells */
long oc_check_deltas_2d(int w, int h, const int16_t* delta, int cw);
long oc_check_deltas_3d(int w, int h, int d, const int16_t* delta, int cw);

/* continuous seeds (extension): nearest seed index + distance, float64 */
void oc_seed_nearest_2d(int w, int h, int n, const float* seeds, int32_t* idx, double* dist);
void oc_seed_nearest_2d_mt(int w, int h, int n, const float* seeds, int32_t* idx, double* dist,
                           int threads);
void oc_seed_grad_2d(int w, int h, int n, const float* seeds, const int32_t* idx,
                     const float* weight, double* grad);

/* raster-order methods off the data-parallel path (Lib:702-778, 912-1013,
 * 2264-2394, 2773-2876): Danielsson, Chamfer (returns its divisor 3), the
 * occlusion init and the shadow sweep */
void oc_danielsson_2d(int w, int h, int16_t* delta, int cw);
void oc_danielsson_3d(int w, int h, int d, int16_t* delta, int cw);
int  oc_chamfer_i16(int w, int h, int16_t* dist);
int  oc_chamfer_i32(int w, int h, int32_t* dist);
void oc_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale);
void oc_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow);

/* batched glyphs (extension): threshold, border init, JumpFlood(cw = 2) and
 * the signed byte image of each ragged mask */
void oc_glyph_sdf_batch(int n, const int32_t* w, const int32_t* h, const int64_t* offset,
                        const uint8_t* bytes, int threshold, int less_than, uint8_t* out,
                        int threads);

/* image writers of the tool (DistanceOcclusionTool.cpp:678-698, 1240-1331, 1367-1469) */
void oc_border_image(int w, int h, int d, int dims, const uint32_t* mask, const int16_t* delta,
                     float max_len, int type, uint8_t* out);
void oc_distance_u8(size_t n, int dims, const int16_t* delta, float max_len, uint8_t* out);
void oc_scalar_u8_f32(size_t n, const float* v, float max_len, uint8_t* out);
void oc_scalar_u8_i32(size_t n, const int32_t* v, float max_len, uint8_t* out);
void oc_scalar_u8_i16(size_t n, const int16_t* v, float max_len, uint8_t* out);
void oc_sqrt_f32(size_t n, const float* v, float* out);

/* CheckBorderDeltas with the x / y loop swap fixed (DistanceOcclusionTool.cpp:441-599):
 * number of failed expectations */
long oc_check_border_deltas_2d(int w, int h, const uint32_t* mask, const int16_t* delta);
long oc_check_border_deltas_3d(int w, int h, int d, const uint32_t* mask, const int16_t* delta);

#ifdef __cplusplus
}
#endif

#endif
