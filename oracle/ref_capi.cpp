// ref_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI face over the reference library, DistanceOcclusionLib.cpp, compiled
// from its own sources under /root/reference by oracle/Makefile into
// oracle/_ref/libdolref.so.  Tests use it to pin the oracle restatement (and,
// through the oracle, the HIP library) to the reference bit for bit; bench.py
// may time it as the CPU baseline.  Nothing here is product code.
//
// Each entry point forwards to the DOL:: function it is named after
// (DistanceOcclusionLib.h:21-98); delta arrays are passed as int16 triples or
// pairs, which is the cCellDelta2/cCellDelta3 layout (DistanceOcclusionLib.h:41-53).
#include "DistanceOcclusionLib.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace DOL;

extern "C" {

// threads of the reference's OpenMP loops in the -fopenmp build
// (libdolref_omp.so: JumpFlood's row loops, DistanceOcclusionLib.cpp:784, 2402);
// returns the count in effect, 1 in the plain build
int ref_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
    return n > 0 ? n : 1;
#else
    (void)n;
    return 1;
#endif
}

void ref_mask_add_points_2d(int w, int h, int n, uint32_t seed, uint32_t* mask)
{ BitMaskAddPoints(w, h, n, seed, mask); }
void ref_mask_add_points_3d(int w, int h, int d, int n, uint32_t seed, uint32_t* mask)
{ BitMaskAddPoints(w, h, d, n, seed, mask); }
void ref_mask_add_block_2d(int w, int h, int sides, uint32_t* mask)
{ BitMaskAddBlock(w, h, sides, mask); }
void ref_mask_add_block_3d(int w, int h, int d, int sides, uint32_t* mask)
{ BitMaskAddBlock(w, h, d, sides, mask); }

void ref_init_deltas_2d(int w, int h, const uint32_t* mask, int16_t* delta, int invert)
{ InitDeltasFromBitMask(w, h, mask, reinterpret_cast<cCellDelta2*>(delta), invert != 0); }
void ref_init_deltas_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta, int invert)
{ InitDeltasFromBitMask(w, h, d, mask, reinterpret_cast<cCellDelta3*>(delta), invert != 0); }
void ref_border_2d(int w, int h, const uint32_t* mask, int16_t* delta)
{ InitBorderDeltasFromBitMask(w, h, mask, reinterpret_cast<cCellDelta2*>(delta)); }
void ref_border_3d(int w, int h, int d, const uint32_t* mask, int16_t* delta)
{ InitBorderDeltasFromBitMask(w, h, d, mask, reinterpret_cast<cCellDelta3*>(delta)); }
void ref_init_distances_f32(int w, int h, const uint32_t* mask, float* dist, float max_d)
{ InitDistancesFromBitMask(w, h, mask, dist, max_d); }
void ref_init_distances_i32(int w, int h, const uint32_t* mask, int32_t* dist, int32_t max_d)
{ InitDistancesFromBitMask(w, h, mask, dist, max_d); }
void ref_init_distances_i16(int w, int h, const uint32_t* mask, int16_t* dist, int16_t max_d)
{ InitDistancesFromBitMask(w, h, mask, dist, max_d); }

void ref_jfa_2d(int w, int h, int16_t* delta, int cw)
{ JumpFlood(w, h, reinterpret_cast<cCellDelta2*>(delta), cw); }
void ref_jfa_3d(int w, int h, int d, int16_t* delta, int cw)
{ JumpFlood(w, h, d, reinterpret_cast<cCellDelta3*>(delta), cw); }
void ref_fast_sweep_2d(int w, int h, int16_t* delta, int cw)
{ FastSweep(w, h, reinterpret_cast<cCellDelta2*>(delta), cw); }
void ref_fast_sweep_3d(int w, int h, int d, int16_t* delta, int cw)
{ FastSweep(w, h, d, reinterpret_cast<cCellDelta3*>(delta), cw); }
void ref_init_dirw_2d(int w, int h, const uint32_t* mask, float* dirw)
{ InitDirWFromBitMask(w, h, mask, dirw); }
void ref_init_dirw_3d(int w, int h, int d, const uint32_t* mask, float* dirw)
{ InitDirWFromBitMask(w, h, d, mask, dirw); }
void ref_occlusion_sweep_2d(int w, int h, float* dirw)
{ OcclusionSweep(w, h, dirw); }
void ref_occlusion_sweep_3d(int w, int h, int d, float* dirw)
{ OcclusionSweep(w, h, d, dirw); }
void ref_occlusion_sweep_3d_from(int w, int h, int d, const float* in, float* out)
{ OcclusionSweep(w, h, d, in, out); }
void ref_danielsson_2d(int w, int h, int16_t* delta, int cw)
{ Danielsson(w, h, reinterpret_cast<cCellDelta2*>(delta), cw); }
void ref_danielsson_3d(int w, int h, int d, int16_t* delta, int cw)
{ Danielsson(w, h, d, reinterpret_cast<cCellDelta3*>(delta), cw); }
int ref_chamfer_i16(int w, int h, int16_t* dist) { return Chamfer(w, h, dist); }
int ref_chamfer_i32(int w, int h, int32_t* dist) { return Chamfer(w, h, dist); }
void ref_init_occ(int w, int h, const uint32_t* mask, float* occ, float scale)
{ InitOccFromBitMask(w, h, mask, occ, scale); }
void ref_shadow_sweep(float gx, float gy, int w, int h, const float* occ, float* shadow)
{ ShadowSweep(gx, gy, w, h, occ, shadow); }
void ref_felzenszwalb_2d(int w, int h, float* dist)
{ Felzenszwalb(w, h, dist); }
void ref_brute_force_2d(int w, int h, int16_t* delta, int cw)
{ BruteForce(w, h, reinterpret_cast<cCellDelta2*>(delta), cw); }
void ref_brute_force_3d(int w, int h, int d, int16_t* delta, int cw)
{ BruteForce(w, h, d, reinterpret_cast<cCellDelta3*>(delta), cw); }

uint8_t ref_encode_u8(float v) { return EncodeU8(v); }

// the header's per-cell AO encoders over n cells of 4 or 8 planes
void ref_encode_ao(size_t n, int planes, const float* dirw, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) {
        float c[8];
        for (int p = 0; p < planes; p++) c[p] = dirw[p * n + i];
        out[i] = planes == 4 ? EncodeAO4(c) : EncodeAO8(c);
    }
}
void ref_encode_dirw(size_t n, int planes, const float* dirw, uint8_t* out)
{
    for (size_t i = 0; i < n; i++) {
        float c[8];
        for (int p = 0; p < planes; p++) c[p] = dirw[p * n + i];
        if (planes == 4) EncodeDirW4(c, out + 4 * i);
        else EncodeDirW8(c, out + 4 * i);
    }
}

}
