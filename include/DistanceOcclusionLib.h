/*
 * DistanceOcclusionLib.h -- drop-in C++ interface for the distance-field path
 * of DistanceOcclusionLib, backed by libdolhip (gfx950 kernels).
 *
 * Declares, with the reference's signatures, types and constants, the
 * functions of DistanceOcclusionLib.h that lie on the data-parallel distance
 * path (mask creation and generators :21-35, cell deltas :41-57, distance
 * initialisation and Felzenszwalb :61-71, delta initialisation and border
 * initialisation :75-82, FastSweep :84-85, JumpFlood :92-94, direction fields
 * and occlusion sweeps :104-112, MaskSize/EncodeU8 :124-141).  Danielsson,
 * Chamfer, BruteForce, the shadow sweep and the AO encoders are not part of
 * this path and are not declared here.
 *
 * Semantics are the reference's: buffers are host memory owned by the
 * caller and updated in place; every call is synchronous; results are
 * bit-identical.  The reference has no error channel, so a device failure
 * (or no gfx950 device) prints the library's message and aborts; there is no
 * CPU fallback.  Extensions (exact integer EDT in 2D/3D, two-field signed
 * EDT) are declared at the end.
 */
#ifndef DISTANCE_OCCLUSION_H
#define DISTANCE_OCCLUSION_H

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

namespace DOL
{
    // --- Bit masks (DistanceOcclusionLib.h:16-36) ----------------------------
    int MaskSize(int w);
    int MaskSize(int w, int h);
    int MaskSize(int w, int h, int d);

    uint32_t* CreateBitMask (int w, int h);
    uint32_t* CreateBitMask (int w, int h, int d);
    void      DestroyBitMask(uint32_t*& mask);

    void BitMaskAddPoints(int w, int h,        int n, uint32_t seed, uint32_t mask[]);
    void BitMaskAddPoints(int w, int h, int d, int n, uint32_t seed, uint32_t mask[]);
    void BitMaskAddBlock (int w, int h,        int sides, uint32_t mask[]);
    void BitMaskAddBlock (int w, int h, int d, int sides, uint32_t mask[]);

    // --- Cell deltas (DistanceOcclusionLib.h:41-57) --------------------------
    struct cCellDelta2
    {
        int16_t x;
        int16_t y;
    };

    struct cCellDelta3
    {
        int16_t x;
        int16_t y;
        int16_t z;
    };

    const int16_t kMaxDelta2 = INT16_MAX;
    const int16_t kMaxDelta3 = 25000;

    // --- Distance (DistanceOcclusionLib.h:61-71) -----------------------------
    void InitDistancesFromBitMask(int w, int h, const uint32_t mask[], int16_t distances[], int16_t maxD = INT16_MAX);
    void InitDistancesFromBitMask(int w, int h, const uint32_t mask[], int32_t distances[], int32_t maxD = INT32_MAX);
    void InitDistancesFromBitMask(int w, int h, const uint32_t mask[], float   distances[], float   maxD = FLT_MAX);

    void Felzenszwalb(int w, int h, float distances[]);

    // --- Delta-based (DistanceOcclusionLib.h:75-94) --------------------------
    void InitDeltasFromBitMask(int w, int h,        const uint32_t mask[], cCellDelta2 delta[], bool invert = false);
    void InitDeltasFromBitMask(int w, int h, int d, const uint32_t mask[], cCellDelta3 delta[], bool invert = false);

    void InitBorderDeltasFromBitMask(int w, int h,        const uint32_t mask[], cCellDelta2 delta[]);
    void InitBorderDeltasFromBitMask(int w, int h, int d, const uint32_t mask[], cCellDelta3 delta[]);

    void JumpFlood(int w, int h,        cCellDelta2 delta[], int cw = 1);
    void JumpFlood(int w, int h, int d, cCellDelta3 delta[], int cw = 1);

    void FastSweep(int w, int h,        cCellDelta2 delta[], int cw = 1);
    void FastSweep(int w, int h, int d, cCellDelta3 delta[], int cw = 1);

    // --- Occlusion (DistanceOcclusionLib.h:104-112) ----------------------------
    void InitDirWFromBitMask(int w, int h,        const uint32_t mask[], float dirW4[]);
    void InitDirWFromBitMask(int w, int h, int d, const uint32_t mask[], float dirW8[]);
    void OcclusionSweep(int w, int h,        float dirW4[]);
    void OcclusionSweep(int w, int h, int d, float dirW8[]);
    void OcclusionSweep(int w, int h, int d, const float inDirW8[], float outDirW8[]);

    // --- Extensions (no counterpart in the reference) ------------------------
    // Exact squared Euclidean distance to the nearest set bit (clear bit if
    // invert); INT32_MAX where there is none.  dist (optional) gets sqrtf of it.
    void EDT(int w, int h,        const uint32_t mask[], int32_t d2[], float dist[] = 0, bool invert = false);
    void EDT(int w, int h, int d, const uint32_t mask[], int32_t d2[], float dist[] = 0, bool invert = false);
    // sqrt(EDT to set bits) - sqrt(EDT to clear bits): negative inside
    void SignedEDT(int w, int h,        const uint32_t mask[], float sdf[]);
    void SignedEDT(int w, int h, int d, const uint32_t mask[], float sdf[]);

    // --- Inlines (DistanceOcclusionLib.h:124-141) ----------------------------
    inline int MaskSize(int w)               { return (w + 31) / 32; }
    inline int MaskSize(int w, int h)        { return ((w + 31) / 32) * h; }
    inline int MaskSize(int w, int h, int d) { return ((w + 31) / 32) * h * d; }

    inline uint8_t EncodeU8(float v)
    {
        v = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
        return (uint8_t) lrintf(v * 255);
    }
}

#endif
