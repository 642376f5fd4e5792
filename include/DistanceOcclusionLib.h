/*
 * DistanceOcclusionLib.h -- drop-in C++ interface for the distance-field path
 * of DistanceOcclusionLib, backed by libdolhip (gfx950 kernels).
 *
 * Declares, with the reference's signatures, types, constants and default
 * arguments, every function of DistanceOcclusionLib.h (:16-117): masks and
 * generators, cell deltas, distance initialisation, Chamfer, Felzenszwalb,
 * delta and border initialisation, FastSweep, Danielsson, JumpFlood,
 * BruteForce, the occlusion image, shadow sweep, direction fields and
 * occlusion sweeps, and the inline encoders (:121-199), so a program written
 * against the reference header builds against this one unchanged.
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

    int Chamfer(int w, int h, int16_t distances[]);
    int Chamfer(int w, int h, int32_t distances[]);
    // returns the multiplier: divide by it to get distances in cells

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

    void Danielsson(int w, int h,        cCellDelta2 delta[], int cw = 1);
    void Danielsson(int w, int h, int d, cCellDelta3 delta[], int cw = 1);

    void BruteForce(int w, int h,        cCellDelta2 delta[], int cw = 1);
    void BruteForce(int w, int h, int d, cCellDelta3 delta[], int cw = 1);

    // --- Occlusion (DistanceOcclusionLib.h:101-117) ----------------------------
    void InitOccFromBitMask(int w, int h, const uint32_t mask[], float occlusion[], float scale = 1.0f);
    void ShadowSweep(float gx, float gy, int w, int h, const float* occlusion, float* shadow);

    void InitDirWFromBitMask(int w, int h,        const uint32_t mask[], float dirW4[]);
    void InitDirWFromBitMask(int w, int h, int d, const uint32_t mask[], float dirW8[]);
    void OcclusionSweep(int w, int h,        float dirW4[]);
    void OcclusionSweep(int w, int h, int d, float dirW8[]);
    void OcclusionSweep(int w, int h, int d, const float inDirW8[], float outDirW8[]);

    uint8_t EncodeAO4  (const float dirW[4]);
    uint8_t EncodeAO8  (const float dirW[8]);
    void    EncodeDirW4(const float dirW[4], uint8_t aoDir[4]);
    void    EncodeDirW8(const float dirW[8], uint8_t aoDir[4]);

    // --- Extensions (no counterpart in the reference) ------------------------
    // Exact squared Euclidean distance to the nearest set bit (clear bit if
    // invert); INT32_MAX where there is none.  dist (optional) gets sqrtf of it.
    void EDT(int w, int h,        const uint32_t mask[], int32_t d2[], float dist[] = 0, bool invert = false);
    void EDT(int w, int h, int d, const uint32_t mask[], int32_t d2[], float dist[] = 0, bool invert = false);
    // sqrt(EDT to set bits) - sqrt(EDT to clear bits): negative inside
    void SignedEDT(int w, int h,        const uint32_t mask[], float sdf[]);
    void SignedEDT(int w, int h, int d, const uint32_t mask[], float sdf[]);

    // --- Inlines (DistanceOcclusionLib.h:121-199) ----------------------------
    inline int MaskSize(int w)               { return (w + 31) / 32; }
    inline int MaskSize(int w, int h)        { return ((w + 31) / 32) * h; }
    inline int MaskSize(int w, int h, int d) { return ((w + 31) / 32) * h * d; }

    inline uint8_t EncodeU8(float v)
    {
        v = (v < 0.0f) ? 0.0f : ((v > 1.0f) ? 1.0f : v);
        return (uint8_t) lrintf(v * 255);
    }

    // The direction encoders sum the n = 4 or 8 direction weights left to
    // right, each one added, or taken away where bit `axis` of its index is
    // set (the quadrant or octant lies on the negative side of that axis):
    // the same float operations in the same order as the reference's sums.
    inline float DirWSum(const float dirW[], int n, int axis)
    {
        float s = dirW[0];
        for (int i = 1; i < n; i++)
            s = (axis >= 0 && ((i >> axis) & 1)) ? s - dirW[i] : s + dirW[i];
        return s;
    }

    // overall occlusion: 1 - mean weight
    inline uint8_t EncodeAO4(const float dirW[4]) { return EncodeU8(1.0f - 0.25f  * DirWSum(dirW, 4, -1)); }
    inline uint8_t EncodeAO8(const float dirW[8]) { return EncodeU8(1.0f - 0.125f * DirWSum(dirW, 8, -1)); }

    // directed occlusion: the unit direction of the signed sums, scaled by
    // the mean weight and mapped to [0, 1], then the mean weight (2D: in the
    // third byte, the fourth 255; 3D: in the fourth)
    inline void EncodeDirW4(const float dirW[4], uint8_t aoDir[4])
    {
        float x = DirWSum(dirW, 4, 0);
        float y = DirWSum(dirW, 4, 1);
        const float inv = 1.0f / sqrtf(x * x + y * y);
        x *= inv;
        y *= inv;
        const float mean = DirWSum(dirW, 4, -1) * 0.25f;
        aoDir[0] = EncodeU8(mean * x * 0.5f + 0.5f);
        aoDir[1] = EncodeU8(mean * y * 0.5f + 0.5f);
        aoDir[2] = EncodeU8(mean);
        aoDir[3] = 255;
    }

    inline void EncodeDirW8(const float dirW[8], uint8_t aoDir[4])
    {
        float x = DirWSum(dirW, 8, 0);
        float y = DirWSum(dirW, 8, 1);
        float z = DirWSum(dirW, 8, 2);
        const float inv = 1.0f / sqrtf(x * x + y * y + z * z);
        x *= inv;
        y *= inv;
        z *= inv;
        const float mean = DirWSum(dirW, 8, -1) * 0.125f;
        aoDir[0] = EncodeU8(mean * x * 0.5f + 0.5f);
        aoDir[1] = EncodeU8(mean * y * 0.5f + 0.5f);
        aoDir[2] = EncodeU8(mean * z * 0.5f + 0.5f);
        aoDir[3] = EncodeU8(mean);
    }
}

#endif
