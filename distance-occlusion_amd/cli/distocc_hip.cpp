// distocc_hip -- the distance modes of the reference tool (distocc) through
// the DOL:: drop-in interface, for checking the C++ boundary end to end.
//
//   distocc_hip (-d | -s | -S) [-m method] [-w size] [-W size] [-p count [seed]]
//               [-b sides] [-l maxLen] [-c] [-o file]
//
// Flags as the reference's (DistanceOcclusionTool.cpp:1782-1883): -d distance,
// -s signed distance with border initialisation, -S signed distance from
// exterior and interior fields, -m 1 FastSweep, 2 JumpFlood (default here)
// or 5 Felzenszwalb (-d only; FindDeltas, :711-739 and 937-965), -w image size, -W volume size, -p random points
// (default seed 12345), -b block sides, -l maxLen, -c check.  Flows follow
// GenerateDistance (:832-883), GenerateDistanceFelzenszwalb (:741-782),
// GenerateSignedDistance (:1082-1124, 1197-1232) and
// GenerateSignedDistanceBorder (:1333-1365, 1471-1498).  Instead of PNG
// images the encoded bytes are written raw to -o.  -c compares distances
// with the exact EDT (max error < 1 cell, :265-311) and exits with -1 on
// failure, as the tool does.
#include "DistanceOcclusionLib.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace DOL;

namespace {

uint8_t EncodeU8Signed(float v)   // DistanceOcclusionTool.cpp:71-75
{
    v = (v < -1.0f) ? -1.0f : ((v > 1.0f) ? 1.0f : v);
    return (uint8_t)lrintf(128 + v * 127);
}

float Dist(int x, int y, int z = 0) { return sqrtf((float)(x * x + y * y + z * z)); }

// FindDeltas (DistanceOcclusionTool.cpp:711-739, 937-965) for the methods
// on the GPU path: 1 FastSweep, 2 JumpFlood
void FindDeltas(int method, int w, int h, cCellDelta2 delta[], int cw = 1)
{
    if (method == 1) FastSweep(w, h, delta, cw);
    else JumpFlood(w, h, delta, cw);
}
void FindDeltas(int method, int w, int h, int d, cCellDelta3 delta[], int cw = 1)
{
    if (method == 1) FastSweep(w, h, d, delta, cw);
    else JumpFlood(w, h, d, delta, cw);
}

bool write_raw(const char* path, const void* p, size_t n)
{
    if (!path) return true;
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    bool ok = std::fwrite(p, 1, n, f) == n;
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char** argv)
{
    int mode = 0, method = 2, dim = 0, sides = 0, points = -1;
    bool volume = false, check = false;
    uint32_t seed = 12345;
    float maxLen = 0.0f;
    const char* out = nullptr;
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        if (!std::strcmp(a, "-d")) mode = 'd';
        else if (!std::strcmp(a, "-s")) mode = 's';
        else if (!std::strcmp(a, "-S")) mode = 'S';
        else if (!std::strcmp(a, "-c")) check = true;
        else if (!std::strcmp(a, "-m") && i + 1 < argc) method = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "-w") && i + 1 < argc) dim = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "-W") && i + 1 < argc) { dim = std::atoi(argv[++i]); volume = true; }
        else if (!std::strcmp(a, "-b") && i + 1 < argc) sides = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "-l") && i + 1 < argc) maxLen = (float)std::atof(argv[++i]);
        else if (!std::strcmp(a, "-o") && i + 1 < argc) out = argv[++i];
        else if (!std::strcmp(a, "-p") && i + 1 < argc) {
            points = std::atoi(argv[++i]);
            if (i + 1 < argc && argv[i + 1][0] != '-') seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        } else {
            std::fprintf(stderr, "Unknown argument: %s\n", a);
            return -1;
        }
    }
    if (!mode) {
        std::fprintf(stderr, "Please specify -d, -s or -S\n");
        return -1;
    }
    const int w = dim ? dim : (volume ? 32 : 256), h = w, d = volume ? w : 1;
    if (points < 0) points = (sides == 0) ? (w + h + (volume ? d : 0)) / 8 : 0;   // :1900-1901
    if (maxLen == 0.0f) maxLen = 0.25f * w;                                        // :1973-1974

    uint32_t* mask = volume ? CreateBitMask(w, h, d) : CreateBitMask(w, h);
    if (volume) {
        if (sides) BitMaskAddBlock(w, h, d, sides, mask);
        if (points) BitMaskAddPoints(w, h, d, points, seed, mask);
    } else {
        if (sides) BitMaskAddBlock(w, h, sides, mask);
        if (points) BitMaskAddPoints(w, h, points, seed, mask);
    }
    const size_t n = (size_t)w * h * d;
    std::vector<float> dist(n);
    bool ok = true;

    if (mode == 'd' && method == 5 && !volume) {
        InitDistancesFromBitMask(w, h, mask, dist.data());
        Felzenszwalb(w, h, dist.data());
        for (float& v : dist) v = sqrtf(v);
    } else if (mode == 'd' && (method == 1 || method == 2)) {
        if (volume) {
            std::vector<cCellDelta3> delta(n);
            InitDeltasFromBitMask(w, h, d, mask, delta.data());
            FindDeltas(method, w, h, d, delta.data());
            for (size_t i = 0; i < n; i++) dist[i] = Dist(delta[i].x, delta[i].y, delta[i].z);
        } else {
            std::vector<cCellDelta2> delta(n);
            InitDeltasFromBitMask(w, h, mask, delta.data());
            FindDeltas(method, w, h, delta.data());
            for (size_t i = 0; i < n; i++) dist[i] = Dist(delta[i].x, delta[i].y);
        }
    } else if (mode == 's' && (method == 1 || method == 2)) {
        std::vector<uint8_t> sdf(n);
        const int ws = MaskSize(w);
        if (volume) {
            std::vector<cCellDelta3> delta(n);
            InitBorderDeltasFromBitMask(w, h, d, mask, delta.data());
            FindDeltas(method, w, h, d, delta.data(), 2);
            for (size_t i = 0; i < n; i++) {
                const size_t row = i / w;
                const int x = (int)(i - row * w);
                const bool inside = (mask[row * ws + (x >> 5)] >> (x & 31)) & 1;
                const float v = Dist(delta[i].x, delta[i].y, delta[i].z) / (maxLen * 2);
                sdf[i] = EncodeU8Signed(inside ? -v : v);
            }
        } else {
            std::vector<cCellDelta2> delta(n);
            InitBorderDeltasFromBitMask(w, h, mask, delta.data());
            FindDeltas(method, w, h, delta.data(), 2);
            for (size_t i = 0; i < n; i++) {
                const size_t row = i / w;
                const int x = (int)(i - row * w);
                const bool inside = (mask[row * ws + (x >> 5)] >> (x & 31)) & 1;
                const float v = Dist(delta[i].x, delta[i].y) / (maxLen * 2);
                sdf[i] = EncodeU8Signed(inside ? -v : v);
            }
        }
        if (!write_raw(out, sdf.data(), n)) ok = false;
        delete[] mask;
        return ok ? 0 : -1;
    } else if (mode == 'S' && (method == 1 || method == 2)) {
        // GenerateSignedDistance (:1082-1124, 1197-1232): exterior and
        // interior fields, then (Dist(d0) - Dist(d1)) / maxLen (:1035-1036)
        std::vector<uint8_t> sdf(n);
        if (volume) {
            std::vector<cCellDelta3> d0(n), d1(n);
            InitDeltasFromBitMask(w, h, d, mask, d0.data(), false);
            InitDeltasFromBitMask(w, h, d, mask, d1.data(), true);
            FindDeltas(method, w, h, d, d0.data());
            FindDeltas(method, w, h, d, d1.data());
            for (size_t i = 0; i < n; i++)
                sdf[i] = EncodeU8Signed((Dist(d0[i].x, d0[i].y, d0[i].z) - Dist(d1[i].x, d1[i].y, d1[i].z)) / maxLen);
        } else {
            std::vector<cCellDelta2> d0(n), d1(n);
            InitDeltasFromBitMask(w, h, mask, d0.data(), false);
            InitDeltasFromBitMask(w, h, mask, d1.data(), true);
            FindDeltas(method, w, h, d0.data());
            FindDeltas(method, w, h, d1.data());
            for (size_t i = 0; i < n; i++)
                sdf[i] = EncodeU8Signed((Dist(d0[i].x, d0[i].y) - Dist(d1[i].x, d1[i].y)) / maxLen);
        }
        if (!write_raw(out, sdf.data(), n)) ok = false;
        delete[] mask;
        return ok ? 0 : -1;
    } else {
        std::fprintf(stderr, "Unknown method for this technique\n");
        delete[] mask;
        return -1;
    }

    if (check) {
        std::vector<int32_t> d2(n);
        if (volume) EDT(w, h, d, mask, d2.data());
        else EDT(w, h, mask, d2.data());
        float maxErr = 0.0f;
        for (size_t i = 0; i < n; i++) {
            const float exact = d2[i] == INT32_MAX ? dist[i] : sqrtf((float)d2[i]);
            maxErr = fmaxf(maxErr, fabsf(dist[i] - exact));
        }
        std::fprintf(stderr, "Max error: %g\n", maxErr);
        if (maxErr >= 1.0f) {
            std::fprintf(stderr, "FAILED\n");
            ok = false;
        }
    }
    if (!write_raw(out, dist.data(), n * sizeof(float))) ok = false;
    delete[] mask;
    return ok ? 0 : -1;
}
