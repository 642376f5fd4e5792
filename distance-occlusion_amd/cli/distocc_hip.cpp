// distocc_hip -- the distance modes of the reference tool (distocc) on the
// GPU library: the algorithms through the DOL:: drop-in interface, masks from
// bytes and every output image through the C ABI (the images are made on the
// device, not by host loops).
//
//   distocc_hip (-d | -s | -S) [-m method] [-w w [h]] [-W w [h [d]]]
//               [-p [count [seed]]] [-b [sides]] [-f file] [-t threshold] [-r]
//               [-l maxLen] [-i type] [-o file] [-v] [-c]
//
// Flags as the reference's (DistanceOcclusionTool.cpp:1782-1883): -d
// distance, -s signed distance with border initialisation, -S signed distance
// from exterior and interior fields; -m 0 Danielsson, 1 FastSweep, 2
// JumpFlood (default), 3 BruteForce, 4 Chamfer and 5 Felzenszwalb (-d, 2D);
// -w / -W image or volume sizes (h defaults to w, d to h), -p random points
// (default count (w + h + d) / 8, seed 12345), -b block sides (default 2),
// -l maxLen (default 0.25 max(w, h); 127 with -i 1), -i image type (0 sdf,
// 1 sdf-delta, 2 sdf-rg, 3 sdf-angles).  -f reads a mask as raw bytes (w x h,
// or w x h x d), occupied where (byte >= threshold) XOR lessThan, the tool's
// BitMaskFromImageFile test (:646-665) with threshold 128 and dark occupied
// by default, -r flipping it.  The image the tool would write as PNG
// (:678-698, 1029-1080, 1240-1331, 1367-1469) is written raw to -o; -v logs
// each stage and its time; -c runs the tool's checks and exits with -1 on a
// failure: CheckBorderDeltas of the border init (with its x / y loop swap,
// :474-475, fixed) and CheckDeltas of the result (:371-439) for -s, and the
// largest error against the exact EDT (< 1 cell, :265-311) for -d and -S.
#include "DistanceOcclusionLib.h"
#include "dol_hip.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace DOL;

namespace {

struct Log {
    bool on = false;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void stage(const char* what)
    {
        const auto now = std::chrono::steady_clock::now();
        if (on)
            std::printf("%-40s %9.3f ms\n", what, std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

void require(int rc, const char* what)
{
    if (rc != DOLHIP_OK) {
        std::fprintf(stderr, "%s failed: %s\n", what, dolhip_last_error());
        std::exit(-1);
    }
}

// FindDeltas (DistanceOcclusionTool.cpp:711-739, 937-965)
template <class Cell>
bool find_deltas(int method, int w, int h, int d, Cell* delta, int cw)
{
    const bool vol = d > 0;
    switch (method) {
    case 0: vol ? Danielsson(w, h, d, (cCellDelta3*)delta, cw) : Danielsson(w, h, (cCellDelta2*)delta, cw); break;
    case 1: vol ? FastSweep(w, h, d, (cCellDelta3*)delta, cw) : FastSweep(w, h, (cCellDelta2*)delta, cw); break;
    case 2: vol ? JumpFlood(w, h, d, (cCellDelta3*)delta, cw) : JumpFlood(w, h, (cCellDelta2*)delta, cw); break;
    case 3: vol ? BruteForce(w, h, d, (cCellDelta3*)delta, cw) : BruteForce(w, h, (cCellDelta2*)delta, cw); break;
    default: return false;
    }
    return true;
}

int bit(const uint32_t* mask, int w, long long row, int x)
{
    return (mask[row * MaskSize(w) + (x >> 5)] >> (x & 31)) & 1;
}

// CheckBorderDeltas (DistanceOcclusionTool.cpp:441-599) with x over the width
// and y over the height.  2D reads clamp to the image; 3D neighbours outside
// the volume pass.  Returns the number of failed expectations.
long check_border(int w, int h, int d, const uint32_t* mask, const int16_t* delta)
{
    long bad = 0;
    if (d == 0) {
        auto occ = [&](int x, int y) {
            x = x < 0 ? 0 : (x >= w ? w - 1 : x);
            y = y < 0 ? 0 : (y >= h ? h - 1 : y);
            return bit(mask, w, y, x);
        };
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const int16_t* c = delta + 2 * ((size_t)y * w + x);
                const int o = occ(x, y);
                if (std::abs(c[0]) < 2 && std::abs(c[1]) < 2) {
                    bad += occ(x + c[0], y + c[1]) == o;
                    if (c[0] && c[1])
                        bad += (occ(x - 1, y) != o) + (occ(x + 1, y) != o) + (occ(x, y - 1) != o) + (occ(x, y + 1) != o);
                } else {
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) bad += occ(x + dx, y + dy) != o;
                }
            }
        return bad;
    }
    auto same = [&](int x, int y, int z, int v) {
        if (x < 0 || x >= w || y < 0 || y >= h || z < 0 || z >= d) return true;
        return bit(mask, w, (long long)z * h + y, x) == v;
    };
    for (int z = 0; z < d; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const int16_t* c = delta + 3 * (((size_t)z * h + y) * w + x);
                const int o = bit(mask, w, (long long)z * h + y, x);
                if (std::abs(c[0]) < 2 && std::abs(c[1]) < 2 && std::abs(c[2]) < 2) {
                    bad += !same(x + c[0], y + c[1], z + c[2], !o);
                    // the neighbours that must not hold a closer edge: the face
                    // diagonals of a box diagonal, the axis neighbours in the
                    // plane of each diagonal pair of components
                    if (c[0] && c[1] && c[2])
                        for (int a = -1; a <= 1; a += 2)
                            for (int b = -1; b <= 1; b += 2)
                                bad += !same(x + a, y + b, z, o) + !same(x + a, y, z + b, o) + !same(x, y + a, z + b, o);
                    auto axes = [&](int ax, int ay, int az, int bx, int by, int bz) {
                        bad += !same(x - ax, y - ay, z - az, o) + !same(x + ax, y + ay, z + az, o);
                        bad += !same(x - bx, y - by, z - bz, o) + !same(x + bx, y + by, z + bz, o);
                    };
                    if (c[0] && c[1]) axes(1, 0, 0, 0, 1, 0);
                    if (c[1] && c[2]) axes(0, 1, 0, 0, 0, 1);
                    if (c[2] && c[0]) axes(1, 0, 0, 0, 0, 1);
                } else {
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) bad += !same(x + dx, y + dy, z + dz, o);
                }
            }
    return bad;
}

// CheckDeltas (DistanceOcclusionTool.cpp:371-439): every delta lands inside the
// field on a boundary cell
long check_deltas(int w, int h, int d, const int16_t* delta, int cw)
{
    const int nc = d ? 3 : 2, dd = d ? d : 1;
    long bad = 0;
    for (int k = 0; k < dd; k++)
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                const int16_t* c = delta + nc * (((size_t)k * h + j) * w + i);
                const int cx = (int16_t)(c[0] + i * cw), cy = (int16_t)(c[1] + j * cw);
                const int cz = d ? (int16_t)(c[2] + k * cw) : 0;
                if (cx < 0 || cx >= cw * w || cy < 0 || cy >= cw * h || (d && (cz < 0 || cz >= cw * d))) {
                    bad++;
                    continue;
                }
                const int16_t* r = delta + nc * (((size_t)(cz / cw) * h + cy / cw) * w + cx / cw);
                bool boundary = std::abs(r[0]) < cw && std::abs(r[1]) < cw && (!d || std::abs(r[2]) < cw);
                bad += !boundary;
            }
    return bad;
}

bool write_raw(const char* path, const void* p, size_t n)
{
    if (!path) return true;
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, n, f) == n;
    std::fclose(f);
    return ok;
}

bool next_is_arg(int i, int argc, char** argv) { return i + 1 < argc && argv[i + 1][0] != '-'; }

}  // namespace

int main(int argc, char** argv)
{
    int mode = 0, method = 2, dim[3] = { 0, 0, 0 }, sides = 0, points = 0, itype = 0, threshold = 128;
    bool volume = false, check = false, less_than = true, have_points = false;
    uint32_t seed = 12345;
    float maxLen = 0.0f;
    const char* out = nullptr;
    const char* src = nullptr;
    Log log;
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        if (a[0] != '-' || !a[1] || a[2]) {
            std::fprintf(stderr, "Unknown argument: %s\n", a);
            return -1;
        }
        switch (a[1]) {
        case 'd': case 's': case 'S': mode = a[1]; break;
        case 'c': check = true; break;
        case 'v': log.on = true; break;
        case 'r': less_than = !less_than; break;
        case 'm': if (next_is_arg(i, argc, argv)) method = std::atoi(argv[++i]); break;
        case 'i':
            if (next_is_arg(i, argc, argv)) {
                itype = std::atoi(argv[++i]);
                if (maxLen == 0.0f && itype == 1) maxLen = 127;   // :1821-1822
            }
            break;
        case 't': if (next_is_arg(i, argc, argv)) threshold = std::atoi(argv[++i]); break;
        case 'l': if (next_is_arg(i, argc, argv)) maxLen = (float)std::atof(argv[++i]); break;
        case 'o': if (next_is_arg(i, argc, argv)) out = argv[++i]; break;
        case 'f': if (next_is_arg(i, argc, argv)) src = argv[++i]; break;
        case 'b': sides = next_is_arg(i, argc, argv) ? std::atoi(argv[++i]) : 2; break;
        case 'w': case 'W':
            for (int k = 0; k < 3 && next_is_arg(i, argc, argv); k++) dim[k] = std::atoi(argv[++i]);
            volume = a[1] == 'W';
            break;
        case 'p':
            have_points = true;
            points = next_is_arg(i, argc, argv) ? std::atoi(argv[++i]) : -1;
            if (next_is_arg(i, argc, argv)) seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
            break;
        default: std::fprintf(stderr, "Unknown option: %s\n", a); return -1;
        }
    }
    if (!mode) {
        std::fprintf(stderr, "Please specify -d, -s or -S\n");
        return -1;
    }
    // sizes and defaults as the tool's (:1886-1901, 1973-1974)
    if (!dim[0]) dim[0] = volume ? 32 : 256;
    if (!dim[1]) dim[1] = dim[0];
    if (volume && !dim[2]) dim[2] = dim[1];
    const int w = dim[0], h = dim[1], d = volume ? dim[2] : 0, dd = volume ? d : 1;
    if (have_points && points < 0) points = (w + h + d) / 8;
    if (!src && !sides && !have_points) points = (w + h + d) / 8;
    if (maxLen == 0.0f) maxLen = 0.25f * (w >= h ? w : h);
    const size_t n = (size_t)w * h * dd;

    uint32_t* mask = volume ? CreateBitMask(w, h, d) : CreateBitMask(w, h);
    if (src) {
        std::vector<uint8_t> bytes(n);
        FILE* f = std::fopen(src, "rb");
        if (!f || std::fread(bytes.data(), 1, n, f) != n) {
            std::fprintf(stderr, "cannot read %zu bytes from %s\n", n, src);
            if (f) std::fclose(f);
            return -1;
        }
        std::fclose(f);
        require(dolhip_mask_from_bytes(w, h * dd, bytes.data(), threshold, less_than, mask), "mask from bytes");
    }
    if (volume) {
        if (sides) BitMaskAddBlock(w, h, d, sides, mask);
        if (points) BitMaskAddPoints(w, h, d, points, seed, mask);
    } else {
        if (sides) BitMaskAddBlock(w, h, sides, mask);
        if (points) BitMaskAddPoints(w, h, points, seed, mask);
    }
    log.stage("mask");

    bool ok = true;
    std::vector<float> dist;          // -d / -S: distances for the -c error check
    std::vector<uint8_t> image;       // what the tool writes
    if (mode == 'd' && !volume && (method == 4 || method == 5)) {
        dist.resize(n);
        image.resize(n);
        if (method == 5) {   // GenerateDistanceFelzenszwalb, :741-782
            InitDistancesFromBitMask(w, h, mask, dist.data());
            Felzenszwalb(w, h, dist.data());
            log.stage("Felzenszwalb");
            require(dolhip_sqrt_f32(w, h, dist.data(), dist.data()), "sqrt");
            require(dolhip_encode_scalar_u8_f32(w, h, dist.data(), maxLen, image.data()), "encode");
        } else {             // GenerateDistanceChamfer, :784-830
            std::vector<int16_t> cd(n);
            InitDistancesFromBitMask(w, h, mask, cd.data(), (int16_t)(INT16_MAX - 16));
            const int cm = Chamfer(w, h, cd.data());
            log.stage("Chamfer");
            for (size_t i = 0; i < n; i++) dist[i] = (float)cd[i] / cm;
            require(dolhip_encode_scalar_u8_i16(w, h, cd.data(), maxLen * cm, image.data()), "encode");
        }
    } else if (mode == 'd' && method <= 3) {   // GenerateDistance, :832-883, 967-1014
        dist.resize(n);
        image.resize(n);
        std::vector<int16_t> delta(n * (volume ? 3 : 2));
        if (volume) InitDeltasFromBitMask(w, h, d, mask, (cCellDelta3*)delta.data());
        else InitDeltasFromBitMask(w, h, mask, (cCellDelta2*)delta.data());
        log.stage("init");
        find_deltas(method, w, h, d, delta.data(), 1);
        log.stage("propagate");
        for (size_t i = 0; i < n; i++) {
            const int16_t* c = delta.data() + (volume ? 3 : 2) * i;
            const int z = volume ? c[2] : 0;
            dist[i] = sqrtf((float)(c[0] * c[0] + c[1] * c[1] + z * z));
        }
        if (volume) require(dolhip_encode_distance_u8_3d(w, h, d, delta.data(), maxLen, image.data()), "encode");
        else require(dolhip_encode_distance_u8_2d(w, h, delta.data(), maxLen, image.data()), "encode");
    } else if (mode == 's' && method <= 3) {   // GenerateSignedDistanceBorder, :1333-1365, 1471-1498
        std::vector<int16_t> delta(n * (volume ? 3 : 2));
        if (volume) InitBorderDeltasFromBitMask(w, h, d, mask, (cCellDelta3*)delta.data());
        else InitBorderDeltasFromBitMask(w, h, mask, (cCellDelta2*)delta.data());
        log.stage("border init");
        if (check) {
            const long bad = check_border(w, h, d, mask, delta.data());
            if (bad) {
                std::fprintf(stderr, "CheckBorderDeltas: %ld failed expectations\n", bad);
                ok = false;
            }
        }
        find_deltas(method, w, h, d, delta.data(), 2);
        log.stage("propagate");
        if (check) {
            const long bad = check_deltas(w, h, d, delta.data(), 2);
            if (bad) {
                std::fprintf(stderr, "CheckDeltas: %ld cells\n", bad);
                ok = false;
            }
        }
        image.resize(n * (itype ? 4 : 1));
        if (volume)
            require(dolhip_encode_border_image_3d(w, h, d, mask, delta.data(), maxLen * 2, itype, image.data()), "encode");
        else
            require(dolhip_encode_border_image_2d(w, h, mask, delta.data(), maxLen * 2, itype, image.data()), "encode");
    } else if (mode == 'S' && method <= 3) {   // GenerateSignedDistance, :1082-1124, 1197-1232
        const int nc = volume ? 3 : 2;
        std::vector<int16_t> d0(n * nc), d1(n * nc);
        if (volume) {
            InitDeltasFromBitMask(w, h, d, mask, (cCellDelta3*)d0.data(), false);
            InitDeltasFromBitMask(w, h, d, mask, (cCellDelta3*)d1.data(), true);
        } else {
            InitDeltasFromBitMask(w, h, mask, (cCellDelta2*)d0.data(), false);
            InitDeltasFromBitMask(w, h, mask, (cCellDelta2*)d1.data(), true);
        }
        find_deltas(method, w, h, d, d0.data(), 1);
        find_deltas(method, w, h, d, d1.data(), 1);
        log.stage("two fields");
        dist.resize(n);
        for (size_t i = 0; i < n; i++) {
            const int16_t* a = d0.data() + nc * i;
            const int16_t* b = d1.data() + nc * i;
            const int za = volume ? a[2] : 0, zb = volume ? b[2] : 0;
            dist[i] = sqrtf((float)(a[0] * a[0] + a[1] * a[1] + za * za)) -
                      sqrtf((float)(b[0] * b[0] + b[1] * b[1] + zb * zb));
        }
        image.resize(n * (itype ? 4 : 1));
        if (volume)
            require(dolhip_encode_two_field_3d(w, h, d, d0.data(), d1.data(), maxLen, itype, image.data()), "encode");
        else
            require(dolhip_encode_two_field_2d(w, h, d0.data(), d1.data(), maxLen, itype, image.data()), "encode");
    } else {
        std::fprintf(stderr, "Unknown method for this technique\n");
        delete[] mask;
        return -1;
    }
    log.stage("image");

    if (check && mode != 's') {
        // against the exact EDT to the set cells (and, for -S, to the clear ones)
        std::vector<int32_t> e0(n), e1;
        if (volume) EDT(w, h, d, mask, e0.data());
        else EDT(w, h, mask, e0.data());
        if (mode == 'S') {
            e1.resize(n);
            if (volume) EDT(w, h, d, mask, e1.data(), nullptr, true);
            else EDT(w, h, mask, e1.data(), nullptr, true);
        }
        auto root = [](int32_t v) { return v == INT32_MAX ? 0.0f : sqrtf((float)v); };
        float maxErr = 0.0f;
        for (size_t i = 0; i < n; i++) {
            if (mode == 'd' && e0[i] == INT32_MAX) continue;
            const float exact = mode == 'S' ? root(e0[i]) - root(e1[i]) : root(e0[i]);
            maxErr = fmaxf(maxErr, fabsf(dist[i] - exact));
        }
        std::fprintf(stderr, "Max error: %g\n", maxErr);
        if (!(maxErr < 1.0f)) {
            std::fprintf(stderr, "FAILED\n");
            ok = false;
        }
        log.stage("check");
    }
    if (!write_raw(out, image.data(), image.size())) ok = false;
    delete[] mask;
    return ok ? 0 : -1;
}
