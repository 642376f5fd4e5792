// dol_shim.cpp -- the DOL:: drop-in interface (include/DistanceOcclusionLib.h)
// over the C ABI (include/dol_hip.h).
//
// Each function keeps the reference's signature, its in-place host-memory
// contract and its synchronous behaviour (DistanceOcclusionLib.h:21-117).  The
// reference reports nothing, so a failure here -- a bad argument the reference
// would have run into undefined behaviour with, or no usable device -- prints
// the library's message and aborts.
#include "../../include/DistanceOcclusionLib.h"
#include "../../include/dol_hip.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

static_assert(sizeof(DOL::cCellDelta2) == 4, "cCellDelta2 must be 2 x int16");
static_assert(sizeof(DOL::cCellDelta3) == 6, "cCellDelta3 must be 3 x int16, packed");

namespace {

void require(int rc, const char* what)
{
    if (rc != DOLHIP_OK) {
        std::fprintf(stderr, "DOL::%s failed: %s\n", what, dolhip_last_error());
        std::abort();
    }
}

int16_t* raw(DOL::cCellDelta2* p) { return reinterpret_cast<int16_t*>(p); }
int16_t* raw(DOL::cCellDelta3* p) { return reinterpret_cast<int16_t*>(p); }

}  // namespace

// DistanceOcclusionLib.cpp:79-93, 1214-1222: new[]-allocated, zeroed, so
// callers may delete[] them directly
uint32_t* DOL::CreateBitMask(int w, int h)
{
    size_t n = (size_t)MaskSize(w, h);
    uint32_t* m = new uint32_t[n];
    std::memset(m, 0, n * sizeof(uint32_t));
    return m;
}

uint32_t* DOL::CreateBitMask(int w, int h, int d)
{
    size_t n = (size_t)MaskSize(w, h, d);
    uint32_t* m = new uint32_t[n];
    std::memset(m, 0, n * sizeof(uint32_t));
    return m;
}

void DOL::DestroyBitMask(uint32_t*& mask)
{
    delete[] mask;
    mask = nullptr;
}

void DOL::BitMaskAddPoints(int w, int h, int n, uint32_t seed, uint32_t mask[])
{ dolhip_mask_add_points_2d(w, h, n, seed, mask); }
void DOL::BitMaskAddPoints(int w, int h, int d, int n, uint32_t seed, uint32_t mask[])
{ dolhip_mask_add_points_3d(w, h, d, n, seed, mask); }
void DOL::BitMaskAddBlock(int w, int h, int sides, uint32_t mask[])
{ dolhip_mask_add_block_2d(w, h, sides, mask); }
void DOL::BitMaskAddBlock(int w, int h, int d, int sides, uint32_t mask[])
{ dolhip_mask_add_block_3d(w, h, d, sides, mask); }

void DOL::InitDistancesFromBitMask(int w, int h, const uint32_t mask[], int16_t distances[], int16_t maxD)
{ require(dolhip_init_distances_i16(w, h, mask, distances, maxD), "InitDistancesFromBitMask"); }
void DOL::InitDistancesFromBitMask(int w, int h, const uint32_t mask[], int32_t distances[], int32_t maxD)
{ require(dolhip_init_distances_i32(w, h, mask, distances, maxD), "InitDistancesFromBitMask"); }
void DOL::InitDistancesFromBitMask(int w, int h, const uint32_t mask[], float distances[], float maxD)
{ require(dolhip_init_distances_f32(w, h, mask, distances, maxD), "InitDistancesFromBitMask"); }

void DOL::Felzenszwalb(int w, int h, float distances[])
{ require(dolhip_felzenszwalb_2d(w, h, distances), "Felzenszwalb"); }

void DOL::InitDeltasFromBitMask(int w, int h, const uint32_t mask[], cCellDelta2 delta[], bool invert)
{ require(dolhip_init_deltas_2d(w, h, mask, raw(delta), invert), "InitDeltasFromBitMask"); }
void DOL::InitDeltasFromBitMask(int w, int h, int d, const uint32_t mask[], cCellDelta3 delta[], bool invert)
{ require(dolhip_init_deltas_3d(w, h, d, mask, raw(delta), invert), "InitDeltasFromBitMask"); }

void DOL::InitBorderDeltasFromBitMask(int w, int h, const uint32_t mask[], cCellDelta2 delta[])
{ require(dolhip_init_border_deltas_2d(w, h, mask, raw(delta)), "InitBorderDeltasFromBitMask"); }
void DOL::InitBorderDeltasFromBitMask(int w, int h, int d, const uint32_t mask[], cCellDelta3 delta[])
{ require(dolhip_init_border_deltas_3d(w, h, d, mask, raw(delta)), "InitBorderDeltasFromBitMask"); }

void DOL::JumpFlood(int w, int h, cCellDelta2 delta[], int cw)
{ require(dolhip_jump_flood_2d(w, h, raw(delta), cw), "JumpFlood"); }
void DOL::JumpFlood(int w, int h, int d, cCellDelta3 delta[], int cw)
{ require(dolhip_jump_flood_3d(w, h, d, raw(delta), cw), "JumpFlood"); }
void DOL::FastSweep(int w, int h, cCellDelta2 delta[], int cw)
{ require(dolhip_fast_sweep_2d(w, h, raw(delta), cw), "FastSweep"); }
void DOL::FastSweep(int w, int h, int d, cCellDelta3 delta[], int cw)
{ require(dolhip_fast_sweep_3d(w, h, d, raw(delta), cw), "FastSweep"); }
void DOL::Danielsson(int w, int h, cCellDelta2 delta[], int cw)
{ require(dolhip_danielsson_2d(w, h, raw(delta), cw), "Danielsson"); }
void DOL::Danielsson(int w, int h, int d, cCellDelta3 delta[], int cw)
{ require(dolhip_danielsson_3d(w, h, d, raw(delta), cw), "Danielsson"); }
void DOL::BruteForce(int w, int h, cCellDelta2 delta[], int cw)
{ require(dolhip_brute_force_2d(w, h, raw(delta), cw), "BruteForce"); }
void DOL::BruteForce(int w, int h, int d, cCellDelta3 delta[], int cw)
{ require(dolhip_brute_force_3d(w, h, d, raw(delta), cw), "BruteForce"); }

int DOL::Chamfer(int w, int h, int16_t distances[])
{
    int m = 0;
    require(dolhip_chamfer_i16(w, h, distances, &m), "Chamfer");
    return m;
}
int DOL::Chamfer(int w, int h, int32_t distances[])
{
    int m = 0;
    require(dolhip_chamfer_i32(w, h, distances, &m), "Chamfer");
    return m;
}

void DOL::InitOccFromBitMask(int w, int h, const uint32_t mask[], float occlusion[], float scale)
{ require(dolhip_init_occ(w, h, mask, occlusion, scale), "InitOccFromBitMask"); }
void DOL::ShadowSweep(float gx, float gy, int w, int h, const float* occlusion, float* shadow)
{ require(dolhip_shadow_sweep(gx, gy, w, h, occlusion, shadow), "ShadowSweep"); }

void DOL::InitDirWFromBitMask(int w, int h, const uint32_t mask[], float dirW4[])
{ require(dolhip_init_dirw_2d(w, h, mask, dirW4), "InitDirWFromBitMask"); }
void DOL::InitDirWFromBitMask(int w, int h, int d, const uint32_t mask[], float dirW8[])
{ require(dolhip_init_dirw_3d(w, h, d, mask, dirW8), "InitDirWFromBitMask"); }
void DOL::OcclusionSweep(int w, int h, float dirW4[])
{ require(dolhip_occlusion_sweep_2d(w, h, dirW4), "OcclusionSweep"); }
void DOL::OcclusionSweep(int w, int h, int d, float dirW8[])
{ require(dolhip_occlusion_sweep_3d(w, h, d, dirW8), "OcclusionSweep"); }
void DOL::OcclusionSweep(int w, int h, int d, const float inDirW8[], float outDirW8[])
{ require(dolhip_occlusion_sweep_3d_from(w, h, d, inDirW8, outDirW8), "OcclusionSweep"); }

void DOL::EDT(int w, int h, const uint32_t mask[], int32_t d2[], float dist[], bool invert)
{ require(dolhip_edt_2d(w, h, mask, invert, d2, dist), "EDT"); }
void DOL::EDT(int w, int h, int d, const uint32_t mask[], int32_t d2[], float dist[], bool invert)
{ require(dolhip_edt_3d(w, h, d, mask, invert, d2, dist), "EDT"); }

void DOL::SignedEDT(int w, int h, const uint32_t mask[], float sdf[])
{ require(dolhip_signed_edt_2d(w, h, mask, sdf), "SignedEDT"); }
void DOL::SignedEDT(int w, int h, int d, const uint32_t mask[], float sdf[])
{ require(dolhip_signed_edt_3d(w, h, d, mask, sdf), "SignedEDT"); }
