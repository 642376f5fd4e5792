// ref_mesh_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI face over the reference's mesh support, MeshSupport.cpp, compiled
// from its own sources under /root/reference by oracle/Makefile into
// oracle/_ref/libmeshref.so.  As in the reference's tool
// (DistanceOcclusionTool.cpp:33-35), the source is included into a unit that
// has the fixed-width integer types already declared.  Tests use it to pin the oracle's restatement of
// the voxeliser (MeshSupport.cpp:369-475) bit for bit and to read the
// reference's sample mesh.  Nothing here is product code.
#include <cstdint>
#include <cstdio>

#include "MeshSupport.cpp"

using namespace MSL;

extern "C" {

// ReadObjFile (MeshSupport.cpp:795-831).  Returns the position and
// position-index counts; copies them out when the arrays are given.
int ref_read_obj(const char* path, int* n_positions, int* n_indices, float* positions, int* indices)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return 0;
    cMesh mesh;
    const bool ok = ReadObjFile(f, &mesh);
    std::fclose(f);
    if (!ok) return 0;
    *n_positions = (int)mesh.mPositions.size();
    *n_indices = (int)mesh.mPositionIndices.size();
    if (positions)
        for (size_t i = 0; i < mesh.mPositions.size(); i++) {
            positions[3 * i] = mesh.mPositions[i].x;
            positions[3 * i + 1] = mesh.mPositions[i].y;
            positions[3 * i + 2] = mesh.mPositions[i].z;
        }
    if (indices)
        for (size_t i = 0; i < mesh.mPositionIndices.size(); i++) indices[i] = mesh.mPositionIndices[i];
    return 1;
}

// FindBounds (MeshSupport.cpp:833-860) of a position array: {min xyz, max xyz}
void ref_find_bounds(int n_positions, const float* positions, float* bbox)
{
    cMesh mesh;
    for (int i = 0; i < n_positions; i++)
        mesh.mPositions.push_back(Vec3f{positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]});
    const Bounds3f b = FindBounds(mesh);
    bbox[0] = b.mMin.x; bbox[1] = b.mMin.y; bbox[2] = b.mMin.z;
    bbox[3] = b.mMax.x; bbox[4] = b.mMax.y; bbox[5] = b.mMax.z;
}

// CreateBitMaskFromTriangles (MeshSupport.cpp:369-475); ORs into mask
void ref_mask_from_triangles(int tri_count, const int* indices, const float* vertices,
                             const float* bbox, int w, int h, int d, uint32_t* mask)
{
    const Bounds3f b = {{bbox[0], bbox[1], bbox[2]}, {bbox[3], bbox[4], bbox[5]}};
    MSL::CreateBitMaskFromTriangles(tri_count, indices, reinterpret_cast<const Vec3f*>(vertices), b, w, h, d, mask);
}

}
