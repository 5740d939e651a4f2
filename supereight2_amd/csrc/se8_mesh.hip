// Single-resolution marching cubes over the block pool and the mesh file writers.
//
// se::Map::mesh (include/se/map/impl/map_impl.hpp:452-458) -> se::Octree::mesh
// (include/se/map/octree/impl/octree_impl.hpp:121-133) -> algorithms::marching_cube, single-res
// branch (include/se/map/algorithms/impl/marching_cube_impl.hpp:832-889, 1175-1203).
//
// The reference walks every allocated block on the host (OpenMP over blocks), and for each of the
// cells x,y,z in [block, min(block + 8, size - 1)) gathers the 8 corner voxels (through the octree
// at block borders), takes the case index of the `occupied` isosurface (0 if any corner has weight
// 0, else bit k = corner k has tsdf < 0), and appends the case's triangles, each vertex linearly
// interpolated on its cell edge; a face with a vertex outside (0, size] is dropped (checkVertex).
//
// Here one 256-thread workgroup meshes one block: it stages the block's 8^3 voxels and the 217
// voxels of the 7 neighbour blocks that the +1 corners reach (a 9^3 corner grid) into LDS, then
// every thread meshes two cells from LDS.  Two launches: the count pass writes each block's face
// count, one workgroup scans the counts into offsets, and the emit pass recomputes the cells and
// writes every face at its block's offset plus its rank inside the block.  The output order is
// therefore fixed for a given pool (block slot, then cell index x + 8y + 64z, then the case's
// triangle order); the reference's order depends on its OpenMP schedule, so faces are compared as
// a set.
#include "se8_internal.hpp"
#include "se8_mc_table.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

using namespace se8;

namespace {

constexpr int kMeshThreads = 256;
constexpr int kGrid = 9 * 9 * 9; // corner grid of one block: 8^3 voxels + the +1 faces
constexpr int kHalo = kGrid - 512;

// LDS index of corner k of the cell whose lowest corner is at grid index g
__device__ __forceinline__ int corner_at(int g, int k)
{
    return g + ((SE8_MC_CORNER_DX >> k) & 1) + 9 * ((SE8_MC_CORNER_DY >> k) & 1) + 81 * ((SE8_MC_CORNER_DZ >> k) & 1);
}

__device__ __forceinline__ int corner_off(int k, int axis)
{
    const unsigned mask = axis == 0 ? SE8_MC_CORNER_DX : axis == 1 ? SE8_MC_CORNER_DY : SE8_MC_CORNER_DZ;
    return (mask >> k) & 1;
}

// Vertex on edge e of the cell with lowest voxel c (map voxel units), from the tsdf values at its
// two corners (marching_cube_impl.hpp:83-95, compute_intersection): the sample points are the voxel
// centres c + 0.5, and the vertex is p0 + ((0 - v0) / (v1 - v0)) * (p1 - p0), per component.
__device__ __forceinline__ void edge_vertex(const float* s_tsdf, int g, const int c[3], int e, float p[3])
{
    const int a = (int) ((SE8_MC_EDGE_A >> (4 * e)) & 0xF);
    const int b = (int) ((SE8_MC_EDGE_B >> (4 * e)) & 0xF);
    const float v0 = s_tsdf[corner_at(g, a)];
    const float v1 = s_tsdf[corner_at(g, b)];
    const float t = (0.f - v0) / (v1 - v0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float p0 = (float) (c[i] + corner_off(a, i)) + 0.5f;
        const float p1 = (float) (c[i] + corner_off(b, i)) + 0.5f;
        p[i] = p0 + t * (p1 - p0);
    }
}

// checkVertex (marching_cube_impl.hpp:814-818): true if the vertex lies outside (0, dim]^3.  A NaN
// coordinate compares false everywhere and keeps the face, as in the reference.
__device__ __forceinline__ bool vertex_outside(const float p[3], float dim)
{
    return p[0] <= 0.f || p[1] <= 0.f || p[2] <= 0.f || p[0] > dim || p[1] > dim || p[2] > dim;
}

// Exclusive prefix of v over the workgroup and the workgroup total (64-wide waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d)
            inc += o;
    }
    if (lane == 63)
        s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kMeshThreads / 64; w++) {
        const uint32_t s = s_wave[w];
        before += w < wave ? s : 0u;
        total += s;
    }
    return before + inc - v;
}

// Count pass (Emit = false): block_faces[slot] = number of faces of the block.
// Emit pass (Emit = true): block_faces[slot] = offset of the block's first face; its faces are
// written to faces (9 floats each: three vertices in the output frame) in the fixed order.
template <bool Emit>
__global__ __launch_bounds__(kMeshThreads) void k_mesh(MapView m, Iso T_OV, uint64_t* block_faces, float* faces)
{
    __shared__ float s_tsdf[kGrid];
    __shared__ uint8_t s_valid[kGrid];
    __shared__ uint64_t s_tri[256];
    __shared__ int s_nb[8];
    __shared__ uint32_t s_wave[kMeshThreads / 64];

    const int slot = blockIdx.x;
    const int t = threadIdx.x;
    s_tri[t] = se8_mc_tri[t];
    int bc[3];
    key_to_coord(m.block_key[slot], bc);
    if (t < 8) {
        // neighbour t = (dx, dy, dz) bits; none beyond the map's upper face (no wrap-around): the
        // cells that would read it are outside the meshed range anyway
        const int nx = bc[0] + 8 * (t & 1), ny = bc[1] + 8 * ((t >> 1) & 1), nz = bc[2] + 8 * ((t >> 2) & 1);
        int nb = slot;
        if (t != 0)
            nb = (nx < m.size && ny < m.size && nz < m.size) ? lookup_block_nb(m, nx, ny, nz) : -1;
        s_nb[t] = nb;
    }
#pragma unroll
    for (int i = t; i < 512; i += kMeshThreads) {
        float ts, w;
        load_voxel(m, slot, i, ts, w);
        const int g = (i & 7) + 9 * ((i >> 3) & 7) + 81 * (i >> 6);
        s_tsdf[g] = ts;
        s_valid[g] = w != 0.f;
    }
    __syncthreads();
    if (t < kHalo) {
        // halo point t -> grid (x, y, z) with at least one coordinate 8: the x = 8 face, the y = 8
        // face, the z = 8 face (64 each), the three edges (8 each) and the far corner
        int x, y, z;
        if (t < 64) {
            x = 8, y = t & 7, z = t >> 3;
        }
        else if (t < 128) {
            x = (t - 64) & 7, y = 8, z = (t - 64) >> 3;
        }
        else if (t < 192) {
            x = (t - 128) & 7, y = (t - 128) >> 3, z = 8;
        }
        else if (t < 200) {
            x = 8, y = 8, z = t - 192;
        }
        else if (t < 208) {
            x = 8, y = t - 200, z = 8;
        }
        else if (t < 216) {
            x = t - 208, y = 8, z = 8;
        }
        else {
            x = 8, y = 8, z = 8;
        }
        const int nb = s_nb[(x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2)];
        float ts = 1.f, w = 0.f; // getData of an unallocated block: the default {1, 0}
        if (nb >= 0)
            load_voxel(m, nb, (x & 7) + 8 * (y & 7) + 64 * (z & 7), ts, w);
        const int g = x + 9 * y + 81 * z;
        s_tsdf[g] = ts;
        s_valid[g] = w != 0.f;
    }
    __syncthreads();

    const float dim = (float) m.size;
    uint32_t n_faces = 0;
    uint32_t keep[2];
    uint64_t word[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int ci = t + h * kMeshThreads;
        const int lc[3] = {ci & 7, (ci >> 3) & 7, ci >> 6};
        const int c[3] = {bc[0] + lc[0], bc[1] + lc[1], bc[2] + lc[2]};
        const int g = lc[0] + 9 * lc[1] + 81 * lc[2];
        uint32_t idx = 0;
        bool all_valid = true;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int gk = corner_at(g, k);
            all_valid = all_valid && s_valid[gk];
            idx |= (s_tsdf[gk] < 0.f ? 1u : 0u) << k;
        }
        const bool in_range = c[0] < m.size - 1 && c[1] < m.size - 1 && c[2] < m.size - 1;
        word[h] = s_tri[(all_valid && in_range) ? idx : 0u];
        keep[h] = 0;
        const int ntri = (int) (word[h] >> 60);
        for (int j = 0; j < ntri; j++) {
            bool out = false;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float p[3];
                edge_vertex(s_tsdf, g, c, (int) ((word[h] >> (4 * (3 * j + v))) & 0xF), p);
                out = out || vertex_outside(p, dim);
            }
            keep[h] |= (out ? 0u : 1u) << j;
        }
        n_faces += __builtin_popcount(keep[h]);
    }

    uint32_t total;
    const uint32_t rank = block_exclusive_scan(n_faces, s_wave, total);
    if (!Emit) {
        if (t == 0)
            block_faces[slot] = total;
        return;
    }
    uint64_t out = block_faces[slot] + rank;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (!keep[h])
            continue;
        const int ci = t + h * kMeshThreads;
        const int lc[3] = {ci & 7, (ci >> 3) & 7, ci >> 6};
        const int c[3] = {bc[0] + lc[0], bc[1] + lc[1], bc[2] + lc[2]};
        const int g = lc[0] + 9 * lc[1] + 81 * lc[2];
        for (int j = 0; j < 5; j++) {
            if (!((keep[h] >> j) & 1))
                continue;
            float* f = faces + out * 9;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float p[3], o[3];
                edge_vertex(s_tsdf, g, c, (int) ((word[h] >> (4 * (3 * j + v))) & 0xF), p);
                iso_apply(T_OV, p, o);
                f[3 * v] = o[0];
                f[3 * v + 1] = o[1];
                f[3 * v + 2] = o[2];
            }
            out++;
        }
    }
}

// In-place exclusive scan of the n per-block face counts with one 1024-thread workgroup;
// data[n] receives the total.
__global__ __launch_bounds__(1024) void k_mesh_scan(uint64_t* data, int n)
{
    __shared__ uint64_t sums[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int start = min(t * chunk, n);
    const int end = min(start + chunk, n);
    uint64_t s = 0;
    for (int i = start; i < end; i++)
        s += data[i];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint64_t v = t >= off ? sums[t - off] : 0ull;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint64_t run = sums[t] - s;
    for (int i = start; i < end; i++) {
        const uint64_t v = data[i];
        data[i] = run;
        run += v;
    }
    if (t == 1023)
        data[n] = sums[1023];
}

} // namespace

// T_OV = T_OW * T_WM * Scaling(res) (map_impl.hpp:452-458): T_WM translates by -t_MW, and the
// scaling multiplies the linear part.
static Iso mesh_T_OV(const se8_map* m, const float* T_OW16)
{
    static const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const Iso T_OW = iso_from16(T_OW16 ? T_OW16 : kIdentity);
    Iso T_WM;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            T_WM.R[i][j] = i == j ? 1.f : 0.f;
        T_WM.t[i] = -m->view.t_MW[i];
    }
    Iso A = iso_mul(T_OW, T_WM);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            A.R[i][j] = A.R[i][j] * m->view.res;
    return A;
}

extern "C" int se8_mesh(const se8_map* m, const float T_OW[16], float* faces_dev, int64_t max_faces,
                        int64_t* num_faces, se8_stream stream)
{
    SE8_REQUIRE(m && num_faces, "se8_mesh: null argument");
    SE8_REQUIRE(max_faces >= 0 && (max_faces == 0 || faces_dev), "se8_mesh: bad output buffer");
    hipStream_t s = (hipStream_t) stream;
    int host[kNumCounters];
    SE8_HIP_CHECK(hipMemcpyAsync(host, m->counters.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t n = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    *num_faces = 0;
    if (n == 0)
        return SE8_OK;
    int rc = dev_alloc(m->mesh_offsets, sizeof(uint64_t) * (size_t) (n + 1));
    if (rc != SE8_OK)
        return rc;
    uint64_t* offs = (uint64_t*) m->mesh_offsets.ptr;
    const Iso T_OV = mesh_T_OV(m, T_OW);
    k_mesh<false><<<(unsigned) n, kMeshThreads, 0, s>>>(m->view, T_OV, offs, nullptr);
    k_mesh_scan<<<1, 1024, 0, s>>>(offs, (int) n);
    uint64_t total = 0;
    SE8_HIP_CHECK(hipMemcpyAsync(&total, offs + n, sizeof(total), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    SE8_HIP_CHECK(hipGetLastError());
    *num_faces = (int64_t) total;
    if (max_faces == 0)
        return SE8_OK; // count only
    if ((int64_t) total > max_faces) {
        set_error("se8_mesh: the mesh has " + std::to_string(total) + " faces, the buffer holds " +
                  std::to_string(max_faces));
        return SE8_ERR_CAPACITY;
    }
    k_mesh<true><<<(unsigned) n, kMeshThreads, 0, s>>>(m->view, T_OV, offs, faces_dev);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

// ----------------------------------------------------------------------------------------------
// mesh files: include/se/map/io/impl/mesh_io_impl.hpp:14-29 (dispatch on the extension), 31-113
// (VTK), 116-190 (PLY), 194-230 (OBJ); the same layouts, with the vertices written as triangle soup
// (3 per face, face f uses vertices 3f..3f+2) and C++ stream formatting of floats.
// ----------------------------------------------------------------------------------------------
namespace {

bool ends_with(const std::string& s, const char* suffix)
{
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// se::colours::scale[0] (include/se/common/colour_utils.hpp:23): every single-resolution face has
// scale 0.
constexpr int kScale0Rgb[3] = {102, 194, 165};

void apply_T(const float* T, const float* v, float* o)
{
    if (!T) {
        o[0] = v[0], o[1] = v[1], o[2] = v[2];
        return;
    }
    iso_apply(iso_from16(T), v, o);
}

} // namespace

extern "C" int se8_has_supported_mesh_extension(const char* filename)
{
    if (!filename)
        return 0;
    const std::string f(filename);
    return ends_with(f, ".ply") || ends_with(f, ".vtk") || ends_with(f, ".obj");
}

extern "C" int se8_save_mesh(const float* faces, int64_t n, const char* filename, const float T_OM[16])
{
    SE8_REQUIRE(filename && n >= 0 && (n == 0 || faces), "se8_save_mesh: null argument");
    const std::string fname(filename);
    const int kind = ends_with(fname, ".ply") ? 0 : ends_with(fname, ".vtk") ? 1 : ends_with(fname, ".obj") ? 2 : -1;
    SE8_REQUIRE(kind >= 0, "unknown file extension: " + fname);
    std::ofstream file(fname);
    if (!file.is_open()) {
        set_error("Error writing mesh file " + fname);
        return SE8_ERR_IO;
    }
    const int64_t nv = 3 * n;
    auto vertex = [&](int64_t i, float o[3]) { apply_T(T_OM, faces + 3 * i, o); };
    if (kind == 0) {
        file << "ply\nformat ascii 1.0\ncomment Mesh generated by supereight2_amd\n";
        file << "element vertex " << nv << "\nproperty float x\nproperty float y\nproperty float z\n";
        file << "element face " << n << "\nproperty list uchar int vertex_index\n";
        file << "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n";
        for (int64_t i = 0; i < nv; i++) {
            float o[3];
            vertex(i, o);
            file << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        for (int64_t f = 0; f < n; f++)
            file << 3 << " " << 3 * f << " " << 3 * f + 1 << " " << 3 * f + 2 << " " << kScale0Rgb[0] << " "
                 << kScale0Rgb[1] << " " << kScale0Rgb[2] << "\n";
    }
    else if (kind == 1) {
        file << "# vtk DataFile Version 3.0\nMesh generated by supereight2_amd\nASCII\nDATASET POLYDATA\n\n";
        file << "POINTS " << nv << " float\n";
        for (int64_t i = 0; i < nv; i++) {
            float o[3];
            vertex(i, o);
            file << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        file << "\nPOLYGONS " << n << " " << n * 4 << "\n";
        for (int64_t f = 0; f < n; f++)
            file << 3 << " " << 3 * f << " " << 3 * f + 1 << " " << 3 * f + 2 << "\n";
        file << "\nCELL_DATA " << n << "\n\n";
        file << "SCALARS scale unsigned_char\nLOOKUP_TABLE default\n";
        for (int64_t f = 0; f < n; f++)
            file << 0 << "\n";
        file << "COLOR_SCALARS scale_colour 3\n";
        for (int64_t f = 0; f < n; f++)
            file << kScale0Rgb[0] / 255.0f << " " << kScale0Rgb[1] / 255.0f << " " << kScale0Rgb[2] / 255.0f << "\n";
        file << "\n";
    }
    else {
        file << "# Mesh generated by supereight2_amd\n# " << nv << " vertices\n# " << n << " faces\n";
        for (int64_t i = 0; i < nv; i++) {
            float o[3];
            vertex(i, o);
            file << "v " << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        for (int64_t f = 0; f < n; f++)
            file << "f " << 3 * f + 1 << " " << 3 * f + 2 << " " << 3 * f + 3 << "\n";
    }
    file.close();
    if (!file) {
        set_error("Error writing mesh file " + fname);
        return SE8_ERR_IO;
    }
    return SE8_OK;
}
