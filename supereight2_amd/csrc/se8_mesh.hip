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
// every thread meshes two cells from LDS.  The launches: a neighbour table (the 7 neighbour slots of
// every block, so that a workgroup's halo gather is two memory round trips: table, then voxels);
// the count pass (faces per block); a two-level scan of the counts, packed with a non-empty flag,
// giving each block its face offset and its rank among the non-empty blocks; a compaction of the
// non-empty blocks; and the emit pass over those only, which recomputes the cells and writes the
// faces through an LDS stage so that the stores are contiguous across the workgroup.  The host reads
// the total back between the scan and the emit (the capacity check).  The output order is fixed
// for a given pool (block slot, then cell index x + 8y + 64z, then the case's triangle order); the
// reference's order depends on its OpenMP schedule, so faces are compared as a set.
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

// checkVertex (marching_cube_impl.hpp:814-818) drops a face with a vertex outside (0, size]^3.  In
// the single-resolution mesher it never fires, so every triangle of a cell's case is kept: the
// two corners of a vertex's edge differ in inside-ness (tsdf < 0), so t = (0 - v0) / (v1 - v0) is
// in [0, 1] (|v1 - v0| >= |v0| holds after rounding, and a division or sum that rounds monotonically
// stays in that range; an overflow of v1 - v0 gives t = 0) or is NaN (a NaN or infinite corner);
// with p1 - p0 = +-1 exactly, each coordinate lies between the two voxel centres c + 0.5 and
// c + 1.5 (c <= size - 2), or is NaN, which checkVertex keeps.  The oracle applies checkVertex
// literally; the parity tests cover NaN and infinite voxels.

// Exclusive prefix of v over the workgroup and the workgroup total (64-wide waves, 32-bit halves
// shuffled separately).
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* s_wave, uint64_t& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t) inc, d, 64);
        const uint32_t hi = __shfl_up((uint32_t) (inc >> 32), d, 64);
        if (lane >= d)
            inc += ((uint64_t) hi << 32) | lo;
    }
    if (lane == 63)
        s_wave[wave] = inc;
    __syncthreads();
    uint64_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kMeshThreads / 64; w++) {
        const uint64_t s = s_wave[w];
        before += w < wave ? s : 0ull;
        total += s;
    }
    __syncthreads(); // s_wave may be reused by the caller
    return before + inc - v;
}

// Per-block scan value: face count in the high bits, 1 for a block with faces in the low bits, so
// that one exclusive scan yields both the face offset and the rank among non-empty blocks.
constexpr int kPackShift = 28;
constexpr uint64_t kPackMask = (1ull << kPackShift) - 1;

// Neighbour table: nb[8 slot + t] = slot of the block at (bc + 8 (t & 1, t >> 1 & 1, t >> 2 & 1)),
// -1 if unallocated or beyond the map's upper faces (no wrap-around: the cells that would read it
// are outside the meshed range), nb[8 slot] = slot.
__global__ __launch_bounds__(256) void k_mesh_neighbours(MapView m, int n, int* nb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 8 * n)
        return;
    const int slot = i >> 3, t = i & 7;
    int bc[3];
    key_to_coord(m.block_key[slot], bc);
    const int nx = bc[0] + 8 * (t & 1), ny = bc[1] + 8 * ((t >> 1) & 1), nz = bc[2] + 8 * ((t >> 2) & 1);
    int r = slot;
    if (t != 0)
        r = (nx < m.size && ny < m.size && nz < m.size) ? lookup_block_nb(m, nx, ny, nz) : -1;
    nb[i] = r;
}

struct MeshLds {
    float tsdf[kGrid];
    uint8_t valid[kGrid];
    uint8_t cell_case[512]; // case index of each cell (0 for no face)
    uint64_t tri[256];
    uint64_t wave[kMeshThreads / 64];
    int nb[8];
};

// Stages the 9^3 corner grid of block `slot` (its 512 voxels and the +1 faces of its neighbours)
// and the case table into LDS.
__device__ __forceinline__ void stage_block(const MapView& m, int slot, const int* nb_tab, MeshLds& L)
{
    const int t = threadIdx.x;
    L.tri[t] = se8_mc_tri[t];
    if (t < 8)
        L.nb[t] = nb_tab[8 * slot + t];
#pragma unroll
    for (int i = t; i < 512; i += kMeshThreads) {
        float ts, w;
        load_voxel(m, slot, i, ts, w);
        const int g = (i & 7) + 9 * ((i >> 3) & 7) + 81 * (i >> 6);
        L.tsdf[g] = ts;
        L.valid[g] = w != 0.f;
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
        const int s = L.nb[(x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2)];
        float ts = 1.f, w = 0.f; // getData of an unallocated block: the default {1, 0}
        if (s >= 0)
            load_voxel(m, s, (x & 7) + 8 * (y & 7) + 64 * (z & 7), ts, w);
        const int g = x + 9 * y + 81 * z;
        L.tsdf[g] = ts;
        L.valid[g] = w != 0.f;
    }
    __syncthreads();
}

// The two cells of this thread (cell index t and t + 256 of the block): their case words.
struct ThreadCells {
    uint64_t word[2];
};

__device__ __forceinline__ void cell_geometry(int ci, const int bc[3], int lc[3], int c[3], int& g)
{
    lc[0] = ci & 7, lc[1] = (ci >> 3) & 7, lc[2] = ci >> 6;
    c[0] = bc[0] + lc[0], c[1] = bc[1] + lc[1], c[2] = bc[2] + lc[2];
    g = lc[0] + 9 * lc[1] + 81 * lc[2];
}

// Case words of the thread's cells (isosurface::occupied: 0 if any corner is unknown; cells at
// x, y or z = size - 1 are outside the reference's loop range) and their number of faces; the case
// indices are also kept in L.cell_case.
__device__ __forceinline__ uint32_t classify_cells(const MapView& m, MeshLds& L, const int bc[3], ThreadCells& tc)
{
    uint32_t n_faces = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        int lc[3], c[3], g;
        cell_geometry(threadIdx.x + h * kMeshThreads, bc, lc, c, g);
        uint32_t idx = 0;
        bool all_valid = true;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int gk = corner_at(g, k);
            all_valid = all_valid && L.valid[gk];
            idx |= (L.tsdf[gk] < 0.f ? 1u : 0u) << k;
        }
        const bool in_range = c[0] < m.size - 1 && c[1] < m.size - 1 && c[2] < m.size - 1;
        const uint32_t cs = (all_valid && in_range) ? idx : 0u;
        L.cell_case[threadIdx.x + h * kMeshThreads] = (uint8_t) cs;
        tc.word[h] = L.tri[cs];
        n_faces += (uint32_t) (tc.word[h] >> 60);
    }
    return n_faces;
}

// Count pass: scan[slot] = (faces << kPackShift) | (faces > 0) for every block of the pool.
__global__ __launch_bounds__(kMeshThreads) void k_mesh_count(MapView m, const int* nb_tab, uint64_t* scan)
{
    __shared__ MeshLds L;
    const int slot = blockIdx.x;
    int bc[3];
    key_to_coord(m.block_key[slot], bc);
    stage_block(m, slot, nb_tab, L);
    ThreadCells tc;
    const uint32_t n_faces = classify_cells(m, L, bc, tc);
    uint64_t total;
    (void) block_exclusive_scan(n_faces, L.wave, total);
    if (threadIdx.x == 0)
        scan[slot] = (total << kPackShift) | (total > 0 ? 1ull : 0ull);
}

constexpr int kScanTile = 2048; // scan values per workgroup of the tile scan (8 per thread)

// Exclusive scan of the packed values within tiles of kScanTile; tile_sums[tile] = the tile total.
__global__ __launch_bounds__(256) void k_mesh_scan_tiles(uint64_t* data, int n, uint64_t* tile_sums)
{
    __shared__ uint64_t s_wave[4];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 8;
    uint64_t v[8], s = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        v[j] = base + j < n ? data[base + j] : 0ull;
        s += v[j];
    }
    uint64_t total;
    uint64_t run = block_exclusive_scan(s, s_wave, total);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (base + j < n)
            data[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0)
        tile_sums[blockIdx.x] = total;
}

// Exclusive scan of the n_tiles tile totals with one workgroup; tile_sums[n_tiles] = grand total.
__global__ __launch_bounds__(1024) void k_mesh_scan_top(uint64_t* tile_sums, int n_tiles)
{
    __shared__ uint64_t sums[1024];
    const int t = threadIdx.x;
    const int chunk = (n_tiles + 1023) / 1024;
    const int start = min(t * chunk, n_tiles);
    const int end = min(start + chunk, n_tiles);
    uint64_t s = 0;
    for (int i = start; i < end; i++)
        s += tile_sums[i];
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
        const uint64_t v = tile_sums[i];
        tile_sums[i] = run;
        run += v;
    }
    if (t == 1023)
        tile_sums[n_tiles] = sums[1023];
}

__device__ __forceinline__ uint64_t scanned(const uint64_t* scan, const uint64_t* tile_sums, int slot)
{
    return scan[slot] + tile_sums[slot / kScanTile];
}

// list[rank among non-empty blocks] = slot, for every block with faces.
__global__ __launch_bounds__(256) void k_mesh_compact(const uint64_t* scan, const uint64_t* tile_sums, int n,
                                                      int* list)
{
    const int slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= n)
        return;
    const uint64_t here = scanned(scan, tile_sums, slot);
    const uint64_t next = slot + 1 < n ? scanned(scan, tile_sums, slot + 1) : tile_sums[(n + kScanTile - 1) / kScanTile];
    if ((next & kPackMask) != (here & kPackMask))
        list[here & kPackMask] = slot;
}

constexpr int kStageFaces = kMeshThreads; // faces per write round: one per thread
constexpr int kMaxBlockFaces = 512 * 5;

// Emit pass over the non-empty blocks.  The block's faces are listed in LDS in the fixed order
// (cell index << 3 | triangle), then computed one face per thread in rounds of kStageFaces, each
// round staged in LDS and stored contiguously at the block's offset.
__global__ __launch_bounds__(kMeshThreads) void k_mesh_emit(MapView m, Iso T_OV, const int* nb_tab,
                                                             const uint64_t* scan, const uint64_t* tile_sums,
                                                             const int* list, float* faces)
{
    __shared__ MeshLds L;
    __shared__ uint16_t s_face[kMaxBlockFaces];
    __shared__ float s_out[kStageFaces * 9];
    const int slot = list[blockIdx.x];
    int bc[3];
    key_to_coord(m.block_key[slot], bc);
    const uint64_t first = scanned(scan, tile_sums, slot) >> kPackShift;
    stage_block(m, slot, nb_tab, L);
    ThreadCells tc;
    const uint32_t n_faces = classify_cells(m, L, bc, tc);
    uint64_t total64;
    uint32_t r = (uint32_t) block_exclusive_scan(n_faces, L.wave, total64);
    const uint32_t total = (uint32_t) total64;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int ci = threadIdx.x + h * kMeshThreads;
        const int ntri = (int) (tc.word[h] >> 60);
        for (int j = 0; j < ntri; j++)
            s_face[r++] = (uint16_t) ((ci << 3) | j);
    }
    __syncthreads();
    float* dst = faces + first * 9;
    for (uint32_t round = 0; round < total; round += kStageFaces) {
        const uint32_t f = round + threadIdx.x;
        if (f < total) {
            const int ci = s_face[f] >> 3, j = s_face[f] & 7;
            int lc[3], c[3], g;
            cell_geometry(ci, bc, lc, c, g);
            const uint64_t word = L.tri[L.cell_case[ci]];
            float* o9 = s_out + threadIdx.x * 9;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float p[3], o[3];
                edge_vertex(L.tsdf, g, c, (int) ((word >> (4 * (3 * j + v))) & 0xF), p);
                iso_apply(T_OV, p, o);
                o9[3 * v] = o[0];
                o9[3 * v + 1] = o[1];
                o9[3 * v + 2] = o[2];
            }
        }
        __syncthreads();
        const int nf = (int) min((uint32_t) kStageFaces, total - round);
        for (int i = threadIdx.x; i < nf * 9; i += kMeshThreads)
            dst[(size_t) round * 9 + i] = s_out[i];
        __syncthreads();
    }
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
    SE8_REQUIRE(!m || !is_multires(m), "se8_mesh: not available for a multi-resolution map");
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
    // scratch: scan values (n), tile sums (n_tiles + 1), neighbour table (8 n ints), list (n ints)
    const int64_t n_tiles = (n + kScanTile - 1) / kScanTile;
    const size_t off_tiles = sizeof(uint64_t) * (size_t) n;
    const size_t off_nb = off_tiles + sizeof(uint64_t) * (size_t) (n_tiles + 1);
    const size_t off_list = off_nb + sizeof(int) * 8 * (size_t) n;
    int rc = dev_alloc(m->mesh_offsets, off_list + sizeof(int) * (size_t) n);
    if (rc != SE8_OK)
        return rc;
    char* base = (char*) m->mesh_offsets.ptr;
    uint64_t* scan = (uint64_t*) base;
    uint64_t* tile_sums = (uint64_t*) (base + off_tiles);
    int* nb = (int*) (base + off_nb);
    int* list = (int*) (base + off_list);
    const Iso T_OV = mesh_T_OV(m, T_OW);
    k_mesh_neighbours<<<(unsigned) ((8 * n + 255) / 256), 256, 0, s>>>(m->view, (int) n, nb);
    k_mesh_count<<<(unsigned) n, kMeshThreads, 0, s>>>(m->view, nb, scan);
    k_mesh_scan_tiles<<<(unsigned) n_tiles, 256, 0, s>>>(scan, (int) n, tile_sums);
    k_mesh_scan_top<<<1, 1024, 0, s>>>(tile_sums, (int) n_tiles);
    uint64_t total = 0;
    SE8_HIP_CHECK(hipMemcpyAsync(&total, tile_sums + n_tiles, sizeof(total), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    SE8_HIP_CHECK(hipGetLastError());
    const int64_t n_faces = (int64_t) (total >> kPackShift);
    const int64_t n_nonempty = (int64_t) (total & kPackMask);
    *num_faces = n_faces;
    if (max_faces == 0 || n_faces == 0)
        return SE8_OK; // count only, or nothing to write
    if (n_faces > max_faces) {
        set_error("se8_mesh: the mesh has " + std::to_string(n_faces) + " faces, the buffer holds " +
                  std::to_string(max_faces));
        return SE8_ERR_CAPACITY;
    }
    k_mesh_compact<<<(unsigned) ((n + 255) / 256), 256, 0, s>>>(scan, tile_sums, (int) n, list);
    k_mesh_emit<<<(unsigned) n_nonempty, kMeshThreads, 0, s>>>(m->view, T_OV, nb, scan, tile_sums, list, faces_dev);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

// ----------------------------------------------------------------------------------------------
// mesh files: include/se/map/io/impl/mesh_io_impl.hpp:14-29 (dispatch on the extension), 31-113
// (VTK), 116-190 (PLY), 194-230 (OBJ); the same layouts for faces of 3 (surface) or 4 (octree
// structure) vertices, written as a vertex soup (face f uses vertices nv*f .. nv*f + nv - 1) with C++
// stream formatting of floats, and each face coloured by its scale.
// ----------------------------------------------------------------------------------------------
namespace {

bool ends_with(const std::string& s, const char* suffix)
{
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// se::colours::scale (include/se/common/colour_utils.hpp:23-30); scale_colour clamps to the last
// entry (colour_utils_impl.hpp:40-47).
constexpr int kScaleRgb[8][3] = {{102, 194, 165}, {252, 141, 98}, {141, 160, 203}, {231, 138, 195},
                                 {166, 216, 84},  {255, 217, 47}, {229, 196, 148}, {179, 179, 179}};

const int* scale_rgb(const int8_t* scales, int64_t f)
{
    const int s = scales ? scales[f] : 0;
    return kScaleRgb[s < 0 ? 0 : (s > 7 ? 7 : s)];
}

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

extern "C" int se8_save_mesh(const float* vertices, int64_t n, int nv, const int8_t* scales, const char* filename,
                             const float T_OM[16])
{
    SE8_REQUIRE(filename && n >= 0 && (n == 0 || vertices), "se8_save_mesh: null argument");
    SE8_REQUIRE(nv == 3 || nv == 4, "se8_save_mesh: faces have 3 or 4 vertices");
    const std::string fname(filename);
    const int kind = ends_with(fname, ".ply") ? 0 : ends_with(fname, ".vtk") ? 1 : ends_with(fname, ".obj") ? 2 : -1;
    SE8_REQUIRE(kind >= 0, "unknown file extension: " + fname);
    std::ofstream file(fname);
    if (!file.is_open()) {
        set_error("Error writing mesh file " + fname);
        return SE8_ERR_IO;
    }
    const int64_t n_vert = nv * n;
    auto vertex = [&](int64_t i, float o[3]) { apply_T(T_OM, vertices + 3 * i, o); };
    auto indices = [&](int64_t f, int64_t base) {
        for (int v = 0; v < nv; v++)
            file << " " << nv * f + v + base;
    };
    if (kind == 0) {
        file << "ply\nformat ascii 1.0\ncomment Mesh generated by supereight2_amd\n";
        file << "element vertex " << n_vert << "\nproperty float x\nproperty float y\nproperty float z\n";
        file << "element face " << n << "\nproperty list uchar int vertex_index\n";
        file << "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n";
        for (int64_t i = 0; i < n_vert; i++) {
            float o[3];
            vertex(i, o);
            file << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        for (int64_t f = 0; f < n; f++) {
            const int* rgb = scale_rgb(scales, f);
            file << nv;
            indices(f, 0);
            file << " " << rgb[0] << " " << rgb[1] << " " << rgb[2] << "\n";
        }
    }
    else if (kind == 1) {
        file << "# vtk DataFile Version 3.0\nMesh generated by supereight2_amd\nASCII\nDATASET POLYDATA\n\n";
        file << "POINTS " << n_vert << " float\n";
        for (int64_t i = 0; i < n_vert; i++) {
            float o[3];
            vertex(i, o);
            file << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        file << "\nPOLYGONS " << n << " " << n * (1 + nv) << "\n";
        for (int64_t f = 0; f < n; f++) {
            file << nv;
            indices(f, 0);
            file << "\n";
        }
        file << "\nCELL_DATA " << n << "\n\n";
        file << "SCALARS scale unsigned_char\nLOOKUP_TABLE default\n";
        for (int64_t f = 0; f < n; f++)
            file << (scales ? (int) scales[f] : 0) << "\n";
        file << "COLOR_SCALARS scale_colour 3\n";
        for (int64_t f = 0; f < n; f++) {
            const int* rgb = scale_rgb(scales, f);
            file << rgb[0] / 255.0f << " " << rgb[1] / 255.0f << " " << rgb[2] / 255.0f << "\n";
        }
        file << "\n";
    }
    else {
        file << "# Mesh generated by supereight2_amd\n# " << n_vert << " vertices\n# " << n << " faces\n";
        for (int64_t i = 0; i < n_vert; i++) {
            float o[3];
            vertex(i, o);
            file << "v " << o[0] << " " << o[1] << " " << o[2] << "\n";
        }
        for (int64_t f = 0; f < n; f++) {
            file << "f";
            indices(f, 1);
            file << "\n";
        }
    }
    file.close();
    if (!file) {
        set_error("Error writing mesh file " + fname);
        return SE8_ERR_IO;
    }
    return SE8_OK;
}
