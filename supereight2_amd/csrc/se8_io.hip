// Debug views of the map: field slices and the octree structure.
//
// se::Map::saveFieldSlices (include/se/map/impl/map_impl.hpp:272-308) writes get_field(getData) over
// the three axis-aligned planes of the octree through a point's voxel with
// se::io::save_3d_slice_vtk (include/se/map/io/impl/octree_io_impl.hpp:19-80).  Here one thread per
// voxel of a plane gathers the values on the device (an octree-size square, e.g. 1024^2 lookups);
// the host writes the VTK text in the reference's layout.
//
// se::Map::structure / saveStructure (map_impl.hpp:473-490, octree_impl.hpp:148-167) ->
// octree_structure_mesh (include/se/map/algorithms/impl/structure_meshing_impl.hpp:15-60): six quads
// per octant (every node and block, or only the leaves), the vertices mapped by
// T_OV = T_OW * T_WM * Scaling(res).  Here one thread per octant writes its 24 vertices; octants
// are ordered nodes (by pool index) then blocks (by slot), the reference's by its octree iterator, so
// the quads are compared as a set.
#include "se8_internal.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace se8;

namespace {

// Element i of plane `axis` = index, in the order save_3d_slice_vtk visits it (z outer, then y,
// then x): the plane's two free coordinates, the slower one outer.
__global__ __launch_bounds__(256) void k_field_slice(MapView m, int axis, int index, float* out)
{
    const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t) m.size * m.size;
    if (i >= n)
        return;
    const int inner = (int) (i % m.size), outer = (int) (i / m.size);
    int v[3];
    if (axis == 0)
        v[0] = index, v[1] = inner, v[2] = outer;
    else if (axis == 1)
        v[0] = inner, v[1] = index, v[2] = outer;
    else
        v[0] = inner, v[1] = outer, v[2] = index;
    float tsdf = 1.f, w = 0.f; // getData outside the octree: the default data
    if (index >= 0 && index < m.size)
        get_voxel(m, v[0], v[1], v[2], tsdf, w);
    out[i] = tsdf;
}

// The 8 corners of an octant in structure_meshing_impl.hpp's order (bit 0: +x, bit 1: +y, bit 2:
// +z) and its six faces as corner indices.
__constant__ int kQuadCorners[6][4] = {{0, 2, 3, 1}, {1, 3, 7, 5}, {5, 7, 6, 4}, {0, 4, 6, 2}, {0, 1, 5, 4}, {2, 6, 7, 3}};

// Octant o < n_listed: node node_list[o] (or node o when node_list is null); o >= n_listed: block
// o - n_listed.  Quads of octant o are 6 o .. 6 o + 5.
__global__ __launch_bounds__(256) void k_structure(MapView m, const int* node_list, int n_listed, int n_blocks,
                                                   Iso T_OV, float* quads, int8_t* scales)
{
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= n_listed + n_blocks)
        return;
    uint64_t key;
    int8_t scale;
    if (o < n_listed) {
        key = m.node_key[node_list ? node_list[o] : o];
        scale = 7; // the reference's grey for nodes
    }
    else {
        key = m.block_key[o - n_listed];
        // the block's current scale (always 0 in a single-resolution map)
        scale = m.storage == SE8_STORAGE_MULTIRES ? (int8_t) m.block_scale[2 * (o - n_listed)] : (int8_t) 0;
    }
    int c[3];
    key_to_coord(key, c);
    const int size = 1 << (int) (key & 0x1F);
    float corner[8][3];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float p[3] = {(float) (c[0] + ((k & 1) ? size : 0)), (float) (c[1] + ((k & 2) ? size : 0)),
                            (float) (c[2] + ((k & 4) ? size : 0))};
        iso_apply(T_OV, p, corner[k]);
    }
    float* q = quads + (size_t) o * 6 * 12;
#pragma unroll
    for (int f = 0; f < 6; f++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = kQuadCorners[f][v];
            q[12 * f + 3 * v] = corner[k][0];
            q[12 * f + 3 * v + 1] = corner[k][1];
            q[12 * f + 3 * v + 2] = corner[k][2];
        }
        if (scales)
            scales[(size_t) o * 6 + f] = scale;
    }
}

int read_counts(const se8_map* m, hipStream_t s, int64_t& n_blocks, int64_t& n_nodes)
{
    int host[kNumCounters];
    SE8_HIP_CHECK(hipMemcpyAsync(host, m->counters.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    SE8_HIP_CHECK(hipStreamSynchronize(s));
    n_blocks = std::min<int64_t>(host[kCntBlocks], m->view.max_blocks);
    n_nodes = std::min<int64_t>(host[kCntNodes], m->view.max_nodes);
    return SE8_OK;
}

// T_OV = T_OW * T_WM * Scaling(res) (map_impl.hpp:452-458, 473-479).
Iso octree_to_output(const se8_map* m, const float* T_OW16)
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

// save_3d_slice_vtk's layout for the plane [lower, upper) with the values in visiting order.
bool write_slice_vtk(const std::string& filename, const int lower[3], const int upper[3], const float* values)
{
    std::ofstream file(filename.c_str());
    if (!file.is_open())
        return false;
    int dim[3];
    for (int i = 0; i < 3; i++)
        dim[i] = std::max(1, upper[i] - lower[i]);
    file << "# vtk DataFile Version 1.0" << std::endl;
    file << "vtk mesh generated from KFusion" << std::endl;
    file << "ASCII" << std::endl;
    file << "DATASET RECTILINEAR_GRID" << std::endl;
    file << "DIMENSIONS " << dim[0] << " " << dim[1] << " " << dim[2] << std::endl;
    const char* names[3] = {"X_COORDINATES ", "Y_COORDINATES ", "Z_COORDINATES "};
    std::ostringstream coords[3];
    for (int a = 0; a < 3; a++)
        for (int c = lower[a]; c < upper[a]; c++)
            coords[a] << c << " ";
    for (int a = 0; a < 3; a++) {
        file << names[a] << dim[a] << " int " << std::endl;
        file << coords[a].str() << std::endl;
    }
    const int64_t n = (int64_t) dim[0] * dim[1] * dim[2];
    file << "POINT_DATA " << n << std::endl;
    file << "SCALARS scalars float 1" << std::endl;
    file << "LOOKUP_TABLE default" << std::endl;
    std::ostringstream sc;
    for (int64_t i = 0; i < n; i++)
        sc << values[i] << "\n";
    file << sc.str() << std::endl;
    file.close();
    return (bool) file;
}

} // namespace

extern "C" int se8_field_slice(const se8_map* m, int axis, int index, float* out_dev, se8_stream stream)
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_field_slice: not available for a multi-resolution map");
    SE8_REQUIRE(m && out_dev && axis >= 0 && axis < 3, "se8_field_slice: bad argument");
    const int64_t n = (int64_t) m->view.size * m->view.size;
    k_field_slice<<<(unsigned) ((n + 255) / 256), 256, 0, (hipStream_t) stream>>>(m->view, axis, index, out_dev);
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

extern "C" int se8_save_field_slices(const se8_map* m, const char* filename_x, const char* filename_y,
                                     const char* filename_z, const float point_W[3])
{
    SE8_REQUIRE(!m || !is_multires(m), "se8_save_field_slices: not available for a multi-resolution map");
    SE8_REQUIRE(m && point_W, "se8_save_field_slices: null argument");
    // pointToVoxel (map_impl.hpp:541-550): (-1, -1, -1) for a point outside the map
    int v[3] = {-1, -1, -1};
    float pM[3];
    for (int i = 0; i < 3; i++)
        pM[i] = point_W[i] + m->view.t_MW[i];
    const float dim = m->view.dim;
    if (pM[0] >= 0.f && pM[0] < dim && pM[1] >= 0.f && pM[1] < dim && pM[2] >= 0.f && pM[2] < dim)
        for (int i = 0; i < 3; i++)
            v[i] = (int) (pM[i] / m->view.res);
    const int size = m->view.size;
    const int64_t n = (int64_t) size * size;
    DeviceBuffer d;
    int rc = dev_alloc(d, sizeof(float) * (size_t) n);
    if (rc != SE8_OK)
        return rc;
    std::vector<float> host((size_t) n);
    const char* names[3] = {filename_x, filename_y, filename_z};
    for (int axis = 0; axis < 3 && rc == SE8_OK; axis++) {
        if (!names[axis] || !names[axis][0])
            continue;
        rc = se8_field_slice(m, axis, v[axis], (float*) d.ptr, nullptr);
        if (rc == SE8_OK && hipMemcpy(host.data(), d.ptr, sizeof(float) * (size_t) n, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("se8_save_field_slices: copy failed");
            rc = SE8_ERR_HIP;
        }
        if (rc != SE8_OK)
            break;
        int lower[3] = {0, 0, 0}, upper[3] = {size, size, size};
        lower[axis] = v[axis];
        upper[axis] = v[axis] + 1;
        if (!write_slice_vtk(names[axis], lower, upper, host.data())) {
            set_error(std::string("Unable to write file ") + names[axis]);
            rc = SE8_ERR_IO;
        }
    }
    dev_free(d);
    return rc;
}

extern "C" int se8_structure(const se8_map* m, const float T_OW[16], int only_leaves, float* quads_dev,
                             int8_t* scales_dev, int64_t max_quads, int64_t* num_quads, se8_stream stream)
{
    SE8_REQUIRE(m && num_quads, "se8_structure: null argument");
    SE8_REQUIRE(max_quads >= 0 && (max_quads == 0 || quads_dev), "se8_structure: bad output buffer");
    hipStream_t s = (hipStream_t) stream;
    int64_t n_blocks, n_nodes;
    int rc = read_counts(m, s, n_blocks, n_nodes);
    if (rc != SE8_OK)
        return rc;
    // leaves (isLeaf: no child) among the nodes; blocks are always leaves
    std::vector<int> leaf_nodes;
    if (only_leaves && n_nodes > 0) {
        std::vector<int> child((size_t) n_nodes * 8);
        SE8_HIP_CHECK(hipMemcpy(child.data(), m->node_child.ptr, sizeof(int) * child.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n_nodes; i++)
            if (std::all_of(&child[8 * i], &child[8 * i + 8], [](int c) { return c < 0; }))
                leaf_nodes.push_back((int) i);
    }
    const int64_t n_listed = only_leaves ? (int64_t) leaf_nodes.size() : n_nodes;
    const int64_t total = 6 * (n_listed + n_blocks);
    *num_quads = total;
    if (max_quads == 0 || total == 0)
        return SE8_OK;
    if (total > max_quads) {
        set_error("se8_structure: the structure has " + std::to_string(total) + " quads, the buffer holds " +
                  std::to_string(max_quads));
        return SE8_ERR_CAPACITY;
    }
    DeviceBuffer list;
    if (only_leaves && !leaf_nodes.empty()) {
        rc = dev_alloc(list, sizeof(int) * leaf_nodes.size());
        if (rc != SE8_OK)
            return rc;
        SE8_HIP_CHECK(hipMemcpy(list.ptr, leaf_nodes.data(), sizeof(int) * leaf_nodes.size(), hipMemcpyHostToDevice));
    }
    const int n_oct = (int) (n_listed + n_blocks);
    k_structure<<<(unsigned) ((n_oct + 255) / 256), 256, 0, s>>>(m->view, only_leaves ? (const int*) list.ptr : nullptr,
                                                                 (int) n_listed, (int) n_blocks, octree_to_output(m, T_OW),
                                                                 quads_dev, scales_dev);
    SE8_HIP_CHECK(hipGetLastError());
    if (list.ptr) {
        SE8_HIP_CHECK(hipStreamSynchronize(s)); // the list is freed on return
        dev_free(list);
    }
    return SE8_OK;
}
