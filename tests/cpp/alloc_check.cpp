// CPU bounds and structure checks of libse8's allocation pipeline (supereight2_amd/csrc/se8_alloc.hpp).
// The kernels and the host code that enqueues them are compiled unchanged as host C++ against
// tests/cpp/simt (lanes as threads, see simt_host.hpp), with AddressSanitizer: every buffer is a
// separate heap allocation of exactly the size the library allocates, so an index outside a
// buffer, or outside a workgroup's __shared__ array, is reported where it happens.
//
// usage: alloc_check <input> <output>
// input (little-endian binary, written by tests/test_alloc_host_sim.py):
//   int32 dim_cm, res_um, max_blocks (0 = default), cus (compute units the grids are sized for)
//   int32 n_ops, then per op:
//     int32 kind = 0 (integrate): int32 w, h; float fx, fy, cx, cy, near, far; float T_WS[16];
//                                 float depth[w * h]
//     int32 kind = 1 (allocate):  int32 n; int32 voxel_coords[3 n]   (se8_allocate_blocks' lists,
//                                 passed unsorted, with repeats and allocated blocks)
// output (text): per op a line "op <i> blocks <n> nodes <n> overflow <0|1> update <n>", then the
//   sorted block keys ("b <key>") and, for integrations, the sorted update-list keys ("u <key>").
// After every op the map's structure is checked (keys, tables and links agree; see check_map);
// a violation prints "FAIL ..." and exits with status 2.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "se8_alloc.hpp"

using namespace se8;

namespace se8 {
void set_error(const std::string& msg)
{
    if (!msg.empty())
        std::fprintf(stderr, "se8 error: %s\n", msg.c_str());
}
int dev_alloc(DeviceBuffer& b, size_t bytes)
{
    if (b.ptr && b.bytes >= bytes)
        return SE8_OK;
    dev_free(b);
    if (bytes == 0)
        bytes = 16;
    b.ptr = std::malloc(bytes);
    if (!b.ptr)
        return SE8_ERR_OOM;
    if (bytes <= ((size_t) 64 << 20))
        std::memset(b.ptr, 0xA5, bytes); // not zero: device memory starts with arbitrary contents
    b.bytes = bytes;
    return SE8_OK;
}
void dev_free(DeviceBuffer& b)
{
    std::free(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}
static int g_cus = 4;
int device_count_sm()
{
    return g_cus;
}
void graph_note_launch(int, hipStream_t) {}
} // namespace se8

static void fill(DeviceBuffer& b, int64_t n, int v)
{
    int* p = (int*) b.ptr;
    for (int64_t i = 0; i < n; i++)
        p[i] = v;
}

static void fail(const std::string& what)
{
    std::printf("FAIL %s\n", what.c_str());
    std::fflush(stdout);
    std::exit(2);
}

// The buffers se8_map_create allocates for a single-resolution FP32 map (voxels only when an
// explicit allocation initialises them), sized by plan_map_view, initialised as it does.
static se8_map* create_map(float dim, float res, int64_t max_blocks, bool voxels)
{
    se8_map* m = new se8_map();
    se8_map_config cfg{{dim, dim, dim}, res, {dim / 2, dim / 2, dim / 2}};
    se8_tsdf_config t{8.f, 100.f, SE8_STORAGE_F32, max_blocks};
    m->cfg = cfg;
    m->tsdf = t;
    MapView& v = m->view;
    int64_t node_dense_total = 0;
    if (!plan_map_view(cfg, t, v, node_dense_total))
        fail("plan_map_view");
    const int64_t bricks_side3 = (int64_t) v.bricks_side * v.bricks_side * v.bricks_side;
    const int64_t dense_side = v.size / 8;
    dev_alloc(m->brick_top, sizeof(int) * bricks_side3);
    dev_alloc(m->brick_slots, sizeof(int) * 512 * v.max_bricks);
    if (v.size <= kDenseMaxSize) {
        dev_alloc(m->block_dense, sizeof(int) * dense_side * dense_side * dense_side);
        dev_alloc(m->node_dense, sizeof(int) * std::max<int64_t>(node_dense_total, 1));
    }
    dev_alloc(m->node_child, sizeof(int) * 8 * v.max_nodes);
    dev_alloc(m->node_key, sizeof(uint64_t) * v.max_nodes);
    dev_alloc(m->node_ts, sizeof(int) * v.max_nodes);
    dev_alloc(m->block_key, sizeof(uint64_t) * v.max_blocks);
    dev_alloc(m->block_parent, sizeof(int) * v.max_blocks);
    dev_alloc(m->block_ts, sizeof(int) * v.max_blocks);
    if (voxels)
        dev_alloc(m->voxels, 8 * 512 * (size_t) v.max_blocks);
    dev_alloc(m->counters, sizeof(int) * kNumCounters);
    dev_alloc(m->update_list, sizeof(int) * v.max_blocks);
    dev_alloc(m->frame_poses, sizeof(FramePoses));
    v.brick_top = (int*) m->brick_top.ptr;
    v.brick_slots = (int*) m->brick_slots.ptr;
    v.block_dense = (int*) m->block_dense.ptr;
    v.node_dense = (int*) m->node_dense.ptr;
    v.node_child = (int*) m->node_child.ptr;
    v.node_key = (uint64_t*) m->node_key.ptr;
    v.node_ts = (int*) m->node_ts.ptr;
    v.block_key = (uint64_t*) m->block_key.ptr;
    v.block_parent = (int*) m->block_parent.ptr;
    v.block_ts = (int*) m->block_ts.ptr;
    v.voxels = m->voxels.ptr;
    v.counters = (int*) m->counters.ptr;
    v.block_scale = nullptr;
    fill(m->brick_top, bricks_side3, -1);
    fill(m->brick_slots, 512 * v.max_bricks, -1);
    if (m->block_dense.ptr) {
        fill(m->block_dense, dense_side * dense_side * dense_side, -1);
        fill(m->node_dense, std::max<int64_t>(node_dense_total, 1), -1);
    }
    fill(m->node_child, 8 * v.max_nodes, -1);
    int* c = v.counters;
    std::memset(c, 0, sizeof(int) * kNumCounters);
    c[kCntNodes] = 1;
    for (int a = 0; a < 3; a++) {
        c[kCntAabb + a] = 0x7fffffff;
        c[kCntAabb + 3 + a] = -1;
    }
    if (v.brick_shift + kBlockScale == v.max_scale) {
        c[kCntBricks] = 1;
        v.brick_top[0] = 0;
    }
    v.node_key[0] = (uint64_t) v.max_scale;
    v.node_ts[0] = -1;
    return m;
}

static uint64_t key_at(uint64_t key, int scale)
{
    const uint64_t code = key >> kScaleOffset;
    const uint64_t mask = (0x1ffffffffffffffull >> (3 * scale)) << (3 * scale);
    return ((code & mask) << kScaleOffset) | (uint64_t) scale;
}

// Structure of the map after an op.  Without an overflow: every block's key is a scale-3 key inside
// the octree, reached from the root through node_child (each node's key is its parent's child key
// at the child index), found by the brick index and the dense tables, and the block/node key sets
// have no repeats; the dense tables hold no claim marks and name only octants with their key; the
// AABB counters bound the blocks.  After an overflow the pools' contents are checked the same way
// except for links that the overflow may have left out.
static void check_map(const se8_map* m)
{
    const MapView& v = m->view;
    const int* c = v.counters;
    const bool overflow = c[kCntOverflow] != 0;
    const int64_t nb = std::min<int64_t>(c[kCntBlocks], v.max_blocks);
    const int64_t nn = std::min<int64_t>(c[kCntNodes], v.max_nodes);
    if (!overflow && (c[kCntBlocks] > v.max_blocks || c[kCntNodes] > v.max_nodes))
        fail("counter past its pool without the overflow flag");
    std::vector<uint64_t> bkeys(v.block_key, v.block_key + nb), nkeys(v.node_key, v.node_key + nn);
    std::vector<uint64_t> sb = bkeys, sn = nkeys;
    std::sort(sb.begin(), sb.end());
    std::sort(sn.begin(), sn.end());
    if (std::adjacent_find(sb.begin(), sb.end()) != sb.end())
        fail("repeated block key");
    if (std::adjacent_find(sn.begin(), sn.end()) != sn.end())
        fail("repeated node key");
    if (nn < 1 || v.node_key[0] != (uint64_t) v.max_scale)
        fail("root key");
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
    for (int64_t i = 0; i < nb; i++) {
        const uint64_t k = bkeys[i];
        int cc[3];
        key_to_coord(k, cc);
        if ((k & 0x1f) != (uint64_t) kBlockScale || (k >> kScaleOffset) & 0x1ff)
            fail("block key not at scale 3: slot " + std::to_string(i));
        for (int a = 0; a < 3; a++) {
            if (cc[a] < 0 || cc[a] >= v.size)
                fail("block outside the octree");
            lo[a] = std::min(lo[a], cc[a]);
            hi[a] = std::max(hi[a], cc[a] + 7);
        }
        // walk from the root
        int node = 0;
        bool linked = true;
        for (int s = v.max_scale - 1; s > kBlockScale && linked; s--) {
            const int ch = v.node_child[node * 8 + (int) ((k >> (kScaleOffset + 3 * s)) & 7)];
            if (ch < 0 || ch >= nn) {
                linked = false;
                break;
            }
            if (v.node_key[ch] != key_at(k, s))
                fail("node key is not the child key on the path of slot " + std::to_string(i));
            node = ch;
        }
        if (linked) {
            const int leaf = v.node_child[node * 8 + (int) ((k >> (kScaleOffset + 9)) & 7)];
            if (leaf != (int) i)
                linked = false;
            else if (v.block_parent[i] != node)
                fail("block_parent of slot " + std::to_string(i));
        }
        if (!linked && !overflow)
            fail("block not reachable from the root: slot " + std::to_string(i));
        const int bs = v.brick_shift + kBlockScale;
        const int brick = v.brick_top[((cc[2] >> bs) * v.bricks_side + (cc[1] >> bs)) * v.bricks_side + (cc[0] >> bs)];
        const int msk = (1 << v.brick_shift) - 1;
        if (brick >= 0) {
            if (brick >= v.max_bricks)
                fail("brick id past the brick pool");
            const int e = v.brick_slots[brick * 512 + ((cc[0] >> 3) & msk) + (((cc[1] >> 3) & msk) << 3)
                                        + (((cc[2] >> 3) & msk) << 6)];
            if (e != (int) i && !overflow)
                fail("brick index misses slot " + std::to_string(i));
        }
        else if (!overflow)
            fail("no brick for slot " + std::to_string(i));
        if (v.block_dense) {
            const int side = v.size >> 3;
            const int e = v.block_dense[((cc[2] >> 3) * side + (cc[1] >> 3)) * side + (cc[0] >> 3)];
            if (e != (int) i)
                fail("dense block table misses slot " + std::to_string(i));
        }
    }
    if (nb > 0 && !overflow)
        for (int a = 0; a < 3; a++)
            if (c[kCntAabb + a] != lo[a] || c[kCntAabb + 3 + a] != hi[a])
                fail("AABB counters");
    // nodes: scale in (3, max_scale], node_dense names them
    for (int64_t i = 1; i < nn; i++) {
        const uint64_t k = nkeys[i];
        const int s = (int) (k & 0x1f);
        if (s <= kBlockScale || s >= v.max_scale || key_at(k, s) != k)
            fail("node key " + std::to_string(i));
        if (v.node_dense) {
            int cc[3];
            key_to_coord(k, cc);
            if (v.node_dense[v.node_dense_off[s] + dense_index(v, cc, s)] != (int) i)
                fail("dense node table misses node " + std::to_string(i));
        }
        for (int ch = 0; ch < 8; ch++) {
            const int e = v.node_child[i * 8 + ch];
            if (e < -1 || (s > kBlockScale + 1 && e >= nn) || (s == kBlockScale + 1 && e >= nb))
                fail("child entry out of range at node " + std::to_string(i));
        }
    }
    if (v.block_dense) {
        const int64_t side = v.size >> 3;
        int64_t n_dense = 0;
        for (int64_t j = 0; j < side * side * side; j++) {
            const int e = v.block_dense[j];
            if (e < -1 || e >= nb)
                fail("dense block entry out of range");
            n_dense += e >= 0;
        }
        if (n_dense != nb)
            fail("dense block table count");
        for (int s = kBlockScale + 1; s < v.max_scale; s++) {
            const int64_t sd = (int64_t) v.size >> s;
            for (int64_t j = 0; j < sd * sd * sd; j++) {
                const int e = v.node_dense[v.node_dense_off[s] + j];
                if (e < -1 || e >= nn || (e > 0 && (int) (v.node_key[e] & 0x1f) != s))
                    fail("dense node entry at scale " + std::to_string(s));
            }
        }
    }
}

template<class T>
static T rd(FILE* f)
{
    T x;
    if (std::fread(&x, sizeof(T), 1, f) != 1)
        fail("short input");
    return x;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: alloc_check <input> <output>\n");
        return 1;
    }
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "w");
    if (!in || !out)
        fail("cannot open files");
    const float dim = (float) rd<int32_t>(in) / 100.f;
    const float res = (float) rd<int32_t>(in) / 1e6f;
    const int64_t max_blocks = rd<int32_t>(in);
    g_cus = rd<int32_t>(in);
    const int n_ops = rd<int32_t>(in);
    se8_map* m = create_map(dim, res, max_blocks, true);
    for (int op = 0; op < n_ops; op++) {
        const int kind = rd<int32_t>(in);
        if (kind == 0) {
            se8_pinhole p;
            p.width = rd<int32_t>(in);
            p.height = rd<int32_t>(in);
            p.fx = rd<float>(in);
            p.fy = rd<float>(in);
            p.cx = rd<float>(in);
            p.cy = rd<float>(in);
            p.near_plane = rd<float>(in);
            p.far_plane = rd<float>(in);
            float T[16];
            for (float& x : T)
                x = rd<float>(in);
            DeviceBuffer depth;
            dev_alloc(depth, sizeof(float) * p.width * p.height);
            if (std::fread(depth.ptr, sizeof(float), (size_t) p.width * p.height, in) != (size_t) p.width * p.height)
                fail("short depth");
            const PinholeSensor sensor{make_cam(p, 1, false)};
            // integrate_impl: the candidate buffers of the sorted-key path, then allocate_frame
            if (!(m->view.block_dense && m->view.node_dense)) {
                const float band = 2.0f * m->view.res * m->tsdf.truncation_boundary_factor;
                const int num_steps = (int) std::ceil(band / (2.f * m->view.res));
                if (ensure_candidates(m, (size_t) sensor.width() * sensor.height() * num_steps) != SE8_OK)
                    fail("ensure_candidates");
            }
            if (allocate_frame(m, iso_from16(T), nullptr, (const float*) depth.ptr, sensor, nullptr, nullptr,
                               nullptr) != SE8_OK)
                fail("allocate_frame");
            dev_free(depth);
        }
        else {
            const int n = rd<int32_t>(in);
            std::vector<int32_t> vc(3 * (size_t) n);
            if (n > 0 && std::fread(vc.data(), sizeof(int32_t), vc.size(), in) != vc.size())
                fail("short coordinates");
            // se8_allocate_blocks: the blocks' keys through the sorted-key build (mark_new = 0)
            if (ensure_candidates(m, (size_t) std::max(n, 1)) != SE8_OK)
                fail("ensure_candidates");
            uint64_t* cand = (uint64_t*) m->candidates.ptr;
            for (int i = 0; i < n; i++)
                cand[i] = block_key_of_voxel(vc[3 * i], vc[3 * i + 1], vc[3 * i + 2]);
            m->view.counters[kCntCandidates] = n;
            if (build_from_candidates(m, cand, m->view.counters + kCntCandidates, n, nullptr, 0, nullptr) != SE8_OK)
                fail("build_from_candidates");
        }
        check_map(m);
        const int* c = m->view.counters;
        const int64_t nb = std::min<int64_t>(c[kCntBlocks], m->view.max_blocks);
        const int n_upd = kind == 0 ? (int) std::min<int64_t>(c[kCntUpdate], m->view.max_blocks) : 0;
        std::fprintf(out, "op %d blocks %lld nodes %lld overflow %d update %d\n", op, (long long) nb,
                     (long long) std::min<int64_t>(c[kCntNodes], m->view.max_nodes), c[kCntOverflow] ? 1 : 0, n_upd);
        std::vector<uint64_t> keys(m->view.block_key, m->view.block_key + nb);
        std::sort(keys.begin(), keys.end());
        for (uint64_t k : keys)
            std::fprintf(out, "b %llu\n", (unsigned long long) k);
        if (kind == 0) {
            const int* ul = (const int*) m->update_list.ptr;
            std::vector<uint64_t> uk;
            for (int i = 0; i < n_upd; i++) {
                const int slot = ul[i] & 0x7fffffff;
                if (slot >= nb)
                    fail("update-list entry past the blocks");
                uk.push_back(m->view.block_key[slot]);
            }
            std::sort(uk.begin(), uk.end());
            if (std::adjacent_find(uk.begin(), uk.end()) != uk.end())
                fail("repeated update-list entry");
            for (uint64_t k : uk)
                std::fprintf(out, "u %llu\n", (unsigned long long) k);
        }
        std::fflush(out);
        std::fprintf(stderr, "op %d done: %lld blocks\n", op, (long long) nb);
    }
    std::fclose(in);
    std::fclose(out);
    std::printf("OK\n");
    return 0;
}
