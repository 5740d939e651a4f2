// supereight.hpp - header-only C++ facade with the reference's API for the single-resolution
// TSDF path, over the C-ABI of libse8.so (include/se8.h).
//
// Reference call sites (app/src/main.cpp:128, 137-147, 154-160) compile against these names:
//   se::TSDFMap<se::Res::Single>          include/se/map/map.hpp:82-95, 621-622
//   se::TSDFMap<se::Res::Multi>           include/se/map/map.hpp:623 (integration and structure)
//   se::MapIntegrator<MapT>::integrateDepth include/se/integrator/map_integrator.hpp:66-81
//   se::OctantBase (updated_octants)       include/se/map/octant/octant.hpp:19-120
//   se::raycaster::raycast_volume          include/se/map/raycaster.hpp:56-64
//   se::Tracker<MapT, SensorT>::track      include/se/tracker/tracker.hpp:42-73
//   se::PinholeCamera, se::Image<T>        include/se/sensor/pinhole_camera.hpp:74-78,
//                                          include/se/image/image.hpp:18-111
//   se::OusterLidar                        include/se/sensor/ouster_lidar.hpp:12-80
//   se::Map::mesh / saveMesh, se::Triangle include/se/map/map.hpp:396-413,
//                                          include/se/map/algorithms/mesh.hpp:48-103
//   se::preprocessor::downsample_depth     src/map/preprocessor.cpp:19-84
//   se::Reader / create_reader (app)       app/include/reader_base.hpp:45-120, reader.hpp:22
// Images live in host memory as in the reference; each call stages them through device buffers
// owned by the object (or a per-thread scratch for raycast_volume).  Value types below provide the
// subset of Eigen's interface the call sites use (x(), y(), z(), translation(), matrix data).
// Device-side failures throw std::runtime_error carrying se8_last_error().  Calls that take a HIP
// stream take it as se::Stream (default: the null stream), a type no other argument converts to.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <set>
#include <stdexcept>
#include <utility>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../se8.h"

namespace se {

// A HIP stream (hipStream_t) for the calls that accept one; explicit so that no pointer argument
// converts to it by accident.
struct Stream {
    se8_stream s = nullptr;
    Stream() = default;
    explicit Stream(se8_stream st) : s(st) {}
};

namespace detail {
inline void check(int rc, const char* what)
{
    if (rc != SE8_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + se8_last_error());
}

class DeviceBuffer {
    public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer()
    {
        if (ptr_)
            se8_device_free(ptr_);
    }
    void* get(size_t bytes)
    {
        if (bytes > bytes_) {
            if (ptr_)
                check(se8_device_free(ptr_), "se8_device_free");
            ptr_ = nullptr;
            check(se8_device_alloc(bytes, &ptr_), "se8_device_alloc");
            bytes_ = bytes;
        }
        return ptr_;
    }

    private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};
} // namespace detail

struct Vector3f {
    float v[3] = {0.f, 0.f, 0.f};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    static Vector3f Zero() { return Vector3f(); }
    static Vector3f Constant(float c) { return Vector3f(c, c, c); }
    float& x() { return v[0]; }
    float& y() { return v[1]; }
    float& z() { return v[2]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    float maxCoeff() const { return std::fmax(v[0], std::fmax(v[1], v[2])); }
    float norm() const { return std::sqrt(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2])); }
    const float* data() const { return v; }
    bool operator==(const Vector3f& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};
static_assert(sizeof(Vector3f) == 3 * sizeof(float), "Vector3f must be 3 packed floats");

struct Vector3i {
    int v[3] = {0, 0, 0};
    Vector3i() = default;
    Vector3i(int x, int y, int z) : v{x, y, z} {}
    int x() const { return v[0]; }
    int y() const { return v[1]; }
    int z() const { return v[2]; }
};

// Rigid transform stored as a row-major 4x4 matrix.
struct Isometry3f {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    static Isometry3f Identity() { return Isometry3f(); }
    float& operator()(int r, int c) { return m[4 * r + c]; }
    float operator()(int r, int c) const { return m[4 * r + c]; }
    Vector3f translation() const { return Vector3f(m[3], m[7], m[11]); }
    float* data() { return m; }
    const float* data() const { return m; }
};

template<typename T>
class Image {
    public:
    Image(int width, int height, const T& init = T()) : w_(width), h_(height), data_((size_t) width * height, init) {}
    int width() const { return w_; }
    int height() const { return h_; }
    size_t size() const { return data_.size(); }
    T& operator[](size_t i) { return data_[i]; }
    const T& operator[](size_t i) const { return data_[i]; }
    T& operator()(int x, int y) { return data_[x + (size_t) y * w_]; }
    const T& operator()(int x, int y) const { return data_[x + (size_t) y * w_]; }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }

    private:
    int w_, h_;
    std::vector<T> data_;
};

// se::MeshFace<3> / se::Triangle (include/se/map/algorithms/mesh.hpp:48-66, 88-103): three
// vertices and the octree scale the face came from (always 0 for a single-resolution map).
struct Triangle {
    Vector3f vertexes[3];
    std::int8_t scale = 0;
    static constexpr size_t num_vertexes = 3;
};
typedef std::vector<Triangle> TriangleMesh;

// se::Quad (mesh.hpp:79-82): the octree structure's faces (scale 0 for blocks, 7 for nodes).
struct Quad {
    Vector3f vertexes[4];
    std::int8_t scale = 0;
    static constexpr size_t num_vertexes = 4;
};
typedef std::vector<Quad> QuadMesh;

namespace io {
// se::io::has_supported_mesh_extension (include/se/map/io/mesh_io.hpp:24-29)
inline bool has_supported_mesh_extension(const std::string& filename)
{
    return se8_has_supported_mesh_extension(filename.c_str()) != 0;
}

// se::io::save_mesh (include/se/map/io/mesh_io.hpp:44-46, impl/mesh_io_impl.hpp:14-230): .ply, .vtk
// or .obj by the extension, each vertex mapped by T_OM; 0 on success, 1 if the file cannot be
// written, std::invalid_argument for another extension.
template<typename FaceT>
inline int save_mesh(const std::vector<FaceT>& mesh, const std::string& filename,
                     const Isometry3f& T_OM = Isometry3f::Identity())
{
    if (!has_supported_mesh_extension(filename))
        throw std::invalid_argument("unknown file extension: " + filename);
    constexpr int nv = (int) FaceT::num_vertexes;
    std::vector<float> flat(3 * nv * mesh.size());
    for (size_t f = 0; f < mesh.size(); f++)
        for (int v = 0; v < nv; v++)
            std::memcpy(&flat[3 * nv * f + 3 * v], mesh[f].vertexes[v].data(), sizeof(float) * 3);
    std::vector<int8_t> scales(mesh.size());
    for (size_t f = 0; f < mesh.size(); f++)
        scales[f] = mesh[f].scale;
    const int rc = se8_save_mesh(flat.data(), (int64_t) mesh.size(), nv, scales.data(), filename.c_str(), T_OM.data());
    if (rc == SE8_ERR_IO)
        return 1;
    detail::check(rc, "se8_save_mesh");
    return 0;
}
} // namespace io

enum class Res { Single, Multi };

// se::OctantBase (octant.hpp:19-120) as integrateDepth reports it in updated_octants: an
// allocated block's coordinates in voxels (its vertex closest to the origin), its key (code << 5 |
// scale, key_util_impl.hpp:74-101) and its timestamp when reported.  The map owns one handle per
// block for its lifetime, so a pointer names the same block across integrations, as the
// reference's octant pointers do.
class OctantBase {
    public:
    explicit OctantBase(uint64_t key) : coord(decode(key, 0), decode(key, 1), decode(key, 2)), key_(key) {}
    OctantBase(const OctantBase&) = delete;
    OctantBase& operator=(const OctantBase&) = delete;

    const Vector3i coord;
    int timestamp = -1;
    const bool is_block = true;

    uint64_t getKey() const { return key_; }
    int getScale() const { return (int) (key_ & 31u); }
    int getSize() const { return 1 << getScale(); }

    private:
    static int decode(uint64_t key, int axis)
    {
        const uint64_t code = key >> 5;
        int v = 0;
        for (int b = 0; b < 21; b++)
            v |= (int) ((code >> (3 * b + axis)) & 1u) << b;
        return v;
    }
    uint64_t key_;
};

template<typename MapT>
class MapIntegrator;

// se::FieldData<Field::TSDF>::Config plus the storage precision and pool capacity
struct TSDFDataConfig {
    float truncation_boundary_factor = 8.f;
    float max_weight = 100.f;
    int storage = SE8_STORAGE_F32;
    int64_t max_blocks = 0;
};

class PinholeCamera {
    public:
    struct Config {
        int width = 0;
        int height = 0;
        float near_plane = 0.01f;
        float far_plane = 10.f;
        float fx = std::nanf("");
        float fy = std::nanf("");
        float cx = std::nanf("");
        float cy = std::nanf("");
    };
    // PinholeCamera(Config, downsampling) (pinhole_camera.cpp:29-39, 88-92)
    explicit PinholeCamera(const Config& c, float dsf = 1.f)
    {
        p_.width = (int) ((float) c.width / dsf);
        p_.height = (int) ((float) c.height / dsf);
        p_.fx = c.fx / dsf;
        p_.fy = c.fy / dsf;
        p_.cx = dsf == 1.f ? c.cx : (c.cx + 0.5f) / dsf - 0.5f;
        p_.cy = dsf == 1.f ? c.cy : (c.cy + 0.5f) / dsf - 0.5f;
        p_.near_plane = c.near_plane;
        p_.far_plane = c.far_plane;
    }
    int imageWidth() const { return p_.width; }
    int imageHeight() const { return p_.height; }
    const se8_pinhole& c() const { return p_; }

    private:
    se8_pinhole p_{};
};

// se::OusterLidar (ouster_lidar.hpp:12-80, ouster_lidar.cpp:59-85): one elevation and one azimuth
// offset per beam in degrees (row 0 the highest beam, elevations strictly descending).  The beam
// tables live on the device for the object's lifetime.
class OusterLidar {
    public:
    struct Config {
        int width = 0;
        int height = 0;
        float near_plane = 0.01f;
        float far_plane = 10.f;
        std::vector<float> beam_elevation_angles;
        std::vector<float> beam_azimuth_angles;
    };
    explicit OusterLidar(const Config& c) : w_(c.width), h_(c.height)
    {
        if ((int) c.beam_elevation_angles.size() != c.height || (int) c.beam_azimuth_angles.size() != c.height)
            throw std::invalid_argument("OusterLidar: one elevation and one azimuth angle per beam are required");
        const se8_lidar_config lc{c.width, c.height, c.near_plane, c.far_plane, c.beam_elevation_angles.data(),
                                  c.beam_azimuth_angles.data()};
        detail::check(se8_lidar_create(&lc, &l_), "se8_lidar_create");
    }
    OusterLidar(const OusterLidar&) = delete;
    OusterLidar& operator=(const OusterLidar&) = delete;
    ~OusterLidar() { se8_lidar_destroy(l_); }
    int imageWidth() const { return w_; }
    int imageHeight() const { return h_; }
    const se8_lidar* handle() const { return l_; }

    private:
    int w_, h_;
    se8_lidar* l_ = nullptr;
};

namespace detail {
inline int integrate(se8_map* m, int ts, const float* dev, const PinholeCamera& s, const float* T, se8_stream st)
{
    return se8_integrate_depth(m, ts, dev, &s.c(), T, st);
}
inline int integrate(se8_map* m, int ts, const float* dev, const OusterLidar& s, const float* T, se8_stream st)
{
    return se8_integrate_range(m, ts, dev, s.handle(), T, st);
}
} // namespace detail

// TSDFMap<Res::Single> and TSDFMap<Res::Multi> (map.hpp:621-623).  The multi-resolution map
// (TSDFDataConfig::storage SE8_STORAGE_MULTIRES, set by the constructor) is integrated with a
// pinhole camera and read through structure(), saveStructure() and se8_dump_blocks_multires on
// handle(); the members that read the field (getData, getFieldInterp, getFieldGrad, mesh,
// saveFieldSlices) and the raycaster and tracker are single-resolution only and do not compile
// for it.
template<Res ResT = Res::Single>
class TSDFMap {
    public:
    static constexpr Res res = ResT;

    struct Config {
        Vector3f dim = Vector3f::Constant(10.f);
        float res = 0.1f;
        Vector3f t_MW = Vector3f::Constant(5.f); // T_MW is a translation (dim/2 by default)
    };

    TSDFMap(const Vector3f& dim, float res, const TSDFDataConfig& data_cfg = TSDFDataConfig())
        : TSDFMap(Config{dim, res, Vector3f(dim.x() / 2, dim.y() / 2, dim.z() / 2)}, data_cfg)
    {
    }

    TSDFMap(const Config& cfg, const TSDFDataConfig& data_cfg = TSDFDataConfig()) : res_(cfg.res)
    {
        int storage = data_cfg.storage;
        if (ResT == Res::Multi) {
            if (storage != SE8_STORAGE_F32 && storage != SE8_STORAGE_MULTIRES)
                throw std::invalid_argument("TSDFMap<Res::Multi>: the multi-resolution map stores FP32 data");
            storage = SE8_STORAGE_MULTIRES;
        }
        else if (storage == SE8_STORAGE_MULTIRES) {
            throw std::invalid_argument("TSDFMap<Res::Single>: use TSDFMap<Res::Multi> for a multi-resolution map");
        }
        se8_map_config m{{cfg.dim.x(), cfg.dim.y(), cfg.dim.z()}, cfg.res, {cfg.t_MW.x(), cfg.t_MW.y(), cfg.t_MW.z()}};
        se8_tsdf_config t{data_cfg.truncation_boundary_factor, data_cfg.max_weight, storage, data_cfg.max_blocks};
        detail::check(se8_map_create(&m, &t, &h_), "se8_map_create");
        dim_ = cfg.dim;
    }
    TSDFMap(const TSDFMap&) = delete;
    TSDFMap& operator=(const TSDFMap&) = delete;
    ~TSDFMap() { se8_map_destroy(h_); }

    float getRes() const { return res_; }
    Vector3f getDim() const
    {
        const se8_map_info i = info();
        return Vector3f::Constant(i.dim);
    }
    int getOctreeSize() const { return info().octree_size; }
    se8_map_info info() const
    {
        se8_map_info i{};
        detail::check(se8_map_get_info(h_, &i), "se8_map_get_info");
        return i;
    }
    // getData (map_impl.hpp:68-86): {tsdf, weight}
    std::pair<float, float> getData(const Vector3f& p) const
    {
        static_assert(ResT == Res::Single, "getData: single-resolution maps only");
        float t = 1.f, w = 0.f;
        detail::check(se8_get_data(h_, p.data(), 1, &t, &w), "se8_get_data");
        return {t, w};
    }
    // getFieldInterp (map_impl.hpp:154-172)
    std::optional<float> getFieldInterp(const Vector3f& p) const
    {
        static_assert(ResT == Res::Single, "getFieldInterp: single-resolution maps only");
        float v = 0.f;
        uint8_t ok = 0;
        detail::check(se8_field_interp(h_, p.data(), 1, &v, &ok), "se8_field_interp");
        return ok ? std::optional<float>(v) : std::nullopt;
    }
    // getFieldGrad (map_impl.hpp:239-268)
    std::optional<Vector3f> getFieldGrad(const Vector3f& p) const
    {
        static_assert(ResT == Res::Single, "getFieldGrad: single-resolution maps only");
        Vector3f g;
        uint8_t ok = 0;
        detail::check(se8_field_grad(h_, p.data(), 1, g.v, &ok), "se8_field_grad");
        return ok ? std::optional<Vector3f>(g) : std::nullopt;
    }
    typedef TriangleMesh SurfaceMesh;

    // mesh (map.hpp:396-404, map_impl.hpp:452-458): the surface in frame O (metres), vertices
    // mapped by T_OW; min_desired_scale is unused by a single-resolution map.  The faces are
    // meshed on the device and copied back.
    SurfaceMesh mesh(const Isometry3f& T_OW = Isometry3f::Identity(), const int min_desired_scale = 0) const
    {
        static_assert(ResT == Res::Single, "mesh: single-resolution maps only");
        (void) min_desired_scale;
        int64_t n = 0;
        detail::check(se8_mesh(h_, T_OW.data(), nullptr, 0, &n, nullptr), "se8_mesh");
        SurfaceMesh out((size_t) n);
        if (n == 0)
            return out;
        void* dev = mesh_buf_.get(sizeof(float) * 9 * (size_t) n);
        detail::check(se8_mesh(h_, T_OW.data(), (float*) dev, n, &n, nullptr), "se8_mesh");
        std::vector<float> host(9 * (size_t) n);
        detail::check(se8_copy(host.data(), dev, sizeof(float) * host.size(), SE8_COPY_D2H, nullptr), "se8_copy");
        detail::check(se8_stream_synchronize(nullptr), "se8_stream_synchronize");
        for (size_t f = 0; f < out.size(); f++)
            for (int v = 0; v < 3; v++)
                out[f].vertexes[v] = Vector3f(host[9 * f + 3 * v], host[9 * f + 3 * v + 1], host[9 * f + 3 * v + 2]);
        return out;
    }

    // saveMesh (map.hpp:406-413, map_impl.hpp:462-469): .ply, .vtk or .obj by the extension;
    // 0 on success, non-zero if the file cannot be written; std::invalid_argument for another
    // extension, as io::save_mesh throws (mesh_io_impl.hpp:14-29).
    int saveMesh(const std::string& filename, const Isometry3f& T_OW = Isometry3f::Identity(),
                 const int min_desired_scale = 0) const
    {
        if (!io::has_supported_mesh_extension(filename))
            throw std::invalid_argument("unknown file extension: " + filename);
        return io::save_mesh(mesh(T_OW, min_desired_scale), filename);
    }

    typedef QuadMesh StructureMesh;

    // structure (map.hpp:415-421, map_impl.hpp:473-479): six quads per octant (only the leaves by
    // default), vertices mapped by T_OW, metres.
    StructureMesh structure(const Isometry3f& T_OW = Isometry3f::Identity(), const bool only_leaves = true) const
    {
        int64_t n = 0;
        detail::check(se8_structure(h_, T_OW.data(), only_leaves, nullptr, nullptr, 0, &n, nullptr), "se8_structure");
        StructureMesh out((size_t) n);
        if (n == 0)
            return out;
        void* dq = mesh_buf_.get(sizeof(float) * 12 * (size_t) n + (size_t) n);
        int8_t* ds = (int8_t*) ((char*) dq + sizeof(float) * 12 * (size_t) n);
        detail::check(se8_structure(h_, T_OW.data(), only_leaves, (float*) dq, ds, n, &n, nullptr), "se8_structure");
        std::vector<float> q(12 * (size_t) n);
        std::vector<int8_t> sc((size_t) n);
        detail::check(se8_copy(q.data(), dq, sizeof(float) * q.size(), SE8_COPY_D2H, nullptr), "se8_copy");
        detail::check(se8_copy(sc.data(), ds, sc.size(), SE8_COPY_D2H, nullptr), "se8_copy");
        detail::check(se8_stream_synchronize(nullptr), "se8_stream_synchronize");
        for (size_t f = 0; f < out.size(); f++) {
            for (int v = 0; v < 4; v++)
                out[f].vertexes[v] = Vector3f(q[12 * f + 3 * v], q[12 * f + 3 * v + 1], q[12 * f + 3 * v + 2]);
            out[f].scale = sc[f];
        }
        return out;
    }

    // saveStructure (map.hpp:422-430, map_impl.hpp:484-490)
    int saveStructure(const std::string& filename, const Isometry3f& T_OW = Isometry3f::Identity(),
                      const bool only_leaves = true) const
    {
        if (!io::has_supported_mesh_extension(filename))
            throw std::invalid_argument("unknown file extension: " + filename);
        return io::save_mesh(structure(T_OW, only_leaves), filename);
    }

    // saveFieldSlices (map.hpp:309-325, map_impl.hpp:272-308): empty names skip a plane.
    int saveFieldSlices(const std::string& filename_x, const std::string& filename_y, const std::string& filename_z,
                        const Vector3f& point_W) const
    {
        static_assert(ResT == Res::Single, "saveFieldSlices: single-resolution maps only");
        detail::check(se8_save_field_slices(h_, filename_x.c_str(), filename_y.c_str(), filename_z.c_str(),
                                            point_W.data()),
                      "se8_save_field_slices");
        return 0;
    }

    se8_map* handle() const { return h_; }

    private:
    template<typename MapT>
    friend class MapIntegrator;

    // Inserts the handles of the blocks the last integration updated (se8_read_update_list) into
    // out, stamped with the integration's timestamp.
    void reportUpdated(int timestamp, std::set<const OctantBase*>& out)
    {
        int64_t n = 0;
        detail::check(se8_read_update_list(h_, nullptr, 0, &n), "se8_read_update_list");
        std::vector<uint64_t> keys((size_t) n);
        if (n > 0)
            detail::check(se8_read_update_list(h_, keys.data(), n, &n), "se8_read_update_list");
        for (const uint64_t k : keys) {
            std::unique_ptr<OctantBase>& h = handles_[k];
            if (!h)
                h = std::make_unique<OctantBase>(k);
            h->timestamp = timestamp;
            out.insert(h.get());
        }
    }

    se8_map* h_ = nullptr;
    float res_;
    Vector3f dim_;
    mutable detail::DeviceBuffer mesh_buf_;
    std::unordered_map<uint64_t, std::unique_ptr<OctantBase>> handles_;
};

namespace preprocessor {
// downsample_depth (preprocessor.cpp:19-84): median of each ratio x ratio region of valid depths
// (>= 1e-5) on the device; returns the index map (input pixel of each output pixel).
inline Image<size_t> downsample_depth(const Image<float>& input_depth_img, Image<float>& output_depth_img)
{
    const size_t n_in = input_depth_img.size(), n_out = output_depth_img.size();
    detail::DeviceBuffer din, dout, dmap;
    void* pin = din.get(sizeof(float) * n_in);
    void* pout = dout.get(sizeof(float) * n_out);
    void* pmap = dmap.get(sizeof(uint64_t) * n_out);
    detail::check(se8_copy(pin, input_depth_img.data(), sizeof(float) * n_in, SE8_COPY_H2D, nullptr), "se8_copy");
    detail::check(se8_downsample_depth((const float*) pin, input_depth_img.width(), input_depth_img.height(),
                                       (float*) pout, output_depth_img.width(), output_depth_img.height(),
                                       (uint64_t*) pmap, nullptr),
                  "se8_downsample_depth");
    Image<size_t> map(output_depth_img.width(), output_depth_img.height());
    static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t must be 64-bit");
    detail::check(se8_copy(output_depth_img.data(), pout, sizeof(float) * n_out, SE8_COPY_D2H, nullptr), "se8_copy");
    detail::check(se8_copy(map.data(), pmap, sizeof(uint64_t) * n_out, SE8_COPY_D2H, nullptr), "se8_copy");
    detail::check(se8_stream_synchronize(nullptr), "se8_stream_synchronize");
    return map;
}
} // namespace preprocessor

// Dataset readers (app/include/reader_base.hpp, reader_raw.hpp, reader_tum.hpp) over the library's
// host readers: depth and ground-truth poses (colour is not read).
enum class ReaderType { RAW = SE8_READER_RAW, TUM = SE8_READER_TUM };
enum class ReaderStatus : int { ok = SE8_READER_OK, skip = SE8_READER_SKIP, eof = SE8_READER_EOF, error = SE8_READER_ERROR };

class Reader {
    public:
    struct Config {
        ReaderType reader_type = ReaderType::RAW;
        std::string sequence_path;
        std::string ground_truth_file;
        float inverse_scale = 0.0f;
    };

    explicit Reader(const Config& c) : cfg_(c)
    {
        se8_reader_config rc{(int) c.reader_type, cfg_.sequence_path.c_str(), cfg_.ground_truth_file.c_str(),
                             c.inverse_scale};
        if (se8_reader_create(&rc, &h_) != SE8_OK) {
            h_ = nullptr;
            status_ = ReaderStatus::error; // as the reference's reader starts on an unreadable dataset
            return;
        }
        int64_t n = 0;
        detail::check(se8_reader_info(h_, &w_, &hgt_, &n, nullptr), "se8_reader_info");
        num_frames_ = (size_t) n;
    }
    Reader(const Reader&) = delete;
    Reader& operator=(const Reader&) = delete;
    ~Reader() { se8_reader_destroy(h_); }

    bool good() const { return status_ == ReaderStatus::ok || status_ == ReaderStatus::skip; }
    ReaderStatus status() const { return status_; }
    size_t numFrames() const { return num_frames_; }
    int depthWidth() const { return w_; }
    int depthHeight() const { return hgt_; }

    ReaderStatus nextData(Image<float>& depth_image) { return next(depth_image, nullptr); }
    ReaderStatus nextData(Image<float>& depth_image, Isometry3f& T_WB) { return next(depth_image, &T_WB); }
    void restart()
    {
        if (h_) {
            detail::check(se8_reader_restart(h_), "se8_reader_restart");
            status_ = ReaderStatus::ok;
        }
    }

    private:
    ReaderStatus next(Image<float>& depth_image, Isometry3f* T)
    {
        if (!h_)
            return status_;
        if (depth_image.width() != w_ || depth_image.height() != hgt_)
            depth_image = Image<float>(w_, hgt_);
        int st = SE8_READER_ERROR;
        detail::check(se8_reader_next(h_, depth_image.data(), T ? T->data() : nullptr, &st), "se8_reader_next");
        status_ = (ReaderStatus) st;
        return status_;
    }

    Config cfg_;
    se8_reader* h_ = nullptr;
    int w_ = 0, hgt_ = 0;
    size_t num_frames_ = 0;
    ReaderStatus status_ = ReaderStatus::ok;
};

// create_reader (app/include/reader.hpp:22)
inline Reader* create_reader(const Reader::Config& config)
{
    return new Reader(config);
}

template<typename SensorT>
struct Measurement {
    const Image<float>& image;
    const SensorT& sensor;
    const Isometry3f T_WC;
};

// deduction guides as in measurement.hpp:26-27 and :52-58
template<typename SensorT>
Measurement(const Image<float>&, const SensorT&, const Isometry3f&) -> Measurement<SensorT>;

template<typename SensorT>
struct Measurements {
    Measurement<SensorT> depth;
};

template<typename SensorT>
Measurements(Measurement<SensorT>) -> Measurements<SensorT>;

template<typename MapT>
class MapIntegrator {
    public:
    explicit MapIntegrator(MapT& map) : map_(map) {}

    // integrateDepth(timestamp, measurements, updated_octants) (map_integrator.hpp:78-81): when
    // updated_octants is non-null the handles of the blocks this integration updated (the frustum
    // blocks and the blocks it allocated, map_integrator_impl.hpp:77-80) are inserted into it.
    template<typename SensorT>
    void integrateDepth(int timestamp, const Measurements<SensorT>& m,
                        std::set<const OctantBase*>* const updated_octants = nullptr)
    {
        integrateDepth(timestamp, m, Stream(), updated_octants);
    }

    // The same on a HIP stream (returns once the integration has finished on it).
    template<typename SensorT>
    void integrateDepth(int timestamp, const Measurements<SensorT>& m, Stream stream,
                        std::set<const OctantBase*>* const updated_octants = nullptr)
    {
        static_assert(MapT::res == Res::Single || std::is_same<SensorT, PinholeCamera>::value,
                      "the multi-resolution map is integrated with a pinhole camera");
        const Image<float>& d = m.depth.image;
        const size_t bytes = sizeof(float) * d.size();
        void* dev = depth_.get(bytes);
        detail::check(se8_copy(dev, d.data(), bytes, SE8_COPY_H2D, stream.s), "se8_copy");
        detail::check(detail::integrate(map_.handle(), timestamp, (const float*) dev, m.depth.sensor,
                                        m.depth.T_WC.data(), stream.s),
                      "integrateDepth");
        detail::check(se8_stream_synchronize(stream.s), "se8_stream_synchronize");
        if (updated_octants)
            map_.reportUpdated(timestamp, *updated_octants);
    }

    private:
    MapT& map_;
    detail::DeviceBuffer depth_;
};

namespace raycaster {
// raycast_volume(map, sensor, T_WS, points, normals, scale) (raycaster.hpp:56-64)
template<typename MapT, typename SensorT>
void raycast_volume(const MapT& map, const SensorT& sensor, const Isometry3f& T_WS, Image<Vector3f>& points_W,
                    Image<Vector3f>& normals_W, Image<int8_t>& scale, Stream stream = Stream())
{
    static_assert(MapT::res == Res::Single, "raycast_volume: single-resolution maps only");
    thread_local detail::DeviceBuffer pts, nrm, scl;
    const size_t n = points_W.size();
    void* dp = pts.get(sizeof(float) * 3 * n);
    void* dn = nrm.get(sizeof(float) * 3 * n);
    void* ds = scl.get(n);
    detail::check(se8_copy(ds, scale.data(), n, SE8_COPY_H2D, stream.s), "se8_copy");
    detail::check(se8_raycast(map.handle(), &sensor.c(), T_WS.data(), (float*) dp, (float*) dn, (int8_t*) ds, stream.s),
                  "se8_raycast");
    detail::check(se8_copy(points_W.data(), dp, sizeof(float) * 3 * n, SE8_COPY_D2H, stream.s), "se8_copy");
    detail::check(se8_copy(normals_W.data(), dn, sizeof(float) * 3 * n, SE8_COPY_D2H, stream.s), "se8_copy");
    detail::check(se8_copy(scale.data(), ds, n, SE8_COPY_D2H, stream.s), "se8_copy");
    detail::check(se8_stream_synchronize(stream.s), "se8_stream_synchronize");
}
} // namespace raycaster

// TrackerConfig (tracker.hpp:22-33) + the compensated-reduction switch
struct TrackerConfig {
    std::vector<int> iterations{10, 5, 4};
    float dist_threshold = 0.1f;
    float normal_threshold = 0.8f;
    float track_threshold = 0.15f;
    float icp_threshold = 0.00001f;
    bool compensated = false;
};

template<typename MapT, typename SensorT>
class Tracker {
    public:
    static_assert(MapT::res == Res::Single, "se::Tracker: single-resolution maps only");

    Tracker(MapT& map, const SensorT& sensor, const TrackerConfig config = TrackerConfig())
        : map_(map), sensor_(sensor)
    {
        // one pyramid level per entry of iterations (tracker_impl.hpp:50-84)
        if (config.iterations.empty() || config.iterations.size() > SE8_TRACKER_MAX_LEVELS)
            throw std::invalid_argument("se::Tracker: 1 to " + std::to_string(SE8_TRACKER_MAX_LEVELS)
                                        + " pyramid levels are supported");
        se8_tracker_config c{};
        for (size_t l = 0; l < config.iterations.size(); l++)
            c.iterations[l] = config.iterations[l];
        c.num_levels = (int) config.iterations.size();
        c.dist_threshold = config.dist_threshold;
        c.normal_threshold = config.normal_threshold;
        c.track_threshold = config.track_threshold;
        c.icp_threshold = config.icp_threshold;
        c.compensated = config.compensated ? 1 : 0;
        detail::check(se8_tracker_create(&sensor.c(), &c, &h_), "se8_tracker_create");
    }
    Tracker(const Tracker&) = delete;
    Tracker& operator=(const Tracker&) = delete;
    ~Tracker() { se8_tracker_destroy(h_); }

    // track(depth, T_WS): raycasts the map at T_WS first (tracker_impl.hpp:17-26)
    bool track(const Image<float>& depth, Isometry3f& T_WS, Stream stream = Stream())
    {
        void* dd = stage_depth(depth, stream);
        int ok = 0;
        detail::check(se8_track_raycast(h_, map_.handle(), (const float*) dd, T_WS.data(), &ok, stream.s),
                      "se8_track_raycast");
        return ok != 0;
    }

    // track(depth, T_WS, points_W, normals_W) (tracker_impl.hpp:30-116)
    bool track(const Image<float>& depth, Isometry3f& T_WS, Image<Vector3f>& points_W, Image<Vector3f>& normals_W,
               Stream stream = Stream())
    {
        void* dd = stage_depth(depth, stream);
        const size_t n = points_W.size();
        void* dp = pts_.get(sizeof(float) * 3 * n);
        void* dn = nrm_.get(sizeof(float) * 3 * n);
        detail::check(se8_copy(dp, points_W.data(), sizeof(float) * 3 * n, SE8_COPY_H2D, stream.s), "se8_copy");
        detail::check(se8_copy(dn, normals_W.data(), sizeof(float) * 3 * n, SE8_COPY_H2D, stream.s), "se8_copy");
        int ok = 0;
        detail::check(se8_track(h_, (const float*) dd, T_WS.data(), (const float*) dp, (const float*) dn, &ok,
                                nullptr, stream.s),
                      "se8_track");
        return ok != 0;
    }

    private:
    void* stage_depth(const Image<float>& depth, Stream stream)
    {
        const size_t bytes = sizeof(float) * depth.size();
        void* dd = depth_.get(bytes);
        detail::check(se8_copy(dd, depth.data(), bytes, SE8_COPY_H2D, stream.s), "se8_copy");
        return dd;
    }

    MapT& map_;
    const SensorT& sensor_;
    se8_tracker* h_ = nullptr;
    detail::DeviceBuffer depth_, pts_, nrm_;
};

} // namespace se
