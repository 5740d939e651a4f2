// Exercises include/se8/supereight.hpp the way the reference app drives the path
// (app/src/main.cpp:126-160): track -> integrateDepth -> raycast_volume per frame, on synthetic
// depth of a room corner (three planes, so point-to-plane ICP is fully constrained) seen by a
// camera that slides 1 cm per frame along x.
// Prints one line per frame: frame, tracking ok, estimated x, ground-truth x, centre hit depth;
// the blocks integrateDepth reported as updated at frames 4 and 5 (the second on se::Stream);
// then integrates the first frame into a TSDFMap<Res::Multi>, and two scans of a spinning lidar
// through se::OusterLidar and prints the field at a ground point and a wall point.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>

#include "se8/supereight.hpp"

namespace {
struct V3 {
    double x, y, z;
};
V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
V3 unit(V3 a)
{
    const double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / n, a.y / n, a.z / n};
}
} // namespace

// The reference app's input steps (main.cpp:104-120) on a TUM sequence: read, downsample by 2,
// integrate at the ground-truth pose; prints the frames read and the valid fraction.
static int run_tum(const std::string& seq)
{
    se::Reader::Config rc;
    rc.reader_type = se::ReaderType::TUM;
    rc.sequence_path = seq;
    rc.ground_truth_file = seq + "/groundtruth.txt";
    se::Reader* reader = se::create_reader(rc);
    se::PinholeCamera::Config cfg;
    cfg.width = 640;
    cfg.height = 480;
    cfg.fx = 481.2f;
    cfg.fy = -480.f;
    cfg.cx = 319.5f;
    cfg.cy = 239.5f;
    cfg.near_plane = 0.4f;
    cfg.far_plane = 6.f;
    const se::PinholeCamera sensor(cfg, 2.f);
    se::TSDFMap<se::Res::Single> map(se::Vector3f::Constant(10.24f), 0.01f);
    se::MapIntegrator integrator(map);
    se::Image<float> input(640, 480), processed(320, 240);
    se::Isometry3f T_WB;
    int frames = 0;
    double valid = 0.0;
    while (reader->nextData(input, T_WB) == se::ReaderStatus::ok) {
        const se::Image<size_t> ds_map = se::preprocessor::downsample_depth(input, processed);
        integrator.integrateDepth(frames, se::Measurements{se::Measurement{processed, sensor, T_WB}});
        size_t n = 0;
        for (size_t i = 0; i < processed.size(); i++)
            n += processed[i] > 0.f;
        valid += (double) n / processed.size();
        (void) ds_map;
        frames++;
    }
    const auto mesh = map.mesh();
    std::printf("tum frames %d status %d valid %.4f blocks %lld faces %zu\n", frames, (int) reader->status(),
                valid / std::max(frames, 1), (long long) map.info().num_blocks, mesh.size());
    delete reader;
    return frames == 3 ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc > 1)
        return run_tum(argv[1]);
    se::PinholeCamera::Config cfg;
    cfg.width = 640;
    cfg.height = 480;
    cfg.fx = cfg.fy = 525.f;
    cfg.cx = 319.5f;
    cfg.cy = 239.5f;
    cfg.near_plane = 0.4f;
    cfg.far_plane = 6.f;
    const se::PinholeCamera sensor(cfg);
    se::TSDFMap<se::Res::Single> map(se::Vector3f::Constant(10.24f), 0.01f);
    se::MapIntegrator integrator(map);
    se::Tracker tracker(map, sensor);
    se::Image<se::Vector3f> points(640, 480), normals(640, 480);
    se::Image<int8_t> scale(640, 480, -1);

    // camera frame: z forward, x right, y down; looking into the (+x, +y, floor) corner of the
    // room [-3,3] x [-2,2] x [0,3]
    const V3 f = unit({1.0, 0.7, -0.35});
    const V3 xr = unit(cross(f, {0.0, 0.0, 1.0}));
    const V3 yd = cross(f, xr);
    const double R[3][3] = {{xr.x, yd.x, f.x}, {xr.y, yd.y, f.y}, {xr.z, yd.z, f.z}};
    const double lo[3] = {-3.0, -2.0, 0.0}, hi[3] = {3.0, 2.0, 3.0};

    se::Isometry3f T_WS = se::Isometry3f::Identity(), T_0;
    se::Image<float> depth_0(640, 480);
    std::set<const se::OctantBase*> updated_4, updated_5;
    int failures = 0;
    double worst = 0.0;
    for (int frame = 0; frame < 6; frame++) {
        const double p[3] = {0.01 * frame, 0.0, 1.4};
        se::Image<float> depth(640, 480);
        for (int v = 0; v < 480; v++)
            for (int u = 0; u < 640; u++) {
                const double dc[3] = {(u - 319.5) / 525.0, (v - 239.5) / 525.0, 1.0};
                double t = 1e9;
                for (int i = 0; i < 3; i++) {
                    const double dw = R[i][0] * dc[0] + R[i][1] * dc[1] + R[i][2] * dc[2];
                    if (dw > 0)
                        t = std::fmin(t, (hi[i] - p[i]) / dw);
                    else if (dw < 0)
                        t = std::fmin(t, (lo[i] - p[i]) / dw);
                }
                depth(u, v) = (float) t; // z-depth, since dc.z = 1
            }
        bool ok = true;
        if (frame > 0)
            ok = tracker.track(depth, T_WS, points, normals);
        else {
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++)
                    T_WS(i, j) = (float) R[i][j];
                T_WS(i, 3) = (float) p[i];
            }
        }
        if (frame == 0) {
            depth_0 = depth;
            T_0 = T_WS;
        }
        if (frame == 4)
            integrator.integrateDepth(frame, se::Measurements{se::Measurement{depth, sensor, T_WS}}, &updated_4);
        else if (frame == 5)
            integrator.integrateDepth(frame, se::Measurements{se::Measurement{depth, sensor, T_WS}}, se::Stream(),
                                      &updated_5);
        else
            integrator.integrateDepth(frame, se::Measurements{se::Measurement{depth, sensor, T_WS}});
        se::raycaster::raycast_volume(map, sensor, T_WS, points, normals, scale);
        const se::Vector3f c = points(320, 240);
        const double dz = std::sqrt((c.x() - p[0]) * (c.x() - p[0]) + (c.y() - p[1]) * (c.y() - p[1])
                                    + (c.z() - p[2]) * (c.z() - p[2]));
        std::printf("frame %d ok %d x %.5f gt %.5f hit_dist %.5f depth_dist %.5f scale %d\n", frame, (int) ok,
                    T_WS(0, 3), p[0], dz,
                    depth(320, 240) * std::sqrt(1.0 + std::pow(0.5 / 525.0, 2) * 2), (int) scale(320, 240));
        worst = std::fmax(worst, std::fabs(T_WS(0, 3) - p[0]));
        if (!ok)
            failures++;
    }
    // updated_octants: blocks (size 8, coordinates on the block grid) stamped with the frame; a block
    // reported at both frames is the same handle
    int bad = 0;
    size_t shared = 0;
    for (const se::OctantBase* o : updated_5) {
        bad += !o->is_block || o->getSize() != 8 || o->timestamp != 5 || o->coord.x() % 8 || o->coord.y() % 8
               || o->coord.z() % 8;
        shared += updated_4.count(o);
    }
    std::printf("updated frame4 %zu frame5 %zu shared %zu bad %d\n", updated_4.size(), updated_5.size(), shared, bad);

    // the multi-resolution map: frame 0 integrated, every block new and updated
    se::TSDFMap<se::Res::Multi> mmap(se::Vector3f::Constant(10.24f), 0.02f);
    se::MapIntegrator mint(mmap);
    std::set<const se::OctantBase*> updated_multi;
    mint.integrateDepth(0, se::Measurements{se::Measurement{depth_0, sensor, T_0}}, &updated_multi);
    std::printf("multires blocks %lld updated %zu quads %zu\n", (long long) mmap.info().num_blocks,
                updated_multi.size(), mmap.structure().size());

    const auto f_interp = map.getFieldInterp(se::Vector3f(points(320, 240).x(), points(320, 240).y(),
                                                          points(320, 240).z()));
    std::printf("interp %d %.5f blocks %lld worst_x_err %.6f\n", f_interp ? 1 : 0, f_interp ? *f_interp : 0.f,
                (long long) map.info().num_blocks, worst);

    // surface mesh of the room: every vertex within 2 voxels of a wall, floor or ceiling
    const se::TSDFMap<se::Res::Single>::SurfaceMesh mesh = map.mesh();
    double mesh_worst = 0.0;
    for (const se::Triangle& t : mesh)
        for (const se::Vector3f& v : t.vertexes) {
            double d = 1e9;
            for (int i = 0; i < 3; i++)
                d = std::fmin(d, std::fmin(std::fabs(v(i) - lo[i]), std::fabs(v(i) - hi[i])));
            mesh_worst = std::fmax(mesh_worst, d);
        }
    const int saved = map.saveMesh("facade_app_mesh.ply") + se::io::save_mesh(mesh, "facade_app_mesh.obj");
    bool threw = false;
    try {
        map.saveMesh("facade_app_mesh.png");
    }
    catch (const std::invalid_argument&) {
        threw = true;
    }
    std::printf("mesh faces %zu worst_plane_dist %.5f saved %d bad_ext_threw %d\n", mesh.size(), mesh_worst, saved,
                (int) threw);
    const auto structure = map.structure();
    const int saved_structure = map.saveStructure("facade_app_structure.vtk", se::Isometry3f::Identity(), false);
    const int saved_slices =
        map.saveFieldSlices("", "", "facade_app_slice_z.vtk", se::Vector3f(0.f, 0.f, 1.4f));
    std::printf("structure quads %zu leaves %lld saved %d slices %d\n", structure.size(),
                (long long) map.info().num_blocks, saved_structure, saved_slices);

    // A 512 x 32 beam lidar at 2 m over the ground plane z = 0 with a wall at x = 12, two scans
    // 0.5 m apart, into a 51.2 m map at 4 cm voxels; the field at a ground point is then ~0.
    se::OusterLidar::Config lc;
    lc.width = 512;
    lc.height = 32;
    lc.near_plane = 0.5f;
    lc.far_plane = 40.f;
    for (int v = 0; v < lc.height; v++) {
        lc.beam_elevation_angles.push_back(10.f - 30.f * (float) v / (float) (lc.height - 1));
        lc.beam_azimuth_angles.push_back(0.f);
    }
    const se::OusterLidar lidar(lc);
    se::TSDFDataConfig dc;
    dc.truncation_boundary_factor = 4.f;
    se::TSDFMap<se::Res::Single> lmap(se::Vector3f::Constant(51.2f), 0.04f, dc);
    se::MapIntegrator lint(lmap);
    const double kPi = 3.14159265358979323846;
    for (int scan = 0; scan < 2; scan++) {
        const double o[3] = {0.5 * scan, 0.0, 2.0};
        se::Image<float> range(lc.width, lc.height);
        for (int v = 0; v < lc.height; v++)
            for (int u = 0; u < lc.width; u++) {
                const double e = lc.beam_elevation_angles[v] * kPi / 180.0;
                const double a = 2.0 * kPi * u / lc.width - kPi;
                const double d[3] = {std::cos(e) * std::cos(a), std::cos(e) * std::sin(a), std::sin(e)};
                double t = 1e9;
                if (d[2] < 0)
                    t = -o[2] / d[2];
                if (d[0] > 0)
                    t = std::fmin(t, (12.0 - o[0]) / d[0]);
                range(u, v) = t <= lc.far_plane ? (float) t : 0.f;
            }
        se::Isometry3f T = se::Isometry3f::Identity();
        T(0, 3) = (float) o[0];
        T(2, 3) = (float) o[2];
        lint.integrateDepth(scan, se::Measurements{se::Measurement{range, lidar, T}});
    }
    const auto g = lmap.getFieldInterp(se::Vector3f(-6.f, 1.f, 0.f));
    const auto w = lmap.getFieldInterp(se::Vector3f(12.f, -2.f, 1.f));
    std::printf("lidar ground %d %.5f wall %d %.5f blocks %lld\n", g ? 1 : 0, g ? *g : 0.f, w ? 1 : 0, w ? *w : 0.f,
                (long long) lmap.info().num_blocks);
    return failures == 0 ? 0 : 1;
}
