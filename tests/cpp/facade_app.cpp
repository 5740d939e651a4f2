// Exercises include/se8/supereight.hpp the way the reference app drives the path
// (app/src/main.cpp:126-160): track -> integrateDepth -> raycast_volume per frame, on a synthetic
// fronto-parallel wall at 2 m seen by a camera that slides 1 cm per frame along x.
// Prints one line per frame: frame, tracking ok, estimated x, ground-truth x, centre hit z.
#include <cstdio>

#include "se8/supereight.hpp"

int main()
{
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
    se::Isometry3f T_WS = se::Isometry3f::Identity();
    int failures = 0;
    for (int frame = 0; frame < 6; frame++) {
        const float gt_x = 0.01f * frame;
        // a wall tilted about y so that x-motion is observable: z = 2 + 0.3 (x_W)
        se::Image<float> depth(640, 480);
        for (int v = 0; v < 480; v++)
            for (int u = 0; u < 640; u++) {
                const float dx = (u - cfg.cx) / cfg.fx;
                // ray x_W = gt_x + t dx, z = t ; plane z = 2 + 0.3 x_W  ->  t = (2 + 0.3 gt_x) / (1 - 0.3 dx)
                depth(u, v) = (2.f + 0.3f * gt_x) / (1.f - 0.3f * dx);
            }
        bool ok = true;
        if (frame > 0)
            ok = tracker.track(depth, T_WS, points, normals);
        else
            T_WS(0, 3) = gt_x;
        integrator.integrateDepth(frame, se::Measurements{se::Measurement{depth, sensor, T_WS}});
        se::raycaster::raycast_volume(map, sensor, T_WS, points, normals, scale);
        const se::Vector3f c = points(320, 240);
        std::printf("frame %d ok %d x %.5f gt %.5f hit_z %.5f scale %d\n", frame, (int) ok, T_WS(0, 3), gt_x, c.z(),
                    (int) scale(320, 240));
        if (!ok)
            failures++;
    }
    const auto f = map.getFieldInterp(se::Vector3f(0.02f, 0.f, 2.006f));
    std::printf("interp %d %.5f blocks %lld\n", f ? 1 : 0, f ? *f : 0.f, (long long) map.info().num_blocks);
    return failures == 0 ? 0 : 1;
}
