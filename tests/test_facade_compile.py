"""Compile-time behaviour of the C++ facade (include/se8/supereight.hpp), checked with g++
-fsyntax-only on the CPU: the reference's call forms compile, and the misuses the facade rules out
by type do not.

* integrateDepth(ts, measurements[, updated_octants]) (map_integrator.hpp:78-81) and the stream
  overload integrateDepth(ts, measurements, se::Stream, updated_octants) compile; a raw pointer in
  the third position that is not a std::set<const se::OctantBase*>* (a hipStream_t, a void*) does not.
* TSDFMap<Res::Multi> (map.hpp:623) compiles with integrateDepth and structure(); its field reads,
  mesh, raycast_volume and the tracker do not.
* TSDFDataConfig::storage SE8_STORAGE_MULTIRES on a TSDFMap<Res::Single> is a run-time error, not a
  compile error (it is a value).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")

PRELUDE = """
#include <set>
#include "se8/supereight.hpp"
void f(se::TSDFMap<se::Res::Single>& map, se::TSDFMap<se::Res::Multi>& mmap, const se::PinholeCamera& sensor,
       const se::Image<float>& depth, se::Image<se::Vector3f>& pts, se::Image<se::Vector3f>& nrm,
       se::Image<int8_t>& scale, se8_stream raw_stream)
{
    se::Isometry3f T;
    se::MapIntegrator integrator(map);
    se::MapIntegrator mint(mmap);
    std::set<const se::OctantBase*> updated;
    (void) raw_stream; (void) pts; (void) nrm; (void) scale; (void) updated;
"""


def _compiles(body, tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PRELUDE + body + "\n}\n")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                         capture_output=True, text=True, timeout=120)
    return res.returncode == 0, res.stderr


GOOD = {
    "two_args": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}});",
    "updated_octants": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, &updated);",
    "nullptr_updated": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, nullptr);",
    "stream_overload": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, "
                       "se::Stream(raw_stream), &updated);",
    "stream_only": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, "
                   "se::Stream(raw_stream));",
    "raycast_stream": "se::raycaster::raycast_volume(map, sensor, T, pts, nrm, scale, se::Stream(raw_stream));",
    "multi_integrate": "mint.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, &updated);",
    "multi_structure": "auto q = mmap.structure(); (void) q;",
    "octant_fields": "for (const se::OctantBase* o : updated) { int x = o->coord.x() + o->timestamp + o->getSize();"
                     " bool b = o->is_block; (void) x; (void) b; }",
}

BAD = {
    "raw_stream_third": "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, raw_stream);",
    "void_ptr_third": "void* p = nullptr; "
                      "integrator.integrateDepth(0, se::Measurements{se::Measurement{depth, sensor, T}}, p);",
    "raycast_raw_stream": "se::raycaster::raycast_volume(map, sensor, T, pts, nrm, scale, raw_stream);",
    "multi_getdata": "auto d = mmap.getData(se::Vector3f()); (void) d;",
    "multi_mesh": "auto m = mmap.mesh(); (void) m;",
    "multi_raycast": "se::raycaster::raycast_volume(mmap, sensor, T, pts, nrm, scale);",
    "multi_tracker": "se::Tracker tracker(mmap, sensor);",
}


@pytest.mark.parametrize("name", sorted(GOOD))
def test_reference_call_forms_compile(name, tmp_path):
    ok, err = _compiles(GOOD[name], tmp_path)
    assert ok, err


@pytest.mark.parametrize("name", sorted(BAD))
def test_misuse_does_not_compile(name, tmp_path):
    ok, err = _compiles(BAD[name], tmp_path)
    assert not ok
    print(err.splitlines()[0] if err else "")
