"""The header-only C++ facade (include/se8/supereight.hpp) driven like the reference app:
track -> integrateDepth -> raycast_volume per frame on a tilted synthetic wall."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "cpp", "_build", "facade_app")


def test_facade_app_tracks_and_maps():
    assert os.path.exists(APP), "build with __graft_entry__.build() first"
    print("running facade_app", flush=True)
    res = subprocess.run([APP], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    frames = re.findall(r"frame (\d+) ok (\d) x (\S+) gt (\S+) hit_z (\S+) scale (-?\d+)", res.stdout)
    assert len(frames) == 6
    for f, ok, x, gt, hz, sc in frames:
        assert ok == "1"
        assert abs(float(x) - float(gt)) < 2e-3
        assert abs(float(hz) - (2.0 + 0.3 * float(gt))) < 5e-3
        assert sc == "0"
    m = re.search(r"interp (\d) (\S+) blocks (\d+)", res.stdout)
    assert m and m.group(1) == "1" and int(m.group(3)) > 1000
