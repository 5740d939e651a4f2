"""The raycaster's and the carver's voxel-position division (div_by_recip / div_rn in
supereight2_amd/csrc/se8_internal.hpp: a reciprocal multiply and two fma corrections) equals IEEE
division, which the oracle and the reference use (pointToVoxel: point_M / res).  tests/cpp/div_check.cpp
compiles the header as host C++ and compares bit patterns: every x significand against 16 map
resolutions, 8 extreme and 256 random divisors; x over the exponent range [-60, 60] of both signs;
and div_rn's guarded inputs (zero, tiny, subnormal, huge, infinite, NaN), which take the division.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "div_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "div_check")
DEPS = [SRC, os.path.join(ROOT, "supereight2_amd", "csrc", "se8_internal.hpp")]


def test_division_by_reciprocal_equals_ieee_division():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(os.path.getmtime(d) for d in DEPS):
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-Wall", "-Wno-unknown-pragmas",
               "-I", os.path.join(ROOT, "tests", "cpp", "simt"), "-I", os.path.join(ROOT, "supereight2_amd", "csrc"),
               "-I", os.path.join(ROOT, "include"), SRC, "-o", BIN]
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert res.returncode == 0, res.stderr
    res = subprocess.run([BIN, "256"], capture_output=True, text=True, timeout=600)
    print(res.stdout[-2000:])
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout[-2000:]
