"""DroneRace raceline projection (dr_drace.h fast_project) against the oracle's
restatement of projectOnTrack / findMinAlongDirection / fastProjectOnTrack
(include/environments/env_dronerace.h:69-209).

The device header's walk (the rollout kernels and the world step both use it)
must equal the oracle's bit for bit.  Cases: positions near three
closed racelines (circle, a wavy loop whose distance along the line has local
minima that the 10-sample margin finds, and a loop with repeated samples, so
that distances tie), the previous progress placed 0..40 samples behind or
ahead of the nearest sample (walks of up to 40 samples), indices that wrap past 0 and 511, the circle's centre (every
distance equal) and the full scan (previous progress < 0).  Host build of the
device header, no GPU.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NS = 512


@pytest.fixture(scope="module")
def hd(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("drace") / "libdrace_project.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip", "-fPIC", "-shared",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(ROOT, "tests", "cpp", "drace_project_host.cpp"), "-o", out], check=True)
    lib = C.CDLL(out)
    lib.hd_drace_project.restype = C.c_float
    lib.hd_drace_project.argtypes = [C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float]
    return lib


def racelines():
    u = np.arange(NS) / NS * 2 * np.pi
    circle = np.stack([10 * np.cos(u), 10 * np.sin(u)], 1)
    r = 10 + 2.5 * np.sin(9 * u)
    wavy = np.stack([r * np.cos(u), 0.7 * r * np.sin(u)], 1)
    rep = circle[(np.arange(NS) // 3) * 3]  # runs of three equal samples
    return {"circle": circle.astype(np.float32), "wavy": wavy.astype(np.float32), "repeated": rep.astype(np.float32)}


def test_projection_matches_the_walk(hd, oracle):
    rng = np.random.default_rng(41)
    n_long = 0
    for name, line in racelines().items():
        tp = np.ascontiguousarray(line.reshape(-1))
        ptr = tp.ctypes.data_as(C.POINTER(C.c_float))
        cases = []
        for _ in range(1500):
            i = int(rng.integers(NS))
            p = line[i] + rng.normal(0, 0.6, 2)
            off = int(rng.integers(-40, 41)) if rng.random() < 0.5 else int(rng.integers(-4, 5))
            n_long += abs(off) > 12
            cases.append((float(p[0]), float(p[1]), ((i + off) % NS) / NS))
        for i in (0, 1, 2, 509, 510, 511):  # wrap-around
            for off in (-15, -3, 0, 3, 15):
                p = line[i] + rng.normal(0, 0.3, 2)
                cases.append((float(p[0]), float(p[1]), ((i + off) % NS) / NS))
        cases.append((0.0, 0.0, 0.25))
        cases.append((0.0, 0.0, -1.0))
        cases += [(float(x), float(y), -1.0) for x, y in rng.normal(0, 8, (50, 2))]
        for x, y, s in cases:
            s32 = float(np.float32(s))
            want = np.float32(oracle.drace_project(0, line, np.float32(x), np.float32(y), s32))
            got = np.float32(hd.hd_drace_project(ptr, x, y, s32))
            assert got == want, (name, x, y, s32, got, want)
    assert n_long > 1000
    print(f"4,000+ projections equal; {n_long} start more than 12 samples from the nearest sample")


def test_ties_and_margin_edges(hd, oracle):
    """Designed distance sequences along the raceline from the previous progress
    (sample 100, the position at the origin, every other sample 50 m away):
    the walk stops at a tie or a rise at sample j*, and a nearer sample sits at
    j* + 9 (inside the margin) or j* + 10 (just outside), in both directions."""
    far = 50 * np.stack([np.cos(np.arange(NS) / NS * 2 * np.pi), np.sin(np.arange(NS) / NS * 2 * np.pi)], 1)
    n = 0
    for direction in (+1, -1):
        for j_stop in range(1, 8):
            for tie in (True, False):
                for extra in (9, 10, 11):
                    r = [5.0 - 0.25 * j for j in range(j_stop)]  # strictly decreasing walk 0 .. j* - 1
                    r.append(r[-1] if tie else r[-1] + 0.1)     # the stop
                    r += [r[-1] + 0.5] * 12
                    r[j_stop + extra] = 0.5                     # a nearer sample at j* + extra
                    line = far.copy()
                    for j, v in enumerate(r):
                        line[(100 + direction * j) % NS] = (v, 0.0)
                    line = line.astype(np.float32)
                    tp = np.ascontiguousarray(line.reshape(-1))
                    ptr = tp.ctypes.data_as(C.POINTER(C.c_float))
                    s32 = float(np.float32(100 / NS))
                    want = np.float32(oracle.drace_project(0, line, 0.0, 0.0, s32))
                    got = np.float32(hd.hd_drace_project(ptr, 0.0, 0.0, s32))
                    assert got == want, (direction, j_stop, tie, extra, got, want)
                    n += 1
    print(f"{n} designed walks equal")
