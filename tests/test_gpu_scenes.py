"""The reference's own demo scenes (tests/configs/*.json, copied as data fixtures under
tests/golden/scenes/) read by physkit_amd.scene_json and stepped on the GPU against the
oracle, bit for bit, every step."""
import glob
import os

import numpy as np
import pytest

from test_gpu_parity import _compare_world

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = sorted(glob.glob(os.path.join(HERE, "golden", "scenes", "*.json")))


@pytest.mark.parametrize("path", SCENES, ids=[os.path.basename(p) for p in SCENES])
def test_reference_scene_bitexact(pk, oracle, path):
    from physkit_amd import scene_json
    s = scene_json.load(path)
    steps = 150
    print(f"[scenes] {os.path.basename(path)}: {len(s['shape'])} bodies, {steps} steps", flush=True)
    o = oracle.OracleWorld(gravity=s["gravity"], iterations=s["iterations"], solver_mode=1)
    o.load_scene(s)
    g = pk.GpuWorld(gravity=s["gravity"], iterations=s["iterations"])
    g.load_scene(s)
    contacts = 0
    for k in range(steps):
        g.step(1 / 60)
        o.step(1 / 60)
        _compare_world(k, g, o)
        contacts += g.counters()["contacts"]
    assert contacts > 0
    g.close()


@pytest.mark.parametrize("path", SCENES[:4], ids=[os.path.basename(p) for p in SCENES[:4]])
def test_cpp_scene_reader_through_the_facade(pk, path):
    """The same files read by the C++ reader (include/physkit_amd/scene_json.hpp) into
    the C++ facade world (tests/cpp/scene_run.cpp) and stepped 90 steps reach the poses
    the Python path reaches, bit for bit."""
    import subprocess
    from physkit_amd import scene_json
    exe = os.path.join(HERE, "cpp", "_build", "scene_run")
    r = subprocess.run([exe, path, "90", repr(1 / 60)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().split("\n")])
    s = scene_json.load(path)
    g = pk.GpuWorld(gravity=s["gravity"], iterations=s["iterations"])
    g.load_scene(s)
    g.step(1 / 60, 90)
    b = g.bodies()
    np.testing.assert_array_equal(got[:, :3], b["pos"])
    np.testing.assert_array_equal(got[:, 3:], b["quat"])
    g.close()
