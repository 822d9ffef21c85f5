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
