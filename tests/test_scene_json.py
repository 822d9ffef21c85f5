"""Scene files (tests/configs/*.json of the reference, copied as data fixtures under
tests/golden/scenes/) read by physkit_amd.scene_json, on the CPU."""
import glob
import json
import os
from fractions import Fraction
from decimal import Decimal

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = sorted(glob.glob(os.path.join(HERE, "golden", "scenes", "*.json")))


def test_fixtures_present():
    assert len(SCENES) == 11


@pytest.mark.parametrize("path", SCENES, ids=[os.path.basename(p) for p in SCENES])
def test_scene_reads(path):
    from physkit_amd import scene_json
    raw = json.load(open(path))
    s = scene_json.load(path)
    n = len(raw["objects"])
    assert len(s["shape"]) == n and s["pos"].shape == (n, 3) and s["quat"].shape == (n, 4)
    assert s["shape_offsets"][-1] == len(s["verts"])
    assert s["iterations"] == raw.get("solver_iterations", 10)
    for i, o in enumerate(raw["objects"]):
        assert s["is_static"][i] == (o.get("type", "dynam") == "stat")
        np.testing.assert_array_equal(s["pos"][i], o["pos"])
        v = s["verts"][s["shape_offsets"][s["shape"][i]]:s["shape_offsets"][s["shape"][i] + 1]]
        m = o.get("mesh", {"name": "box"})
        if m["name"] == "box":
            h = np.float32(m.get("half_extents", [0.5] * 3)).astype(float)
            np.testing.assert_array_equal(v.max(axis=0), h)
            np.testing.assert_array_equal(v.min(axis=0), -h)
        elif m["name"] == "sphere":
            assert len(v) == 2 + (m.get("stacks", 16) - 1) * m.get("sectors", 32)
            np.testing.assert_allclose(np.linalg.norm(v, axis=1), float(np.float32(m["radius"])), rtol=1e-15)
        else:  # the reader's argument order: "height" is the base half-width
            assert v[4, 1] == float(np.float32(m["base_size"]))
            assert v[1, 0] == float(np.float32(m["height"]))


def _float32_of_decimal(text):
    """Nearest float32 to the exact decimal value (ties to even)."""
    exact = Fraction(Decimal(text))
    f = np.float32(float(text))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - exact), int(c.view(np.uint32)) & 1))
    return best


def test_mesh_sizes_round_as_direct_decimal_to_float():
    """The reader rounds double -> float; the reference rounds decimal -> float. For
    every mesh size in the fixtures the two give the same float."""
    checked = 0
    for p in SCENES:
        raw = json.load(open(p), parse_float=str, parse_int=str)
        for o in raw["objects"]:
            m = o.get("mesh", {})
            vals = list(m.get("half_extents", [])) + [m[k] for k in ("radius", "height", "base_size") if k in m]
            for t in vals:
                assert np.float32(float(t)) == _float32_of_decimal(t), (p, t)
                checked += 1
    assert checked > 1000


def test_bad_scenes_are_rejected():
    from physkit_amd import scene_json
    with pytest.raises(ValueError):
        scene_json.parse({"objects": [{"pos": [0, 0, 0]}]})  # dynamic without mass
    with pytest.raises(ValueError):
        scene_json.parse({"objects": [{"pos": [0, 0, 0], "mass": 1, "mesh": {"name": "cone"}}]})
    with pytest.raises(ValueError):
        scene_json.parse({"objects": [{"mass": 1}]})  # no position
    s = scene_json.parse({"objects": [{"pos": [0, 1, 0], "type": "stat", "mesh": {"name": "box", "half_extents": [1, 2, 3]}},
                                      {"pos": [0, 3, 0], "mass": 2.0, "mesh": {"name": "box", "half_extents": [1, 2, 3]}}]})
    assert list(s["shape"]) == [0, 0] and len(s["shape_offsets"]) == 2  # one shared shape


@pytest.fixture(scope="module")
def scene_dump(tmp_path_factory):
    """tests/cpp/scene_dump.cpp: the C++ reader (include/physkit_amd/scene_json.hpp)."""
    import subprocess
    root = os.path.dirname(HERE)
    exe = str(tmp_path_factory.mktemp("cpp") / "scene_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(root, "include"),
                    os.path.join(HERE, "cpp", "scene_dump.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("path", SCENES, ids=[os.path.basename(p) for p in SCENES])
def test_cpp_reader_matches(scene_dump, path):
    """The C++ reader builds, field for field and vertex for vertex, what the Python
    reader builds from the same file (both follow graphics.cpp:20-135)."""
    import subprocess
    from physkit_amd import scene_json
    s = scene_json.load(path)
    out = subprocess.run([scene_dump, path], capture_output=True, text=True, check=True).stdout.split("\n")
    w = out[0].split()
    assert w[0] == "world"
    assert [float(x) for x in w[1:4]] == list(s["gravity"]) and int(w[4]) == s["iterations"]
    n = int(w[5])
    assert n == len(s["shape"])
    for i in range(n):
        f = out[1 + i].split()
        assert f[0] == "object"
        vals = [float(x) for x in f[2:21]]
        assert int(f[1]) == int(s["is_static"][i])
        # a static body's mass is infinite whatever the file says (object.h:81-85): not compared
        mass = 0.0 if s["is_static"][i] else s["mass"][i]
        exp = (list(s["pos"][i]) + list(s["vel"][i]) + list(s["quat"][i]) + list(s["angvel"][i]) +
               [s["inertia"][i, 0], s["inertia"][i, 4], s["inertia"][i, 8], mass,
                s["restitution"][i], s["friction"][i]])
        assert vals == exp, i
        nv = int(f[21])
        v = np.array([float(x) for x in f[22:22 + 3 * nv]]).reshape(nv, 3)
        sh = s["shape"][i]
        np.testing.assert_array_equal(v, s["verts"][s["shape_offsets"][sh]:s["shape_offsets"][sh + 1]])


def test_cpp_reader_rejects_what_the_schema_does_not_allow(scene_dump, tmp_path):
    import subprocess
    cases = {
        "no_pos": '{"objects": [{"mass": 1}]}',
        "bad_type": '{"objects": [{"pos": [0,0,0], "type": "kinematic", "mass": 1}]}',
        "no_mass": '{"objects": [{"pos": [0,0,0]}]}',
        "bad_mesh": '{"objects": [{"pos": [0,0,0], "mass": 1, "mesh": {"name": "torus"}}]}',
        "thin_sphere": '{"objects": [{"pos": [0,0,0], "mass": 1, "mesh": {"name": "sphere", "radius": 1, "stacks": 1}}]}',
        "broken": '{"objects": [',
    }
    for name, text in cases.items():
        p = tmp_path / f"{name}.json"
        p.write_text(text)
        out = subprocess.run([scene_dump, str(p)], capture_output=True, text=True, check=True).stdout
        assert out.startswith("error "), (name, out)
