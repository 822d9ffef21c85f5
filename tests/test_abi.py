"""CPU checks of the C-ABI library: it loads, exports every symbol include/physkit_amd.h
declares, and its host-side scene generators are deterministic. No device calls."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "physkit_amd.h")).read()
    return sorted(set(re.findall(r"^(?:pkg_status|const char \*)\s*(pkg_\w+)\s*\(", txt, re.M)))


def test_header_matches_python_exports(pk):
    assert sorted(pk.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol(pk):
    L = pk.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_scene_generators_deterministic(pk):
    for kind, n in ((0, 64), (1, 1024), (2, 1)):
        a = pk.scene(kind, None, n)
        b = pk.scene(kind, None, n)
        for k in ("verts", "pos", "vel", "quat", "mass", "inertia"):
            assert np.array_equal(a[k], b[k]), (kind, k)
        c = pk.scene(kind, 12345, n)
        assert not np.array_equal(a["pos"], c["pos"])


def test_scene_shapes(pk):
    s = pk.scene(0)
    assert len(s["shape"]) == 1025 and s["n_static"] == 1
    assert s["is_static"][0] == 1
    q = s["quat"]
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    h = pk.scene(1, None, 2048)
    nv = np.diff(h["shape_offsets"])
    assert nv.min() >= 8 and nv.max() <= 32
    p = pk.scene(2)
    assert len(p["shape"]) == 64 * 120 + 1


def test_batched_scene_partition(pk):
    s = pk.scene(pk.SCENE_BATCHED, 777, 16)
    fb = s["first_body"].astype(np.int64)
    sizes = np.diff(fb) - 1  # one static ground box per world
    assert fb[0] == 0 and fb[-1] == len(s["shape"])
    assert sizes.min() >= 16 and sizes.max() < 4096
    shp = s["shape"].astype(np.int64)
    offs = s["shape_offsets"].astype(np.int64)
    nv = (offs[1:] - offs[:-1])[shp]
    for w in range(len(fb) - 1):
        b0, b1 = fb[w], fb[w + 1]
        assert s["is_static"][b0] == 1 and not s["is_static"][b0 + 1:b1].any()
        # shapes of a world are contiguous and private to it
        assert np.array_equal(np.sort(shp[b0:b1]), np.arange(shp[b0], shp[b0] + (b1 - b0)))
    dyn = ~s["is_static"].astype(bool)
    frac_sphere = np.mean(nv[dyn] == 2 + 15 * 32)
    frac_box = np.mean(nv[dyn] == 8)
    frac_hull = np.mean((nv[dyn] >= 9) & (nv[dyn] <= 16))
    assert abs(frac_sphere - 0.4) < 0.03 and abs(frac_hull - 0.2 * 8 / 9) < 0.03
    assert abs(frac_box - (0.4 + 0.2 / 9)) < 0.03
