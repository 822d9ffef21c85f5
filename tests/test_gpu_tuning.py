"""Launch shapes (pkg_set_tuning / pkg_autotune) change timings, never results: every
combination of GJK lanes per pair, EPA workgroup size and LDS threshold steps the same
world to bit-identical bodies, pairs and manifolds, and pkg_autotune advances the world
exactly as plain steps would."""
import itertools

import numpy as np
import pytest

from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu


def _state(g):
    b = g.bodies()
    c, d = g.manifolds()
    return b, g.pairs(), c, d


def _same(tag, a, b):
    (ba, pa, ca, da), (bb, pb, cb, db) = a, b
    for f in ("pos", "vel", "angvel", "quat", "fat"):
        _assert_same(tag, f, ba[f], bb[f])
    _assert_same(tag, "pairs", pa, pb)
    _assert_same(tag, "manifold counts", ca, cb)
    _assert_same(tag, "manifold data", da, db)


@pytest.mark.parametrize("kind, n", [(3, 4), (1, 4096)], ids=["batched", "hulls"])
def test_every_launch_shape_gives_the_same_results(pk, kind, n):
    s = pk.scene(kind, 901, n)
    steps = 12 if kind == 3 else 2

    def run(lanes=None, block=None, lds=None):
        g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
        g.load_scene(s)
        if len(s["first_body"]) > 2:
            g.set_worlds(s["first_body"])
        if lanes is not None:
            g.set_tuning(lanes, block, lds)
            assert g.tuning() == {"gjk_lanes": lanes, "epa_block": block, "epa_lds_max": lds}
        g.step(s["dt"], steps)
        out = _state(g)
        g.close()
        return out

    ref = run()
    for lanes, block, lds in itertools.product((1, 2, 4, 8), (64, 128, 256), (0, 1 << 20)):
        print(f"[tuning] kind={kind} lanes={lanes} block={block} lds_max={lds}", flush=True)
        _same((lanes, block, lds), run(lanes, block, lds), ref)


def test_autotune_steps_like_plain_steps(pk):
    s = pk.scene(pk.SCENE_BATCHED, 902, 6)
    a = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"] | pk.FLAG_MIXED)
    b = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"] | pk.FLAG_MIXED)
    for g in (a, b):
        g.load_scene(s)
        g.set_worlds(s["first_body"])
        g.step(s["dt"], 3)
    t = a.autotune(s["dt"], 2)
    print(f"[tuning] autotune chose {t}", flush=True)
    assert t["gjk_lanes"] in (1, 2, 4, 8) and t["epa_block"] in (64, 128, 256)
    assert all(x > 0 for x in t["gjk_ms"])
    b.step(s["dt"], 20)  # pkg_autotune took 10 x 2 steps
    _same("autotune", _state(a), _state(b))
    a.close()
    b.close()


def test_set_tuning_rejects_unknown_shapes(pk):
    s = pk.scene(pk.SCENE_BOXES, None, 16)
    g = pk.GpuWorld()
    g.load_scene(s)
    with pytest.raises(pk.PkgError):
        g.set_tuning(3, 64, 0)
    with pytest.raises(pk.PkgError):
        g.set_tuning(2, 96, 0)
    g.close()
