"""Launch shapes (pkg_set_tuning / pkg_autotune) change timings, never results: every
combination of GJK lanes per pair, EPA workgroup size and LDS threshold steps the same
world to bit-identical bodies, pairs and manifolds, and pkg_autotune (which reruns the
GJK and EPA stages of each step it takes with every candidate) advances the world
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
            assert g.tuning() == {"gjk_lanes": lanes, "epa_block": block, "epa_lds_max": lds, "tree_refits": 7}
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
    b.step(s["dt"], 2)  # pkg_autotune took 2 steps, timing the candidates within each
    _same("autotune", _state(a), _state(b))
    a.close()
    b.close()


@pytest.mark.parametrize("kind, n, steps", [(3, 6, 40), (0, 128, 60), (2, 1, 60)], ids=["batched", "boxes", "pyramid"])
def test_tree_refits_give_the_same_results(pk, kind, n, steps):
    """Refitting the last LBVH hierarchy between full builds (pkg_tuning.tree_refits)
    changes neither pairs, manifolds nor bodies, step by step, nor the answers of world
    queries made on a refitted tree."""
    s = pk.scene(kind, 903, n)

    def world(refits):
        g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
        g.load_scene(s)
        if len(s["first_body"]) > 2:
            g.set_worlds(s["first_body"])
        t = g.tuning()
        g.set_tuning(t["gjk_lanes"], t["epa_block"], t["epa_lds_max"], refits)
        assert g.tuning()["tree_refits"] == refits
        return g

    ref = world(0)
    others = [world(r) for r in (1, 5, 64)]
    rng = np.random.default_rng(5)
    pos = ref.bodies()["pos"]
    origins = pos[rng.integers(0, len(pos), 256)] + rng.normal(scale=3.0, size=(256, 3))
    dirs = rng.normal(size=(256, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    for k in range(steps):
        ref.step(s["dt"])
        want = _state(ref)
        for g in others:
            g.step(s["dt"])
            _same(("refits", g.tuning()["tree_refits"], k), _state(g), want)
        if k % 10 == 9:
            hit_ref = ref.raycast(origins, dirs)
            for g in others:
                got = g.raycast(origins, dirs)
                for a, b in zip(got, hit_ref):
                    np.testing.assert_array_equal(a, b)
    for g in [ref] + others:
        g.close()


def test_set_tuning_rejects_unknown_shapes(pk):
    s = pk.scene(pk.SCENE_BOXES, None, 16)
    g = pk.GpuWorld()
    g.load_scene(s)
    with pytest.raises(pk.PkgError):
        g.set_tuning(3, 64, 0)
    with pytest.raises(pk.PkgError):
        g.set_tuning(2, 96, 0)
    with pytest.raises(pk.PkgError):
        g.set_tuning(2, 64, 0, 65)
    g.close()


def test_epa_tiers_and_overflows_bitexact(pk, oracle):
    """Deeply overlapping 16x32 UV spheres (up to 64 expansions and 319 faces per EPA
    solve, more than the 192 faces of the first tier's store): every first EPA tier
    (the private store at each workgroup size, the 896-face LDS store), with its
    overflows rerun in the larger stores, reaches the oracle's manifolds bit for bit."""
    from test_gpu_parity import _compare_world
    rng = np.random.default_rng(29)
    sph = oracle.sphere_verts(0.5, 16, 32)
    n_pairs = 48
    pos, quat = [], []
    for k in range(n_pairs):
        ca = np.array([5.0 * (k % 8), 0.0, 5.0 * (k // 8)]) + rng.uniform(-0.1, 0.1, 3)
        u = rng.normal(size=3)
        pos += [ca, ca + u / np.linalg.norm(u) * rng.uniform(0.05, 0.6)]
        for _ in range(2):
            q = rng.normal(size=4)
            quat.append(q / np.linalg.norm(q))
    n = 2 * n_pairs
    s = {"shape_offsets": np.array([0, len(sph)], np.uint32), "verts": sph,
         "is_static": np.zeros(n, np.uint8), "shape": np.zeros(n, np.uint32),
         "pos": np.array(pos), "vel": np.zeros((n, 3)), "angvel": np.zeros((n, 3)), "quat": np.array(quat),
         "mass": np.ones(n), "inertia": np.tile(np.eye(3).reshape(9) * 0.1, (n, 1)),
         "restitution": np.full(n, 0.5), "friction": np.full(n, 0.5)}
    flags = pk.FLAG_ALL_MOVED | pk.FLAG_NO_SOLVER | pk.FLAG_NO_INTEGRATE
    o = oracle.OracleWorld(iterations=10, solver_mode=1, flags=flags)
    o.load_scene(s)
    o.step(1 / 60)
    assert o.counters()["gjk_hits"] == n_pairs
    for block, lds in itertools.product((64, 128, 256), (0, 1 << 20)):
        print(f"[tuning] deep spheres epa_block={block} lds_max={lds}", flush=True)
        g = pk.GpuWorld(iterations=10, flags=flags)
        g.load_scene(s)
        g.set_tuning(2, block, lds)
        g.step(1 / 60)
        _compare_world((block, lds), g, o)
        if lds == 0:
            assert g.counters()["epa_overflow"] > 0, "no solve outgrew the first tier's store"
        g.close()
