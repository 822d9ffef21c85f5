"""GPU parity: the HIP world step and narrowphase against the oracle (the FP64 CPU
restatement of src/world.cpp:20-56 and its callees) on identical seeded inputs.

The device code fixes the same operation order as the oracle and is built with FMA
contraction off, so every comparison here is exact (bit for bit): pair keys,
GJK/EPA verdicts and outputs, manifold contents, solver colours, events, and all
body state after every step. The oracle runs the contact solver in the same colour
order (solver_mode=1)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene_pair(pk, oracle, kind, n, seed=None, iterations=None, flags=None):
    s = pk.scene(kind, seed, n)
    it = s["iterations"] if iterations is None else iterations
    fl = s["flags"] if flags is None else flags
    o = oracle.OracleWorld(iterations=it, solver_mode=1, flags=fl)
    o.load_scene(s)
    g = pk.GpuWorld(iterations=it, flags=fl)
    g.load_scene(s)
    return s, o, g


def _assert_same(step, name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"step {step} {name}: shape {a.shape} vs {b.shape}"
    # NaN positions must agree (their payload bits differ between x86 and the GPU)
    eq_nan = a.dtype.kind == "f"
    if not np.array_equal(a, b, equal_nan=eq_nan):
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b)))) if eq_nan else np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"step {step} {name}: {len(bad)} mismatches, first at {i}: "
                             f"gpu={a[i]!r} oracle={b[i]!r}")


def _compare_world(step, g, o):
    gb, ob = g.bodies(), o.bodies()
    for k in ("pos", "vel", "angvel", "quat", "fat", "moved"):
        _assert_same(step, k, gb[k], ob[k])
    _assert_same(step, "pairs", g.pairs(), o.pairs())
    gc, gd = g.manifolds()
    oc, od = o.manifolds()
    _assert_same(step, "manifold counts", gc, oc)
    _assert_same(step, "manifold data", gd, od)
    gk, gt = g.events()
    ok, ot = o.events()
    _assert_same(step, "event keys", gk, ok)
    _assert_same(step, "event kinds", gt, ot)
    gcnt, ocnt = g.counters(), o.counters()
    for k in ("moved", "pairs", "gjk_hits", "contacts", "rows", "colours"):
        assert gcnt[k] == ocnt[k], f"step {step} counter {k}: gpu {gcnt[k]} oracle {ocnt[k]}"
    _assert_same(step, "colours", g.colours(), o.colours())


def _run(pk, oracle, kind, n, steps, **kw):
    s, o, g = _scene_pair(pk, oracle, kind, n, **kw)
    print(f"[parity] kind={kind} n={n} steps={steps} bodies={len(s['shape'])}", flush=True)
    totals = {"pairs": 0, "hits": 0, "rows": 0, "still": 0}
    for k in range(steps):
        o.step(s["dt"])
        g.step(s["dt"])
        _compare_world(k, g, o)
        c = o.counters()
        totals["pairs"] += c["pairs"]
        totals["hits"] += c["gjk_hits"]
        totals["rows"] += c["rows"]
        # steps where no leaf left its fat box: the pair list is carried over whole
        totals["still"] += c["moved"] == 0 and c["pairs"] > 0
    g.close()
    return totals


def test_boxes_world_bitexact_small(pk, oracle):
    t = _run(pk, oracle, 0, 128, 60)
    assert t["pairs"] > 0 and t["hits"] > 0 and t["rows"] > 0


def test_boxes_world_bitexact_full(pk, oracle):
    # 1,024 random boxes over the ground box, dt = 1/60, 10 iterations
    t = _run(pk, oracle, 0, 0, 45)
    assert t["rows"] > 0


def test_pyramids_world_bitexact(pk, oracle):
    # two 15-layer pyramids (240 unit boxes), dt = 1/120, 20 iterations, EPA guard on;
    # from the third step on no leaf leaves its fat box, so these steps also cover the
    # carry-over of an unchanged pair list
    t = _run(pk, oracle, 2, 2, 120)
    assert t["rows"] > 0 and t["still"] > 100


def test_pyramids_reference_nan_reproduced(pk, oracle):
    """Without the guard, the reference EPA turns exactly stacked boxes into
    zero-normal, NaN-witness contacts (collision.cpp:284-285, 442-445); the device
    reproduces the same NaN pattern as the oracle."""
    _run(pk, oracle, 2, 1, 6, flags=0)


def test_hulls_narrowphase_bitexact(pk, oracle):
    # convex-hull clusters, every leaf moved, no solver / integration
    t = _run(pk, oracle, 1, 8192, 2)
    assert t["hits"] > 0


def test_boxes_world_sequential_seed_variation(pk, oracle):
    for seed in (1, 2, 3):
        _run(pk, oracle, 0, 64, 30, seed=seed)


def test_zero_gravity_head_on(pk, oracle):
    """Two boxes fly into each other with no gravity (the co_tests scene,
    tests/co/co_tests.cpp:358-429): a begin event fires, then an end event."""
    box = oracle.box_verts((0.5, 0.5, 0.5))
    s = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
         "is_static": np.zeros(2, np.uint8), "shape": np.zeros(2, np.uint32),
         "pos": np.array([[-2.0, 0, 0], [2.0, 0, 0]]), "vel": np.array([[2.0, 0, 0], [-2.0, 0, 0]]),
         "angvel": np.zeros((2, 3)), "quat": np.tile([1.0, 0, 0, 0], (2, 1)),
         "mass": np.ones(2), "inertia": np.tile(np.eye(3).reshape(9) / 6.0, (2, 1)),
         "restitution": np.full(2, 0.5), "friction": np.full(2, 0.5)}
    o = oracle.OracleWorld(gravity=(0, 0, 0), iterations=10, solver_mode=1)
    o.load_scene(s)
    g = pk.GpuWorld(gravity=(0, 0, 0), iterations=10)
    g.load_scene(s)
    kinds = []
    for k in range(240):
        o.step(0.005)
        g.step(0.005)
        _compare_world(k, g, o)
        kinds += g.events()[1].tolist()
    assert 1 in kinds and 2 in kinds and kinds.index(1) < kinds.index(2)
    g.close()


def test_edge_worlds(pk, oracle):
    box = oracle.box_verts((0.5, 0.5, 0.5))
    base = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box}

    def bodies(pos, static):
        n = len(pos)
        return dict(base, is_static=np.array(static, np.uint8), shape=np.zeros(n, np.uint32),
                    pos=np.array(pos, float).reshape(n, 3), vel=np.zeros((n, 3)),
                    angvel=np.zeros((n, 3)), quat=np.tile([1.0, 0, 0, 0], (n, 1)),
                    mass=np.ones(n), inertia=np.tile(np.eye(3).reshape(9), (n, 1)),
                    restitution=np.full(n, 0.5), friction=np.full(n, 0.5))

    cases = [
        bodies([[0, 0, 0]], [0]),                          # one dynamic body
        bodies([[0, 0, 0], [0.5, 0, 0]], [1, 1]),          # only static bodies (no pairs)
        bodies([[0, 3, 0], [0, 0, 0]], [0, 1]),            # a box falling on a static box
        bodies([[0, 0, 0], [0, 0, 0], [0, 0, 0]], [0, 0, 0]),  # coincident boxes
        # slot 2 disabled (a removed body's arena slot) inside the falling box's path
        bodies([[0, 3, 0], [0, 0, 0], [0, 2.5, 0], [0.2, 1.6, 0]], [0, 1, 2, 0]),
    ]
    for ci, s in enumerate(cases):
        o = oracle.OracleWorld(iterations=10, solver_mode=1)
        o.load_scene(s)
        g = pk.GpuWorld(iterations=10)
        g.load_scene(s)
        for k in range(40):
            o.step(1 / 60)
            g.step(1 / 60)
            _compare_world((ci, k), g, o)
            if ci == 4:
                ids = np.concatenate([g.pairs() >> 32, g.pairs() & 0xFFFFFFFF])
                assert not np.any(ids == 2), "a disabled slot paired"
        if ci == 4:
            assert np.array_equal(g.bodies()["pos"][2], [0, 2.5, 0])
        g.close()


def test_set_body_state(pk, oracle):
    s, o, g = _scene_pair(pk, oracle, 0, 64)
    for k in range(5):
        o.step(s["dt"])
        g.step(s["dt"])
    b = g.bodies()
    ids = np.array([3, 7], np.uint32)
    g.set_state(ids, pos=b["pos"][ids] + 0.25, quat=b["quat"][ids])
    nb = g.bodies()
    assert np.array_equal(nb["pos"][ids], b["pos"][ids] + 0.25)
    assert np.array_equal(nb["vel"], b["vel"])
    g.close()


# ---------------- standalone narrowphase (collision.cpp:512-518) ----------------
def _rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _shape_cases(oracle, rng, n):
    verts_box = oracle.box_verts((0.5, 0.3, 0.4))
    verts_sph = oracle.sphere_verts(0.5, 8, 12)
    verts_pyr = oracle.pyramid_verts(0.5, 0.8)
    out = []
    for i in range(n):
        sides = []
        for _ in range(2):
            kind = int(rng.integers(0, 3))
            c = rng.uniform(-0.6, 0.6, 3)
            if kind == 0:
                h = rng.uniform(0.1, 0.6, 3)
                sides.append((0, np.concatenate([c - h, c + h]), None))
            elif kind == 1:
                sides.append((1, np.concatenate([c, _rand_quat(rng), rng.uniform(0.1, 0.6, 3)]), None))
            else:
                v = [verts_box, verts_sph, verts_pyr][i % 3]
                sides.append((2, np.concatenate([c, _rand_quat(rng)]), v))
        out.append(tuple(sides))
    return out


def _oracle_shape(oracle, sh):
    k, p, v = sh
    if k == 0:
        return oracle.aabb(p[:3], p[3:6])
    if k == 1:
        return oracle.obb(p[:3], p[3:7], p[7:10])
    return oracle.mesh_instance(v, p[:3], p[3:7])


def test_gjk_epa_batch_random_bitexact(pk, oracle):
    rng = np.random.default_rng(7)
    cases = _shape_cases(oracle, rng, 3000)
    g = pk.GpuWorld()
    hit, out = g.gjk_epa_batch([c[0] for c in cases], [c[1] for c in cases])
    n_contact = 0
    for i, (a, b) in enumerate(cases):
        r = oracle.gjk_epa(_oracle_shape(oracle, a), _oracle_shape(oracle, b))
        if r is None:
            assert hit[i] == 0, i
        elif r == "degenerate":
            assert hit[i] == 2, i
        else:
            n_contact += 1
            assert hit[i] == 1, i
            exp = np.concatenate([r["normal"], r["world_a"], r["world_b"], [r["depth"]]])
            _assert_same(i, "gjk_epa", out[i], exp)
    assert n_contact > 500
    g.close()


def test_gjk_epa_batch_deep_spheres_bitexact(pk, oracle):
    """Deeply overlapping 16x32 UV spheres: long EPA runs (about 36 expansions per
    solve) over polytopes of many faces, with horizons of up to 8 edges."""
    rng = np.random.default_rng(23)
    sph = oracle.sphere_verts(0.5, 16, 32)
    A, B = [], []
    for _ in range(300):
        ca = rng.uniform(-0.1, 0.1, 3)
        u = rng.normal(size=3)
        cb = ca + u / np.linalg.norm(u) * rng.uniform(0.0, 0.7)
        A.append((2, np.concatenate([ca, _rand_quat(rng)]), sph))
        B.append((2, np.concatenate([cb, _rand_quat(rng)]), sph))
    g = pk.GpuWorld()
    hit, out = g.gjk_epa_batch(A, B)
    n_contact = 0
    for i in range(len(A)):
        r = oracle.gjk_epa(_oracle_shape(oracle, A[i]), _oracle_shape(oracle, B[i]))
        assert r is not None, i  # centres at most 0.7 apart: the hulls (inradius 0.49) overlap
        if r == "degenerate":
            assert hit[i] == 2, i
        else:
            n_contact += 1
            assert hit[i] == 1, i
            exp = np.concatenate([r["normal"], r["world_a"], r["world_b"], [r["depth"]]])
            _assert_same(i, "gjk_epa", out[i], exp)
    assert n_contact > 250
    g.close()


def test_gjk_epa_batch_known_depths(pk, oracle):
    """Axis-aligned boxes: the EPA depth is the smallest face distance of the
    Minkowski box A-B (the reference's tests/epa/epa_test.cpp:57-158 rely on it)."""
    rng = np.random.default_rng(11)
    A, B, exp = [], [], []
    for _ in range(400):
        ca, cb = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        ha, hb = rng.uniform(0.2, 1.0, 3), rng.uniform(0.2, 1.0, 3)
        amin, amax, bmin, bmax = ca - ha, ca + ha, cb - hb, cb + hb
        faces = np.concatenate([amax - bmin, bmax - amin])
        if faces.min() <= 1e-3:
            continue
        A.append((0, np.concatenate([amin, amax]), None))
        B.append((0, np.concatenate([bmin, bmax]), None))
        exp.append(faces.min())
    g = pk.GpuWorld()
    hit, out = g.gjk_epa_batch(A, B)
    assert (hit == 1).all()
    np.testing.assert_allclose(out[:, 9], exp, atol=1e-4)
    g.close()


# ---------------- batched independent worlds ----------------
def test_batched_worlds_bitexact(pk, oracle):
    """Six worlds of 81-3,837 bodies (spheres, boxes, hulls) stepped together on the GPU
    after pkg_set_worlds; each world's bodies, pairs, manifolds, events and colours
    equal its own oracle world's, bit for bit, every step."""
    import batched_util as bu
    s = pk.scene(pk.SCENE_BATCHED, 777, 6)
    fb = s["first_body"]
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    g.set_worlds(fb)
    worlds = []
    for w in range(len(fb) - 1):
        ws, b0, b1 = bu.world_slice(s, w)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        worlds.append((o, b0, b1))
    print(f"[parity] batched worlds sizes={np.diff(fb).tolist()}", flush=True)
    total_rows = 0
    overflows = 0
    for k in range(90):
        g.step(s["dt"])
        overflows += g.counters()["epa_overflow"]
        gb = g.bodies()
        gk = g.pairs()
        gc, gd = g.manifolds()
        ek, et = g.events()
        ck_col, ck_keys = g.colours(with_keys=True)
        for w, (o, b0, b1) in enumerate(worlds):
            o.step(s["dt"])
            ob = o.bodies()
            for f in ("pos", "vel", "angvel", "quat", "fat", "moved"):
                _assert_same((w, k), f, gb[f][b0:b1], ob[f])
            local, m = bu.keys_in(gk, b0, b1)
            _assert_same((w, k), "pairs", local, o.pairs())
            oc, od = o.manifolds()
            _assert_same((w, k), "manifold counts", gc[m], oc)
            _assert_same((w, k), "manifold data", gd[m], od)
            el, em = bu.keys_in(ek, b0, b1)
            ok_, ot_ = o.events()
            _assert_same((w, k), "event keys", el, ok_)
            _assert_same((w, k), "event kinds", et[em], ot_)
            cl, cm = bu.keys_in(ck_keys, b0, b1)
            oc_col, oc_keys = o.colours(with_keys=True)
            _assert_same((w, k), "colour keys", cl, oc_keys)
            _assert_same((w, k), "colours", ck_col[cm], oc_col)
            total_rows += o.counters()["rows"]
    g.close()
    print(f"[parity] batched worlds: EPA solves rerun in larger stores: {overflows}", flush=True)
    assert total_rows > 0


def test_set_worlds_rejects_bad_partition(pk):
    s = pk.scene(pk.SCENE_BATCHED, 777, 3)
    g = pk.GpuWorld()
    g.load_scene(s)
    fb = s["first_body"].copy()
    for bad in (fb[:-1], np.concatenate([[1], fb[1:]]), fb[::-1].copy()):
        with pytest.raises(pk.PkgError):
            g.set_worlds(bad)
    g.set_worlds(fb)
    g.close()


def test_granular_pile_bitexact(pk, oracle):
    """The granular-pile generator (UV spheres and boxes of 1-3 cm on a lattice over
    the 64-box hopper) at 343 bodies, 60 steps of 1/240 s: every step bit-exact."""
    t = _run(pk, oracle, 4, 343, 60)
    assert t["pairs"] > 0


def test_fp32_solve_drift_small(pk):
    """FP32 solve (FLAG_FP32_SOLVE) against the FP64 solve on the pyramids, one step at a
    time from a shared state: after each step the FP32 world's bodies are reset to the
    FP64 world's, so the difference measured is that of one step. It stays within 1e-5
    relative to max(|x|, 1 m) (trajectories of exactly stacked boxes diverge much
    faster, since contacts at zero penetration switch on and off)."""
    s = pk.scene(pk.SCENE_PYRAMIDS, None, 4)
    w64 = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    w32 = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"] | pk.FLAG_FP32_SOLVE)
    w64.load_scene(s)
    w32.load_scene(s)
    ids = np.arange(len(s["shape"]), dtype=np.uint32)
    errs = []
    rows = 0
    for k in range(30):
        w64.step(s["dt"])
        w32.step(s["dt"])
        a, b = w64.bodies(), w32.bodies()
        assert np.all(np.isfinite(b["pos"]))
        errs.append(float(np.max(np.abs(a["pos"] - b["pos"]) / np.maximum(np.abs(a["pos"]), 1.0))))
        rows += w32.counters()["rows"]
        w32.set_state(ids, pos=a["pos"], vel=a["vel"], angvel=a["angvel"], quat=a["quat"])
    print("fp32 solve rel. position error per step:", max(errs))
    assert rows > 0
    assert max(errs) < 1e-5
    w64.close()
    w32.close()
