"""Pins of the oracle (the FP64 CPU restatement) on the CPU, no GPU needed.

Each group checks one oracle function against something independent of the
oracle's own formula: the known answers and invariants the reference's tests hold
for this path (tests/gjk/gjk_test.cpp, tests/epa/epa_test.cpp,
tests/dynamic_bvh/main.cpp, tests/obb/obb_test.cpp, tests/mesh/main.cpp, re-expressed
here, not copied), closed forms, brute force, and libm."""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(2026)


def qaxis(axis, ang):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * axis])


def qrot(q, v):
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, float)


def rand_quat():
    q = RNG.normal(size=4)
    return q / np.linalg.norm(q)


# ---------------- math ----------------
def test_sincos_within_one_ulp_of_libm(oracle):
    x = np.concatenate([RNG.uniform(-100, 100, 20000), RNG.uniform(-1e-3, 1e-3, 5000),
                        np.array([0.0, 1e-300, math.pi / 4, math.pi / 2, 3.0])])
    s, c = oracle.sincos(x)
    ref_s, ref_c = np.sin(x), np.cos(x)
    us = np.abs(s.view(np.int64) - ref_s.view(np.int64))
    uc = np.abs(c.view(np.int64) - ref_c.view(np.int64))
    assert us.max() <= 2 and uc.max() <= 2


def test_exp_map_closed_forms(oracle):
    # rotation about a principal axis: q = (cos(w dt / 2), sin(w dt / 2) e_z)
    for w in (0.3, -2.0, 17.0):
        q = oracle.exp_map([0.0, 0.0, w], 0.01)
        np.testing.assert_allclose(q, [math.cos(w * 0.005), 0, 0, math.sin(w * 0.005)], atol=1e-15)
    # below 1e-12 rad the small-angle branch normalises (1, theta/2)
    q = oracle.exp_map([1e-11, 0, 0], 0.01)
    np.testing.assert_allclose(q, [1.0, 5e-14, 0, 0], rtol=1e-12)
    # unit norm in general
    for _ in range(50):
        q = oracle.exp_map(RNG.normal(size=3) * 5, 1 / 60)
        assert abs(np.linalg.norm(q) - 1) < 1e-15


def test_inverse3(oracle):
    np.testing.assert_allclose(oracle.inverse3(np.diag([2.0, 4.0, 8.0])), np.diag([0.5, 0.25, 0.125]),
                               rtol=1e-15)
    for _ in range(50):
        m = RNG.normal(size=(3, 3)) + 3 * np.eye(3)
        np.testing.assert_allclose(oracle.inverse3(m) @ m, np.eye(3), atol=1e-12)


# ---------------- support (mesh.cpp:341-358, 442-448; bounds.h:539-548) ----------------
def test_mesh_support_first_vertex_wins_ties(oracle):
    box = oracle.box_verts((1, 2, 3))
    # direction (1,0,0): vertices 1, 2, 5, 6 all have x = +1; the first (index 1) wins
    np.testing.assert_array_equal(oracle.instance_support(box, (0, 0, 0), (1, 0, 0, 0), (1, 0, 0)), box[1])
    for _ in range(200):
        d = RNG.normal(size=3)
        dots = box @ d
        np.testing.assert_array_equal(
            oracle.instance_support(box, (0, 0, 0), (1, 0, 0, 0), d), box[int(np.argmax(dots))])


def test_instance_support_and_bounds_under_motion(oracle):
    box = oracle.box_verts((0.5, 1.0, 1.5))
    for _ in range(100):
        q, p = rand_quat(), RNG.uniform(-5, 5, 3)
        world = np.array([qrot(q, v) + p for v in box])
        d = RNG.normal(size=3)
        s = oracle.instance_support(box, p, q, d)
        assert abs(s @ d - (world @ d).max()) < 1e-12
        b = oracle.instance_bounds(box, p, q)  # corners of the local box, rotated
        np.testing.assert_allclose(b[:3], world.min(0), atol=1e-12)
        np.testing.assert_allclose(b[3:], world.max(0), atol=1e-12)


def test_obb_support_analytic(oracle):
    for _ in range(100):
        q, c, h = rand_quat(), RNG.uniform(-2, 2, 3), RNG.uniform(0.1, 2, 3)
        d = RNG.normal(size=3)
        corners = np.array([c + qrot(q, np.array(sg) * h) for sg in np.ndindex(2, 2, 2)
                            for sg in [np.array(sg) * 2 - 1]])
        s = oracle.obb_support(c, q, h, d)
        assert abs(s @ d - (corners @ d).max()) < 1e-12


# ---------------- GJK verdicts (tests/gjk/gjk_test.cpp) ----------------
def A(mn, mx):
    return (0, mn, mx)


def _shape(oracle, s):
    if s[0] == 0:
        return oracle.aabb(s[1], s[2])
    if s[0] == 1:
        return oracle.obb(s[1], s[2], s[3])
    return oracle.mesh_instance(s[1], s[2], s[3] if len(s) > 3 else (1, 0, 0, 0))


GJK_CASES = [
    # (shape a, shape b, intersect?)
    (A((-1, -1, -1), (1, 1, 1)), A((0.5, -1, -1), (2.5, 1, 1)), True),
    (A((-1, -1, -1), (1, 1, 1)), A((2, -1, -1), (4, 1, 1)), False),
    (A((-1, -1, -1), (1, 1, 1)), A((-1, 2, -1), (1, 4, 1)), False),
    (A((-1, -1, -1), (1, 1, 1)), A((-1, -1, 2), (1, 1, 4)), False),
    (A((-3, -3, -3), (3, 3, 3)), A((-1, -1, -1), (1, 1, 1)), True),
    (A((-1, -1, -1), (1, 1, 1)), A((-3, -3, -3), (3, 3, 3)), True),
    (A((-1, -1, -1), (1, 1, 1)), A((100, -1, -1), (102, 1, 1)), False),
    (A((-1, -1, -1), (1, 1, 1)), A((-1, -1, -1), (1, 1, 1)), True),
    (A((0, 0, 0), (2, 2, 2)), A((1, 1, 1), (3, 3, 3)), True),
    (A((-1, -1, -1), (1, 1, 1)), A((5, 5, 5), (7, 7, 7)), False),
    ((1, (0, 0, 0), (1, 0, 0, 0), (1, 1, 1)), (1, (1.5, 0, 0), (1, 0, 0, 0), (1, 1, 1)), True),
    ((1, (0, 0, 0), (1, 0, 0, 0), (1, 1, 1)), (1, (3, 0, 0), (1, 0, 0, 0), (1, 1, 1)), False),
    ((1, (0, 0, 0), (1, 0, 0, 0), (1, 1, 1)), (1, (0, 0, 0), tuple(qaxis((0, 0, 1), math.pi / 4)), (1, 1, 1)), True),
    # B rotated 45 deg about z at x=1: its -x support reaches 1 - sqrt 2 < 1, overlap
    ((1, (0, 0, 0), (1, 0, 0, 0), (1, 1, 1)), (1, (1, 0, 0), tuple(qaxis((0, 0, 1), math.pi / 4)), (1, 1, 1)), True),
    # at x=3 the -x support is 3 - sqrt 2 = 1.586 > 1: separated by 0.586 m
    ((1, (0, 0, 0), (1, 0, 0, 0), (1, 1, 1)), (1, (3, 0, 0), tuple(qaxis((0, 0, 1), math.pi / 4)), (1, 1, 1)), False),
]


@pytest.mark.parametrize("i", range(len(GJK_CASES)))
def test_gjk_verdicts_reference_cases(oracle, i):
    a, b, hit = GJK_CASES[i]
    r = oracle.gjk_epa(_shape(oracle, a), _shape(oracle, b))
    assert (r is not None) == hit
    r2 = oracle.gjk_epa(_shape(oracle, b), _shape(oracle, a))  # symmetry
    assert (r2 is not None) == hit


def _sat_obb(ca, qa, ha, cb, qb, hb):
    """Separating-axis test for two boxes (15 axes), Ericson, Real-Time Collision
    Detection 4.4.1: returns the minimum overlap over all axes (< 0 = separated)."""
    Ra = np.stack([qrot(qa, e) for e in np.eye(3)])
    Rb = np.stack([qrot(qb, e) for e in np.eye(3)])
    axes = list(Ra) + list(Rb) + [np.cross(u, v) for u in Ra for v in Rb]
    t = np.asarray(cb) - np.asarray(ca)
    best = np.inf
    for ax in axes:
        n = np.linalg.norm(ax)
        if n < 1e-9:
            continue
        ax = ax / n
        ra = sum(ha[k] * abs(Ra[k] @ ax) for k in range(3))
        rb = sum(hb[k] * abs(Rb[k] @ ax) for k in range(3))
        best = min(best, ra + rb - abs(t @ ax))
    return best


def test_gjk_matches_sat_for_random_boxes(oracle):
    n_checked = 0
    for _ in range(600):
        ca, cb = RNG.uniform(-1, 1, 3), RNG.uniform(-1, 1, 3)
        qa, qb = rand_quat(), rand_quat()
        ha, hb = RNG.uniform(0.1, 0.8, 3), RNG.uniform(0.1, 0.8, 3)
        sat = _sat_obb(ca, qa, ha, cb, qb, hb)
        if abs(sat) < 1e-6:
            continue
        r = oracle.gjk_epa(oracle.obb(ca, qa, ha), oracle.obb(cb, qb, hb))
        assert (r is not None) == (sat > 0)
        if r is not None and r != "degenerate":
            # EPA depth equals the minimum SAT overlap of two boxes
            assert abs(r["depth"] - sat) < 1e-5
        n_checked += 1
    assert n_checked > 500


def test_gjk_mesh_shapes(oracle):
    box = oracle.box_verts((1, 1, 1))
    sph = oracle.sphere_verts(1.0)
    pyr = oracle.pyramid_verts(1.0, 2.0)
    assert len(sph) == 2 + 15 * 32  # mesh::sphere default 16 x 32
    assert oracle.gjk_epa(oracle.mesh_instance(box), oracle.mesh_instance(box, (1.5, 0, 0))) is not None
    assert oracle.gjk_epa(oracle.mesh_instance(box), oracle.mesh_instance(box, (2.5, 0, 0))) is None
    assert oracle.gjk_epa(oracle.mesh_instance(sph), oracle.mesh_instance(sph, (1.9, 0, 0))) is not None
    assert oracle.gjk_epa(oracle.mesh_instance(sph), oracle.mesh_instance(sph, (2.1, 0, 0))) is None
    assert oracle.gjk_epa(oracle.mesh_instance(pyr), oracle.mesh_instance(pyr, (0, 1.5, 0))) is not None
    assert oracle.gjk_epa(oracle.mesh_instance(pyr), oracle.mesh_instance(pyr, (0, 2.1, 0))) is None


# ---------------- EPA depths (tests/epa/epa_test.cpp) ----------------
def _minkowski_box_depth(amin, amax, bmin, bmax):
    return float(np.concatenate([np.subtract(amax, bmin), np.subtract(bmax, amin)]).min())


@pytest.mark.parametrize("a,b,depth", [
    (((-1, -1, -1), (1, 1, 1)), ((0.5, -1, -1), (2.5, 1, 1)), 0.5),
    (((-1, -1, -1), (1, 1, 1)), ((-1, 0.5, -1), (1, 2.5, 1)), 0.5),
    (((-1, -1, -1), (1, 1, 1)), ((0.9, -1, -1), (2.9, 1, 1)), 0.1),
    (((-1, -1, -1), (1, 1, 1)), ((-0.5, -1, -1), (1.5, 1, 1)), 1.5),
    (((-2, -2, -2), (2, 2, 2)), ((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)), 2.5),
    (((-1, -1, -1), (1, 1, 1)), ((-1, -1, -1), (1, 1, 1)), 2.0),
    (((-1, -1, -1), (1, 1, 1)), ((-0.5, -1, -1), (0.5, 1, 1)), 1.5),
    (((0, 0, 0), (2, 2, 2)), ((1, 1, 1), (3, 3, 3)), 1.0),
])
def test_epa_aabb_known_depths(oracle, a, b, depth):
    r = oracle.gjk_epa(oracle.aabb(*a), oracle.aabb(*b))
    assert r is not None and r != "degenerate"
    assert abs(r["depth"] - depth) < 1e-4
    assert abs(np.linalg.norm(r["normal"]) - 1) < 1e-6


def test_epa_aabb_random_min_face_distance_and_mtv(oracle):
    for _ in range(300):
        ca, cb = RNG.uniform(-1, 1, 3), RNG.uniform(-1, 1, 3)
        ha, hb = RNG.uniform(0.2, 1, 3), RNG.uniform(0.2, 1, 3)
        amin, amax, bmin, bmax = ca - ha, ca + ha, cb - hb, cb + hb
        d = _minkowski_box_depth(amin, amax, bmin, bmax)
        if d < 1e-3:
            continue
        r = oracle.gjk_epa(oracle.aabb(amin, amax), oracle.aabb(bmin, bmax))
        assert r is not None and abs(r["depth"] - d) < 1e-4
        # normal points from B to A: moving A by normal * (depth + 1 mm) separates them
        off = r["normal"] * (r["depth"] + 1e-3)
        assert oracle.gjk_epa(oracle.aabb(amin + off, amax + off), oracle.aabb(bmin, bmax)) is None
        r2 = oracle.gjk_epa(oracle.aabb(bmin, bmax), oracle.aabb(amin, amax))
        assert abs(r2["depth"] - r["depth"]) < 1e-4


def test_epa_obb_cases(oracle):
    ident = (1, 0, 0, 0)
    # slab 10x0.2x10 through a pillar: the cheapest exit is along y, 3.3 m
    r = oracle.gjk_epa(oracle.obb((0, 0, 0), ident, (5, 0.1, 5)), oracle.obb((0, 0, 0), ident, (0.2, 3.2, 0.2)))
    assert r is not None and abs(r["depth"] - 3.3) < 1e-4
    # near touching: 1 cm overlap
    r = oracle.gjk_epa(oracle.obb((0, 0, 0), ident, (1, 1, 1)), oracle.obb((1.99, 0, 0), ident, (1, 1, 1)))
    assert r is not None and abs(r["depth"] - 0.01) < 1e-2
    # very deep containment
    r = oracle.gjk_epa(oracle.obb((0, 0, 0), ident, (100, 100, 100)), oracle.obb((0, 0, 0), ident, (0.1, 0.1, 0.1)))
    assert r is not None and abs(r["depth"] - 100.1) < 0.5
    # rotated boxes: MTV separates
    for _ in range(100):
        ca, cb, qa, qb = RNG.uniform(-0.5, 0.5, 3), RNG.uniform(-0.5, 0.5, 3), rand_quat(), rand_quat()
        ha, hb = RNG.uniform(0.3, 1, 3), RNG.uniform(0.3, 1, 3)
        r = oracle.gjk_epa(oracle.obb(ca, qa, ha), oracle.obb(cb, qb, hb))
        if r is None or r == "degenerate":
            continue
        assert r["depth"] > 0
        off = r["normal"] * (r["depth"] + 1e-3)
        assert oracle.gjk_epa(oracle.obb(ca + off, qa, ha), oracle.obb(cb, qb, hb)) is None


def test_epa_mesh_boxes_and_spheres(oracle):
    box = oracle.box_verts((1, 1, 1))
    r = oracle.gjk_epa(oracle.mesh_instance(box), oracle.mesh_instance(box, (1.5, 0, 0)))
    assert abs(r["depth"] - 0.5) < 1e-4 and abs(r["normal"][0]) > 0.9
    big = oracle.box_verts((3, 3, 3))
    r = oracle.gjk_epa(oracle.mesh_instance(big), oracle.mesh_instance(box, (0.5, 0, 0)))
    assert abs(r["depth"] - 3.5) < 1e-4
    sph = oracle.sphere_verts(1.0)
    r = oracle.gjk_epa(oracle.mesh_instance(sph), oracle.mesh_instance(sph, (1.5, 0, 0)))
    assert abs(r["depth"] - 0.5) < 0.15 and abs(r["normal"][0]) > 0.8
    r = oracle.gjk_epa(oracle.mesh_instance(box), oracle.mesh_instance(sph, (1.5, 0, 0)))
    assert abs(r["depth"] - 0.5) < 0.15
    # witness points lie on the shapes' surfaces within EPA tolerance
    for _ in range(50):
        pa, pb = RNG.uniform(-0.5, 0.5, 3), RNG.uniform(-0.5, 0.5, 3)
        qa, qb = rand_quat(), rand_quat()
        r = oracle.gjk_epa(oracle.mesh_instance(box, pa, qa), oracle.mesh_instance(box, pb, qb))
        if r is None or r == "degenerate":
            continue
        la = qrot(np.array([qa[0], -qa[1], -qa[2], -qa[3]]), r["world_a"] - pa)
        assert np.abs(la).max() < 1 + 1e-5


# ---------------- world pieces ----------------
def _world(oracle, pk, kind, n, **kw):
    s = pk.scene(kind, None, n)
    fl = kw.pop("flags", s["flags"])
    w = oracle.OracleWorld(iterations=s["iterations"], solver_mode=kw.pop("solver_mode", 1), flags=fl, **kw)
    w.load_scene(s)
    return s, w


def test_first_step_discovers_no_pairs(oracle, pk):
    """dynamic_bvh::add stores exact bounds (bvh.h:294-302) and step 1 integrates only
    velocities before the leaf update, so no leaf is reinserted and no pair exists."""
    s, w = _world(oracle, pk, 0, 256)
    w.step(s["dt"])
    assert w.counters()["moved"] == 0 and len(w.pairs()) == 0
    w.step(s["dt"])
    assert w.counters()["moved"] > 0 and len(w.pairs()) > 0


def _brute_pairs(b, static, prev):
    fat = b["fat"]
    lo, hi = fat[:, :3], fat[:, 3:]
    ov = np.all((lo[:, None, :] <= hi[None, :, :]) & (hi[:, None, :] >= lo[None, :, :]), axis=2)
    moved = b["moved"].astype(bool) & ~static
    n = len(fat)
    prevset = set(int(k) for k in prev)
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            if not ov[i, j] or (static[i] and static[j]):
                continue
            k = (i << 32) | j
            if moved[i] or moved[j] or k in prevset:
                out.append(k)
    return np.array(sorted(out), dtype=np.uint64)


def test_pair_set_closed_form_brute_force(oracle, pk):
    """collision_phases.h:377-436: the incremental pair set equals the closed form
    over fat bounds, moved flags and the previous set, checked by O(n^2) brute force."""
    s, w = _world(oracle, pk, 0, 300)
    static = s["is_static"].astype(bool)
    prev = w.pairs()
    for k in range(40):
        w.step(s["dt"])
        b = w.bodies()
        np.testing.assert_array_equal(w.pairs(), _brute_pairs(b, static, prev))
        prev = w.pairs()


def test_disabled_slot_never_pairs(oracle, pk):
    """A removed body's slot (is_static=2) has no broadphase leaf (dynamic_bvh::remove_leaf,
    collision/bvh.h:307): with every other body identical, the pair sets equal those of
    the world without that body after renumbering, and the slot never moves."""
    s, _ = _world(oracle, pk, 0, 200)
    d = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in s.items()}
    victim = 17
    d["is_static"][victim] = 2
    keep = np.array([i for i in range(len(s["shape"])) if i != victim])
    r = dict(s)
    for k in ("is_static", "shape", "pos", "vel", "angvel", "quat", "mass", "inertia", "restitution", "friction"):
        r[k] = s[k][keep]
    wd = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    wd.load_scene(d)
    wr = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    wr.load_scene(r)
    remap = np.where(np.arange(len(s["shape"])) > victim, np.arange(len(s["shape"])) - 1, np.arange(len(s["shape"])))
    p0 = d["pos"][victim].copy()
    for k in range(40):
        wd.step(s["dt"])
        wr.step(s["dt"])
        pd = wd.pairs()
        a, b = pd >> 32, pd & 0xFFFFFFFF
        assert not np.any(a == victim) and not np.any(b == victim)
        mapped = np.sort((remap[a].astype(np.uint64) << 32) | remap[b].astype(np.uint64))
        np.testing.assert_array_equal(mapped, wr.pairs())
        np.testing.assert_array_equal(wd.bodies()["pos"][keep], wr.bodies()["pos"])
    assert np.array_equal(wd.bodies()["pos"][victim], p0)


def test_all_moved_pairs_are_all_overlaps(oracle, pk):
    s, w = _world(oracle, pk, 1, 1500)
    w.step(s["dt"])
    b = w.bodies()
    b["moved"][:] = 1
    np.testing.assert_array_equal(w.pairs(), _brute_pairs(b, s["is_static"].astype(bool), []))


def test_fat_bounds_margin_and_prediction(oracle, pk):
    """update_leaf (bvh.cpp:475-514): a reinserted leaf is the true bounds grown by
    0.1 m and stretched one-sidedly by vel*dt; an unmoved leaf keeps its bounds."""
    # free-flying boxes far apart: no contacts, so the velocity after the step is the
    # velocity update_leaf saw (integrate_vel runs before the leaf update)
    n = 24
    pos = np.stack([np.arange(n) * 10.0, np.full(n, 50.0), np.zeros(n)], 1)
    vel = RNG.normal(size=(n, 3)) * 3
    quat = np.array([rand_quat() for _ in range(n)])
    s = _free_body_scene(oracle, pos, vel, np.zeros((n, 3)), quat)
    s["angvel"] = RNG.normal(size=(n, 3))
    w = oracle.OracleWorld(iterations=10, solver_mode=1)
    w.load_scene(s)
    dt = 1 / 60
    checked = 0
    b0 = w.bodies()
    for k in range(30):
        w.step(dt)
        b1 = w.bodies()
        for i in range(n):
            tb = oracle.instance_bounds(s["verts"], b0["pos"][i], b0["quat"][i])
            if b1["moved"][i]:
                disp = b1["vel"][i] * dt
                lo, hi = tb[:3] - 0.1, tb[3:] + 0.1
                lo = np.where(disp < 0, lo + disp, lo)
                hi = np.where(disp < 0, hi, hi + disp)
                np.testing.assert_array_equal(b1["fat"][i], np.concatenate([lo, hi]))
                checked += 1
            else:  # the fat bounds still contain the true bounds: nothing changes
                np.testing.assert_array_equal(b1["fat"][i], b0["fat"][i])
                assert np.all(b1["fat"][i, :3] <= tb[:3]) and np.all(b1["fat"][i, 3:] >= tb[3:])
        b0 = b1
    assert checked > n


def _free_body_scene(oracle, pos, vel, angvel, quat, mass=1.0, inertia=None):
    box = oracle.box_verts((0.2, 0.3, 0.4))
    n = len(pos)
    return {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
            "is_static": np.zeros(n, np.uint8), "shape": np.zeros(n, np.uint32),
            "pos": np.array(pos, float), "vel": np.array(vel, float), "angvel": np.array(angvel, float),
            "quat": np.array(quat, float), "mass": np.full(n, mass),
            "inertia": np.tile((np.eye(3) if inertia is None else inertia).reshape(9), (n, 1)),
            "restitution": np.full(n, 0.5), "friction": np.full(n, 0.5)}


def test_ballistic_closed_form(oracle):
    """semi_implicit_euler (integrate.h:36-46): v_n = v_0 + n g dt, x_n = x_0 + dt sum v_k."""
    s = _free_body_scene(oracle, [[0, 100, 0]], [[1.0, 2.0, -0.5]], [[0, 0, 0]], [[1, 0, 0, 0]])
    w = oracle.OracleWorld(iterations=10, solver_mode=1)
    w.load_scene(s)
    dt, g = 1 / 60, -9.81
    v = np.array([1.0, 2.0, -0.5])
    x = np.array([0, 100.0, 0])
    for k in range(100):
        w.step(dt)
        v = v + np.array([0, g, 0]) * dt
        x = x + v * dt
        b = w.bodies()
        np.testing.assert_allclose(b["vel"][0], v, rtol=0, atol=1e-12)
        np.testing.assert_allclose(b["pos"][0], x, rtol=0, atol=1e-10)


def test_torque_free_spin_closed_form(oracle):
    """With unit inertia the gyroscopic term vanishes: omega stays constant and
    q_n = exp(omega dt)^n q_0 (integrate.h:20-31, 42-46)."""
    w0 = np.array([0.0, 0.0, 2.0])
    s = _free_body_scene(oracle, [[0, 0, 0]], [[0, 0, 0]], [w0], [[1, 0, 0, 0]])
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=10, solver_mode=1)
    w.load_scene(s)
    dt = 0.01
    for k in range(1, 200):
        w.step(dt)
        b = w.bodies()
        np.testing.assert_array_equal(b["angvel"][0], w0)
        ang = 2.0 * dt * k
        np.testing.assert_allclose(b["quat"][0], [math.cos(ang / 2), 0, 0, math.sin(ang / 2)], atol=1e-12)


def test_solver_conserves_momentum_and_respects_cones(oracle):
    """Two dynamic boxes collide with no gravity: impulses are equal and opposite, so
    the total linear momentum is unchanged; normal impulses are >= 0 and friction
    impulses stay inside the cone |lambda_t| <= mu lambda_n (constraint.h:1107-1201)."""
    box = oracle.box_verts((0.5, 0.5, 0.5))
    s = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
         "is_static": np.zeros(2, np.uint8), "shape": np.zeros(2, np.uint32),
         "pos": np.array([[-1.2, 0, 0], [1.2, 0.3, 0.1]]), "vel": np.array([[3.0, 0, 0], [-2.0, 0.4, 0]]),
         "angvel": np.zeros((2, 3)), "quat": np.array([qaxis((0, 1, 1), 0.3), qaxis((1, 0, 0), 0.2)]),
         "mass": np.array([1.0, 2.5]), "inertia": np.array([np.eye(3).reshape(9) / 6, np.eye(3).reshape(9) * 0.4]),
         "restitution": np.array([0.5, 0.2]), "friction": np.array([0.6, 0.4])}
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=10, solver_mode=0)
    w.load_scene(s)
    p0 = s["mass"] @ s["vel"]
    touched = False
    mu = math.sqrt(0.6 * 0.4)
    for k in range(120):
        w.step(1 / 120)
        b = w.bodies()
        np.testing.assert_allclose(s["mass"] @ b["vel"], p0, atol=1e-12)
        cnt, d = w.manifolds()
        for p in range(len(cnt)):
            for c in range(cnt[p]):
                ln, t1, t2 = d[p, c, 10], d[p, c, 11], d[p, c, 12]
                touched |= ln > 0
                assert ln >= 0
                assert math.hypot(t1, t2) <= mu * ln * (1 + 1e-12) + 1e-15
    assert touched


def test_resting_box_settles(oracle):
    """A box dropped flat onto a static box comes to rest."""
    box = oracle.box_verts((0.5, 0.5, 0.5))
    ground = oracle.box_verts((5, 0.5, 5))
    s = {"shape_offsets": np.array([0, 8, 16], np.uint32), "verts": np.concatenate([ground, box]),
         "is_static": np.array([1, 0], np.uint8), "shape": np.array([0, 1], np.uint32),
         "pos": np.array([[0, -0.5, 0], [0, 0.6, 0]]), "vel": np.zeros((2, 3)), "angvel": np.zeros((2, 3)),
         "quat": np.tile([1.0, 0, 0, 0], (2, 1)), "mass": np.array([0.0, 1.0]),
         "inertia": np.tile(np.eye(3).reshape(9) / 6, (2, 1)), "restitution": np.zeros(2),
         "friction": np.full(2, 0.5)}
    w = oracle.OracleWorld(iterations=20, solver_mode=1, flags=8)
    w.load_scene(s)
    for k in range(400):
        w.step(1 / 120)
    b = w.bodies()
    assert np.abs(b["vel"][1]).max() < 1e-2
    assert 0.45 < b["pos"][1, 1] < 0.52


def test_colours_are_proper(oracle, pk):
    """No two manifolds of one colour share a dynamic body, so a colour's manifolds can
    be solved in any order (or in parallel) with identical results."""
    s, w = _world(oracle, pk, 2, 2)
    static = s["is_static"].astype(bool)
    for k in range(30):
        w.step(s["dt"])
    colours, keys = w.colours(with_keys=True)
    c = w.counters()
    assert c["rows"] > 0 and len(colours) > 0 and c["colours"] == colours.max() + 1
    seen = {}
    for col, k in zip(colours.tolist(), keys.tolist()):
        for body in (k >> 32, k & 0xFFFFFFFF):
            if static[body]:
                continue
            assert (col, body) not in seen, (col, body)
            seen[(col, body)] = k


def test_solver_order_modes_agree_closely(oracle, pk):
    """Sequential and coloured Gauss-Seidel orders differ only in summation order; over
    a few steps of a box pile the body states agree closely."""
    s = pk.scene(0, None, 128)
    a = oracle.OracleWorld(iterations=10, solver_mode=0)
    b = oracle.OracleWorld(iterations=10, solver_mode=1)
    a.load_scene(s)
    b.load_scene(s)
    for k in range(30):
        a.step(s["dt"])
        b.step(s["dt"])
    np.testing.assert_allclose(a.bodies()["pos"], b.bodies()["pos"], atol=5e-2)
