"""Pins of the oracle's contact path on the CPU: contact rows and their solve
(include/physkit/collision/constraint.h:874-953, 1052-1201), manifold persistence and
reduction (collision_phases.h:124-198, 244-320), EPA witness points
(src/collision.cpp:424-454) and the gyroscopic term of integrate_vel
(include/physkit/core/particle.h:72-76).

Each expected value comes from what the mechanics or the geometry fixes, not from the
oracle's own arithmetic:
  * one application of a contact's normal row leaves the contact point's relative
    normal velocity equal to -bias (the row's constraint equation), whatever the
    inertia or lever arm; bias is the reference's Baumgarte term
    -(0.1/dt) max(0, pen - 0.005) or the restitution term e v_n, e = max(e_a, e_b),
    taken when v_n < -2|g|dt (constraint.h:880-881, 906-924; the threshold is
    2 |g| dt, constraint.h:1071);
  * one solve of the 2x2 friction block with an unsaturated cone stops the contact
    point's tangential motion exactly;
  * a point mass sliding on a static box loses mu g dt of speed per step along its
    direction of motion (Coulomb friction on a circular cone), mu = sqrt(mu_a mu_b)
    (constraint.h:1076), and sticks once its speed is below that;
  * manifold persistence follows the reference's stated thresholds (5 mm warm-start
    radius on either body, breaking depth 0.05 m, tangential drift^2 0.06 m^2,
    collision_phases.h:212-213, 267-305), scripted case by case;
  * add_reduce keeps the deepest point, the point farthest from it, the point making
    the largest triangle with those two and the point farthest from that third point
    (collision_phases.h:139-198), checked on hand-built pools whose answer is plain;
  * EPA's witness points satisfy world_b - world_a = normal * depth (the relation
    narrow_phase::calculate relies on, collision_phases.h:289-291), and each lies on
    its body's supporting plane;
  * a torque-free body with anisotropic inertia keeps its world angular momentum
    I_w omega to first order in dt (Euler's equations in the world frame).

Planted mistakes each of these tests was run against (each fails at least one test;
tests/test_oracle_mutants.py rebuilds the oracle with each one and checks):
  Baumgarte factor 0.1 -> 0.2; slop 0.005 -> 0.01; Baumgarte without the slop;
  restitution max(e_a, e_b) -> min; restitution threshold 2|g|dt -> 0;
  friction mu sqrt(mu_a mu_b) -> (mu_a + mu_b)/2; friction cone hypot -> max(|l1|,|l2|);
  2x2 friction block off-diagonal sign; normal impulse clamp at 0 removed;
  warm-start radius 5 mm -> 20 mm; drift^2 0.06 -> 0.01; breaking depth 0.05 -> 0.01;
  add_reduce fourth point measured from best[0] instead of best[2]; add_reduce
  deepest-point tie taken by the last index; EPA world_b barycentric weights u <-> w;
  gyroscopic term sign flipped.
"""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(4242)
ALL_MOVED = 1


def qaxis(axis, ang):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * axis])


def qmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def scene(oracle, bodies):
    """bodies: list of dicts with half (box half extents), pos, and optionally static,
    vel, angvel, quat, mass, inertia (3x3), e, mu. Body i has shape i."""
    verts, offs = [], [0]
    for b in bodies:
        v = oracle.box_verts(b["half"])
        verts.append(v)
        offs.append(offs[-1] + len(v))
    n = len(bodies)
    return {"shape_offsets": np.array(offs, np.uint32), "verts": np.concatenate(verts),
            "is_static": np.array([1 if b.get("static") else 0 for b in bodies], np.uint8),
            "shape": np.arange(n, dtype=np.uint32),
            "pos": np.array([b["pos"] for b in bodies], float),
            "vel": np.array([b.get("vel", (0, 0, 0)) for b in bodies], float),
            "angvel": np.array([b.get("angvel", (0, 0, 0)) for b in bodies], float),
            "quat": np.array([b.get("quat", (1, 0, 0, 0)) for b in bodies], float),
            "mass": np.array([b.get("mass", 1.0) for b in bodies], float),
            "inertia": np.array([np.asarray(b.get("inertia", np.eye(3)), float).reshape(9) for b in bodies]),
            "restitution": np.array([b.get("e", 0.0) for b in bodies], float),
            "friction": np.array([b.get("mu", 0.0) for b in bodies], float)}


GROUND = {"half": (20.0, 0.5, 20.0), "pos": (0.0, -0.5, 0.0), "static": True, "mass": 0.0}


def world(oracle, bodies, gravity=(0.0, 0.0, 0.0), iterations=1, flags=ALL_MOVED):
    s = scene(oracle, bodies)
    w = oracle.OracleWorld(gravity=gravity, iterations=iterations, solver_mode=0, flags=flags)
    w.load_scene(s)
    return w, s


def point_velocity(b, i, r):
    return b["vel"][i] + np.cross(b["angvel"][i], r)


def contact_frame(s, cnt, data, p=0, c=0):
    """Lever arms, unit normal and penetration of contact c of manifold p, from the
    manifold's body-frame points and the bodies' poses before integrate_pos."""
    k = data[p, c]
    n = k[0:3] / np.linalg.norm(k[0:3])
    ra = qmat(s["quat"][0]) @ k[3:6]
    rb = qmat(s["quat"][1]) @ k[6:9]
    pen = float(((s["pos"][1] + rb) - (s["pos"][0] + ra)) @ n)
    return ra, rb, n, pen


def rel_normal_velocity(b, ra, rb, n):
    return float((point_velocity(b, 0, ra) - point_velocity(b, 1, rb)) @ n)


def stepped_once(oracle, bodies, gravity=(0.0, 0.0, 0.0), iterations=1, dt=1 / 60):
    """Runs integrate_vel, pairs, narrowphase and solve (no integrate_pos, so the poses
    are the ones the solve saw); returns the world, scene and the single manifold."""
    w, s = world(oracle, bodies, gravity, iterations)
    for st in ("integrate_vel", "pairs", "narrow", "solve"):
        w.run_stage(st, dt)
    cnt, data = w.manifolds()
    assert len(cnt) == 1 and cnt[0] == 1, cnt
    return w, s, data


# ---------------- normal row: Baumgarte and slop ----------------
@pytest.mark.parametrize("pen", [0.02, 0.003, 0.0052])
@pytest.mark.parametrize("tilt", [0.0, 0.35])
def test_normal_row_solves_baumgarte_with_slop(oracle, pen, tilt):
    """A dynamic box resting in a static box with penetration pen, at rest, no gravity,
    mu = 0, one solver iteration: afterwards the contact point separates at
    (0.1/dt) max(0, pen - 0.005). Tilted boxes put the contact at a corner, off the
    centre, with anisotropic inertia, so the lever arm and world inertia matter."""
    dt = 1 / 60
    half = np.array([0.3, 0.2, 0.25])
    q = qaxis((1, 0.3, -0.5), tilt)
    lowest = (np.abs(qmat(q)) @ half)[1]  # half height of the rotated box
    box = {"half": half, "pos": (0.1, lowest - pen, -0.2), "quat": q, "mass": 2.0,
           "inertia": np.diag([0.5, 1.2, 0.8])}
    w, s, data = stepped_once(oracle, [GROUND, box], dt=dt)
    ra, rb, n, p = contact_frame(s, *w.manifolds())
    assert p == pytest.approx(pen, abs=1e-9)  # EPA depth of a box resting in a box
    v = rel_normal_velocity(w.bodies(), ra, rb, n)
    assert v == pytest.approx((0.1 / dt) * max(0.0, p - 0.005), abs=1e-12)
    assert data[0, 0, 10] >= 0.0


def test_normal_row_between_two_dynamic_bodies(oracle):
    """Both bodies dynamic: the same constraint equation holds at the contact point and
    the impulse is equal and opposite (linear momentum unchanged)."""
    dt = 1 / 60
    a = {"half": (0.5, 0.5, 0.5), "pos": (0.0, 0.0, 0.0), "mass": 3.0, "inertia": np.diag([1.0, 0.7, 1.4]),
         "quat": qaxis((0, 0, 1), 0.1)}
    b = {"half": (0.4, 0.3, 0.2), "pos": (0.2, 0.75, 0.1), "mass": 1.5, "inertia": np.diag([0.2, 0.3, 0.4]),
         "quat": qaxis((1, 1, 0), 0.2)}
    w, s, _ = stepped_once(oracle, [a, b], dt=dt)
    ra, rb, n, p = contact_frame(s, *w.manifolds())
    assert p > 0.02
    bd = w.bodies()
    assert rel_normal_velocity(bd, ra, rb, n) == pytest.approx((0.1 / dt) * (p - 0.005), rel=1e-10)
    np.testing.assert_allclose(s["mass"] @ bd["vel"], 0.0, atol=1e-13)


def test_normal_impulse_only_pushes(oracle):
    """Bodies overlapping while already separating fast: the row would pull them
    together, so its impulse is clamped at zero and nothing changes."""
    box = {"half": (0.3, 0.2, 0.25), "pos": (0.0, 0.19, 0.0), "vel": (0.0, 3.0, 0.0)}
    w, s, data = stepped_once(oracle, [GROUND, box])
    assert data[0, 0, 10] == 0.0
    np.testing.assert_array_equal(w.bodies()["vel"][1], [0.0, 3.0, 0.0])


# ---------------- normal row: restitution ----------------
@pytest.mark.parametrize("v_over_gdt, bounce", [(-12.0, True), (-2.5, True), (-1.5, False), (-0.5, False)])
def test_restitution_rule_and_threshold(oracle, v_over_gdt, bounce):
    """A box hits a static box at normal speed v (after gravity) with penetration inside
    the slop: above 2|g|dt it leaves at max(e_a, e_b) |v|, below it stops."""
    g, dt = 9.81, 1 / 60
    v_after_gravity = v_over_gdt * g * dt
    ground = dict(GROUND, e=0.2)
    box = {"half": (0.3, 0.2, 0.25), "pos": (0.0, 0.198, 0.0), "vel": (0.0, v_after_gravity + g * dt, 0.0),
           "e": 0.7, "inertia": np.diag([0.3, 0.5, 0.4])}
    w, s, _ = stepped_once(oracle, [ground, box], gravity=(0.0, -g, 0.0), dt=dt)
    ra, rb, n, p = contact_frame(s, *w.manifolds())
    assert p < 0.005
    v_n0 = float((np.zeros(3) - np.array([0.0, v_after_gravity, 0.0])) @ n)  # box at rest angularly
    assert v_n0 < 0
    v = rel_normal_velocity(w.bodies(), ra, rb, n)
    assert v == pytest.approx(-0.7 * v_n0 if bounce else 0.0, abs=1e-12)


# ---------------- friction ----------------
def test_friction_block_stops_tangential_motion(oracle):
    """A single contact with a large friction coefficient, one iteration: the 2x2
    friction solve runs last and leaves no tangential velocity at the contact point,
    with lever arms and anisotropic inertia coupling the two tangent directions."""
    box = {"half": (0.3, 0.2, 0.25), "pos": (0.0, 0.17, 0.0), "quat": qaxis((0.2, 1.0, 0.4), 0.6),
           "vel": (1.3, -0.2, -0.7), "angvel": (0.4, -1.1, 0.9), "inertia": np.diag([0.3, 0.9, 0.5]),
           "mu": 10.0}
    half = np.array(box["half"])
    box["pos"] = (0.0, (np.abs(qmat(box["quat"])) @ half)[1] - 0.01, 0.0)
    ground = dict(GROUND, mu=10.0)
    w, s, _ = stepped_once(oracle, [ground, box])
    ra, rb, n, p = contact_frame(s, *w.manifolds())
    b = w.bodies()
    vrel = point_velocity(b, 0, ra) - point_velocity(b, 1, rb)
    vt = vrel - (vrel @ n) * n
    assert np.linalg.norm(vt) < 1e-12
    assert abs(vrel @ n) > 0.05  # the normal row was undone in part: the block is coupled


@pytest.mark.parametrize("direction", [(1.0, 0.0, 0.0), (0.6, 0.0, 0.8), (-0.28, 0.0, 0.96)])
def test_sliding_box_decelerates_at_sqrt_mu_g(oracle, direction):
    """A box (very large inertia: a point mass) slides on a static box with gravity:
    the normal impulses carry its weight and the friction impulses sit on the cone, so
    its speed drops by mu g dt per step along its direction of motion, with
    mu = sqrt(mu_ground mu_box); once slower than that it sticks."""
    g, dt = 9.81, 1 / 60
    mu = math.sqrt(0.3 * 0.8)
    d = np.asarray(direction)
    v0 = 4.0
    box = {"half": (0.3, 0.2, 0.25), "pos": (0.0, 0.199, 0.0), "vel": tuple(v0 * d),
           "inertia": np.eye(3) * 1e12, "mass": 1.5, "mu": 0.8}
    w, s = world(oracle, [dict(GROUND, mu=0.3), box], gravity=(0.0, -g, 0.0), iterations=10)
    steps = int(v0 / (mu * g * dt))
    for k in range(1, steps + 1):
        w.step(dt)
        v = w.bodies()["vel"][1]
        np.testing.assert_allclose(v, (v0 - k * mu * g * dt) * d, atol=1e-9)
    w.step(dt)
    np.testing.assert_allclose(w.bodies()["vel"][1], 0.0, atol=1e-9)
    w.step(dt)
    np.testing.assert_allclose(w.bodies()["vel"][1], 0.0, atol=1e-9)


# ---------------- manifold persistence ----------------
def contact_row(normal, la, lb, depth=0.0, ln=0.0, lt1=0.0, lt2=0.0):
    return np.concatenate([normal, la, lb, [depth, ln, lt1, lt2]])


class Pair:
    """A static ground box and an axis-aligned dynamic box resting in it 1 cm deep
    (both unrotated, so body-frame offsets are world offsets). new = the contact the
    narrowphase finds for the current poses."""

    def __init__(self, oracle, separated=False):
        box = {"half": (0.3, 0.2, 0.25), "pos": (0.05, 0.19, -0.1)}
        self.w, self.s = world(oracle, [GROUND, box])
        for st in ("integrate_vel", "pairs", "narrow"):
            self.w.run_stage(st, 1 / 60)
        self.key = int(self.w.pairs()[0])
        cnt, data = self.w.manifolds()
        assert cnt[0] == 1
        self.new = data[0, 0].copy()
        self.xa, self.xb = self.s["pos"][0], self.s["pos"][1].copy()
        if separated:  # lift the box 2 cm clear: the pair stays listed, GJK finds no contact
            self.xb = self.xb + np.array([0.0, 0.03, 0.0])
            self.w.set_bodies([1], pos=self.xb[None])

    def world_points(self, k):
        return self.xa + k[3:6], self.xb + k[6:9]

    def old_at(self, wa, wb, normal, ln=0.0, lt1=0.0, lt2=0.0):
        return contact_row(normal, wa - self.xa, wb - self.xb, 0.0, ln, lt1, lt2)

    def run(self, old):
        self.w.set_manifold(self.key, np.array(old).reshape(-1, 13))
        self.w.run_stage("narrow")
        cnt, data = self.w.manifolds()
        return data[0, :cnt[0]].copy()


@pytest.mark.parametrize("offset_a, offset_b, transfer", [
    ((0.004, 0, 0), (0.3, 0, 0), True),      # 4 mm on body a
    ((0.3, 0, 0), (0, 0, 0.004), True),      # 4 mm on body b
    ((0.0035, 0, 0.0035), (0.3, 0, 0), True),  # 4.95 mm
    ((0.006, 0, 0), (0.006, 0, 0), False),   # 6 mm on both
    ((0, 0.0036, 0.0036), (0.0036, 0, 0.0036), False),  # 5.09 mm on both
])
def test_warm_start_radius(oracle, offset_a, offset_b, transfer):
    """An old point within 5 mm of the new one, on either body, hands its impulses to
    the new point and is dropped (collision_phases.h:271-284); farther on both, the new
    point starts from zero and the old one is judged on depth and drift."""
    pr = Pair(oracle)
    new = pr.new
    old = new.copy()
    old[3:6] += offset_a
    old[6:9] += offset_b
    old[10:13] = (1.5, 0.25, -0.5)
    out = pr.run([old])
    if transfer:
        assert len(out) == 1
        np.testing.assert_array_equal(out[0, :10], new[:10])
        np.testing.assert_array_equal(out[0, 10:13], (1.5, 0.25, -0.5))
    else:
        assert len(out) == 2  # the old point is kept (its points still coincide up to the offsets)
        np.testing.assert_array_equal(out[1], new)
        np.testing.assert_array_equal(out[0, 10:13], (1.5, 0.25, -0.5))
        np.testing.assert_array_equal(out[0, 0:3], new[0:3])


@pytest.mark.parametrize("separated", [False, True])
@pytest.mark.parametrize("depth, drift2, kept", [
    (-0.049, 0.0, True), (-0.051, 0.0, False), (0.01, 0.059, True), (0.01, 0.061, False),
    (-0.0499, 0.0599, True),
])
def test_old_point_kept_by_depth_and_drift(oracle, separated, depth, drift2, kept):
    """An old point far from the new one stays while its depth along the current normal
    exceeds -0.05 m and its tangential drift^2 is below 0.06 m^2 (collision_phases.h:
    286-305), taking the new contact's normal, or its own when there is none."""
    pr = Pair(oracle, separated)
    n_old = np.array([0.0, -1.0, 0.0]) if separated else np.array([0.3, -0.9, 0.1]) / math.sqrt(0.91)
    n = n_old if separated else pr.new[0:3] / np.linalg.norm(pr.new[0:3])
    t = np.cross(n, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    wa = pr.xa + np.array([0.2, 0.5, 0.15])  # 0.2+ m from the new point on both bodies
    wb = wa + n * depth + t * math.sqrt(drift2)
    out = pr.run([pr.old_at(wa, wb, n_old, ln=2.0)])
    n_new = 0 if separated else 1
    assert len(out) == (1 if kept else 0) + n_new
    if kept:
        assert out[0, 9] == pytest.approx(depth, abs=1e-12)
        np.testing.assert_array_equal(out[0, 0:3], n_old if separated else pr.new[0:3])
        assert out[0, 10] == 2.0


def test_events_follow_the_empty_state(oracle):
    """Begin when a manifold gains its first point, end when it loses its last
    (collision_phases.h:309-318); nothing while it stays non-empty."""
    pr = Pair(oracle)
    pr.run([])  # empty -> new point
    k, t = pr.w.events()
    assert list(k) == [pr.key] and list(t) == [1]
    pr.run([pr.new])
    k, t = pr.w.events()
    assert len(k) == 0
    sep = Pair(oracle, separated=True)
    far = sep.old_at(sep.xa + np.array([0, 0.5, 0]), sep.xa + np.array([0, 0.5, 0]) + np.array([0, 0.06, 0]),
                     np.array([0.0, -1.0, 0.0]))
    out = sep.run([far])  # depth -0.06: dropped, no new point
    k, t = sep.w.events()
    assert len(out) == 0 and list(k) == [sep.key] and list(t) == [2]


# ---------------- add_reduce ----------------
def pool_point(x, z, depth, tag):
    return contact_row((0.0, 1.0, 0.0), (x, 0.0, z), (tag, 0.0, 0.0), depth)


def test_add_reduce_hand_pool(oracle):
    """Pool in local_a (x, z): P0 (0,0) deepest; P1 (2,0) farthest from P0; P2 (1,1) the
    largest triangle with P0, P1 (area^2 4 against 3.24 and 0.16); then P4 (-0.3,-0.2)
    is farthest from P2 (3.13 against 0.05 for P3) while P3 (1.2,0.9) is the farther
    from P0. The result is [P0, P1, P2, P4] in that order."""
    P = [pool_point(0, 0, 0.05, 0), pool_point(2, 0, 0.01, 1), pool_point(1, 1, 0.02, 2),
         pool_point(1.2, 0.9, 0.04, 3), pool_point(-0.3, -0.2, 0.03, 4)]
    for old_idx, new_idx in (([3, 0, 4, 1], 2), ([0, 1, 2, 3], 4), ([4, 3, 2, 1], 0)):
        out = oracle.manifold_add([P[i] for i in old_idx], P[new_idx])
        assert list(out[:, 6].astype(int)) == [0, 1, 2, 4]


def test_add_reduce_keeps_the_first_of_equal_depths(oracle):
    """Comparisons are strict, so among equal depths the earliest pool entry is the
    deepest point; the new point (last in the pool) wins only when strictly deeper."""
    P = [pool_point(0, 0, 0.03, 0), pool_point(1, 0, 0.03, 1), pool_point(0, 1, 0.01, 2),
         pool_point(1, 1, 0.01, 3), pool_point(0.5, 0.5, 0.03, 4)]
    out = oracle.manifold_add(P[:4], P[4])
    assert int(out[0, 6]) == 0
    out = oracle.manifold_add(P[:4], pool_point(0.5, 0.5, 0.0300001, 4))
    assert int(out[0, 6]) == 4


def test_add_reduce_invariants_random(oracle):
    """Random pools: four distinct pool members, the deepest first, the second the
    farthest from it, and add_contact below four points just appends."""
    for _ in range(300):
        P = [contact_row(RNG.normal(size=3), RNG.normal(size=3), (i, 0, 0), RNG.uniform(-0.05, 0.1),
                         RNG.uniform(0, 2), RNG.normal(), RNG.normal()) for i in range(5)]
        out = oracle.manifold_add(P[:4], P[4])
        tags = out[:, 6].astype(int)
        assert len(set(tags)) == 4
        depths = np.array([p[9] for p in P])
        assert tags[0] == int(np.argmax(depths))
        d2 = [np.sum((P[i][3:6] - P[tags[0]][3:6]) ** 2) if i != tags[0] else -1 for i in range(5)]
        assert tags[1] == int(np.argmax(d2))
        for t, row in zip(tags, out):
            np.testing.assert_array_equal(row, P[t])
        for m in range(4):
            out = oracle.manifold_add(P[:m], P[4])
            np.testing.assert_array_equal(out, np.array(P[:m] + [P[4]]))


# ---------------- EPA witness points ----------------
def _random_hull(n):
    v = RNG.normal(size=(n, 3)) * RNG.uniform(0.1, 0.6, 3)
    return v


@pytest.mark.parametrize("kind", ["box", "hull"])
def test_epa_witnesses_span_normal_times_depth(oracle, kind):
    """For overlapping convex shapes, world_b - world_a = normal * depth (the EPA
    witness pair straddles the penetration along the normal), and world_a lies on A's
    supporting plane for -normal, world_b on B's for +normal (each face vertex of the
    Minkowski difference is a support point within the 1e-6 m convergence margin)."""
    checked = 0
    for _ in range(200):
        if kind == "box":
            va, vb = oracle.box_verts(RNG.uniform(0.1, 0.8, 3)), oracle.box_verts(RNG.uniform(0.1, 0.8, 3))
        else:
            va, vb = _random_hull(int(RNG.integers(6, 20))), _random_hull(int(RNG.integers(6, 20)))
        qa, qb = RNG.normal(size=4), RNG.normal(size=4)
        qa, qb = qa / np.linalg.norm(qa), qb / np.linalg.norm(qb)
        pa, pb = RNG.normal(size=3) * 0.3, RNG.normal(size=3) * 0.3
        r = oracle.gjk_epa(oracle.mesh_instance(va, pa, qa), oracle.mesh_instance(vb, pb, qb))
        if not isinstance(r, dict):
            continue
        n, d = r["normal"], r["depth"]
        assert abs(np.linalg.norm(n) - 1.0) < 1e-9
        np.testing.assert_allclose(r["world_b"] - r["world_a"], n * d, atol=2e-6)
        wa_verts = (qmat(qa) @ va.T).T + pa
        wb_verts = (qmat(qb) @ vb.T).T + pb
        assert abs(r["world_a"] @ -n - np.max(wa_verts @ -n)) < 2e-6
        assert abs(r["world_b"] @ n - np.max(wb_verts @ n)) < 2e-6
        checked += 1
    assert checked > 60


# ---------------- gyroscopic term ----------------
def _spin(oracle, dt, T, inertia, w0, q0):
    box = {"half": (0.2, 0.3, 0.4), "pos": (0.0, 50.0, 0.0), "angvel": w0, "quat": q0, "inertia": inertia}
    w, s = world(oracle, [box], flags=0)
    iw0, _, _ = w.derived()
    L0 = iw0[0].reshape(3, 3) @ np.asarray(w0, float)
    E0 = 0.5 * np.asarray(w0) @ iw0[0].reshape(3, 3) @ np.asarray(w0)
    w.step(dt, int(round(T / dt)))
    iw, _, _ = w.derived()
    om = w.bodies()["angvel"][0]
    L = iw[0].reshape(3, 3) @ om
    E = 0.5 * om @ iw[0].reshape(3, 3) @ om
    return np.linalg.norm(L - L0) / np.linalg.norm(L0), abs(E - E0) / E0


def test_torque_free_anisotropic_body_keeps_angular_momentum(oracle):
    """omega' = I_w^-1 (tau - omega x (I_w omega)) with tau = 0 is Euler's equation in
    the world frame: d(I_w omega)/dt = 0. Semi-implicit Euler keeps L = I_w omega to
    first order, so the error is small and halves with dt; the opposite sign of the
    gyroscopic term makes L turn at twice the body's rate instead."""
    inertia = np.diag([1.0, 2.0, 3.0])
    w0 = (1.0, 0.5, 0.2)
    q0 = qaxis((0.3, 1.0, -0.2), 0.7)
    e1, k1 = _spin(oracle, 2e-3, 0.5, inertia, w0, q0)
    e2, k2 = _spin(oracle, 1e-3, 0.5, inertia, w0, q0)
    assert e1 < 5e-3 and e2 < 3e-3, (e1, e2)
    assert 1.6 < e1 / e2 < 2.4, (e1, e2)
    assert k1 < 5e-3 and k2 < 3e-3, (k1, k2)
