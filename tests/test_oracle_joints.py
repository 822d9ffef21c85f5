"""Pins of the oracle's joint constraints (include/physkit/collision/constraint.h:139-834,
solved in constraint_solver::solve_constraints :1107-1201), on the CPU.

The reference's tests hold no known answers for joints, so each check here comes from
the mechanics rather than from the oracle's formulas:
  * projected Gauss-Seidel over equality rows converges to the solution of the
    velocity-level system  (J M^-1 J^T + Gamma) lambda = -(J u0 + b),  u = u0 + M^-1 J^T lambda.
    The test builds J from the physical meaning of each joint (relative velocity of the
    anchor points along a direction, relative angular velocity about an axis), solves
    the system with numpy and compares the oracle's velocities after a converged solve;
  * the soft spring equals an implicit-Euler damped spring (closed form);
  * impulses are equal and opposite: linear momentum is conserved, and so is angular
    momentum for a distance joint (its impulse acts along the line of the anchors);
  * a single distance row is solved exactly in one Gauss-Seidel sweep.
"""
import math

import numpy as np
import pytest

RNG = np.random.default_rng(77)

DISTANCE, BALL, HINGE, SLIDER, WELD, SPRING = range(6)


def qaxis(axis, ang):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * axis])


def qmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def qmul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


def qconj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def rand_quat():
    q = RNG.normal(size=4)
    return q / np.linalg.norm(q)


def rand_spd():
    a = RNG.normal(size=(3, 3))
    return a @ a.T + 0.5 * np.eye(3)


def duff_basis(n):
    """Duff et al. 2017, "Building an Orthonormal Basis, Revisited" (JCGT 6(1))."""
    s = math.copysign(1.0, n[2])
    a = -1.0 / (s + n[2])
    b = n[0] * n[1] * a
    return (np.array([1.0 + s * n[0] * n[0] * a, s * b, -s * n[0]]),
            np.array([b, s + n[1] * n[1] * a, -n[1]]))


def scene(oracle, pos, quat, mass, inertia, vel=None, angvel=None, is_static=None):
    n = len(pos)
    box = oracle.box_verts((0.2, 0.3, 0.4))
    return {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
            "is_static": np.zeros(n, np.uint8) if is_static is None else np.asarray(is_static, np.uint8),
            "shape": np.zeros(n, np.uint32), "pos": np.array(pos, float),
            "vel": np.zeros((n, 3)) if vel is None else np.array(vel, float),
            "angvel": np.zeros((n, 3)) if angvel is None else np.array(angvel, float),
            "quat": np.array(quat, float), "mass": np.array(mass, float),
            "inertia": np.array([np.asarray(I, float).reshape(9) for I in inertia]),
            "restitution": np.full(n, 0.5), "friction": np.full(n, 0.5)}


# ------------------------------------------------------------ the velocity-level system
def physical_rows(j, st, s):
    """Constraint rows of one joint from its physical meaning. st: the bodies' pose at
    the solve (pos, quat); j carries the frames fixed at construction (computed here
    from the construction pose kept in j['_pose0']). Returns rows
    (ia, ib, lin_dir, r_a, r_b, ang_dir, bias, gamma): a point row constrains
    (v_a + w_a x r_a - v_b - w_b x r_b) . lin_dir, an angular row (w_a - w_b) . ang_dir."""
    a, b = j["a"], j["b"]
    xa, xb = st["pos"][a], st["pos"][b]
    Ra, Rb = qmat(st["quat"][a]), qmat(st["quat"][b])
    x0a, q0a, x0b, q0b = j["_pose0"]
    R0a, R0b = qmat(q0a), qmat(q0b)
    h = j["_dt"]
    rows = []
    t = j["type"]
    if t in (DISTANCE, SPRING, HINGE):
        la, lb = np.asarray(j["local_a"], float), np.asarray(j["local_b"], float)
    else:  # world anchor, fixed in each body at construction
        la = R0a.T @ (np.asarray(j["anchor"], float) - x0a)
        lb = R0b.T @ (np.asarray(j["anchor"], float) - x0b)
    ra, rb = Ra @ la, Rb @ lb
    pa, pb = xa + ra, xb + rb
    d = pa - pb
    if t in (DISTANCE, SPRING):
        L = np.linalg.norm(d)
        u = d / L
        if t == DISTANCE:
            rows.append((a, b, u, ra, rb, None, (0.1 / h) * (L - j["distance"]), 0.0))
        else:
            k, c = j["stiffness"], j["damping"]
            gamma = 1.0 / (h * (c + h * k))
            beta = h * k / (c + h * k)
            rows.append((a, b, u, ra, rb, None, (beta / h) * (L - j["distance"]), gamma))
        return rows
    if t in (BALL, HINGE, WELD):
        bf = 0.2 if t == HINGE else 0.1
        for e in np.eye(3):
            rows.append((a, b, e, ra, rb, None, (bf / h) * (d @ e), 0.0))
    if t == HINGE:
        axa = np.asarray(j["local_axis_a"], float)
        axb = np.asarray(j["local_axis_b"], float)
        ua, ub = Ra @ (axa / np.linalg.norm(axa)), Rb @ (axb / np.linalg.norm(axb))
        err = -np.cross(ua, ub)
        # any orthonormal pair spanning the plane normal to the hinge axis
        _, _, vt = np.linalg.svd(ua.reshape(1, 3))
        for n in (vt[1], vt[2]):
            rows.append((a, b, None, None, None, n, (0.2 / h) * (err @ n), 0.0))
    if t in (SLIDER, WELD):
        qrel = qmul(qconj(q0a), q0b)
        cur = qmul(qconj(st["quat"][a]), st["quat"][b])
        oe = qmul(qrel, qconj(cur))
        if oe[0] < 0:
            oe = -oe
        ang_err = Ra @ (2.0 * oe[1:])
    if t == WELD:
        for e in np.eye(3):
            rows.append((a, b, None, None, None, e, (0.1 / h) * (ang_err @ e), 0.0))
    if t == SLIDER:
        axn = np.asarray(j["axis"], float)
        axn = axn / np.linalg.norm(axn)
        t1w, t2w = duff_basis(axn)
        axis_a = Ra @ (R0a.T @ axn)
        t1, t2 = Ra @ (R0a.T @ t1w), Ra @ (R0a.T @ t2w)
        lever = pb - xa  # B's anchor seen from A's centre
        for n in (t1, t2):  # gamma: 1 % of the row's angular term, filled in by the solve
            e = float(np.clip(d @ n, -0.2, 0.2))
            rows.append((a, b, n, lever, rb, None, (0.1 / h) * e, "reg"))
        for n in (axis_a, t1, t2):
            rows.append((a, b, None, None, None, n, (0.1 / h) * (ang_err @ n), 0.0))
    return rows


def solve_velocity_system(rows, st, props, u0v, u0w):
    """Exact solution of the converged projected Gauss-Seidel (no bound is active)."""
    nb = len(st["pos"])
    Minv = np.zeros((6 * nb, 6 * nb))
    for i in range(nb):
        if props["static"][i]:
            continue
        R = qmat(props["quat_solve"][i])
        Minv[6 * i:6 * i + 3, 6 * i:6 * i + 3] = np.eye(3) / props["mass"][i]
        Minv[6 * i + 3:6 * i + 6, 6 * i + 3:6 * i + 6] = R @ np.linalg.inv(props["inertia"][i]) @ R.T
    J = np.zeros((len(rows), 6 * nb))
    bias = np.zeros(len(rows))
    gam = np.zeros(len(rows))
    for k, (a, b, n, ra, rb, m, bk, g) in enumerate(rows):
        if n is not None:
            J[k, 6 * a:6 * a + 3] = n
            J[k, 6 * a + 3:6 * a + 6] = np.cross(ra, n)
            J[k, 6 * b:6 * b + 3] -= n
            J[k, 6 * b + 3:6 * b + 6] -= np.cross(rb, n)
        else:
            J[k, 6 * a + 3:6 * a + 6] = m
            J[k, 6 * b + 3:6 * b + 6] -= m
        bias[k] = bk
        if g == "reg":  # slider: regularization of 1 % of the angular term (constraint.h:664-668)
            Ja = np.zeros(6 * nb)
            Ja[6 * a + 3:6 * a + 6] = J[k, 6 * a + 3:6 * a + 6]
            Ja[6 * b + 3:6 * b + 6] = J[k, 6 * b + 3:6 * b + 6]
            gam[k] = 0.01 * (Ja @ Minv @ Ja)
        else:
            gam[k] = g
    u0 = np.concatenate([np.concatenate([u0v[i], u0w[i]]) for i in range(nb)])
    A = J @ Minv @ J.T + np.diag(gam)
    lam = np.linalg.lstsq(A, -(J @ u0 + bias), rcond=None)[0]
    u = u0 + Minv @ J.T @ lam
    return u.reshape(nb, 6)[:, :3], u.reshape(nb, 6)[:, 3:]


def add_joints(w, joints, st, dt):
    for j in joints:
        a, b = j["a"], j["b"]
        j["_pose0"] = (st["pos"][a].copy(), st["quat"][a].copy(), st["pos"][b].copy(), st["quat"][b].copy())
        j["_dt"] = dt
        assert w.add_joint(j)


def chain_joints(s):
    """One joint of every type along a chain of 7 bodies (body 0 static), anchors at
    the midpoints, axes and rest lengths chosen at the construction pose."""
    P, Q = s["pos"], s["quat"]

    def mid(a, b):
        return 0.5 * (P[a] + P[b])

    def local(i, p):
        return qmat(Q[i]).T @ (p - P[i])

    ax = np.array([0.3, 0.2, 1.0])
    m12, m45, m56 = mid(1, 2), mid(4, 5), mid(5, 6)
    return [
        {"type": BALL, "slot": 0, "a": 0, "b": 1, "anchor": mid(0, 1)},
        {"type": HINGE, "slot": 0, "a": 1, "b": 2, "local_a": local(1, m12), "local_b": local(2, m12),
         "local_axis_a": qmat(Q[1]).T @ ax, "local_axis_b": qmat(Q[2]).T @ ax},
        {"type": SLIDER, "slot": 0, "a": 2, "b": 3, "anchor": mid(2, 3), "axis": P[3] - P[2]},
        {"type": WELD, "slot": 0, "a": 3, "b": 4, "anchor": mid(3, 4)},
        {"type": DISTANCE, "slot": 0, "a": 4, "b": 5, "local_a": local(4, m45 + 0.1),
         "local_b": local(5, m45 - 0.1), "distance": 0.5},
        {"type": SPRING, "slot": 0, "a": 5, "b": 6, "local_a": local(5, m56 + [0.2, 0, 0]),
         "local_b": local(6, m56 - [0.2, 0, 0]),
         "distance": 0.8, "stiffness": 40.0, "damping": 2.0},
    ]


def chain_scene(oracle, isotropic):
    n = 7
    pos = [[0.0, 0.0, 0.0]] + [[1.4 * k, -0.9 * k, 0.4 * k] for k in range(1, n)]
    quat = [rand_quat() for _ in range(n)]
    mass = [1.0] + list(RNG.uniform(0.5, 3.0, n - 1))
    inertia = [np.eye(3) * RNG.uniform(0.2, 1.0) if isotropic else rand_spd() for _ in range(n)]
    return scene(oracle, pos, quat, mass, inertia, vel=RNG.normal(size=(n, 3)) * 0.5,
                 is_static=[1] + [0] * (n - 1))


@pytest.mark.parametrize("mode", [0, 1])
def test_joint_chain_matches_velocity_system_first_step(oracle, mode):
    """Non-isotropic inertia and zero angular velocity (so integrate_vel adds only
    gravity): after one converged solve the velocities equal the exact solution."""
    s = chain_scene(oracle, isotropic=False)
    dt = 1 / 60
    g = np.array([0.0, -9.81, 0.0])
    w = oracle.OracleWorld(gravity=g, iterations=3000, solver_mode=mode)
    w.load_scene(s)
    st = {"pos": s["pos"].copy(), "quat": s["quat"].copy()}
    joints = chain_joints(s)
    add_joints(w, joints, st, dt)
    rows = [r for j in joints for r in physical_rows(j, st, s)]
    dyn = s["is_static"] == 0
    u0v = s["vel"] + np.where(dyn[:, None], g * dt, 0.0)
    props = {"static": ~dyn, "quat_solve": st["quat"], "mass": s["mass"],
             "inertia": [I.reshape(3, 3) for I in s["inertia"]]}
    vx, wx = solve_velocity_system(rows, st, props, u0v, s["angvel"])
    w.step(dt)
    b = w.bodies()
    scale = np.abs(np.concatenate([vx, wx])).max()
    np.testing.assert_allclose(b["vel"][dyn], vx[dyn], rtol=0, atol=1e-7 * scale)
    np.testing.assert_allclose(b["angvel"][dyn], wx[dyn], rtol=0, atol=1e-7 * scale)
    keys, acc, nrows = w.joints()
    assert list(nrows) == [1, 3, 5, 5, 6, 1]  # key order: distance, ball, hinge, slider, weld, spring


def test_joint_chain_matches_velocity_system_over_steps(oracle):
    """Isotropic inertia keeps integrate_vel free of the gyroscopic term, so every
    step's pre-solve velocity is the last step's plus g dt; warm starts change only
    the first iterate, so each step still lands on the exact solution."""
    s = chain_scene(oracle, isotropic=True)
    dt = 1 / 60
    g = np.array([0.0, -9.81, 0.0])
    w = oracle.OracleWorld(gravity=g, iterations=3000, solver_mode=1)
    w.load_scene(s)
    st = {"pos": s["pos"].copy(), "quat": s["quat"].copy()}
    joints = chain_joints(s)
    add_joints(w, joints, st, dt)
    dyn = s["is_static"] == 0
    inertia = [I.reshape(3, 3) for I in s["inertia"]]
    v, om = s["vel"].copy(), s["angvel"].copy()
    for k in range(30):
        st = {"pos": st["pos"], "quat": st["quat"]}
        rows = [r for j in joints for r in physical_rows(j, st, s)]
        props = {"static": ~dyn, "quat_solve": st["quat"], "mass": s["mass"], "inertia": inertia}
        vx, wx = solve_velocity_system(rows, st, props, v + np.where(dyn[:, None], g * dt, 0.0), om)
        w.step(dt)
        b = w.bodies()
        scale = max(1.0, np.abs(np.concatenate([vx, wx])).max())
        np.testing.assert_allclose(b["vel"][dyn], vx[dyn], rtol=0, atol=1e-7 * scale, err_msg=f"step {k}")
        np.testing.assert_allclose(b["angvel"][dyn], wx[dyn], rtol=0, atol=1e-7 * scale, err_msg=f"step {k}")
        v, om = b["vel"], b["angvel"]
        st = {"pos": b["pos"], "quat": b["quat"]}


@pytest.mark.parametrize("mode", [0, 1])
def test_joint_chain_large_errors_match_velocity_system(oracle, mode):
    """The chain is built, then every body is turned by up to 0.6 rad and shifted by up
    to 0.25 m (a caller moving objects between steps), so every joint starts the step
    with a large positional and angular error and the slider's clamp is active."""
    s = chain_scene(oracle, isotropic=False)
    dt = 1 / 60
    g = np.array([0.0, -9.81, 0.0])
    w = oracle.OracleWorld(gravity=g, iterations=3000, solver_mode=mode)
    w.load_scene(s)
    st0 = {"pos": s["pos"].copy(), "quat": s["quat"].copy()}
    joints = chain_joints(s)
    add_joints(w, joints, st0, dt)
    dyn = s["is_static"] == 0
    ids = np.nonzero(dyn)[0]
    q = s["quat"].copy()
    x = s["pos"].copy()
    for i in ids:
        q[i] = qmul(qaxis(RNG.normal(size=3), RNG.uniform(0.2, 0.6)), q[i])
        x[i] = x[i] + RNG.uniform(-0.25, 0.25, 3)
    w.set_bodies(ids, pos=x[ids], quat=q[ids])
    st = {"pos": x, "quat": q}
    rows = [r for j in joints for r in physical_rows(j, st, s)]
    props = {"static": ~dyn, "quat_solve": q, "mass": s["mass"], "inertia": [I.reshape(3, 3) for I in s["inertia"]]}
    vx, wx = solve_velocity_system(rows, st, props, s["vel"] + np.where(dyn[:, None], g * dt, 0.0), s["angvel"])
    w.step(dt)
    b = w.bodies()
    scale = np.abs(np.concatenate([vx, wx])).max()
    np.testing.assert_allclose(b["vel"][dyn], vx[dyn], rtol=0, atol=1e-7 * scale)
    np.testing.assert_allclose(b["angvel"][dyn], wx[dyn], rtol=0, atol=1e-7 * scale)


@pytest.mark.parametrize("jt", [WELD, SLIDER])
def test_orientation_error_takes_the_shorter_way(oracle, jt):
    """A body turned by 3.6 rad (more than half a turn) from its welded / slid
    orientation: the error quaternion's sign is flipped so the joint corrects along the
    shorter rotation (constraint.h:691-693, 814-817)."""
    s = scene(oracle, [[0, 0, 0], [1.0, 0.2, -0.1]], [[1, 0, 0, 0], rand_quat()], [1.0, 1.7],
              [np.eye(3), rand_spd()], is_static=[1, 0])
    dt = 1 / 60
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=3000, solver_mode=0)
    w.load_scene(s)
    j = {"type": jt, "a": 0, "b": 1, "anchor": [0.5, 0.1, 0.0], "axis": [1.0, 0.2, 0.0]}
    add_joints(w, [j], {"pos": s["pos"], "quat": s["quat"]}, dt)
    q = s["quat"].copy()
    q[1] = qmul(qaxis([0.2, 1.0, 0.3], 3.6), q[1])
    w.set_bodies([1], quat=q[1:2])
    st = {"pos": s["pos"], "quat": q}
    rows = physical_rows(j, st, s)
    props = {"static": np.array([True, False]), "quat_solve": q, "mass": s["mass"],
             "inertia": [I.reshape(3, 3) for I in s["inertia"]]}
    vx, wx = solve_velocity_system(rows, st, props, s["vel"], s["angvel"])
    w.step(dt)
    b = w.bodies()
    scale = np.abs(np.concatenate([vx, wx])).max()
    np.testing.assert_allclose(b["vel"][1], vx[1], rtol=0, atol=1e-7 * scale)
    np.testing.assert_allclose(b["angvel"][1], wx[1], rtol=0, atol=1e-7 * scale)


def test_spring_is_implicit_euler_damped_spring(oracle):
    """A body on a soft spring to a static body (no gravity, 1-D motion):
    m (v' - v) = -h (k x' + c v') with x' = x + h v', i.e.
    v' = (m v - h k x) / (m + h c + h^2 k), x the stretch at the start of the step."""
    m, k, c, h, rest = 2.0, 50.0, 3.0, 0.01, 1.0
    s = scene(oracle, [[0, 0, 0], [1.7, 0, 0]], [[1, 0, 0, 0]] * 2, [1.0, m], [np.eye(3)] * 2,
              vel=[[0, 0, 0], [0.4, 0, 0]], is_static=[1, 0])
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=4, solver_mode=0)
    w.load_scene(s)
    assert w.add_joint({"type": SPRING, "a": 1, "b": 0, "distance": rest, "stiffness": k, "damping": c})
    x, v = 1.7, 0.4
    for _ in range(300):
        v = (m * v - h * k * (x - rest)) / (m + h * c + h * h * k)
        x = x + h * v
        w.step(h)
        b = w.bodies()
        assert abs(b["vel"][1, 0] - v) <= 1e-12 * max(1.0, abs(v))
        assert abs(b["pos"][1, 0] - x) <= 1e-12
        np.testing.assert_array_equal(b["vel"][1, 1:], [0.0, 0.0])


def test_distance_row_solved_in_one_sweep(oracle):
    """One row, one iteration: afterwards the anchors' relative velocity along the
    joint line equals -(0.1/dt)(|p_a - p_b| - d) (Baumgarte, constraint.h:218-219)."""
    for trial in range(20):
        qa, qb = rand_quat(), rand_quat()
        Ia, Ib = rand_spd(), rand_spd()
        s = scene(oracle, [[0, 0, 0], [1.5, 0.4, -0.3]], [qa, qb], [1.3, 0.7], [Ia, Ib],
                  vel=RNG.normal(size=(2, 3)))
        la, lb = RNG.normal(size=3) * 0.3, RNG.normal(size=3) * 0.3
        dt = 1 / 120
        w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=1, solver_mode=0)
        w.load_scene(s)
        assert w.add_joint({"type": DISTANCE, "a": 0, "b": 1, "local_a": la, "local_b": lb, "distance": 1.2})
        ra, rb = qmat(qa) @ la, qmat(qb) @ lb
        d = (s["pos"][0] + ra) - (s["pos"][1] + rb)
        L = np.linalg.norm(d)
        u = d / L
        w.step(dt)
        b = w.bodies()
        rel = (b["vel"][0] + np.cross(b["angvel"][0], ra)) - (b["vel"][1] + np.cross(b["angvel"][1], rb))
        assert abs(rel @ u + (0.1 / dt) * (L - 1.2)) < 1e-11 * (1 + abs((0.1 / dt) * (L - 1.2)))


@pytest.mark.parametrize("jt", [DISTANCE, BALL, HINGE, SLIDER, WELD, SPRING])
def test_joint_impulses_conserve_momentum(oracle, jt):
    """Two dynamic bodies, no gravity: every joint applies equal and opposite linear
    impulses, so the total linear momentum is unchanged; a distance joint's impulse acts
    along the anchor line, so the angular momentum about the origin is unchanged too
    (evaluated with the world inertia the solve uses)."""
    qa, qb = rand_quat(), rand_quat()
    # isotropic inertia: integrate_vel adds no gyroscopic term, so the solve alone changes momentum
    Ia, Ib = np.eye(3) * 0.3, np.eye(3) * 0.8
    xa, xb = np.array([0.1, 0.2, 0.0]), np.array([1.6, -0.5, 0.7])
    s = scene(oracle, [xa, xb], [qa, qb], [1.3, 0.7], [Ia, Ib], vel=RNG.normal(size=(2, 3)))
    j = {"type": jt, "a": 0, "b": 1, "local_a": [0.1, 0.2, -0.1], "local_b": [-0.2, 0.0, 0.1],
         "local_axis_a": [0, 0, 1], "local_axis_b": [0.1, 0, 1], "anchor": 0.5 * (xa + xb),
         "axis": [1, 0.2, 0], "distance": 0.9, "stiffness": 30.0, "damping": 1.0}
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=10, solver_mode=0)
    w.load_scene(s)
    assert w.add_joint(j)
    p0 = s["mass"] @ s["vel"]
    L0 = sum(np.cross(s["pos"][i], s["mass"][i] * s["vel"][i]) for i in range(2))
    for k in range(20):
        before = w.bodies()
        iw, _, _ = w.derived()
        w.step(1 / 60)
        b = w.bodies()
        np.testing.assert_allclose(s["mass"] @ b["vel"], p0, rtol=0, atol=1e-12 * (1 + np.abs(p0).max()))
        if jt == DISTANCE:
            La = sum(iw[i].reshape(3, 3) @ before["angvel"][i] + np.cross(before["pos"][i], s["mass"][i] * before["vel"][i])
                     for i in range(2))
            Lb = sum(iw[i].reshape(3, 3) @ b["angvel"][i] + np.cross(before["pos"][i], s["mass"][i] * b["vel"][i])
                     for i in range(2))
            np.testing.assert_allclose(Lb, La, rtol=0, atol=1e-10 * (1 + np.abs(La).max()))


def test_coincident_distance_anchors_build_no_row(oracle):
    """Anchors closer than 1e-6 m: no row (constraint.h:204-205), the body moves freely;
    one step later they are 1 cm apart and the row is built and solved."""
    dt = 0.01
    s = scene(oracle, [[0, 0, 0], [0.5, 0, 0]], [[1, 0, 0, 0]] * 2, [1, 1], [np.eye(3)] * 2,
              vel=[[0, 0, 0], [-1.0, 0, 0]], is_static=[1, 0])
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=10, solver_mode=0)
    w.load_scene(s)
    assert w.add_joint({"type": DISTANCE, "a": 1, "b": 0, "local_a": [-0.5, 0, 0], "distance": 0.2})
    w.step(dt)
    _, acc, rows = w.joints()
    assert rows[0] == 0 and acc[0, 0] == 0
    np.testing.assert_array_equal(w.bodies()["vel"][1], [-1.0, 0, 0])
    w.step(dt)
    _, acc, rows = w.joints()
    assert rows[0] == 1 and acc[0, 0] != 0
    # the anchor line points along -x, length 0.01: the row drives u . v_a to -(0.1/dt)(0.01 - 0.2)
    assert abs(-w.bodies()["vel"][1, 0] - (0.1 / dt) * 0.19) < 1e-12


def test_joint_keys_order_and_bookkeeping(oracle):
    s = scene(oracle, [[0, 0, 0], [3, 0, 0], [6, 0, 0]], [[1, 0, 0, 0]] * 3, [1, 1, 1], [np.eye(3)] * 3)
    w = oracle.OracleWorld(gravity=(0, 0, 0), iterations=5, solver_mode=1)
    w.load_scene(s)
    assert w.add_joint({"type": WELD, "slot": 2, "a": 0, "b": 1, "anchor": [1.5, 0, 0]})
    assert w.add_joint({"type": DISTANCE, "slot": 7, "a": 1, "b": 2, "distance": 3.0})
    assert w.add_joint({"type": DISTANCE, "slot": 1, "a": 0, "b": 2, "distance": 6.0})
    assert not w.add_joint({"type": DISTANCE, "slot": 1, "a": 0, "b": 1})  # key taken
    assert not w.add_joint({"type": 6, "slot": 0, "a": 0, "b": 1})  # contact is not a joint type
    assert not w.add_joint({"type": BALL, "slot": 0, "a": 0, "b": 3})  # no body 3
    assert not w.add_joint({"type": BALL, "slot": 0, "a": 1, "b": 1})  # one body at both ends
    keys, _, _ = w.joints()
    assert list(keys) == [1, 7, (WELD << 24) | 2]
    assert w.remove_joint(7) and not w.remove_joint(7)
    keys, _, _ = w.joints()
    assert list(keys) == [1, (WELD << 24) | 2]


def test_joint_colours_are_proper_and_orders_agree_closely(oracle):
    """A hanging net of ball joints: no colour holds two joints on one dynamic body, and
    the coloured order stays close to the reference's sequential order."""
    nx = 6
    pos = [[0.0, 0.0, 0.0]] + [[1.0 * (i % nx), -1.0 - 1.0 * (i // nx), 0.0] for i in range(nx * 4)]
    n = len(pos)
    s = scene(oracle, pos, [[1, 0, 0, 0]] * n, [1.0] * n, [np.eye(3) * 0.2] * n, is_static=[1] + [0] * (n - 1))
    joints = [{"type": BALL, "slot": 0, "a": 0, "b": 1 + i, "anchor": [float(i), -0.5, 0.0]} for i in range(nx)]
    for i in range(nx * 4):
        if i % nx + 1 < nx:
            joints.append({"type": BALL, "a": 1 + i, "b": 2 + i, "anchor": np.add(pos[1 + i], [0.5, 0, 0])})
        if i + nx < nx * 4:
            joints.append({"type": BALL, "a": 1 + i, "b": 1 + i + nx, "anchor": np.add(pos[1 + i], [0, -0.5, 0])})
    for k, j in enumerate(joints):
        j["slot"] = k
    ws = []
    for mode in (0, 1):
        w = oracle.OracleWorld(iterations=20, solver_mode=mode)
        w.load_scene(s)
        for j in joints:
            assert w.add_joint(j)
        ws.append(w)
    for _ in range(60):
        for w in ws:
            w.step(1 / 60)
    col = ws[1].joint_colours()
    assert len(col) == len(joints) and col.max() >= 3
    for c in range(col.max() + 1):
        bodies = [b for k, j in enumerate(joints) if col[k] == c for b in (j["a"], j["b"]) if b != 0]
        assert len(bodies) == len(set(bodies)), c
    p0, p1 = ws[0].bodies()["pos"], ws[1].bodies()["pos"]
    assert np.abs(p0 - p1).max() < 0.05
    assert np.abs(p0 - p1).max() > 0  # the orders differ
