"""GPU parity of the joint constraints (include/physkit/collision/constraint.h:139-834)
against the oracle, bit for bit: the device builds the joint frames from its own body
state, colours the joints with the same rounds as the oracle (solver_mode=1), and
sweeps joint colours before contact colours in every pass, as the reference solves
joint rows before contact rows (constraint.h:1123-1187)."""
import numpy as np
import pytest

from test_gpu_parity import _assert_same, _compare_world

pytestmark = pytest.mark.gpu

BALL, HINGE, SLIDER, WELD, DISTANCE, SPRING = 1, 2, 3, 4, 0, 5


def _qmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _boxes_scene(oracle, halves, pos, quat, static, vel=None, mass=None):
    """Bodies with box meshes (one shape per distinct half extent)."""
    n = len(pos)
    shapes, shape_of = [], []
    for h in halves:
        h = tuple(float(x) for x in h)
        if h not in shapes:
            shapes.append(h)
        shape_of.append(shapes.index(h))
    verts = np.concatenate([oracle.box_verts(h) for h in shapes])
    m = np.ones(n) if mass is None else np.asarray(mass, float)
    inertia = []
    for i, h in enumerate(halves):
        hx, hy, hz = h
        inertia.append(np.diag([m[i] / 3 * (hy * hy + hz * hz), m[i] / 3 * (hx * hx + hz * hz),
                                m[i] / 3 * (hx * hx + hy * hy)]).reshape(9))
    return {"shape_offsets": np.arange(len(shapes) + 1, dtype=np.uint32) * 8, "verts": verts,
            "is_static": np.asarray(static, np.uint8), "shape": np.array(shape_of, np.uint32),
            "pos": np.array(pos, float), "vel": np.zeros((n, 3)) if vel is None else np.array(vel, float),
            "angvel": np.zeros((n, 3)), "quat": np.array(quat, float), "mass": m,
            "inertia": np.array(inertia), "restitution": np.full(n, 0.3), "friction": np.full(n, 0.6)}


def _pair(pk, oracle, s, joints, iterations=20, gravity=(0.0, -9.81, 0.0), flags=0):
    o = oracle.OracleWorld(gravity=gravity, iterations=iterations, solver_mode=1, flags=flags)
    o.load_scene(s)
    g = pk.GpuWorld(gravity=gravity, iterations=iterations, flags=flags)
    g.load_scene(s)
    for j in joints:
        assert o.add_joint(j), j
    if joints:
        g.add_joints(joints)
    return o, g


def _compare_joints(step, g, o):
    gk, gi = g.joints()
    ok, oi, _ = o.joints()
    _assert_same(step, "joint keys", gk, ok)
    _assert_same(step, "joint impulses", gi, oi)


def _chain(oracle, rng):
    """Seven far-apart boxes (body 0 static) joined by one joint of every type."""
    n = 7
    pos = [[0.0, 0.0, 0.0]] + [[1.4 * k, -0.9 * k, 0.4 * k] for k in range(1, n)]
    quat = [_rand_quat(rng) for _ in range(n)]
    s = _boxes_scene(oracle, [(0.2, 0.3, 0.1)] * n, pos, quat, [1] + [0] * (n - 1),
                     vel=rng.normal(size=(n, 3)) * 0.5, mass=[1.0] + list(rng.uniform(0.5, 3.0, n - 1)))
    P, Q = s["pos"], s["quat"]

    def mid(a, b):
        return 0.5 * (P[a] + P[b])

    def local(i, p):
        return _qmat(Q[i]).T @ (p - P[i])

    ax = np.array([0.3, 0.2, 1.0])
    m12, m45, m56 = mid(1, 2), mid(4, 5), mid(5, 6)
    joints = [
        {"type": BALL, "slot": 0, "a": 0, "b": 1, "anchor": mid(0, 1)},
        {"type": HINGE, "slot": 0, "a": 1, "b": 2, "local_a": local(1, m12), "local_b": local(2, m12),
         "local_axis_a": _qmat(Q[1]).T @ ax, "local_axis_b": _qmat(Q[2]).T @ ax},
        {"type": SLIDER, "slot": 0, "a": 2, "b": 3, "anchor": mid(2, 3), "axis": P[3] - P[2]},
        {"type": WELD, "slot": 0, "a": 3, "b": 4, "anchor": mid(3, 4)},
        {"type": DISTANCE, "slot": 0, "a": 4, "b": 5, "local_a": local(4, m45 + 0.1),
         "local_b": local(5, m45 - 0.1), "distance": 0.5},
        {"type": SPRING, "slot": 0, "a": 5, "b": 6, "local_a": local(5, m56 + [0.2, 0, 0]),
         "local_b": local(6, m56 - [0.2, 0, 0]), "distance": 0.8, "stiffness": 40.0, "damping": 2.0},
    ]
    return s, joints


def test_joint_chain_every_type_bitexact(pk, oracle):
    rng = np.random.default_rng(11)
    s, joints = _chain(oracle, rng)
    o, g = _pair(pk, oracle, s, joints)
    print("[joints] chain of every type, 90 steps", flush=True)
    for k in range(90):
        g.step(1 / 60)
        o.step(1 / 60)
        _compare_world(k, g, o)
        _compare_joints(k, g, o)
    # the chain hangs: everything stays finite and the ball joint's anchor holds
    b = g.bodies()
    assert np.all(np.isfinite(b["pos"]))
    g.close()


def test_joint_chain_large_errors_bitexact(pk, oracle):
    """Bodies turned and shifted between steps (pkg_set_body_state): large positional
    and angular errors, the slider clamp active, the weld's error past half a turn."""
    rng = np.random.default_rng(12)
    s, joints = _chain(oracle, rng)
    o, g = _pair(pk, oracle, s, joints, iterations=30)
    ids = np.arange(1, 7, dtype=np.uint32)
    print("[joints] large errors via set_state, 20 rounds", flush=True)
    for r in range(20):
        b = g.bodies()
        q = b["quat"][ids].copy()
        for i in range(len(ids)):
            ang = 3.6 if (r % 5 == 4 and ids[i] == 4) else rng.uniform(0.0, 0.6)
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            dq = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])
            w0, x0, y0, z0 = dq
            w1, x1, y1, z1 = q[i]
            q[i] = [w0 * w1 - x0 * x1 - y0 * y1 - z0 * z1, w0 * x1 + x0 * w1 + y0 * z1 - z0 * y1,
                    w0 * y1 - x0 * z1 + y0 * w1 + z0 * x1, w0 * z1 + x0 * y1 - y0 * x1 + z0 * w1]
        x = b["pos"][ids] + rng.uniform(-0.25, 0.25, (len(ids), 3))
        g.set_state(ids, pos=x, quat=q)
        o.set_bodies(ids, pos=x, quat=q)
        for k in range(3):
            g.step(1 / 60)
            o.step(1 / 60)
            _compare_world((r, k), g, o)
            _compare_joints((r, k), g, o)
    g.close()


def _bridge(oracle, rng, planks=14):
    """A plank bridge between two static posts over a ground box, planks joined by
    hinges (ball joints at the posts), with boxes dropped on it: joint rows and
    contact rows share bodies."""
    halves, pos, quat, static = [(12.0, 0.5, 12.0)], [[0.0, -0.5, 0.0]], [[1, 0, 0, 0]], [1]
    L = 0.5
    x0 = -planks * L
    halves += [(0.3, 2.0, 0.3)] * 2
    pos += [[x0 - 0.3, 2.0, 0.0], [-x0 + 0.3, 2.0, 0.0]]
    quat += [[1, 0, 0, 0]] * 2
    static += [1, 1]
    first = len(pos)
    for i in range(planks):
        halves.append((L * 0.98, 0.08, 0.6))
        pos.append([x0 + L * (2 * i + 1), 3.9, 0.0])
        quat.append([1, 0, 0, 0])
        static.append(0)
    drops = len(pos)
    for i in range(6):
        halves.append((0.25, 0.25, 0.25))
        pos.append([x0 + 1.0 + 2.1 * i, 5.0 + 0.3 * i, rng.uniform(-0.2, 0.2)])
        quat.append(_rand_quat(rng))
        static.append(0)
    s = _boxes_scene(oracle, halves, pos, quat, static)
    joints = [{"type": BALL, "slot": 0, "a": 1, "b": first, "anchor": [x0, 3.9, 0.0]},
              {"type": BALL, "slot": 1, "a": 2, "b": first + planks - 1, "anchor": [-x0, 3.9, 0.0]}]
    for i in range(planks - 1):
        a, b = first + i, first + i + 1
        anchor = np.array([x0 + L * 2 * (i + 1), 3.9, 0.0])
        la = anchor - np.array(pos[a])
        lb = anchor - np.array(pos[b])
        joints.append({"type": HINGE, "slot": i, "a": a, "b": b, "local_a": la, "local_b": lb,
                       "local_axis_a": [0, 0, 1], "local_axis_b": [0, 0, 1]})
    return s, joints, drops


def test_joint_bridge_with_contacts_bitexact(pk, oracle):
    rng = np.random.default_rng(13)
    s, joints, drops = _bridge(oracle, rng)
    o, g = _pair(pk, oracle, s, joints, iterations=15)
    print(f"[joints] bridge: {len(s['shape'])} bodies, {len(joints)} joints, 150 steps", flush=True)
    rows = 0
    for k in range(150):
        g.step(1 / 60)
        o.step(1 / 60)
        _compare_world(k, g, o)
        _compare_joints(k, g, o)
        rows = max(rows, g.counters()["rows"])
    assert rows > 0  # the dropped boxes landed on the bridge
    g.close()


def test_joint_add_remove_during_run_bitexact(pk, oracle):
    """Joints added after bodies moved (frames from the device's current poses),
    removed, and a freed slot re-used, between steps."""
    rng = np.random.default_rng(14)
    s, joints, drops = _bridge(oracle, rng, planks=8)
    o, g = _pair(pk, oracle, s, joints[:4], iterations=10)
    print("[joints] add/remove between steps", flush=True)
    for k in range(60):
        if k == 10:
            g.add_joints(joints[4:])
            for j in joints[4:]:
                assert o.add_joint(j)
        if k == 25:
            gone = [pk.joint_key(HINGE, 1), pk.joint_key(BALL, 0)]
            g.remove_joints(gone)
            for key in gone:
                assert o.remove_joint(key)
        if k == 30:
            extra = [{"type": WELD, "slot": 0, "a": drops, "b": drops + 1,
                      "anchor": 0.5 * (g.bodies()["pos"][drops] + g.bodies()["pos"][drops + 1])},
                     {"type": HINGE, "slot": 1, "a": 4, "b": 5, "local_a": [0.5, 0, 0], "local_b": [-0.5, 0, 0],
                      "local_axis_a": [0, 0, 1], "local_axis_b": [0, 0, 1]}]
            g.add_joints(extra)
            for j in extra:
                assert o.add_joint(j)
        g.step(1 / 60)
        o.step(1 / 60)
        _compare_world(k, g, o)
        _compare_joints(k, g, o)
    g.close()


def test_joint_calls_reject_bad_input(pk, oracle):
    rng = np.random.default_rng(15)
    s, joints = _chain(oracle, rng)
    g = pk.GpuWorld(iterations=5)
    g.load_scene(s)
    g.add_joints(joints[:2])
    bad = [
        [dict(joints[0])],                                          # key already live
        [dict(joints[2]), dict(joints[2])],                         # key repeated
        [{"type": 6, "slot": 0, "a": 0, "b": 1}],                   # contact is no joint type
        [{"type": BALL, "slot": 1 << 24, "a": 0, "b": 1}],          # slot out of range
        [{"type": BALL, "slot": 5, "a": 0, "b": 7}],                # no body 7
        [{"type": BALL, "slot": 5, "a": 2, "b": 2}],                # one body at both ends
    ]
    for b in bad:
        with pytest.raises(pk.PkgError):
            g.add_joints(b)
    with pytest.raises(pk.PkgError):
        g.remove_joints([pk.joint_key(WELD, 0)])                  # not live
    with pytest.raises(pk.PkgError):
        g.set_worlds(np.array([0, 3, 7], np.uint32))              # joints need one world
    keys, _ = g.joints()
    assert list(keys) == [pk.joint_key(BALL, 0), pk.joint_key(HINGE, 0)]
    g.remove_joints(keys)
    assert len(g.joints()[0]) == 0
    g.step(1 / 60)
    g.close()


def test_python_world_pendulum(pk):
    """pk.world with a ball joint to a static body: the bob swings and the anchor
    distance stays near its length (Baumgarte drift only)."""
    w = pk.world(pk.world_desc.make().with_solver_iterations(20))
    box = pk.mesh.box((0.1, 0.1, 0.1))
    post = w.create_rigid(pk.object_desc.stat().with_mesh(box).with_pos((0, 5, 0)))
    bob = w.create_rigid(pk.object_desc.dynam().with_mesh(box).with_pos((1.5, 5, 0)))
    h = w.add_constraint(pk.ball_socket_desc.make(post, bob).with_anchor((0, 5, 0)))
    assert h == (pk.JOINT_BALL_SOCKET, 0, 0)
    lows = []
    for _ in range(240):
        w.step(1 / 120)
        p = w.state()["pos"][bob]
        assert abs(np.linalg.norm(p - [0, 5, 0]) - 1.5) < 0.05
        lows.append(p[1])
    assert min(lows) < 3.6  # swung through the bottom
    assert w.remove_constraint(h) and not w.remove_constraint(h)
    w.step(1 / 120)
    assert len(w.gpu.joints()[0]) == 0
    w.gpu.close()
