"""Benchmark of the collision-and-contact step on one MI355X.

Workload (default): 4,096 independent worlds of 16-4,096 bodies (log-uniform; 40% UV
spheres, 40% boxes, 20% hulls of 8-16 points) over static ground boxes, stepped
together as one batch in the mixed-precision form (FP32 broadphase leaf tests rounded
outward, FP32 certified GJK separation, FP32 contact solve; EPA, manifolds and body
state FP64). Beside it, in the same run: the same worlds batched in FP64, the plain
FP64 drop-in (worlds stepped one at a time, one context each) and the oracle on the
host's cores.

Prints one JSON line. value = body-steps/s (dynamic bodies x timed steps / wall time of
the timed steps, inputs resident in HBM); contact-pairs/s (pairs through the GJK stage)
rides along in the config. The roofline is taken for the stage with the largest device
time, from HIP events recorded on the library's own stream, against that stage's
compulsory HBM bytes (stage_bytes below); every stage's figure is listed too.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "body-steps/s + contact-pairs/s, batched FP32 vs plain FP64, 1 MI355X"

WORKLOADS = {
    "hulls": (1, "narrowphase stress: 262,144 convex hulls (8-32 points on ellipsoids, "
                 "semi-axes U[0.05,0.3] m) in 512-body Gaussian clusters in a 200 m cube; "
                 "LBVH rebuild + pair query + GJK/EPA + manifolds per step, FP64"),
    "boxes": (0, "1,024 random boxes over a static ground box, dt=1/60, 10 PGS iterations, FP64"),
    "pyramids": (2, "64 pyramids of 15-layer unit-box stacks (7,680 boxes), dt=1/120, "
                    "20 PGS iterations, FP64"),
    "granular": (4, "granular pile: 2,097,152 bodies (70% UV spheres 16x32 r U[0.01,0.03], 30% boxes "
                     "half extents U[0.01,0.03]) on a jittered 0.07 m lattice falling into a 64-box "
                     "hopper, dt=1/240, 10 PGS iterations, FP64"),
    "granular10": (6, "granular pile at ten times the lengths: 2,097,152 bodies (70% UV spheres 16x32 "
                       "r U[0.1,0.3], 30% boxes half extents U[0.1,0.3]) on a jittered 0.7 m lattice "
                       "falling into a 64-box hopper, dt=1/240, 10 PGS iterations, FP64"),
    "batched": (3, "4,096 independent worlds of 16-4,096 bodies (log-uniform; 40% UV spheres "
                   "16x32, 40% boxes, 20% 8-16-point hulls) over static ground boxes, dt=1/60, "
                   "10 PGS iterations, stepped together as one batch"),
    "rays": (1, "closest-hit raycasts: 4,194,304 rays (origins uniform in the scene box grown by "
                "10 m; half aimed at a random body, half uniform directions; unbounded) into the "
                "262,144-hull cluster scene's broadphase tree, FP64 slab tests, rays resident in HBM"),
}
RAYS = 1 << 22
# LBVH refits between full builds on the measured worlds (pkg_tuning.tree_refits)
TREE_REFITS = 7


def support_vertices(s):
    """Vertices a support search of each shape must read: all of them, except for a
    UV-sphere mesh (2 + (stacks-1) * sectors vertices, poles on the y axis), whose
    search evaluates 12 candidates around the analytic optimum."""
    off = s["shape_offsets"].astype(np.int64)
    v = s["verts"]
    nv = off[1:] - off[:-1]
    cost = nv.astype(np.float64)
    big = np.nonzero(nv > 64)[0]
    if len(big):
        top, bot = v[off[big]], v[off[big + 1] - 1]
        first, second = v[off[big] + 1], v[off[big] + 2]
        uv = ((top[:, 1] > 0) & (top[:, 0] == 0) & (top[:, 2] == 0) & (bot[:, 0] == 0) & (bot[:, 2] == 0)
              & (bot[:, 1] == -top[:, 1]) & (first[:, 1] == second[:, 1]))
        cost[big[uv]] = 12.0
    return cost


def shape_radii(s):
    """Largest vertex distance from each shape's origin."""
    off = s["shape_offsets"].astype(np.int64)
    r2 = (s["verts"] ** 2).sum(1)
    return np.sqrt(np.maximum.reduceat(r2, off[:-1]))


def algorithmic_bytes(stage, s, pairs, hits_frac, pos=None):
    """Compulsory HBM bytes of one launch of the narrowphase stage (FP64).
    GJK stage: a pair whose bounding spheres are apart (the cull's first test) needs its
    key (8 B) and its status byte; the positions and radii of the bodies in such pairs
    (32 B each) are read once per body, however many pairs share it. Any other
    pair also reads each body's quaternion (32), shape id and vertex offsets (12) and the
    support vertices of both shapes once (24 B each; 12 candidates for a UV sphere); a
    hit writes its hit-list entry (4) and simplex (17 x 8: count, then per point p.xyz
    and the packed support features).
    EPA re-reads, per hit, the hit-list entry, key, poses, offsets, support vertices and
    simplex, and writes 10 doubles of contact data and a status byte."""
    cost = support_vertices(s)[s["shape"]]
    a = (pairs >> np.uint64(32)).astype(np.int64)
    b = (pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)
    verts = 24.0 * (cost[a] + cost[b])
    per_pair_fixed = 8 + 2 * (24 + 32 + 12)
    if stage == "gjk":
        rr = shape_radii(s)[s["shape"]]
        p = s["pos"] if pos is None else pos
        d = np.linalg.norm(p[a] - p[b], axis=1)
        r = rr[a] + rr[b]
        apart = d > r + 1e-6 * r + 1e-6
        culled = float(np.count_nonzero(apart))
        bodies = float(len(np.unique(np.concatenate([a[apart], b[apart]]))))
        rest = np.sum((per_pair_fixed + verts + 1)[~apart])
        return float(culled * (8 + 1) + bodies * 32 + rest + hits_frac * len(pairs) * (4 + 17 * 8))
    if stage == "epa":
        return float(hits_frac * np.sum(4 + per_pair_fixed + verts + 17 * 8 + 80 + 1))
    raise ValueError(stage)


def measured_traffic(workload, precision, stage):
    """HBM bytes per launch of a stage's dominant kernel, from the committed FETCH_SIZE /
    WRITE_SIZE counter passes over this workload (profiles/<workload>_<precision>_hbm_bytes.md,
    scaled by a copy kernel of known size); None when no such summary is committed."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_hbm_bytes.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == stage:
                return float(f[2])
    except OSError:
        pass
    return None


def step_bytes(s, c, unique_vertices, scalar=8):
    """Compulsory HBM bytes of one whole step (every persistent array read and written
    once; grid-wide intermediates written and read once):
    N(54s + 157 + 3s V_u) + S(6s + 4) + 24 P + 10s P_c + C(20s + 48)."""
    n_static = int(s["n_static"])
    n_dyn = len(s["shape"]) - n_static
    vu = 0.0
    if unique_vertices:
        off = s["shape_offsets"].astype(np.int64)
        vu = float(np.mean((off[1:] - off[:-1])[s["shape"]]))
    return (n_dyn * (54 * scalar + 157 + 3 * scalar * vu) + n_static * (6 * scalar + 4) + 24 * c["pairs"]
            + 10 * scalar * c["gjk_hits"] + c["contacts"] * (20 * scalar + 48))


def stage_bytes(s, c, pairs, hits_frac, pos, iterations, mixed, tree_refits=0, integrate=True, steps=1):
    """Compulsory HBM bytes of each stage's launches in one step, with the step's own
    counts (n bodies, N dynamic, M moved leaves, K candidate pairs, P pairs, H GJK hits,
    C manifold contacts, R solver rows; s_r = 4 for FP32 rows, else 8), averaged over
    `steps` steps taken in one call:
      integrate_vel     2n + 412 N + 48 M + 48 per shape (k_integrate_vel: mass, inverse
                        mass, force and torque accumulators, velocities, world inertia and
                        its inverse, pose, leaf box read; velocities written, the
                        accumulators only when set (none here), the leaf box when it
                        moves) on a call's first step; 2n + 612 N + 48 M + 48 per shape
                        on the others, whose first kernel also runs the previous step's
                        integrate_pos (k_integrate_pos_vel: local inertia read; pose and
                        world inertia written, and kept on chip instead of read back)
      lbvh_build        253 n (+32 n with FP32 leaf boxes) on a full build: leaf boxes
                        read, Morton keys and ids written and read, leaf records written,
                        2 internal nodes of 32 B written and read per leaf; 229 n on a refit
                        (no keys); averaged over one build and tree_refits refits
      pair_update       48 M + 80 n + 36 K + 20 P: the moving leaves' boxes, the tree read
                        once, candidates written, sorted and matched, the previous list
                        read and the new one with its sources written; n when no leaf
                        moved (the moved flags read; the list is kept as it is, each pair
                        its own source)
      gjk               algorithmic_bytes(); with per-body cull records (P >= 2n) + 6 P:
                        the cull also reads each pair's source (4) and previous contact
                        count (1) and writes its next count (1) and event (1), empty
                        (the manifold update rewrites the listed pairs), and a culled
                        pair writes no status byte (-1; the listed pairs' share of P is
                        a fraction of a percent and is not subtracted)
      epa               algorithmic_bytes()
      manifold          with per-body cull records, 28 L + 208 C + 192 H over the L pairs
                        GJK ran on (and the few culled pairs that had contacts): key,
                        source, status, previous count and slot, new count, slot and
                        event; contacts read and written; hits' contact data and poses.
                        Otherwise 8 P + 28 L + 208 C + 192 H: every pair's source, status,
                        previous and new count and event, then the listed pairs as above
      solver            496 R: contacts read, both bodies' pose, velocity and inverse
                        inertia read for the rows, impulses and velocities written
      solve_iterations  R (38 s_r + 96): every row read once and its bodies' velocities
                        read and written once (rows that stay on chip across passes)
      integrate_pos     n + 448 N (k_integrate_pos: pose, velocities and local inertia
                        read; pose and world inertia written) on a call's last step only
    With integration off the first stage only clears the moved flags (n) and the last
    launches nothing."""
    n = len(s["shape"])
    nd = n - int(s["n_static"])
    n_shapes = len(s["shape_offsets"]) - 1
    sr = 4 if mixed else 8
    M, K, P, H, C, R = (c[k] for k in ("moved", "candidates", "pairs", "gjk_hits", "contacts", "rows"))
    fused = P >= 2 * n  # the cull reads per-body records and settles the culled pairs' manifolds
    out = {
        "integrate_vel": ("k_integrate_vel, then k_integrate_pos_vel",
                          ((2 * n + 412 * nd) + (steps - 1) * (2 * n + 612 * nd)) / steps + 48 * M + 48 * n_shapes),
        "lbvh_build": ("LBVH build or refit (k_morton, k_rs_*, k_build_internal, k_leaf_gather, k_refit, k_span)",
                       ((253 + 229 * tree_refits) / (1 + tree_refits) + (32 if mixed else 0)) * n),
        "pair_update": ("pair-set update (k_pair_query, sort, k_match, k_prev_*, k_merge)",
                        n if M == 0 and K == 0 else 48 * M + 80 * n + 36 * K + 20 * P),
        "gjk": ("GJK stage (k_gjk_cull" + (", k_gjk32" if mixed else "") + ", k_gjk)",
                (algorithmic_bytes("gjk", s, pairs, hits_frac, pos) if len(pairs) else 0.0) + (6 * P if fused else 0)),
        "epa": ("k_epa (+ larger-store reruns)", algorithmic_bytes("epa", s, pairs, hits_frac) if len(pairs) else 0.0),
        "manifold": ("k_manifold" if fused else "k_manifold_live, k_manifold",
                     (0 if fused else 8 * P) + 28 * c["gjk_listed"] + 208 * C + 192 * H),
        "solver": ("contact solver stage", 496 * R),
        "solve_iterations": ("k_solve_islands" + ("<true> (FP32 rows)" if mixed else "<false>"), R * (38 * sr + 96)),
        "integrate_pos": ("k_integrate_pos", (n + 448 * nd) / steps),
    }
    if not integrate:
        out["integrate_vel"] = ("k_fill_u32 (moved flags cleared; integration off)", n)
        out["integrate_pos"] = ("none (integration off)", 0)
    return out


def roofline_of(stage_ms, steps, models):
    """Achieved GB/s and HBM fraction of every stage, and the dominant one: the stage with
    the largest device time, with the solver stage split into its iteration kernels
    (solve_iterations) and the rest (rows, colouring, islands)."""
    ms = dict(stage_ms)
    rest = max(ms.get("solver", 0.0) - ms.get("solve_iterations", 0.0), 0.0)
    table = {}
    for k, (name, b) in models.items():
        t = (rest if k == "solver" else ms.get(k, 0.0)) / steps
        gbs = b / (t / 1e3) / 1e9 if t > 0 else 0.0
        table[k] = {"kernel": name if k != "solver" else "contact solver stage without its iterations "
                    "(k_row_flags, colouring, islands, k_setup, k_writeback)",
                    "avg_launch_ms": round(t, 4), "algorithmic_bytes": float(b), "achieved_GBs": round(gbs, 2),
                    "frac": round(gbs / HBM_PEAK_GBS, 5)}
    dom = max(table, key=lambda k: table[k]["avg_launch_ms"])
    return dom, table


def worlds_prefix(s, k):
    """The first k worlds of a batched scene as a scene of their own (all shapes kept)."""
    fb = s["first_body"]
    b1 = int(fb[k])
    out = {key: s[key][:b1].copy() for key in ("is_static", "shape", "pos", "vel", "angvel", "quat", "mass",
                                               "inertia", "restitution", "friction")}
    for key in ("shape_offsets", "verts", "dt", "iterations", "steps", "flags"):
        out[key] = s[key]
    out["n_static"] = int(np.sum(out["is_static"] != 0))
    out["first_body"] = fb[:k + 1].copy()
    return out


def precision_drift(s, steps, flags_lo):
    """A lower-precision form (flags_lo added) against the FP64 step from the same start.
    per_step: the largest one-step position difference, max over bodies of
    |x_lo - x64| / max(|x64|, 1 m), with the lower-precision world reset to the FP64
    bodies after every step; trajectory: the same difference without resets after step
    1, 10 and `steps`."""
    import physkit_amd as pk

    def pair():
        w64 = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
        wlo = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"] | flags_lo)
        for w in (w64, wlo):
            w.load_scene(s)
            if len(s["first_body"]) > 2:
                w.set_worlds(s["first_body"])
        return w64, wlo

    def err(a, b):
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)))

    out = {"per_step": 0.0, "trajectory": {}}
    ids = np.arange(len(s["shape"]), dtype=np.uint32)
    w64, wlo = pair()
    for k in range(1, steps + 1):
        w64.step(s["dt"])
        wlo.step(s["dt"])
        a = w64.bodies()
        out["per_step"] = max(out["per_step"], err(a["pos"], wlo.bodies()["pos"]))
        wlo.set_state(ids, pos=a["pos"], vel=a["vel"], angvel=a["angvel"], quat=a["quat"])
    w64.close()
    wlo.close()
    w64, wlo = pair()
    for k in range(1, steps + 1):
        w64.step(s["dt"])
        wlo.step(s["dt"])
        if k in (1, 10, steps):
            out["trajectory"][f"step_{k}"] = err(w64.bodies()["pos"], wlo.bodies()["pos"])
    w64.close()
    wlo.close()
    return out


def tuned_shapes(pk, s, flags, warmup):
    """Launch shapes from pkg_autotune (3 steps, every candidate timed on each step's
    own inputs), run on a separate copy of the world after the same warmup steps, so
    the measured world's timed steps are the same steps with or without tuning."""
    g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    if len(s["first_body"]) > 2:
        g.set_worlds(s["first_body"])
    g.step(s["dt"], warmup)
    t = g.autotune(s["dt"], 3)
    g.close()
    return t


def timed_run(pk, s, flags, warmup, steps, autotune=True, fixed=None, tree_refits=0):
    """Loads s into a context with flags, runs warmup untimed steps, then `steps` timed
    steps in one pkg_step_timed call (it returns after its device work has finished, so
    the timed region is bracketed by device synchronisation), with the launch shapes
    pkg_autotune picked on a copy of the world (tuned_shapes), the `fixed` ones
    (gjk_lanes, epa_block, epa_lds_max) or the defaults, and tree_refits LBVH refits
    between full builds. Returns wall seconds, summed
    stage milliseconds, pair and hit totals, the last counters, pairs, positions and the
    launch shapes used."""
    if fixed:
        tuning = dict(zip(("gjk_lanes", "epa_block", "epa_lds_max"), fixed))
    else:
        tuning = tuned_shapes(pk, s, flags, warmup) if autotune else None
    g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    if len(s["first_body"]) > 2:
        g.set_worlds(s["first_body"])
    if tuning is None:
        tuning = g.tuning()
    g.set_tuning(tuning["gjk_lanes"], tuning["epa_block"], tuning["epa_lds_max"], tree_refits)
    tuning = dict(tuning, tree_refits=tree_refits)
    for _ in range(warmup):
        g.step(s["dt"])
    c0 = g.counters()
    # the timed steps in one call (each step's integrate_pos runs fused with the next
    # step's integrate_vel; the call returns with every step integrated and its device
    # work finished)
    t0 = time.perf_counter()
    stage = g.step_timed(s["dt"], steps)
    wall = time.perf_counter() - t0
    c = g.counters()
    pairs_total = c["pairs_total"] - c0["pairs_total"]
    hits_total = c["gjk_hits_total"] - c0["gjk_hits_total"]
    out = {"wall": wall, "stage": stage, "pairs_total": pairs_total, "hits_total": hits_total,
           "counters": g.counters(), "pairs": g.pairs(), "pos": g.bodies()["pos"], "tuning": tuning}
    g.close()
    return out


def per_world_dropin(s, warmup, steps, n_sample=64):
    """The plain FP64 drop-in on the batched workload: each of the first n_sample worlds
    stepped alone through its own context (one world per physkit::world), warmup steps
    untimed, then `steps` timed steps; body-steps/s over the summed wall times."""
    import physkit_amd as pk

    bodies, wall = 0, 0.0
    n = min(n_sample, len(s["first_body"]) - 1)
    for w in range(n):
        ws, b0, b1 = pk.scene_world(s, w)
        g = pk.GpuWorld(iterations=ws["iterations"], flags=ws["flags"])
        g.load_scene(ws)
        g.step(ws["dt"], warmup)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step(ws["dt"])
        wall += time.perf_counter() - t0
        bodies += (b1 - b0 - ws["n_static"]) * steps
        g.close()
    return {"value": round(bodies / wall, 1), "unit": "body-steps/s",
            "sample": f"worlds 0-{n - 1} stepped one at a time, {steps} timed steps each after {warmup} untimed"}


def cpu_baseline_batched(s, warmup, budget_s=12.0):
    """The oracle on the batched workload, its independent worlds stepped by a pool of
    host threads (the oracle releases the interpreter lock inside a step). Worlds are
    taken in order, in chunks: each chunk's worlds are loaded and given the bench's
    warmup steps untimed, then the next step of every world in the chunk (the first
    step the GPU times) is timed as wall time, until about budget_s seconds of timed
    stepping."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    import physkit_amd as pk
    from concurrent.futures import ThreadPoolExecutor

    cores = max(1, min(16, os.cpu_count() or 1))
    fb = s["first_body"]
    n_worlds = len(fb) - 1

    def setup(w):
        ws, b0, b1 = pk.scene_world(s, w)
        o = oracle_ffi.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        o.step(ws["dt"], warmup)
        return o, int(ws["n_static"]), int(b1 - b0)

    bodies, timed, w = 0, 0.0, 0
    with ThreadPoolExecutor(cores) as ex:
        while w < n_worlds and timed < budget_s:
            chunk = list(range(w, min(w + 4 * cores, n_worlds)))
            worlds = list(ex.map(setup, chunk))
            t0 = time.perf_counter()
            list(ex.map(lambda x: x[0].step(s["dt"], 1), worlds))
            timed += time.perf_counter() - t0
            bodies += sum(nb - ns for _, ns, nb in worlds)
            w += len(chunk)
    return {"value": bodies / timed, "unit": "body-steps/s", "cores": cores, "kind": "port",
            "sample": f"step {warmup + 1} of worlds 0-{w - 1} of {n_worlds} "
                      f"({bodies} dynamic bodies), {timed:.1f} s of wall time on {cores} threads"}


def cpu_baseline(s, budget_s=12.0):
    """The oracle (FP64 CPU restatement, one thread) on the same workload, stepped until
    about budget_s seconds have elapsed (at least one step)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi

    o = oracle_ffi.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    o.load_scene(s)
    steps, t = 0, 0.0
    while t < budget_s and steps < 200:
        t += o.step(s["dt"], 1)
        steps += 1
    n = len(s["shape"]) - int(s["n_static"])
    return {"value": n * steps / t, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of the full workload ({n} dynamic bodies), "
                      f"{t:.1f} s on one host core"}


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes (device buffers without torch)."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        path = "/opt/rocm/lib/libamdhip64.so"
        with open("/proc/self/maps") as f:  # the HIP runtime the library already loaded
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        self.L = ctypes.CDLL(path)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        if self.L.hipMalloc(self.C.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        return p

    def h2d(self, dst, a):
        if self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy failed")

    def d2h(self, a, src):
        if self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")

    def free(self, p):
        self.L.hipFree(p)


def ray_batch(s, n, seed=20261203):
    """n rays: origins uniform in the bodies' box grown by 10 m; half aimed at a random
    body (plus N(0, 0.1 m) per axis), half uniform directions; no distance limit."""
    rng = np.random.default_rng(seed)
    lo, hi = s["pos"].min(axis=0) - 10.0, s["pos"].max(axis=0) + 10.0
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    k = n // 2
    d[:k] = s["pos"][rng.integers(len(s["pos"]), size=k)] + rng.normal(0.0, 0.1, (k, 3)) - o[:k]
    return o, d, np.full(n, np.inf)


def ray_algorithmic_bytes(n_rays, n_bodies):
    """Compulsory HBM bytes of one closest-hit launch: per ray the origin, direction and
    limit read (56 B) and the body id and distance written (12 B); the tree read once:
    n-1 internal nodes of two float4 (32 B) and, per leaf, its FP64 box (48 B), body id,
    escape index (4 B each) and the body's static flag (1 B)."""
    return 68.0 * n_rays + 32.0 * (n_bodies - 1) + 57.0 * n_bodies


def bench_rays(args, pk, s):
    import ctypes
    hip = _Hip()
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    n = args.n_rays
    o, d, m = ray_batch(s, n)
    bufs = [hip.alloc(a.nbytes) for a in (o, d, m)]
    for b, a in zip(bufs, (o, d, m)):
        hip.h2d(b, np.ascontiguousarray(a))
    ids_d, dist_d = hip.alloc(4 * n), hip.alloc(8 * n)
    kms = ctypes.c_double(0.0)

    def call():
        g._check(g.L.pkg_raycast_closest_device(g.h, n, bufs[0], bufs[1], bufs[2], None, ids_d, dist_d,
                                                ctypes.byref(kms)), "raycast_closest_device")
        return kms.value

    print(f"[bench] workload=rays bodies={len(s['shape'])} rays={n} steps={args.steps} "
          f"warmup={args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):  # the first call builds the tree (no step has run)
        call()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms += call()
    wall = time.perf_counter() - t0
    ids = np.zeros(n, np.uint32)
    dist = np.zeros(n)
    hip.d2h(ids, ids_d)
    hip.d2h(dist, dist_d)
    t_launch = kernel_ms / 1e3 / args.steps
    alg = ray_algorithmic_bytes(n, len(s["shape"]))
    achieved = alg / t_launch / 1e9
    out = {
        "metric": "rays/s (closest hit)", "value": round(n * args.steps / wall, 1), "unit": "rays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": WORKLOADS["rays"][1], "bodies": len(s["shape"]), "rays_per_step": n,
                   "hit_fraction": round(float(np.mean(ids != 0xFFFFFFFF)), 4),
                   "kernel_rays_per_s": round(n / t_launch, 1), "parallelism": "1 GPU, one world"},
        "roofline": {"bound": "hbm", "kernel": "k_ray_closest", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "algorithmic_bytes_per_launch": alg,
                     "avg_launch_ms": round(t_launch * 1e3, 4)},
    }
    g.close()
    for b in bufs + [ids_d, dist_d]:
        hip.free(b)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_rays(s, o, d, m, ids, dist)
    return out


def cpu_baseline_rays(s, o, d, m, gpu_ids, gpu_dist, budget_s=12.0):
    """The oracle's tree-walking closest-hit raycast (one host thread) over the same
    scene, on leading chunks of the same rays until about budget_s seconds of walking;
    its answers are checked against the GPU's for those rays."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    w = oracle_ffi.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    w.load_scene(s)
    done, sec, chunk = 0, 0.0, 131072
    while sec < budget_s and done < len(o):
        e = min(done + chunk, len(o))
        ids, dist, t = w.raycast_closest_bvh(o[done:e], d[done:e], m[done:e])
        if not (np.array_equal(ids, gpu_ids[done:e]) and np.array_equal(dist, gpu_dist[done:e])):
            raise RuntimeError("CPU and GPU closest hits differ")
        sec += t
        done = e
    return {"value": done / sec, "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": f"rays 0-{done - 1} of the same batch through a CPU tree walk (median-split BVH "
                      f"over the same leaves, build untimed), {sec:.1f} s on one host core; "
                      f"answers equal the GPU's"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="body count override (smaller runs)")
    ap.add_argument("--precision", default="auto", choices=("auto", "fp64", "mixed", "fp32-solve"),
                    help="auto: the mixed-precision form for the batched workload, else FP64 (the drop-in); "
                         "mixed: FP32 leaf tests, certified FP32 GJK separation, FP32 solve; fp32-solve: "
                         "only the solve in FP32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of those pkg_autotune picks on a copy "
                         "of the world")
    ap.add_argument("--tuning", default=None, metavar="LANES,BLOCK,LDS_MAX",
                    help="fixed launch shapes (GJK lanes per pair, EPA workgroup size, LDS-store "
                         "threshold) instead of pkg_autotune's, e.g. to profile the shapes a bench "
                         "line reports")
    ap.add_argument("--tree-refits", type=int, default=TREE_REFITS, metavar="N",
                    help="steps between full LBVH builds that refit the last hierarchy instead "
                         f"(pkg_tuning.tree_refits; default {TREE_REFITS})")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the comparison runs (batched FP64, per-world drop-in, precision drift)")
    ap.add_argument("--n-rays", type=int, default=RAYS, help="rays workload: rays per launch")
    ap.add_argument("--copy-check", type=int, default=1024, metavar="MIB",
                    help="after the timed steps, time a plain device copy of MIB MiB (measured "
                         "copy bandwidth; the calibration kernel for HBM byte counters); 0 skips it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(each GPU would run its own worlds; multi-GPU sharding is not built)", file=sys.stderr)
        sys.exit(2)

    import physkit_amd as pk

    kind, desc = WORKLOADS[args.workload]
    s = pk.scene(kind, None, args.n)
    if args.workload == "rays":
        print(json.dumps(bench_rays(args, pk, s)), flush=True)
        return
    batched = len(s["first_body"]) > 2
    prec = args.precision if args.precision != "auto" else ("mixed" if batched else "fp64")
    extra = {"fp64": 0, "mixed": pk.FLAG_MIXED, "fp32-solve": pk.FLAG_FP32_SOLVE}[prec]
    mixed = prec == "mixed"
    n_dyn = len(s["shape"]) - int(s["n_static"])
    print(f"[bench] workload={args.workload} precision={prec} bodies={len(s['shape'])} steps={args.steps} "
          f"warmup={args.warmup}", file=sys.stderr, flush=True)
    fixed = tuple(int(x) for x in args.tuning.split(",")) if args.tuning else None
    if fixed is not None and len(fixed) != 3:
        ap.error("--tuning takes three integers: LANES,BLOCK,LDS_MAX")
    r = timed_run(pk, s, s["flags"] | extra, args.warmup, args.steps, autotune=not args.no_autotune,
                  fixed=fixed, tree_refits=args.tree_refits)
    wall, stage, c, pairs = r["wall"], r["stage"], r["counters"], r["pairs"]
    print(f"[bench] timed: {wall * 1e3 / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
    hits_frac = c["gjk_hits"] / max(c["pairs"], 1)
    models = stage_bytes(s, c, pairs, hits_frac, r["pos"], s["iterations"], mixed or prec == "fp32-solve",
                         args.tree_refits, integrate=(int(s["flags"]) & pk.FLAG_NO_INTEGRATE) == 0,
                         steps=args.steps)
    dom, table = roofline_of(stage, args.steps, models)
    d = table[dom]
    roof = {"bound": "hbm", "stage": dom, "kernel": d["kernel"], "achieved": d["achieved_GBs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
            "traffic": measured_traffic(args.workload, prec, dom) if not args.n else None,
            "algorithmic_bytes_per_launch": d["algorithmic_bytes"], "avg_launch_ms": d["avg_launch_ms"]}
    step_b = step_bytes(s, c, unique_vertices=(args.workload in ("hulls", "batched")), scalar=4 if mixed else 8)
    dtype = {"fp64": "f64", "mixed": "f32 leaf tests, GJK separation and solve; f64 EPA, manifolds, body state",
             "fp32-solve": "f64 geometry, f32 solver rows"}[prec]
    value = n_dyn * args.steps / wall
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic",
        "config": {"workload": desc, "precision": prec, "bodies": len(s["shape"]), "dynamic_bodies": n_dyn,
                   "pairs_per_step": c["pairs"], "gjk_hits_per_step": c["gjk_hits"],
                   "last_step_counters": {k: c[k] for k in ("contacts", "rows", "colours", "epa_overflow",
                                                            "candidates", "moved", "gjk_listed", "gjk_fp64")},
                   "contact_pairs_per_s": round(r["pairs_total"] / wall, 1),
                   "colliding_pairs_per_s": round(r["hits_total"] / wall, 1),
                   "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in stage.items()},
                   "stage_roofline": table,
                   "launch_shapes": r["tuning"],
                   "step_algorithmic_bytes": step_b,
                   "step_hbm_frac": round(step_b / (wall / args.steps) / (HBM_PEAK_GBS * 1e9), 5),
                   "worlds": len(s["first_body"]) - 1,
                   "parallelism": "1 GPU, " + ("one world" if not batched
                                               else f"{len(s['first_body']) - 1} worlds batched")},
        "roofline": roof,
    }
    if not args.no_compare:
        if batched:
            # the same worlds batched in FP64, and the plain FP64 drop-in: worlds one at a time
            if prec != "fp64":
                r64 = timed_run(pk, s, s["flags"], args.warmup, args.steps, autotune=not args.no_autotune,
                                fixed=fixed, tree_refits=args.tree_refits)
                v64 = n_dyn * args.steps / r64["wall"]
                out["config"]["batched_fp64"] = {"value": round(v64, 1), "unit": "body-steps/s",
                                                 "ms_per_step": round(r64["wall"] * 1e3 / args.steps, 3),
                                                 "stage_ms_per_step": {k: round(v / args.steps, 4)
                                                                       for k, v in r64["stage"].items()}}
                out["config"]["batched_over_batched_fp64"] = round(value / v64, 3)
                print(f"[bench] batched fp64: {r64['wall'] * 1e3 / args.steps:.3f} ms/step", file=sys.stderr,
                      flush=True)
            dropin = per_world_dropin(s, args.warmup, args.steps)
            out["config"]["per_world_dropin"] = dropin
            out["config"]["batched_over_per_world"] = round(value / dropin["value"], 2)
        if prec != "fp64":
            sub = worlds_prefix(s, 256) if batched else s
            drift = precision_drift(sub, min(args.warmup + args.steps, 30), extra)
            drift["sample"] = ("worlds 0-255" if batched else "the whole workload") + ", from the same start"
            out["config"]["precision_drift"] = drift
    if args.copy_check:
        import ctypes
        from physkit_amd import build as pk_build
        cc = ctypes.CDLL(pk_build.COPY_LIB)
        cc.pkc_copy_check.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        ms = ctypes.c_double(0.0)
        nbytes = args.copy_check << 20
        if cc.pkc_copy_check(nbytes, 3, ctypes.byref(ms)) != 0:
            raise RuntimeError("copy check failed")
        out["copy_check"] = {"bytes": nbytes, "ms": round(ms.value, 4),
                             "GB/s": round(2 * nbytes / ms.value / 1e6, 1)}
    if not args.no_cpu:
        out["cpu_baseline"] = (cpu_baseline_batched(s, args.warmup) if batched else cpu_baseline(s))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
