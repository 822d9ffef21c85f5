"""The FP32 certified-separation pass of the mixed-precision narrowphase
(physkit_amd/csrc/pk_gjk32.hpp), run on the CPU through tests/cpp/gjk32_check.hip
(the same header compiled for the host), against the oracle's FP64 GJK
(gjk_collision, src/collision.cpp:165-189).

What must hold: a pair the pass certifies is one the FP64 GJK finds separated; a
certified pair never reaches the FP64 GJK, so a single wrong certificate would change a
verdict. The pass should also certify most pairs that are clearly apart, or it buys
nothing; that rate is checked loosely."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "gjk32_check")
RNG = np.random.default_rng(32)


@pytest.fixture(scope="module")
def checker():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++20", "-ffp-contract=off",
                    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "physkit_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "gjk32_check.hip"), "-o", EXE],
                   check=True, capture_output=True)

    def run(pairs):
        lines = [str(len(pairs))]
        for side_a, side_b in pairs:
            for v, q, p, uv in (side_a, side_b):
                lines.append(f"{len(v)} " + " ".join(repr(float(x)) for x in np.ravel(v)) + " " +
                             " ".join(repr(float(x)) for x in q) + " " + " ".join(repr(float(x)) for x in p) +
                             f" {int(uv)}")
        r = subprocess.run([EXE], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        return np.array([int(x) for x in r.stdout.split()], dtype=np.int32)

    return run


def rand_quat():
    q = RNG.normal(size=4)
    return q / np.linalg.norm(q)


def rand_shape(oracle, kind):
    if kind == "box":
        return oracle.box_verts(RNG.uniform(0.05, 0.8, 3)), False
    if kind == "hull":
        n = int(RNG.integers(8, 33))
        d = RNG.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return d * RNG.uniform(0.05, 0.3, 3), False
    return oracle.sphere_verts(float(RNG.uniform(0.05, 0.5)), 16, 32), True


def oracle_separated(oracle, va, qa, pa, vb, qb, pb):
    return oracle.gjk_epa(oracle.mesh_instance(va, pa, qa), oracle.mesh_instance(vb, pb, qb)) is None


@pytest.mark.parametrize("kinds", [("box", "box"), ("hull", "hull"), ("box", "hull"), ("sphere", "box"),
                                   ("sphere", "sphere"), ("sphere", "hull")])
def test_certified_pairs_are_separated_by_fp64_gjk(oracle, checker, kinds):
    pairs, sep = [], []
    for _ in range(600):
        va, ua = rand_shape(oracle, kinds[0])
        vb, ub = rand_shape(oracle, kinds[1])
        ra = np.linalg.norm(va, axis=1).max()
        rb = np.linalg.norm(vb, axis=1).max()
        qa, qb = rand_quat(), rand_quat()
        pa = RNG.uniform(-50, 50, 3)  # away from the origin: FP32 position rounding matters
        d = RNG.normal(size=3)
        pb = pa + d / np.linalg.norm(d) * (ra + rb) * RNG.uniform(0.2, 1.1)
        pairs.append(((va, qa, pa, ua), (vb, qb, pb, ub)))
        sep.append(oracle_separated(oracle, va, qa, pa, vb, qb, pb))
    cert = checker(pairs)
    sep = np.array(sep)
    wrong = np.nonzero((cert == 1) & ~sep)[0]
    assert len(wrong) == 0, f"certified but FP64 GJK reports contact: pairs {wrong[:10]}"
    assert sep.sum() > 50
    if "sphere" not in kinds or kinds != ("sphere", "sphere"):
        assert cert.sum() >= 0.5 * sep.sum(), (cert.sum(), sep.sum())


def test_face_to_face_gaps_around_the_margin(oracle, checker):
    """Two boxes face to face at gaps from 1e-8 m to 1e-2 m, B slightly turned and its
    face inside A's (so the gap is the distance): nothing below the margin (about
    1e-6 m) is certified, and gaps well above it are."""
    pairs, gaps = [], []
    for gap in np.geomspace(1e-8, 1e-2, 61):
        for k in range(4):
            ha = np.array([RNG.uniform(0.1, 0.6), 1.0, 1.0])
            hb = np.array([RNG.uniform(0.1, 0.6), RNG.uniform(0.1, 0.6), RNG.uniform(0.1, 0.6)])
            va, vb = oracle.box_verts(ha), oracle.box_verts(hb)
            ang = RNG.uniform(-1e-3, 1e-3)
            qa = np.array([1.0, 0, 0, 0])
            qb = np.array([math.cos(ang / 2), 0, math.sin(ang / 2), 0])
            pa = np.array([10.0 * k, 3.0, -7.0])
            # B's extent along x after turning by ang about y
            ext = hb[0] * math.cos(ang) + hb[2] * abs(math.sin(ang))
            pb = pa + np.array([ha[0] + ext + gap, 0.01, 0.02])
            pairs.append(((va, qa, pa, False), (vb, qb, pb, False)))
            gaps.append(gap)
    cert = checker(pairs)
    gaps = np.array(gaps)
    for c, g, (a, b) in zip(cert, gaps, pairs):
        if c:
            assert oracle_separated(oracle, a[0], a[1], a[2], b[0], b[1], b[2])
    assert not cert[gaps < 1e-6].any()
    assert cert[gaps > 1e-4].all()


def test_non_unit_quaternion_is_never_certified(oracle, checker):
    va = oracle.box_verts((0.5, 0.5, 0.5))
    q_bad = np.array([1.0 + 1e-6, 0, 0, 0])
    pairs = [((va, q_bad, (0, 0, 0), False), (va, (1, 0, 0, 0), (5.0, 0, 0), False)),
             ((va, (1, 0, 0, 0), (0, 0, 0), False), (va, (1, 0, 0, 0), (5.0, 0, 0), False))]
    assert list(checker(pairs)) == [0, 1]
