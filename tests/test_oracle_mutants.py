"""Planted mistakes in the oracle, on the CPU: each one is a plausible slip in the
restatement of the reference's contact path. The test rebuilds the oracle with one
mistake at a time (a copy of oracle/*.cpp with one line changed, built into a
temporary directory) and runs the pin tests against that build through PKO_LIB; every
mistake must make at least one pin fail. A mistake no pin catches would be copied bit
for bit by the GPU, whose parity tests compare against this oracle.

The reference lines each mutated statement restates are given beside it."""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PINS = ["tests/test_oracle_contact.py", "tests/test_oracle_pins.py"]

# (name, reference line, text in oracle/pko.cpp, replacement)
MUTANTS = [
    ("baumgarte factor 0.2", "constraint.h:880",
     "const double bias_factor = 0.1, slop = 0.005;", "const double bias_factor = 0.2, slop = 0.005;"),
    ("slop 0.01", "constraint.h:881",
     "const double bias_factor = 0.1, slop = 0.005;", "const double bias_factor = 0.1, slop = 0.01;"),
    ("baumgarte without slop", "constraint.h:920-921",
     "double baum = (-bias_factor / dt) * smax(0.0, pen - slop);",
     "double baum = (-bias_factor / dt) * smax(0.0, pen);"),
    ("restitution min", "constraint.h:914",
     "double e = smax(A.restitution, B.restitution);", "double e = smin(A.restitution, B.restitution);"),
    ("restitution threshold 0", "constraint.h:1071",
     "const double thr = 2.0 * gnorm * dt;", "const double thr = 0.0 * gnorm * dt;"),
    ("friction arithmetic mean", "constraint.h:1076",
     "p.mu = std::sqrt(A.friction * B.friction);", "p.mu = 0.5 * (A.friction + B.friction);"),
    ("friction cone max norm", "constraint.h:1162",
     "double mag = hypot_pk(n1, n2);", "double mag = std::max(std::fabs(n1), std::fabs(n2));"),
    ("friction block off-diagonal sign", "constraint.h:1088",
     "p.i12 = -m12 / det;", "p.i12 = m12 / det;"),
    ("normal impulse unclamped", "constraint.h:898",
     "out.n.lo = 0.0;", "out.n.lo = -DBL_MAX;"),
    ("warm-start radius 20 mm", "collision_phases.h:212",
     "const double distance2_eps = 0.005 * 0.005;", "const double distance2_eps = 0.02 * 0.02;"),
    ("drift^2 0.01", "collision_phases.h:267",
     "const double drift2_eps = 0.06;", "const double drift2_eps = 0.01;"),
    ("breaking depth 0.01", "collision_phases.h:213",
     "const double breaking = 0.05;", "const double breaking = 0.01;"),
    ("add_reduce fourth point from best[0]", "collision_phases.h:182-192",
     "double d2 = sqn(pool[i].local_a - pool[best[2]].local_a);",
     "double d2 = sqn(pool[i].local_a - pool[best[0]].local_a);"),
    ("add_reduce deepest tie to the last", "collision_phases.h:150-151",
     "if (pool[i].depth > pool[best[0]].depth) best[0] = i;",
     "if (pool[i].depth >= pool[best[0]].depth) best[0] = i;"),
    ("EPA world_b weights u<->w", "collision.cpp:450-451",
     "ci.world_b = (p0.pb * u + p1.pb * v) + p2.pb * w;", "ci.world_b = (p0.pb * w + p1.pb * v) + p2.pb * u;"),
    ("gyroscopic sign", "particle.h:75",
     "v3 aa = mul(b.Iw_inv, b.torque - cross(b.angvel, mul(b.Iw, b.angvel)));",
     "v3 aa = mul(b.Iw_inv, b.torque + cross(b.angvel, mul(b.Iw, b.angvel)));"),
]


def _build(tmp, name, old, new):
    d = os.path.join(tmp, name.replace(" ", "_").replace("/", "_").replace("^", "").replace("<->", "_"))
    os.makedirs(d)
    for f in os.listdir(ORACLE):
        if f.endswith((".cpp", ".hpp")):
            shutil.copy(os.path.join(ORACLE, f), d)
    src = open(os.path.join(d, "pko.cpp")).read()
    assert src.count(old) == 1, f"{name}: the statement to mutate is not unique in oracle/pko.cpp"
    open(os.path.join(d, "pko.cpp"), "w").write(src.replace(old, new))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-O1", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-o", lib, os.path.join(d, "pko.cpp"), os.path.join(d, "pko_capi.cpp")],
                   check=True, capture_output=True)
    return lib


def _pins_fail(lib):
    env = dict(os.environ, PKO_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode != 0, r.stdout[-400:]


def test_every_planted_mistake_fails_a_pin():
    # the control: the same build recipe with nothing changed must pass every pin, so a
    # broken build or harness cannot pass for a caught mistake
    control = ("control", "-", MUTANTS[0][2], MUTANTS[0][2])
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1))) as ex:
            libs = list(ex.map(lambda m: _build(tmp, m[0], m[2], m[3]), [control] + MUTANTS))
            results = list(ex.map(_pins_fail, libs))
    failed, out = results[0]
    assert not failed, "the unaltered oracle fails its pins under the mutant harness:\n" + out
    results = results[1:]
    survivors = [(m[0], m[1], out) for m, (failed, out) in zip(MUTANTS, results) if not failed]
    assert not survivors, "planted mistakes no pin catches:\n" + "\n".join(
        f"  {name} ({ref})\n{out}" for name, ref, out in survivors)


@pytest.mark.parametrize("m", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutated_statement_is_in_the_oracle(m):
    """Each mutated statement still appears exactly once, so the mutant stays meaningful
    as the oracle is edited."""
    assert open(os.path.join(ORACLE, "pko.cpp")).read().count(m[2]) == 1
