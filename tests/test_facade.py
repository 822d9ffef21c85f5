"""The physkit-style C++ facade (include/physkit_amd/world.hpp) and the Python mirror.

CPU: the facade header compiles and links against the C ABI library. GPU: the C++
driver runs the reference's co_tests head-on scene (begin then end events), a
resting box and arena slot reuse; the Python world API steps the same head-on scene.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "facade_demo")


def _build():
    import __graft_entry__
    return __graft_entry__.build_facade_demo()


def test_facade_compiles_and_links(pk):
    exe = _build()
    assert os.path.exists(exe)


def test_facade_host_logic_with_test_double(tmp_path):
    """Handles, slot reuse and re-upload keeping the state reached so far, with the
    facade linked against a host-memory test double of the C ABI (tests/cpp/stub_pkg.cpp)."""
    exe = str(tmp_path / "facade_logic")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "facade_logic.cpp"),
                    os.path.join(ROOT, "tests", "cpp", "stub_pkg.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade logic ok" in r.stdout


@pytest.mark.gpu
def test_facade_demo_runs_on_gpu(pk):
    exe = EXE if os.path.exists(EXE) else _build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade ok" in r.stdout


@pytest.mark.gpu
def test_python_world_api_head_on(pk):
    w = pk.world(pk.world_desc.make().with_gravity((0, 0, 0)))
    box = pk.mesh.box((0.5, 0.5, 0.5))
    a = w.create_rigid(pk.object_desc.dynam().with_mesh(box).with_pos((-2, 0, 0)).with_vel((2, 0, 0)))
    b = w.create_rigid(pk.object_desc.dynam().with_mesh(box).with_pos((2, 0, 0)).with_vel((-2, 0, 0)))
    events = []
    w.on_collision = lambda x, y: events.append(("begin", x, y))
    w.on_collision_exit = lambda x, y: events.append(("end", x, y))
    for _ in range(400):
        w.step(0.005)
    kinds = [e[0] for e in events]
    assert "begin" in kinds and "end" in kinds and kinds.index("begin") < kinds.index("end")
    st = w.state()
    assert st["vel"][a][0] < 0 < st["vel"][b][0]
    w.gpu.close()
