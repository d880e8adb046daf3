"""The C++ SPH::Domain drop-in (persiansph_amd/cpp/SPH/Domain.h): a 3-DamBreak.cpp-style driver
compiles against it, and on the GPU its Solve reproduces the same model run through the
Python binding of the C ABI."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "persiansph_amd", "cpp", "examples", "dam_break")
SRC = EXE + ".cpp"


def test_dropin_compiles(tmp_path):
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "persiansph_amd", "cpp"),
                    SRC], check=True)


def _read(path):
    with open(path, "rb") as f:
        n = struct.unpack("N", f.read(8))[0]
        a = np.frombuffer(f.read(32*n), dtype=np.float64).reshape(n, 4)
    return a


@pytest.mark.gpu
def test_dropin_solve_matches_binding(tmp_path):
    from persiansph_amd import Solver, cases
    assert os.path.exists(EXE), "build first (__graft_entry__.build())"
    steps = 40
    out = tmp_path / "out.bin"
    print(f"[dropin] 2-D dx=0.04 steps={steps}", flush=True)
    r = subprocess.run([EXE, "2", "0.04", str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _read(out)
    cfg, p, dt = cases.dam_break_2d(dx=0.04, scheme=1)
    s = Solver(cfg, p, dt)
    st = s.run(10**9, tf=steps*dt*(1.0 - 1e-9))
    q = s.particles()
    assert len(q) == len(got)
    np.testing.assert_array_equal(q["x"], got[:, :3])
    np.testing.assert_array_equal(q["Density"], got[:, 3])
    assert f"{st.time:.9e}" in r.stdout
