"""XDMF output (Domain::WriteXDMF, Source/Domain.cpp:1869-2064).

CPU: the drop-in's host writer puts every field of a set of particles, rounded to float, into
the reference's datasets, and the .xmf describes them as the reference's does.
GPU: the fields packed on the device (psph_output_begin/end) are the downloaded FP64 state
rounded to nearest float, bit for bit, before the first step, at a mid-step output stop, after
steps and across a run started while the output is pending; they agree with the oracle's state
to one float rounding (at an output stop the fixed particles' Acceleration is their pair sums, as
the reference's; between steps the device holds no such sums and gives Gravity); and the
drop-in's Solve writes, at each output time, the files that state gives.
"""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL_SRC = os.path.join(ROOT, "tests", "xdmf_tool.cpp")
EXE = os.path.join(ROOT, "persiansph_amd", "cpp", "examples", "dam_break")

DATASETS = ["NP", "Position", "Velocity", "Acceleration", "Tag", "Pressure", "Density", "Mass", "h",
            "Sigma", "Strain"]
SYM = [0, 1, 2, 4, 5, 8]      # xx xy xz yy yz zz of a row-major 3x3


def _tool(tmp_path):
    import __graft_entry__ as g
    cflags, lflags = g.hdf5_flags()
    exe = str(tmp_path / "xdmf_tool")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "persiansph_amd", "cpp")] + cflags + [TOOL_SRC,
                    "-L" + os.path.join(ROOT, "persiansph_amd", "lib"), "-lpersiansph_amd",
                    "-Wl,-rpath," + os.path.join(ROOT, "persiansph_amd", "lib"),
                    "-Wl,-rpath," + os.path.join(ROOT, "persiansph_amd", "lib", "hdf5")] + lflags + ["-o", exe],
                   check=True)
    return exe


def _read(tool, path):
    out = subprocess.run([tool, "read", str(path)], check=True, capture_output=True, text=True).stdout.split("\n")
    res, k = {}, 0
    while k < len(out) and out[k]:
        name, kind, cnt = out[k].split()
        cnt = int(cnt)
        vals = out[k + 1:k + 1 + cnt]
        if kind == "int":
            res[name] = np.array([int(v) for v in vals], dtype=np.int32)
        else:
            res[name] = np.array([int(v, 16) for v in vals], dtype=np.uint32).view(np.float32)
        k += 1 + cnt
    return res


def _check_xmf(path, key, n, names):
    txt = open(path).read()
    assert txt.startswith('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n<Xdmf Version="2.0">\n')
    root = ET.fromstring(txt.split("\n", 2)[2])
    grid = root.find("Domain/Grid")
    assert grid.get("Name") == "SPHCenter" and grid.get("GridType") == "Uniform"
    assert grid.find("Topology").get("TopologyType") == "Polyvertex"
    assert grid.find("Topology").get("NumberOfElements") == str(n)
    geo = grid.find("Geometry/DataItem")
    assert geo.get("Dimensions") == f"{n} 3" and geo.text.strip() == f"{key}.hdf5:/Position"
    want = [("Tag", "Scalar", f"{n}", "Int"), ("Position", "Vector", f"{n} 3", "Float"),
            ("Velocity", "Vector", f"{n} 3", "Float"), ("Acceleration", "Vector", f"{n} 3", "Float"),
            ("Density", "Scalar", f"{n}", "Float"), ("Pressure", "Scalar", f"{n}", "Float")]
    want += [(nm, "Scalar", f"{n}", "Float") for nm in names]
    want += [("Sigma", "Tensor6", f"{n} 6", "Float"), ("Strain", "Tensor6", f"{n} 6", "Float")]
    got = grid.findall("Attribute")
    assert len(got) == len(want)
    for a, (nm, ty, dims, num) in zip(got, want):
        d = a.find("DataItem")
        assert (a.get("Name"), a.get("AttributeType"), a.get("Center")) == (nm, ty, "Node")
        assert (d.get("Dimensions"), d.get("NumberType"), d.get("Format")) == (dims, num, "HDF")
        assert d.get("Precision") == (None if num == "Int" else "10")
        assert d.text.strip() == f"{key}.hdf5:/{nm}"


def _val(i, field, c):
    return 0.001*i + 0.1*field + 0.0137*c + 1.0/3.0


def test_host_writer_datasets_and_xmf(tmp_path):
    tool = _tool(tmp_path)
    n = 37
    key = str(tmp_path / "case_0001")
    subprocess.run([tool, "write", key, str(n)], check=True, capture_output=True)
    d = _read(tool, key + ".hdf5")
    names = ["ShearRate", "Property2", "Property3"]
    assert sorted(d) == sorted(DATASETS + names)
    i = np.arange(n)
    f32 = np.float32
    assert d["NP"].tolist() == [n]
    pos = np.stack([_val(i, 0, c) for c in range(3)], 1)
    np.testing.assert_array_equal(d["Position"], pos.astype(f32).ravel())
    np.testing.assert_array_equal(d["Velocity"], np.stack([_val(i, 4, c) for c in range(3)], 1).astype(f32).ravel())
    np.testing.assert_array_equal(d["Acceleration"], np.stack([_val(i, 5, c) for c in range(3)], 1).astype(f32).ravel())
    np.testing.assert_array_equal(d["Tag"], (i % 5).astype(np.int32))
    np.testing.assert_array_equal(d["Pressure"], _val(i, 6, 0).astype(f32))
    dens = _val(i, 2, 0)
    np.testing.assert_array_equal(d["Density"], dens.astype(f32))
    np.testing.assert_array_equal(d["Mass"], _val(i, 1, 0).astype(f32))
    np.testing.assert_array_equal(d["h"], _val(i, 3, 0).astype(f32))
    np.testing.assert_array_equal(d["ShearRate"], (2.0*dens).astype(f32))
    np.testing.assert_array_equal(d["Property2"], (i % 5).astype(f32))
    np.testing.assert_array_equal(d["Property3"], (-_val(i, 1, 0)).astype(f32))
    sig = np.stack([_val(i, 7, k) for k in SYM], 1)
    eps = np.stack([_val(i, 8, k) for k in SYM], 1)
    np.testing.assert_array_equal(d["Sigma"], sig.astype(f32).ravel())
    np.testing.assert_array_equal(d["Strain"], eps.astype(f32).ravel())
    _check_xmf(key + ".xmf", key, n, names)


def _as_output(q, gravity=None):
    """psph_output_f32's fields from a downloaded (or oracle) state: FP64 rounded to float."""
    f32 = np.float32
    return {"position": q["x"].astype(f32), "velocity": q["v"].astype(f32), "acceleration": q["a"].astype(f32),
            "pressure": q["Pressure"].astype(f32), "density": q["Density"].astype(f32),
            "mass": q["Mass"].astype(f32), "h": q["h"].astype(f32), "tag": q["ID"].astype(np.int32),
            "sigma": q["Sigma"].reshape(-1, 9)[:, SYM].astype(f32),
            "strain": q["Strain"].reshape(-1, 9)[:, SYM].astype(f32)}


def _equal(out, ref, where):
    for k, v in ref.items():
        assert out[k].shape == v.shape, (where, k)
        bad = out[k] != v
        assert not bad.any(), f"{where} {k}: {bad.sum()} differ, first {np.argwhere(bad)[0]}"


def _oracle_close(out, q, sc, where, acc_rows):
    """The packed floats against the oracle's FP64 state: within one float rounding of it plus
    the FP64 parity bar (accelerations: 1e-12 of |a| + the sum of |pair contributions|; other
    fields: 1e-11 relative, as the trajectory tests)."""
    ref = {"position": q["x"], "velocity": q["v"], "acceleration": q["a"], "pressure": q["Pressure"],
           "density": q["Density"], "mass": q["Mass"], "h": q["h"],
           "sigma": q["Sigma"].reshape(-1, 9)[:, SYM], "strain": q["Strain"].reshape(-1, 9)[:, SYM]}
    assert np.array_equal(out["tag"], q["ID"].astype(np.int32)), where
    for k, v in ref.items():
        a = out[k].astype(np.float64)
        if k == "acceleration":
            a, v = a[acc_rows], v[acc_rows]
            tol64 = 1e-12*(np.abs(v) + sc[acc_rows, 0:3])
        else:
            tol64 = 1e-11*(1.0 + np.abs(v))
        ulp = np.spacing(np.abs(v).astype(np.float32)).astype(np.float64)
        diff = np.abs(a - v)
        assert np.all(diff <= ulp + tol64), f"{where} {k}: max excess {float((diff - ulp - tol64).max())}"


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fluid_leapfrog", "soil_mverlet"])
def test_gpu_packed_output_equals_download(case):
    from persiansph_amd import Solver, cases
    from oracle.pyoracle import Oracle
    if case == "fluid_leapfrog":
        cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=1)
    else:
        cfg, p, dt = cases.granular_column(dim=2, dx=0.01, size=(0.1, 0.2, 0.0), floor_half=0.2, scheme=0)
    print(f"[xdmf] {case} n={len(p)}", flush=True)
    s = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt, track_scales=True)
    # before the first step: the uploaded state with the EOS pressure
    _equal(s.output_f32(), _as_output(s.particles()), "initial")
    free = p["IsFree"] == 1
    _oracle_close(s.output_f32(), o.particles(), o.scales(len(p)), "initial vs oracle", free)
    # mid-step output stop (forces of the step present, Move pending), then after steps
    for tout, where in ((0.0, "mid-step"), (None, "after 7 steps")):
        if tout is None:
            s.run(7)
            o.run(7)
        else:
            s.run(10**9, tout=tout)
            o.run(10**9, tout=tout)
        _equal(s.output_f32(), _as_output(s.particles()), where)
        out = s.output_f32()
        mid = tout is not None
        _oracle_close(out, o.particles(), o.scales(len(p)), where + " vs oracle",
                      np.ones_like(free) if mid else free)
        walls_g = np.all(out["acceleration"][~free] == np.float32(np.array(cfg.gravity)))
        assert walls_g != mid, where
    # an output may be pending while the solver runs on; the packed state is the one at begin
    s.output_begin()
    before = _as_output(s.particles())
    s.run(3)
    _equal(s.output_end(), before, "pending across run")
    s.close()


@pytest.mark.gpu
def test_dropin_solve_writes_xdmf_series(tmp_path):
    from persiansph_amd import Solver, cases
    assert os.path.exists(EXE), "build first (__graft_entry__.build())"
    tool = _tool(tmp_path)
    steps, every = 30, 10
    key = str(tmp_path / "db")
    print(f"[xdmf] drop-in 2-D dx=0.04 steps={steps} every={every}", flush=True)
    r = subprocess.run([EXE, "2", "0.04", str(steps), "-", key, str(every)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Initial Condition has been generated" in r.stdout
    cfg, p, dt = cases.dam_break_2d(dx=0.04, scheme=1)
    s = Solver(cfg, p, dt)
    names = ["Property1", "Property2", "Property3"]
    expect = {"db_Initial": _as_output(s.particles())}
    # the Solve loop of the drop-in: stop at each output time, write, go on
    tf, tout, idx = steps*dt*(1.0 - 1e-9), 0.0, 1
    while True:
        st = s.run(10**9, tf=tf, tout=tout)
        if st.stop_reason != 2:
            break
        expect[f"db_{idx:04d}"] = _as_output(s.particles())
        idx += 1
        tout += every*dt
    assert len(expect) >= 4
    for name, ref in expect.items():
        d = _read(tool, str(tmp_path / (name + ".hdf5")))
        n = len(ref["tag"])
        assert sorted(d) == sorted(DATASETS + names)
        got = {"position": d["Position"].reshape(n, 3), "velocity": d["Velocity"].reshape(n, 3),
               "acceleration": d["Acceleration"].reshape(n, 3), "pressure": d["Pressure"],
               "density": d["Density"], "mass": d["Mass"], "h": d["h"], "tag": d["Tag"],
               "sigma": d["Sigma"].reshape(n, 6), "strain": d["Strain"].reshape(n, 6)}
        _equal(got, ref, name)
        for nm in names:
            assert not d[nm].any()
        _check_xmf(str(tmp_path / (name + ".xmf")), str(tmp_path / name), n, names)
    assert os.path.exists(str(tmp_path / "db_log.dat"))
    s.close()
