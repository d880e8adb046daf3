"""The C ABI: struct layouts agree between include/persiansph_amd.h and the Python mirror, and the
built HIP library exports every entry point the header declares (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from persiansph_amd import abi
from persiansph_amd import solver as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "persiansph_amd.h")
GNN_HEADER = os.path.join(ROOT, "include", "psph_gnn.h")


def _c_offsets(tmp_path, struct, fields):
    src = tmp_path / "off.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', f'#include "{GNN_HEADER}"',
             'int main(void){']
    lines.append(f'printf("sizeof %zu\\n", sizeof({struct}));')
    for f in fields:
        lines.append(f'printf("{f} %zu\\n", offsetof({struct}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.strip().splitlines())}


def test_particle_struct_layout(tmp_path):
    names = [n for n, _ in abi.PARTICLE_FIELDS]
    off = _c_offsets(tmp_path, "psph_particle", names)
    assert off["sizeof"] == abi.PARTICLE_DTYPE.itemsize
    for n in names:
        assert off[n] == abi.PARTICLE_DTYPE.fields[n][1], n


@pytest.mark.parametrize("cls,struct", [(abi.Config, "psph_config"), (abi.Status, "psph_status"),
                                        (abi.Grid, "psph_grid"), (abi.OutputF32, "psph_output_f32")])
def test_small_struct_layouts(tmp_path, cls, struct):
    names = [f[0] for f in cls._fields_]
    off = _c_offsets(tmp_path, struct, names)
    assert off["sizeof"] == ctypes.sizeof(cls)
    for n in names:
        assert off[n] == getattr(cls, n).offset, n


def _declared(header=HEADER):
    txt = open(header).read()
    return sorted(set(re.findall(r"^(?:int|int64_t) (psph_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(S.LIB_PATH), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(S.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 12
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(S.EXPORTS) == decl
    assert L.psph_version() == 102


def test_errors_without_device_or_bad_config():
    """Without a GPU psph_create reports PSPH_ERR_NO_DEVICE; bad enums are rejected first."""
    L = S.lib()
    h = ctypes.c_void_p()
    cfg = abi.default_config(dimension=4)
    assert L.psph_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    cfg = abi.default_config(dimension=2)
    cfg.periodic[2] = 1
    assert L.psph_create(ctypes.byref(cfg), ctypes.byref(h)) == -2
    buf = ctypes.create_string_buffer(256)
    n = L.psph_last_error(buf, 256)
    assert n > 0 and b"periodic" in buf.value
    cfg = abi.default_config(dimension=2)
    rc = L.psph_create(ctypes.byref(cfg), ctypes.byref(h))
    if rc == 0:
        L.psph_destroy(h)           # a GPU is visible here
    else:
        assert rc == -6


def test_particle_defaults_match_constructor():
    """new_particles restates the Particle constructor defaults (Particle.cpp:25-117)."""
    p = abi.new_particles(3)
    assert np.all(p["TIn"] == 4.0) and np.all(p["FPMassC"] == 1.0) and np.all(p["m"] == 300.0)
    assert np.all(p["ShepardStep"] == 40) and np.all(p["FirstStep"] == 1) and np.all(p["IsFree"] == 1)
    assert np.all(p["CSmag"] == 0.17) and np.all(p["Material"] == 0)


def test_surrogate_config_layout(tmp_path):
    from persiansph_amd.gnn import GnnConfig
    names = [f[0] for f in GnnConfig._fields_]
    off = _c_offsets(tmp_path, "psph_gnn_config", names)
    assert off["sizeof"] == ctypes.sizeof(GnnConfig)
    for n in names:
        assert off[n] == getattr(GnnConfig, n).offset, n


def test_library_exports_every_surrogate_symbol():
    from persiansph_amd.gnn import GNN_EXPORTS
    L = ctypes.CDLL(S.LIB_PATH)
    decl = _declared(GNN_HEADER)
    assert len(decl) == 14
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(GNN_EXPORTS) == decl


def test_surrogate_reports_missing_device():
    """Without a visible GPU the surrogate's create fails with the no-device code."""
    from persiansph_amd.gnn import Network
    try:
        Network(2, 1, 0.1).close()
    except S.SolverError as e:
        assert e.code == -6, e
    else:
        pytest.skip("a GPU is visible")
