"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle's C source
is built with -fsanitize=address,undefined (errors fatal) into a temporary directory and a set of
the oracle's own tests runs against that build through PSPH_ORACLE_LIB, in a child process with
the ASan runtime preloaded (the Python interpreter itself is not instrumented; leak checks are
off because the interpreter keeps its allocations to the end).  Covers the domain loop with
one and with several OpenMP threads, in/out-flow recycling, multi-material pairs and the
constitutive updates.  (ThreadSanitizer needs an OpenMP runtime built for it, which this image
does not have; the oracle's locking is checked by the multi-thread runs agreeing with the
single-thread order to rounding.)"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    try:
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    except OSError:
        return None
    return os.path.realpath(p) if p and os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(shutil.which("gcc") is None or _runtime("libasan.so") is None, reason="needs gcc with ASan")
def test_oracle_under_asan_and_ubsan(tmp_path):
    lib = tmp_path / "libpsph_oracle_asan.so"
    subprocess.run(["gcc", "-O1", "-g", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                    "-I", os.path.join(ROOT, "oracle"), "-o", str(lib), os.path.join(ROOT, "oracle", "psph_oracle.c"),
                    "-lm"], check=True)
    env = dict(os.environ, PSPH_ORACLE_LIB=str(lib), LD_PRELOAD=_runtime("libasan.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    tests = ["test_oracle_domain.py", "test_oracle_pins.py", "test_oracle_inflow.py",
             "test_oracle_multimaterial.py", "test_oracle_periodic.py", "test_oracle_math.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-p", "no:cacheprovider"]
                       + [os.path.join(ROOT, "tests", t) for t in tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in out and "runtime error" not in out, out
