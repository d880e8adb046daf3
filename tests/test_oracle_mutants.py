"""The oracle's pins have teeth: each test here plants one plausible mistake in a copy of the
oracle's C source (a transposed contraction, a flipped sign, a dropped factor), builds that
copy on the CPU into a temporary directory and runs tests/test_oracle_pins.py against it through
PSPH_ORACLE_LIB.  Every mutant must make the pins fail; the real oracle is never touched."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "psph_oracle.c")

# (what the mistake is, text in the oracle, replacement)
MUTANTS = [
    ("stress contracted with its second index (Mult(Vec, Mat) transposed)",
     "T[j] = A[0]*M[j] + A[1]*M[3+j] + A[2]*M[6+j];",
     "T[j] = A[0]*M[3*j] + A[1]*M[3*j+1] + A[2]*M[3*j+2];"),
    ("rotation rate made symmetric",
     "    RR[3] = -RR[1];", "    RR[3] = RR[1];"),
    ("strain-rate sign flipped",
     "for (int q = 0; q < 9; ++q) SR[q] = SR[q]*f05;",
     "for (int q = 0; q < 9; ++q) SR[q] = -SR[q]*f05;"),
    ("hydrostatic wall term with the opposite sign",
     "double t = F->Pressure * K + dot3(G, xij)*F->Density*K;",
     "double t = F->Pressure * K - dot3(G, xij)*F->Density*K;"),
    ("wall velocity not kernel weighted",
     "W->NSv[c] += F->v[c] * K;", "W->NSv[c] += F->v[c];"),
    ("stress density factor 1/(rho_i rho_j) in the squared-density form",
     "double fi = 1.0/(di*di), fj = 1.0/(dj*dj);", "double fi = 1.0/(di*di), fj = 1.0/(di*dj);"),
    ("stress density factor 1/rho_i^2 in the multiplied-density form",
     "double f = 1.0/(di*dj);", "double f = 1.0/(di*di);"),
    ("shear modulus not doubled in the deviatoric rate",
     "R[k] = dev*(2.0*p->G) + SRT[k] + RS[k];", "R[k] = dev*(p->G) + SRT[k] + RS[k];"),
    ("Jaumann spin term with the wrong sign",
     "R[k] = dev*(2.0*p->G) + SRT[k] + RS[k];", "R[k] = dev*(2.0*p->G) + SRT[k] - RS[k];"),
    ("von Mises scale-back on sqrt(J2) instead of sqrt(3 J2)",
     "double a = p->Sigmay/sqrt(3.0*J2);", "double a = p->Sigmay/sqrt(J2);"),
    ("Bingham law without the 1/shear-rate",
     "p->Mu = p->MuRef + p->T0*(1-exp(-p->m*p->ShearRate))/p->ShearRate;",
     "p->Mu = p->MuRef + p->T0*(1-exp(-p->m*p->ShearRate));"),
    ("Cross law with MuRef instead of MuRef^2",
     "p->Mu = (1000.0*p->MuRef + p->MuRef*p->MuRef*1000.0/p->T0*p->ShearRate)",
     "p->Mu = (1000.0*p->MuRef + p->MuRef*1000.0/p->T0*p->ShearRate)"),
    ("Smagorinsky rate sqrt(e:e) instead of sqrt(2 e:e)",
     "p->SBar = sqrt(2.0*temp);", "p->SBar = sqrt(temp);"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("what,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(tmp_path, what, old, new):
    src = open(SRC).read()
    assert src.count(old) >= 1, f"mutation site not found: {old}"
    mut = tmp_path / "psph_oracle.c"
    mut.write_text(src.replace(old, new, 1))
    lib = tmp_path / "libpsph_oracle_mutant.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    "-I", os.path.join(ROOT, "oracle"), "-o", str(lib), str(mut), "-lm"], check=True)
    env = dict(os.environ, PSPH_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "failed" in r.stdout, f"mutant survived: {what}\n{r.stdout[-2000:]}"
