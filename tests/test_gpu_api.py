"""GPU tests of the QuEST API surface: the error convention (a QuEST error
message and exit(code), src/QuEST_validation.c:19-88) and a plain C program
linked against libQuEST_f64.so."""
import math
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from dgquest_amd import circuits

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = textwrap.dedent("""
    import sys
    sys.path.insert(0, %r)
    from dgquest_amd import quest as Q
    Q_ = Q.load(2)
    env = Q_.createQuESTEnv()
""" % ROOT)

CASES = [
    ("createQureg", 1, "Invalid number of qubits", "Q_.createQureg(0, env)"),
    ("hadamard", 2, "Invalid target qubit", "q = Q_.createQureg(3, env); Q_.hadamard(q, 5)"),
    ("controlledNot", 3, "Invalid control qubit", "q = Q_.createQureg(3, env); Q_.controlledNot(q, -1, 0)"),
    ("getAmp", 4, "Invalid state index", "q = Q_.createQureg(3, env); Q_.getAmp(q, 8)"),
    ("controlledNot", 7, "Control qubit cannot equal target qubit",
     "q = Q_.createQureg(3, env); Q_.controlledNot(q, 1, 1)"),
    ("multiControlledUnitary", 8, "Control qubits cannot include target qubit",
     "q = Q_.createQureg(4, env); import ctypes as C; a=(C.c_int*2)(1,2); "
     "Q_.multiControlledUnitary(q, a, 2, 2, Q_.matrix([1,0,0,0,0,0,1,0]))"),
    ("swapGate", 9, "must be unique", "q = Q_.createQureg(3, env); Q_.swapGate(q, 1, 1)"),
    ("unitary", 11, "Matrix is not unitary", "q = Q_.createQureg(3, env); Q_.unitary(q, 0, Q_.matrix([1,0,1,0,0,0,1,0]))"),
    ("compactUnitary", 12, "not unitary",
     "q = Q_.createQureg(3, env); Q_.compactUnitary(q, 0, Q_.complex(1, 0), Q_.complex(0.5, 0))"),
    ("rotateAroundAxis", 13, "Invalid axis vector", "q = Q_.createQureg(3, env); Q_.rotateAroundAxis(q, 0, 0.3, Q_.Vector(0,0,0))"),
    ("collapseToOutcome", 15, "zero probability", "q = Q_.createQureg(3, env); Q_.collapseToOutcome(q, 0, 1)"),
    ("calcProbOfOutcome", 16, "Invalid measurement outcome", "q = Q_.createQureg(3, env); Q_.calcProbOfOutcome(q, 0, 2)"),
]


@pytest.mark.parametrize("func,code,msg,body", CASES)
def test_validation_exits_with_quest_error(func, code, msg, body):
    p = subprocess.run([sys.executable, "-c", PRELUDE + body], capture_output=True, text=True, timeout=120)
    assert p.returncode == code, (p.returncode, p.stdout[-500:], p.stderr[-500:])
    assert f"QuEST Error in function {func}: " in p.stdout
    assert msg in p.stdout


def test_c_program_against_closed_form(tmp_path):
    exe = tmp_path / "ghz_qft_c"
    lib = os.path.join(ROOT, "dgquest_amd", "lib")
    subprocess.check_call(["gcc", "-O2", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "programs", "ghz_qft_c.c"), "-o", str(exe), "-L", lib,
                           "-lQuEST_f64", "-lm", f"-Wl,-rpath,{lib}"])
    n = 20
    out = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120, check=True).stdout
    cf = circuits.ghz_qft_closed_form(n)
    idx = np.arange(1 << n)
    for line in out.splitlines():
        f = line.split()
        if f[0] == "p1":
            t = int(f[1])
            p1 = float(np.sum(np.abs(cf[((idx >> t) & 1) == 1]) ** 2))
            assert abs(float(f[2]) - p1) <= 1e-12
        elif f[0] == "total":
            assert abs(float(f[1]) - 1.0) <= 1e-12
        elif f[0] == "amp":
            i = int(f[1])
            assert abs(complex(float(f[2]), float(f[3])) - cf[i]) <= 1e-12
