"""GPU: the reference's own user programs, unchanged, on the drop-in.

tests/programs/Makefile compiles testcases/GHZ_QFT-30.c, bv-32.c, random-32.c
and supremecy-32.c of the reference against include/QuEST.h and
libQuEST_f64.so.  Each binary runs as it is; DGQ_REPORT_FILE makes the
library append a full-precision summary of every register it destroys
(total probability, P(q = 1) per qubit, the first 16 amplitudes), which is
what is checked, together with the lines the programs print themselves:

* GHZ_QFT-30 (30 qubits): every amplitude against the closed form of that
  circuit under its mistyped pi (circuits.ghz_qft_testcase_amps, pinned
  against the oracle and the dense brute force on the CPU) within 1e-12,
  P(q = 1) = 1/2 for every qubit within 1e-12; the printed values agree;
* bv-32 (32 qubits): the data register ends exactly in the hidden string
  of its CNOT pattern (testcases/bv-32.c:41-55), the ancilla in |->:
  P(q = 1) is 1 on the string's bits, 0 on the others, 1/2 on qubit 31;
* random-32 and supremecy-32 (32 qubits, 505 and 828 gates): total
  probability 1 within 1e-12, and the same summary from a second run under a
  different plan (fusion width 4 on a 10-qubit tile, fixed lanes instead of
  the default width-7 rotated plan) within 1e-12.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from dgquest_amd import circuits

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "programs", "_ref")
TOL = 1e-12

# bits set in the hidden string of bv-32 (the controls of its CNOTs onto qubit 31)
BV32_STRING = (2, 4, 6, 7, 8, 16, 17, 19, 20, 21, 22, 24, 25, 27, 28)


def run_program(name, tmp_path, env_extra=None, tag="a"):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (tests/programs/Makefile needs the reference checkout)")
    report = tmp_path / f"{name}.{tag}.jsonl"
    env = dict(os.environ, DGQ_REPORT_FILE=str(report))
    env.update(env_extra or {})
    print(f"running {name} ({tag})", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env, cwd=str(tmp_path))
    print(p.stdout[-1500:], flush=True)
    assert p.returncode == 0, p.stderr[-2000:]
    rows = [json.loads(x) for x in report.read_text().splitlines()]
    assert len(rows) == 1
    return p.stdout, rows[0]


def printed_probs(out):
    return {int(m.group(1)): float(m.group(2))
            for m in re.finditer(r"probability for q\[\s*(\d+)\]==1 : ([-\d.]+)", out)}


def test_ghz_qft_30(tmp_path):
    out, r = run_program("GHZ_QFT-30", tmp_path)
    n = 30
    assert r["n"] == n and r["prec"] == 2
    assert abs(r["total_prob"] - 1.0) <= TOL
    assert np.abs(np.array(r["p1"]) - 0.5).max() <= TOL
    got = np.array([complex(a, b) for a, b in r["amps"]])
    ref = circuits.ghz_qft_testcase_amps(n, np.arange(len(got)))
    err = float(np.abs(got - ref).max())
    print(f"GHZ_QFT-30: max |amp - closed form| = {err:.2e}", flush=True)
    assert err <= TOL
    pp = printed_probs(out)
    assert sorted(pp) == list(range(n))
    assert all(abs(v - 0.5) <= 1e-6 for v in pp.values())
    # the printed amplitudes (6 decimals), including index 2^30 - 1
    for m in re.finditer(r"Amplitude of (\d+)th state vector:\s+([-\d.]+),\s+([-\d.]+)", out):
        i = int(m.group(1))
        a = circuits.ghz_qft_testcase_amps(n, [i])[0]
        assert abs(float(m.group(2)) - a.real) <= 1e-6 and abs(float(m.group(3)) - a.imag) <= 1e-6


def test_bv_32(tmp_path):
    out, r = run_program("bv-32", tmp_path)
    assert "Complete the simulation" in out
    assert r["n"] == 32 and abs(r["total_prob"] - 1.0) <= TOL
    p1 = np.array(r["p1"])
    want = np.zeros(32)
    want[list(BV32_STRING)] = 1.0
    want[31] = 0.5
    assert np.abs(p1 - want).max() <= TOL
    # getAmp(q, 0), the amplitude the program reads: 0, since the string is not 0
    assert np.abs(np.array(r["amps"])).max() <= TOL


@pytest.mark.parametrize("name", ["random-32", "supremecy-32"])
def test_random_programs_plan_independent(tmp_path, name):
    out, a = run_program(name, tmp_path, tag="default")
    _, b = run_program(name, tmp_path, {"DGQ_TUNE": "4,4,0,4096", "DGQ_ROTATE": "0"}, tag="k4")
    for r in (a, b):
        assert r["n"] == 32 and abs(r["total_prob"] - 1.0) <= TOL
    d_p = float(np.abs(np.array(a["p1"]) - np.array(b["p1"])).max())
    d_a = float(np.abs(np.array(a["amps"]) - np.array(b["amps"])).max())
    print(f"{name}: plans differ by {d_p:.2e} in P(q=1), {d_a:.2e} in amplitudes", flush=True)
    assert d_p <= TOL and d_a <= TOL
    pp = printed_probs(out)
    assert sorted(pp) == list(range(32))
    # printed from a float to 6 decimals
    assert max(abs(pp[q] - a["p1"][q]) for q in range(32)) <= 1e-6
