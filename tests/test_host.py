"""CPU checks of the built libraries and host logic (no GPU, no kernel calls)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "dgquest_amd", "lib")


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def exported(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(LIB, lib)], text=True)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_engine_exports_every_declared_symbol():
    decl = declared_functions("dgq_engine.h")
    assert len(decl) > 20
    missing = [d for d in decl if d not in exported("libdgq_engine.so")]
    assert not missing, missing


@pytest.mark.parametrize("lib", ["libQuEST_f64.so", "libQuEST_f32.so"])
def test_quest_exports_every_declared_symbol(lib):
    decl = declared_functions("QuEST.h") + declared_functions("dgquest_ext.h")
    assert "createQureg" in decl and "controlledPhaseShift" in decl and "dgq_plan_only" in decl
    missing = [d for d in decl if d not in exported(lib)]
    assert not missing, missing


def test_libraries_load_and_report_precision():
    for name, prec in (("libQuEST_f64.so", 2), ("libQuEST_f32.so", 1)):
        L = C.CDLL(os.path.join(LIB, name))
        assert L.getQuEST_PREC() == prec


class Rec(C.Structure):
    _fields_ = [("type", C.c_int), ("pad", C.c_int), ("tmask", C.c_ulonglong), ("cmask", C.c_ulonglong),
                ("op", C.c_double * 8)]


def records_of(circ):
    import math

    out = []
    for name, qs, p in circ:
        r = Rec()
        if name == "u3Gate":
            th, ph, la = p
            ct, st = math.cos(th / 2), math.sin(th / 2)
            m = [ct, 0, -math.cos(la) * st, -math.sin(la) * st, math.cos(ph) * st, math.sin(ph) * st,
                 math.cos(ph + la) * ct, math.sin(ph + la) * ct]
            r.type, r.tmask = 18, 1 << qs[0]
            for i in range(8):
                r.op[i] = m[i]
        elif name == "controlledPhaseFlip":
            r.type, r.cmask = 4, (1 << qs[0]) | (1 << qs[1])
        out.append(r)
    return (Rec * len(out))(*out), len(out)


def plan(n, circ, k, R, W, max_ops=100000):
    L = C.CDLL(os.path.join(LIB, "libQuEST_f64.so"))
    arr, cnt = records_of(circ)
    st = (C.c_int * 8)()
    rc = L.dgq_plan_only(n, 2, arr, cnt, k, R, W, max_ops, st)
    return rc, list(st)


def test_planner_counts_on_the_30_qubit_workload():
    from dgquest_amd import circuits

    c = circuits.layered_u3_cz(30, 24, 2)
    sweeps = {}
    for (k, R, W) in [(3, 3, 0), (4, 4, 0), (6, 4, 2), (7, 4, 3)]:
        rc, st = plan(30, c, k, R, W)
        assert rc == 0
        assert st[4] == len(c)          # every recorded gate reached the planner
        assert st[5:8] == [k, R, W]     # shape accepted as asked
        assert st[1] >= st[0]           # at least one phase per sweep
        sweeps[k] = st[0]
    # wider fusion -> fewer sweeps; every qubit is a target, a sweep holds at most 6 + k
    assert sweeps[3] > sweeps[4] > sweeps[6] > sweeps[7]
    assert sweeps[7] >= -(-30 // (6 + 7))


def test_planner_caps_ops_per_sweep():
    from dgquest_amd import circuits

    c = circuits.layered_u3_cz(16, 8, 3)
    rc, st_big = plan(16, c, 6, 4, 2, 100000)
    rc2, st_small = plan(16, c, 6, 4, 2, 10)
    assert rc == rc2 == 0
    assert st_small[0] >= len(c) // 10
    assert st_big[0] < st_small[0]


def test_planner_rejects_bad_records():
    L = C.CDLL(os.path.join(LIB, "libQuEST_f64.so"))
    r = (Rec * 1)()
    r[0].type = 14
    r[0].tmask = 1 << 40  # qubit beyond n
    st = (C.c_int * 8)()
    assert L.dgq_plan_only(10, 2, r, 1, 6, 4, 2, 100, st) == -1
    assert L.dgq_plan_only(10, 3, r, 0, 6, 4, 2, 100, st) == -1


@pytest.mark.parametrize("case", [
    (9, 1 << 2, 0, 0.0),           # ROTATE_Z_FAKE: kernel-internal, never recorded
    (9, 0, 1 << 2, 0.0),
    (4, 1 << 2, 1 << 3, 0.0),      # phase gate with a target
    (2, 0, 0, 0.0),                # phase gate on no qubit
    (8, 1 << 1, 0, 0.5),
    (14, 0, 0, 0.0),               # Hadamard without a target
    (11, 0b11, 0, 0.0),            # two target bits
    (18, 1 << 3, 1 << 3, 0.0),     # target among the controls
    (10, 0, 1 << 2, 0.3),          # Rz with no target
    (10, 0b101, 0, 0.3),
    (19, 1 << 4, 0, 2.0),          # collapse outcome not 0/1
    (19, 0, 0, 1.0),
    (20, 1 << 4, 0, 0.0),          # unknown type
    (-1, 1 << 4, 0, 0.0),
])
def test_planner_rejects_malformed_records(case):
    """Records from outside the API (dgq_plan_only here, server requests
    through the same check) are refused unless they have the shape the API
    records for their type."""
    L = C.CDLL(os.path.join(LIB, "libQuEST_f64.so"))
    typ, tm, cm, op0 = case
    r = (Rec * 1)()
    r[0].type, r[0].tmask, r[0].cmask = typ, tm, cm
    r[0].op[0] = op0
    r[0].op[1] = 1.0
    st = (C.c_int * 8)()
    assert L.dgq_plan_only(10, 2, r, 1, 6, 4, 2, 100, st) == -1


@pytest.mark.parametrize("case", [
    (4, 0, 1 << 2, 0.0), (8, 0, 0b110, 0.4), (0, 0, 1, 0.0), (10, 1 << 2, 0, 0.3), (10, 1 << 2, 1 << 5, 0.3),
    (18, 1 << 3, 0b110000, 0.0), (19, 1 << 4, 0, 1.0), (19, 1 << 4, 0, 0.0), (14, 1 << 9, 0, 0.0),
])
def test_planner_accepts_api_shaped_records(case):
    L = C.CDLL(os.path.join(LIB, "libQuEST_f64.so"))
    typ, tm, cm, op0 = case
    r = (Rec * 1)()
    r[0].type, r[0].tmask, r[0].cmask = typ, tm, cm
    r[0].op[0] = op0
    r[0].op[1] = 1.0
    st = (C.c_int * 8)()
    assert L.dgq_plan_only(10, 2, r, 1, 6, 4, 2, 100, st) == 0


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_reference_testcases_compile_against_the_drop_in(tmp_path):
    """The reference's own user programs compile and link unchanged against
    include/QuEST.h and libQuEST_f64.so (compile/link only, nothing runs)."""
    tc = "/root/reference/testcases"
    if not os.path.isdir(tc):
        pytest.skip("reference sources not present")
    for prog in ("GHZ_QFT-30.c", "bv-32.c", "random-32.c", "supremecy-32.c"):
        out = tmp_path / prog.replace(".c", "")
        cmd = ["gcc", "-O0", "-w", "-I", os.path.join(ROOT, "include"), "-I", tc, os.path.join(tc, prog), "-o",
               str(out), "-L", LIB, "-lQuEST_f64", "-lm", f"-Wl,-rpath,{LIB}"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-2000:]


def _searched(n, circ):
    """(shape, predicted ms, sweeps of the searched plan) for a recording register."""
    from dgquest_amd import quest as Q
    from dgquest_amd.quest import Qureg

    q64 = Q.load(2)
    L = q64.lib
    L.dgq_last_search.argtypes = [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    q = L.dgq_create_recorder(n)
    q64.apply_circuit(q, circ)
    L.searchQuregTaskPartitions(q, Q.QuESTEnv())  # the search needs no device
    sh, pm = (C.c_int * 4)(), C.c_double()
    assert L.dgq_last_search(q, sh, C.byref(pm)) == 0  # the shape came from a search
    return list(sh), pm.value


def test_shape_search_on_the_benchmark_circuit():
    """searchQuregTaskPartitions (src/cpp/QuEST_backend_distributed.cpp:86-121 picks a task
    partition; here a fused-sweep shape): on the 30-qubit layered circuit the search keeps
    the widest tile with rotating lanes -- the shape the bench's autotuner measures fastest
    (24 sweeps against 30 with fixed lanes and 31 at k = 6) -- and every candidate's
    predicted time is at least the chosen one's."""
    from dgquest_amd import circuits

    c = circuits.layered_u3_cz(30, 24, 2)
    shape, pred = _searched(30, c)
    assert shape == [7, 4, 3, 1]
    assert pred > 0
    # the prediction grows with the plan's sweeps: the same circuit at half the depth
    half, pred_half = _searched(30, circuits.layered_u3_cz(30, 12, 2))
    assert half == [7, 4, 3, 1] and 0.35 * pred < pred_half < 0.65 * pred


def test_shape_search_fits_small_registers():
    """A register narrower than a candidate's tile only considers the shapes that fit it."""
    from dgquest_amd import circuits

    for n in (10, 12, 14):
        shape, pred = _searched(n, circuits.layered_u3_cz(n, 6, 1))
        k, R, W, rot = shape
        assert 6 + R + W <= n and k <= R + W and pred > 0
