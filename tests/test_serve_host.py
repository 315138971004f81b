"""CPU: the host side of the request server (no GPU).

* the arena allocator against a brute-force page model (allocation succeeds
  exactly when a long enough free run exists, no overlap, coalescing);
* the sample uniforms against the pure-Python MT19937 and the reference's own
  src/mt19937ar.c (compiled into oracle/_ref when the sources are present);
* the recorder export (the QuEST front end's gate records) and the plans
  made from it;
* the serving workload generator;
* configuration validation (dgs_create refuses bad values before touching a
  device)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from dgquest_amd import circuits as CC
from dgquest_amd import serve as S


@pytest.fixture(scope="module")
def q64():
    from dgquest_amd import quest as Q

    return Q.load(2)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_arena_against_page_model(q64, seed):
    assert S._bind(q64.lib).dgs_arena_selftest(seed, 20000) == 0


def test_uniforms_match_mt19937(q64):
    for seed in [0, 5, 77, 0xFFFFFFFF]:
        u = S.uniforms(q64, seed, 1500)
        m = CC.MT19937([seed])
        ref = np.array([m.genrand_res53() for _ in range(1500)])
        assert np.array_equal(u, ref)
        assert (u >= 0).all() and (u < 1).all()


def test_uniforms_match_reference_mt(q64):
    from oracle.binding import REF_MT, ReferenceMT, build

    build(ref=True)
    if not os.path.exists(REF_MT):
        pytest.skip("reference mt19937ar.c not present to compile")
    r = ReferenceMT()
    r.init_by_array([1234567])
    ref = np.array([r.res53() for _ in range(700)])  # crosses one state refill (624 words)
    assert np.array_equal(S.uniforms(q64, 1234567, 700), ref)


def test_recorder_export_round_trip(q64):
    n = 12
    rec = S.record(q64, n, CC.ghz_qft(n))
    # GHZ + QFT: H, n-1 CNOT, n H, n(n-1)/2 controlled phases; the swaps only relabel qubits
    assert len(rec.gates) == 1 + (n - 1) + n + n * (n - 1) // 2
    assert list(rec.l2p) == list(range(n))[::-1]
    assert rec.init_kind == S.DQ_INIT_CLASSICAL and rec.init_index == 0
    # exported records plan exactly as the recording register does
    ints = (C.c_int * 8)()
    lib = q64.lib
    lib.dgq_plan_only.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_int)]
    assert lib.dgq_plan_only(n, 2, rec.gates.ctypes.data, len(rec.gates), 6, 4, 2, 4096, ints) == 0
    q = lib.dgq_create_recorder(n)
    q64.apply_circuit(q, CC.ghz_qft(n))
    b, p, o, l2p = q64.recorder_plan(q, 6, 4, 2, 4096)
    lib.dgq_destroy_recorder(q)
    assert ints[0] == len(b) and ints[1] == len(p)
    assert l2p == list(rec.l2p)


def test_recorder_export_plus_and_classical(q64):
    rec = S.record(q64, 5, [("hadamard", (1,), ())], init="plus")
    assert rec.init_kind == S.DQ_INIT_PLUS
    assert math.isclose(rec.init_value, 1 / math.sqrt(32))
    rec = S.record(q64, 5, [], init=9)
    assert rec.init_kind == S.DQ_INIT_CLASSICAL and rec.init_index == 9 and len(rec.gates) == 0


def test_serving_mix_shape():
    mix = CC.serving_mix(3000, seed=5, rate=50.0, build=False)
    assert mix == CC.serving_mix(3000, seed=5, rate=50.0, build=False)
    ns = np.array([d["n"] for d in mix])
    assert ns.min() >= 10 and ns.max() <= 30
    # log-uniform: P(n) = ln((n+1)/n) / ln(31/10)
    p10 = math.log(11 / 10) / math.log(3.1)
    assert abs((ns == 10).mean() - p10) < 0.02
    kinds = [d["kind"] for d in mix]
    assert abs(kinds.count("hea") / 3000 - 0.4) < 0.03
    assert abs(kinds.count("ghz_qft") / 3000 - 0.3) < 0.03
    hd = [d["depth"] for d in mix if d["kind"] == "hea"]
    ud = [d["depth"] for d in mix if d["kind"] == "u3_cz"]
    assert min(hd) == 4 and max(hd) == 16 and min(ud) == 8 and max(ud) == 24
    t = np.array([d["arrive_s"] for d in mix])
    assert (np.diff(t) > 0).all()
    assert abs(t[-1] / 3000 - 1 / 50.0) < 0.002  # mean gap of the Poisson stream


def test_hea_circuit():
    c = CC.hea(4, 2, 9)
    names = [g[0] for g in c]
    assert names == (["rotateY"] * 4 + ["rotateZ"] * 4 + ["controlledNot"] * 4) * 2
    assert [g[1] for g in c if g[0] == "controlledNot"][:4] == [(0, 1), (1, 2), (2, 3), (3, 0)]


def test_config_validation(q64):
    lib = S._bind(q64.lib)
    for bad in [dict(prec=3), dict(hbm_budget=1000), dict(planner_threads=0), dict(batch_lanes=0, large_lanes=0),
                dict(lane_depth=0), dict(batch_max=0), dict(recheck_fraction=1.5)]:
        c = S.default_config(q64, **bad)
        assert not lib.dgs_create(C.byref(c)), bad
        assert b"config" in lib.dgs_last_error()


def test_tune_table_lookup(q64, tmp_path):
    """Nearest-size lookup (larger n on a tie), per precision; DGQ_TUNE overrides."""
    lib = q64.lib
    lib.dgq_tune_lookup.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.dgq_tune_reload.argtypes = [C.c_char_p]
    p = tmp_path / "t.txt"
    p.write_text("# comment\n2 12 5 4 1 4096 1.0\n2 20 6 4 2 4096 2.0\n2 30 7 4 3 4096 9.0\n"
                 "1 30 6 3 3 1000 4.0\nbad line\n2 16 9 9 9 1 1\n")
    try:
        assert lib.dgq_tune_reload(str(p).encode()) == 4
        out = (C.c_int * 4)()

        def look(prec, n):
            assert lib.dgq_tune_lookup(prec, n, out) == 0
            return tuple(out)

        assert look(2, 10) == (5, 4, 1, 4096)
        assert look(2, 16) == (6, 4, 2, 4096)  # 16 is 4 from 12 and 20: the larger wins
        assert look(2, 26) == (7, 4, 3, 4096)
        assert look(1, 12) == (6, 3, 3, 1000)
        os.environ["DGQ_TUNE"] = "3,3,0,77"
        try:
            assert look(2, 30) == (3, 3, 0, 77)
        finally:
            del os.environ["DGQ_TUNE"]
        assert lib.dgq_tune_reload(str(tmp_path / "missing.txt").encode()) == -1
        assert look(2, 30) == (6, 4, 2, 4096)  # empty table: planner defaults
    finally:
        lib.dgq_tune_reload(None)


def test_both_precisions_in_one_process():
    """Both API libraries loaded in one process (FP64 first): calls between API
    functions inside the FP32 library (fSim -> controlledPhaseShift with a
    float angle, ...) must stay inside it, so both record the same gates."""
    from dgquest_amd import quest as Q

    q64, q32 = Q.load(2), Q.load(1)
    circ = [("fSim", (0, 1), (0.3, 0.2)), ("rotateAroundAxis", (2,), (0.7, 0.1, 0.2, 0.3)),
            ("controlledRotateAroundAxis", (0, 2), (0.4, 0.3, -0.2, 0.5)), ("u3Gate", (1,), (0.1, 0.2, 0.3))]
    r64 = S.record(q64, 3, circ)
    r32 = S.record(q32, 3, circ)
    assert len(r64.gates) == len(r32.gates)
    for a, b in zip(r64.gates, r32.gates):
        assert a["type"] == b["type"] and a["tmask"] == b["tmask"] and a["cmask"] == b["cmask"]
        assert np.abs(a["op"] - b["op"]).max() < 1e-6  # float arguments, double records
