"""GPU: sampling, the many-state read-out and the request server.

Bars:
* samples of an FP64 state equal the oracle's inverse-CDF index (numpy
  cumulative sum of |a|^2, first index whose sum exceeds u * total) for every
  sample whose target lies more than 1e-12 from a CDF boundary;
* FP64 serving: P(q=0) per qubit within 1e-12 of the oracle's state, samples
  as above;
* FP32 serving: every result rechecked in FP64 on the device, fidelity
  >= 1 - 1e-6, sum <Z> within 2e-5 per qubit of the oracle, FP32 samples
  consistent with the oracle's CDF within 1e-5;
* a request's result is bit-identical alone and inside a mixed batch;
* requests that do not fit the HBM budget wait (all complete, the arena never
  exceeds the budget); a request larger than the whole budget fails alone.
"""
import math

import numpy as np
import pytest

from dgquest_amd import circuits as CC
from dgquest_amd import serve as S

pytestmark = pytest.mark.gpu


def oracle_state(oracle, n, circ):
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return oracle.run(s, circ)


def p0_of(state, n):
    p = np.abs(state) ** 2
    idx = np.arange(len(p))
    return np.array([p[((idx >> q) & 1) == 0].sum() for q in range(n)] + [p.sum()])


def cdf_pick(state, u):
    """Oracle samples: first index whose cumulative probability exceeds u * total,
    plus the distance of the target to the nearest CDF boundary."""
    cdf = np.cumsum(np.abs(state) ** 2)
    t = np.asarray(u) * cdf[-1]
    idx = np.searchsorted(cdf, t, side="right")
    idx = np.minimum(idx, len(cdf) - 1)
    lo = np.where(idx > 0, cdf[np.maximum(idx - 1, 0)], 0.0)
    gap = np.minimum(np.abs(t - lo), np.abs(cdf[idx] - t))
    return idx, gap, cdf, t


def check_exact_samples(got, state, u, tol=1e-12):
    idx, gap, _, _ = cdf_pick(state, u)
    clear = gap > tol
    assert clear.mean() > 0.95
    bad = np.nonzero((got != idx) & clear)[0]
    assert len(bad) == 0, f"{len(bad)} samples differ, first at {bad[:5]}: {got[bad[:5]]} vs {idx[bad[:5]]}"


def check_consistent_samples(got, state, u, delta):
    """got[s] must be an index whose CDF interval contains u*total within delta."""
    cdf = np.cumsum(np.abs(state) ** 2)
    t = np.asarray(u) * cdf[-1]
    g = got.astype(np.int64)
    lo = np.where(g > 0, cdf[np.maximum(g - 1, 0)], 0.0)
    hi = cdf[g]
    assert ((lo - delta <= t) & (t <= hi + delta)).all()


def random_state(rng, n, zero_frac=0.3):
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a[rng.random(1 << n) < zero_frac] = 0.0
    if not a.any():
        a[0] = 1.0
    return a / np.linalg.norm(a)


@pytest.mark.parametrize("n", [1, 4, 10, 11, 15, 21])
def test_sample_fp64(quest, env, oracle, n):
    rng = np.random.default_rng(100 + n)
    st = random_state(rng, n)
    import ctypes as C

    q = quest.createQureg(n, env)
    re = np.ascontiguousarray(st.real)
    im = np.ascontiguousarray(st.imag)
    quest.initStateFromAmps(q, re.ctypes.data_as(C.POINTER(C.c_double)), im.ctypes.data_as(C.POINTER(C.c_double)))
    # a relabelling swap and a gate after it: samples come back in logical order
    circ = [("swapGate", (0, n - 1), ())] if n > 1 else []
    circ += [("hadamard", (n // 2,), ())]
    quest.apply_circuit(q, circ)
    ref = oracle.run(st.copy(), circ)
    u = np.concatenate([[0.0, 1.0 - 2.0 ** -53, 0.5], rng.random(1021)])
    got = S.sample(quest, q, u)
    print(f"n={n} unique samples {len(np.unique(got))}", flush=True)
    check_exact_samples(got, ref, u)
    # zero-probability states are never drawn
    assert (np.abs(ref[got.astype(np.int64)]) > 0).all()
    quest.destroyQureg(q, env)


def test_sample_fp32(quest32, env32, oracle):
    import ctypes as C

    n = 14
    rng = np.random.default_rng(5)
    circ = CC.hea(n, 6, 11)
    q = quest32.createQureg(n, env32)
    quest32.apply_circuit(q, circ)
    ref = oracle_state(oracle, n, circ)
    u = rng.random(2048)
    got = S.sample(quest32, q, u)
    check_consistent_samples(got, ref, u, 1e-5)
    quest32.destroyQureg(q, env32)


def make_requests(quest, mix, shots=1024, flags=None):
    recs = [S.record(quest, d["n"], d["circuit"]) for d in mix]
    return S.RequestSet(recs, [shots] * len(mix), [d["shot_seed"] for d in mix], flags)


def small_mix(count, seed, n_max):
    return CC.serving_mix(count, seed=seed, rate=1000.0, n_max=n_max)


def test_serve_fp64_matches_oracle(quest, oracle):
    mix = small_mix(14, 21, 16)
    rs = make_requests(quest, mix)
    with S.Server(quest, prec=S.DQ_F64, hbm_budget=2 << 30) as srv:
        out = srv.replay(rs)
        st = srv.stats()
    print(st, flush=True)
    for d, r in zip(mix, out):
        assert r["status"] == 0
        ref = oracle_state(oracle, d["n"], d["circuit"])
        p0 = p0_of(ref, d["n"])
        assert np.abs(r["p0"] - p0).max() <= 1e-12, d["kind"]
        assert abs(r["sum_z"] - (2 * p0[:-1] - 1).sum()) <= 1e-11
        u = S.uniforms(quest, d["shot_seed"], 1024)
        check_exact_samples(r["samples"], ref, u)


def test_serve_fp32_rechecked(quest, oracle):
    mix = small_mix(12, 22, 16)
    rs = make_requests(quest, mix)
    with S.Server(quest, prec=S.DQ_F32, hbm_budget=2 << 30, recheck_fraction=1.0) as srv:
        out = srv.replay(rs)
        st = srv.stats()
    assert st["rechecked"] == len(mix)
    for d, r in zip(mix, out):
        assert r["status"] == 0 and r["rechecked"] == 1
        n = d["n"]
        ref = oracle_state(oracle, n, d["circuit"])
        p0 = p0_of(ref, n)
        sz = (2 * p0[:-1] - 1).sum()
        print(f"{d['kind']} n={n} fid={r['fidelity']:.9f} maxerr={r['max_abs_err']:.2e} "
              f"dz32={abs(r['sum_z'] - sz):.2e} dz64={abs(r['sum_z_fp64'] - sz):.2e} mism={r['shot_mismatches']}",
              flush=True)
        assert r["fidelity"] >= 1 - 1e-6
        assert r["max_abs_err"] <= 1e-5
        assert abs(r["sum_z_fp64"] - sz) <= 1e-10
        assert abs(r["sum_z"] - sz) <= 2e-5 * n
        u = S.uniforms(quest, d["shot_seed"], 1024)
        check_consistent_samples(r["samples"], ref, u, 1e-5)


def test_alone_equals_mix(quest):
    mix = small_mix(40, 23, 15) + [dict(d, n=18, circuit=CC.hea(18, 5, 7)) for d in small_mix(2, 24, 10)]
    rs = make_requests(quest, mix, shots=256)
    with S.Server(quest, prec=S.DQ_F32, hbm_budget=4 << 30, recheck_fraction=0.0) as srv:
        together = srv.replay(rs)
        alone = [srv.replay(S.RequestSet([rs.recs[i]], [256], [mix[i]["shot_seed"]]))[0] for i in range(len(mix))]
        st = srv.stats()
    print(st, flush=True)
    assert max(r["batch_size"] for r in together) > 1
    for a, b in zip(together, alone):
        assert a["status"] == 0 and b["status"] == 0 and b["batch_size"] == 1
        assert np.array_equal(a["p0"], b["p0"])
        assert a["sum_z"] == b["sum_z"]
        assert np.array_equal(a["samples"], b["samples"])


def test_admission_waits_for_memory(quest, oracle):
    mix = [dict(kind="hea", n=18, depth=4, circuit_seed=s, shot_seed=s, arrive_s=0.0) for s in range(10)]
    for d in mix:
        d["circuit"] = CC.hea(18, 4, d["circuit_seed"])
    big = dict(kind="ghz_qft", n=22, depth=0, circuit_seed=0, shot_seed=1, arrive_s=0.0, circuit=CC.ghz_qft(22))
    rs = make_requests(quest, mix[:5] + [big] + mix[5:], shots=64)
    budget = 12 << 20
    with S.Server(quest, prec=S.DQ_F32, hbm_budget=budget, recheck_fraction=0.0) as srv:
        out = srv.replay(rs)
        st = srv.stats()
    print(st, flush=True)
    assert out[5]["status"] != 0  # 32 MiB state > 12 MiB budget: refused alone
    rest = out[:5] + out[6:]
    assert all(r["status"] == 0 for r in rest)
    assert st["peak_arena_bytes"] <= budget
    ref = oracle_state(oracle, 18, mix[0]["circuit"])
    sz = (2 * p0_of(ref, 18)[:-1] - 1).sum()
    assert abs(rest[0]["sum_z"] - sz) <= 2e-5 * 18


def test_plain_loop_matches_oracle(quest, oracle):
    mix = small_mix(5, 25, 15)
    rs = make_requests(quest, mix)
    out = S.run_plain(quest, rs)
    for d, r in zip(mix, out):
        assert r["status"] == 0
        ref = oracle_state(oracle, d["n"], d["circuit"])
        assert np.abs(r["p0"] - p0_of(ref, d["n"])).max() <= 1e-12
        check_exact_samples(r["samples"], ref, S.uniforms(quest, d["shot_seed"], 1024))


def test_serve_fp32_rechecked_large(quest):
    """FP32 serving of 24-25-qubit requests, every one rechecked in FP64 on
    the device: fidelity |<a|b>|^2 / (<a|a><b|b>) >= 1 - 1e-6 (a = FP64 run,
    b = served state), the raw overlap consistent with it and the two norms,
    and for GHZ + QFT sum <Z> against the closed form (FP64 run within 1e-10,
    served FP32 within 1e-6 per qubit: P(q=0) comes from the normalised
    state)."""
    mix = [dict(kind="ghz_qft", n=25, circuit=CC.ghz_qft(25), shot_seed=1),
           dict(kind="ghz_qft", n=24, circuit=CC.ghz_qft(24), shot_seed=2),
           dict(kind="hea", n=24, circuit=CC.hea(24, 8, 3), shot_seed=3),
           dict(kind="u3_cz", n=25, circuit=CC.layered_u3_cz(25, 12, 4), shot_seed=4)]
    rs = make_requests(quest, mix, shots=256, flags=[S.DGS_RECHECK] * len(mix))
    with S.Server(quest, prec=S.DQ_F32, hbm_budget=8 << 30) as srv:
        out = srv.replay(rs)
    for d, r in zip(mix, out):
        n = d["n"]
        print(f"{d['kind']} n={n} fid=1-{1 - r['fidelity']:.3e} raw=1{r['fidelity_raw'] - 1:+.3e} "
              f"tot32=1{r['total_prob'] - 1:+.3e} tot64=1{r['norm_fp64'] - 1:+.3e} "
              f"maxerr={r['max_abs_err']:.2e} dz={abs(r['sum_z'] - r['sum_z_fp64']):.2e}", flush=True)
        assert r["status"] == 0 and r["rechecked"] == 1
        assert r["fidelity"] >= 1 - 1e-6
        assert abs(r["norm_fp64"] - 1.0) <= 1e-12
        assert abs(r["fidelity_raw"] - r["fidelity"] * r["norm_fp64"] * r["total_prob"]) <= 1e-12
        assert r["max_abs_err"] <= 1e-5
        assert abs(r["sum_z"] - r["sum_z_fp64"]) <= 1e-6 * n
        if d["kind"] == "ghz_qft":
            p = np.abs(CC.ghz_qft_closed_form(n)) ** 2
            idx = np.arange(len(p))
            sz = sum(2 * p[((idx >> q) & 1) == 0].sum() - 1 for q in range(n))
            del p, idx
            assert abs(r["sum_z_fp64"] - sz) <= 1e-10
            assert abs(r["sum_z"] - sz) <= 1e-6 * n


def test_serve_without_large_lanes(quest, oracle):
    """Batch lanes only: registers above batch_max_qubits and requests
    rechecked in FP64 (neither batchable) still run, one per launch group,
    and every request completes."""
    mix = small_mix(6, 26, 12)
    mix.append(dict(kind="hea", n=17, depth=4, circuit_seed=9, shot_seed=9, arrive_s=0.0, circuit=CC.hea(17, 4, 9)))
    flags = [0, S.DGS_RECHECK, 0, 0, S.DGS_RECHECK, 0, 0]
    rs = make_requests(quest, mix, shots=64, flags=flags)
    with S.Server(quest, prec=S.DQ_F32, hbm_budget=1 << 30, batch_lanes=1, large_lanes=0,
                  batch_max_qubits=14) as srv:
        out = srv.replay(rs)
        st = srv.stats()
    print(st, flush=True)
    assert st["rechecked"] == 2
    for d, r, f in zip(mix, out, flags):
        assert r["status"] == 0
        assert r["rechecked"] == (1 if f == S.DGS_RECHECK else 0)
        ref = oracle_state(oracle, d["n"], d["circuit"])
        sz = (2 * p0_of(ref, d["n"])[:-1] - 1).sum()
        assert abs(r["sum_z"] - sz) <= 2e-5 * d["n"]
