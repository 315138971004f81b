"""Generate the golden fixtures of tests/golden/ (run from the repo root).

Everything here is computed independently of oracle/oracle.c:
  * mt19937_kat.json  -- MT19937 streams from the reference's own
    src/mt19937ar.c, compiled by oracle/Makefile into oracle/_ref/
  * circuits.json + states.npz -- small circuits over the whole gate set and
    their output states from a dense numpy brute force (2^n x 2^n operators
    built by Kronecker products, tests/helpers.py)
  * measure.json -- seeded measurement outcomes: the reference MT stream
    applied to P(0) of the dense states with the outcome rule
    genrand_real1() > P(0) (src/QuEST_common.c:97-115)
"""
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.binding import ReferenceMT, build  # noqa: E402
from tests import helpers as H  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def mt_kat():
    ref = ReferenceMT()
    seeds = [[0x123, 0x234, 0x345, 0x456], [5489], [2], [0xBEEF, 1], [123456789, 987654321, 42]]
    out = []
    for keys in seeds:
        ref.init_by_array(keys)
        ints = [int(ref.int32()) for _ in range(32)]
        reals = [float(ref.real1()) for _ in range(16)]
        res53 = [float(ref.res53()) for _ in range(16)]
        out.append({"keys": keys, "int32": ints, "real1": reals, "res53": res53})
    # long stream for one seed (position 1000..1009 crosses the twist boundary)
    ref.init_by_array([0x123, 0x234, 0x345, 0x456])
    stream = [int(ref.int32()) for _ in range(1010)]
    return {"seeds": out, "long": {"keys": [0x123, 0x234, 0x345, 0x456], "from": 1000, "int32": stream[1000:]}}


def circuits_and_states():
    rng = np.random.default_rng(20261202)
    cases = []
    # every gate alone on a random state, n = 4
    names = sorted(H.GATE_SIGNATURES if hasattr(H, "GATE_SIGNATURES") else [])
    from dgquest_amd.quest import GATE_SIGNATURES

    for name in sorted(GATE_SIGNATURES):
        for trial in range(2):
            c = H.random_circuit(rng, 4, 1, allow={name})
            cases.append({"name": f"{name}_{trial}", "n": 4, "init": "random", "circuit": c})
    for n in (3, 5, 7):
        for trial in range(3):
            cases.append({"name": f"random_n{n}_{trial}", "n": n, "init": "zero",
                          "circuit": H.random_circuit(rng, n, 40)})
    from dgquest_amd import circuits

    cases.append({"name": "ghz_qft_8", "n": 8, "init": "zero", "circuit": circuits.ghz_qft(8)})
    cases.append({"name": "layered_u3_cz_6", "n": 6, "init": "zero", "circuit": circuits.layered_u3_cz(6, 4, 2)})
    cases.append({"name": "grid_2x3", "n": 6, "init": "zero", "circuit": circuits.grid_random(2, 3, 6, 4)})
    bv, s = circuits.bernstein_vazirani(5, 3)
    cases.append({"name": "bv_5", "n": 6, "init": "zero", "circuit": bv, "hidden": s})
    states_in, states_out = {}, {}
    for i, cs in enumerate(cases):
        n = cs["n"]
        if cs["init"] == "random":
            v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
            v /= np.linalg.norm(v)
        else:
            v = np.zeros(1 << n, dtype=complex)
            v[0] = 1
        states_in[f"in_{i}"] = v
        states_out[f"out_{i}"] = H.dense_run(n, cs["circuit"], v.copy())
        cs["circuit"] = [[g[0], list(g[1]), list(g[2])] for g in cs["circuit"]]
    return cases, states_in, states_out


def measurements(cases, states_out):
    ref = ReferenceMT()
    out = []
    for i, cs in enumerate(cases):
        if not cs["name"].startswith("random_n"):
            continue
        n = cs["n"]
        v = states_out[f"out_{i}"].copy()
        keys = [i + 1, 0xBEEF]
        ref.init_by_array(keys)
        order = list(range(n))[::-1]
        outcomes = []
        for q in order:
            idx = np.arange(1 << n)
            p0 = float(np.sum(np.abs(v[((idx >> q) & 1) == 0]) ** 2))
            if p0 < 1e-13:
                o = 1
            elif 1 - p0 < 1e-13:
                o = 0
            else:
                o = int(ref.real1() > p0)
            prob = p0 if o == 0 else 1 - p0
            mask = ((idx >> q) & 1) == o
            v = np.where(mask, v / math.sqrt(prob), 0)
            outcomes.append(o)
        out.append({"case": i, "keys": keys, "order": order, "outcomes": outcomes, "final": f"out_{i}"})
    return out


def main():
    build(ref=True)
    kat = mt_kat()
    with open(os.path.join(OUT, "mt19937_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    cases, sin, sout = circuits_and_states()
    with open(os.path.join(OUT, "circuits.json"), "w") as f:
        json.dump(cases, f)
    np.savez_compressed(os.path.join(OUT, "states.npz"), **sin, **sout)
    meas = measurements(cases, sout)
    with open(os.path.join(OUT, "measure.json"), "w") as f:
        json.dump(meas, f)
    print(f"wrote {len(cases)} circuits, {len(meas)} measurement sequences")


if __name__ == "__main__":
    main()
