"""The sweep engine's environment overrides against its default path.

DGQ_SPLIT_EXCHANGE=0 forces whole-amplitude exchanges on every tile and =1
split (owned-slot) exchanges on every tile; DGQ_SELECT=dfs restores the
depth-first gate selection.  Each setting is read once per process, so a
child process runs the same circuit under each and writes its state.
Exchanges only move amplitudes, so the exchange modes give bit-identical
states; the selection order regroups the gates (1e-12, the FP64 parity bar;
FP32 states within 1e-5).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from dgquest_amd import circuits
from dgquest_amd import quest as Q
n, prec, k, R, W = (int(x) for x in sys.argv[2:7])
qq = Q.load(prec)
env = qq.createQuESTEnv()
q = qq.createQureg(n, env)
qq.dgq_set_tuning(q, k, R, W, 4096)
qq.apply_circuit(q, circuits.layered_u3_cz(n, 6, 3))
np.save(sys.argv[7], qq.state_vector(q))
print("child done", n, prec, k, R, W, flush=True)
"""

# (n, precision, k, R, W): a 13-bit FP64 tile (split by default), a 10-bit
# FP64 tile (whole amplitudes by default), a 13-bit FP32 tile of 32 registers
CASES = [(20, 2, 7, 4, 3), (20, 2, 4, 4, 0), (18, 1, 7, 5, 2)]


def run_child(case, extra_env, out):
    env = dict(os.environ)
    for k in ("DGQ_SPLIT_EXCHANGE", "DGQ_SELECT"):
        env.pop(k, None)
    env.update(extra_env)
    args = [sys.executable, "-c", CHILD, ROOT] + [str(x) for x in case] + [str(out)]
    print("case", case, extra_env, flush=True)
    r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.load(out)


@pytest.mark.parametrize("case", CASES)
def test_env_overrides_match_default(tmp_path, case):
    base = run_child(case, {}, tmp_path / "base.npy")
    norm = float(np.sum(np.abs(base) ** 2))
    assert abs(norm - 1.0) < (1e-12 if case[1] == 2 else 1e-5)
    for i, (extra, exact) in enumerate(((({"DGQ_SPLIT_EXCHANGE": "0"}), True),
                                        (({"DGQ_SPLIT_EXCHANGE": "1"}), True),
                                        (({"DGQ_SELECT": "dfs"}), False))):
        sv = run_child(case, extra, tmp_path / f"alt{i}.npy")
        if exact:
            assert np.array_equal(sv, base), extra
        else:
            err = float(np.abs(sv - base).max())
            assert err <= (1e-12 if case[1] == 2 else 1e-5), (extra, err)
