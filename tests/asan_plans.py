"""Planner under AddressSanitizer (host code only; run by asan_planner.sh,
not collected by pytest): plans of the workload families at their benchmark
sizes and of random circuits, every built tile shape, lane rotation on and
off, diagonal fusion on and off, through a QuEST library built with
-fsanitize=address in the directory given as argv[1]."""
import sys
import os

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np

from dgquest_amd import circuits
from dgquest_amd.quest import QuEST
from tests import helpers as H

lib = QuEST(2, sys.argv[1])
rng = np.random.default_rng(9)
cases = [(30, circuits.layered_u3_cz(30, 24, 2)), (30, circuits.hea(30, 10, 7)), (30, circuits.ghz_qft(30)),
         (28, circuits.grid_circuit(4, 7, 20, 4)), (24, H.random_circuit(rng, 24, 600)),
         (31, circuits.bernstein_vazirani(30, 3)[0]), (16, H.random_circuit(rng, 16, 300)),
         (14, H.random_circuit(rng, 14, 200))]
shapes = [(7, 4, 3), (6, 3, 3), (5, 4, 1), (4, 4, 0), (3, 3, 0), (6, 4, 2), (7, 5, 2), (2, 2, 1), (1, 1, 0)]
blocks = 0
for n, c in cases:
    for sh in shapes:
        for rot in (0, 1):
            for fuse in (True, False):
                r = lib.dgq_create_recorder(n)
                lib.dgq_set_lane_rotation(r, rot)
                lib.apply_circuit(r, c)
                b, _, _, l2p = lib.recorder_plan(r, *sh, 4096, fuse)
                lib.dgq_destroy_recorder(r)
                assert sorted(l2p) == list(range(n))
                blocks += len(b)
print(f"plans ok ({blocks} blocks)")
