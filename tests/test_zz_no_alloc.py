"""The allocation-free route-cost entry point against aco_hip_route_costs,
through the C ABI: the same int64 costs (bit-exact) on two calls in a row,
the second smaller than the first (the array keeps the first call's
contents), and a refusal of a batch larger than its staging array."""
import numpy as np
import pytest

from vehicleroutingproblem_amd import AcoError, HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 4 << 20  # ACO_HIP_NOALLOC_SCRATCH_BYTES


def test_route_costs_noalloc_matches_allocating_entry(gpu):
    inst = Instance.clustered(300, 10, side=10000, sigma=300, dmin=1, dmax=100, slack_pct=105,
                              seed=7)
    with HipSolver(inst, Params(cand_k=32, ants=96, rho=0.1), seeds=[5]) as s:
        s.prepare()
        s.epochs(2)
        tours, lens, cost, _ = s.tours()
        # call 1: every ant's tour of the last epoch (96 tours, stride 320)
        a = s.route_costs(tours, lens)
        b = s.route_costs_noalloc(tours, lens)
        assert np.array_equal(a, b) and np.array_equal(b, cost)
        # call 2: a smaller batch with another stride: random permutations
        # and a tour shorter than its row
        rng = np.random.default_rng(11)
        perms = np.zeros((5, inst.n + 2), np.int32)
        for r in range(5):
            perms[r, 1:-1] = rng.permutation(np.arange(1, inst.n + 1))
        ln = np.full(5, inst.n + 2, np.int32)
        ln[3] = 7
        perms[3, 6] = 0
        a = s.route_costs(perms, ln)
        b = s.route_costs_noalloc(perms, ln)
        assert np.array_equal(a, b)
        assert b[3] == s.route_costs(perms[3:4, :7], ln[3:4])[0]
        # a batch above the staging array is refused before any device work
        big_count = SCRATCH_BYTES // (4 * 1024) + 1
        big = np.zeros((big_count, 1024), np.int32)
        with pytest.raises(AcoError) as err:
            s.route_costs_noalloc(big, np.full(big_count, 2, np.int32))
        assert err.value.status == N.ACO_HIP_ERR_INVALID_INPUT
        # the context still works after the refusal
        assert np.array_equal(s.route_costs_noalloc(tours[:3], lens[:3]), cost[:3])
