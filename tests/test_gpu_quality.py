"""Solution quality against the reference's sequential solver at equal epochs.

The sequential CLI (aco_sequential.c) ran each seeded unit-demand instance
to its 50-epoch stagnation stop in the build container (tests/golden/
make_golden.py, which builds it from the reference sources); its best cost
and epoch count are the fixture tests/golden/ref_seq_quality.json, so no
reference binary runs here.  This solver runs exactly as many epochs with the
same alpha, beta, rho, tau0, Q and ant count, optimising the same metric
(real-distance mode: unrounded lengths, like the reference), and its best
tour is costed in real (double) distances like the reference's
(solution.c:126 computes the printed cost that way).  The claim tested is one-sided: every tour is
feasible and the mean best-cost difference is at most +0.5%, i.e. this solver
is not worse.  The two constructions differ on purpose (the sequential
reference fills each route to its customer cap with a full roulette; this
solver follows the CUDA path's candidate roulette with a depot-close
decision, aco_cuda_kernels.cu:447-590), so the costs are not expected to be
equal.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "ref_seq_quality.json")


def _real_cost(inst, tour):
    x = inst.x.astype(np.float64)[tour]
    y = inst.y.astype(np.float64)[tour]
    return float(np.hypot(np.diff(x), np.diff(y)).sum())


def _feasible(inst, tour):
    t = np.asarray(tour)
    if t[0] != 0 or t[-1] != 0:
        return False
    cust = np.sort(t[t != 0])
    if len(cust) != inst.n or (cust != np.arange(1, inst.n + 1)).any():
        return False
    routes, load = 0, 0
    for a, b in zip(t[:-1], t[1:]):
        if b == 0:
            routes += a != 0
            load = 0
        else:
            load += int(inst.demand[b])
            if load > inst.capacity:
                return False
    return routes <= inst.K


def _digest(inst):
    h = hashlib.sha1()
    for a in (inst.x, inst.y, inst.demand):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def test_quality_not_worse_than_sequential_reference(gpu):
    doc = json.load(open(FIXTURE))
    runs = [r for r in doc["runs"] if r["set"] == "test"]
    assert len(runs) == 8
    ants = doc["ants"]
    diffs = []
    for r in runs:
        inst = Instance.uniform(100, 8, side=1000, dmin=1, dmax=1, slack_pct=120,
                                seed=r["instance_seed"])
        assert _digest(inst) == r["instance_sha1"]  # the instance the reference solved
        ref_cost, epochs, i = r["best_cost"], r["epochs"], r["instance_seed"]
        with HipSolver(inst, Params(cand_k=32, ants=ants, alpha=1.0, beta=2.0, rho=0.5,
                                    tau0=1.0, Q=1.0, dist_mode=N.DIST_REAL),
                       seeds=[r["solver_seed"]]) as s:
            s.prepare()
            s.run(fixed_epochs=epochs, chunk_epochs=16)
            assert s.colony_state()["epoch"] == epochs
            tour, cost = s.best()
        assert _feasible(inst, tour), f"seed {i}: infeasible tour"
        real = _real_cost(inst, tour)
        # the solver's own cost: the exact sum of the float edge lengths
        assert abs(cost * N.REAL_COST_UNIT - real) <= len(tour) * real * 2.0 ** -24
        diffs.append((real - ref_cost) / ref_cost)
    assert float(np.mean(diffs)) <= 0.005, diffs
