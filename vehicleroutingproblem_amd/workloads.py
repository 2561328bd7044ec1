"""The named workloads of the benchmark, shared by bench.py and the parity
tests so that both run exactly the same instances, seeds and shapes.

* c1: plain ACO, n=100, K=8, 64 ants, k=15, rho 0.1, seed 1234.
* c2: the reference's benchmark shape, n=1000, K=8, 256 ants, k=20, rho 0.1,
  uniform [0,1000]^2, demands U[1,10], Q=ceil(1.1 sum/8); seed s sets both
  the instance and the solver seed; the workload runs seeds 1..32.
* c3: n=5000 clustered (25 Gaussian components, centres U[0,10000]^2,
  sigma 300, depot at the centre), demands U[1,100], K=40,
  Q=ceil(1.05 sum/40), 1024 ants, k=32.
* c4: 512 instances n ~ U[50,2000] (half uniform, half clustered sigma 30),
  K=ceil(n/125), demands U[1,10], Q=ceil(1.1 sum/K), 128 ants each,
  candidate lengths from the tuned per-size-class table, solver seeds
  9000+i.
* c5: test instances of the learned prior, n=2000, K=16, 256 ants, k=16.
"""
from __future__ import annotations

import json
import os

from .solver import Instance, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_K_PATH = os.path.join(_HERE, "tuned_k.json")

C1_PARAMS = dict(cand_k=15, ants=64, rho=0.1)
C2_N, C2_K, C2_ANTS, C2_CAND_K, C2_RHO = 1000, 8, 256, 20, 0.1
C2_SEEDS = 32
C3_ANTS, C3_CAND_K = 1024, 32
C4_COUNT, C4_ANTS, C4_SEED, C4_SOLVER_SEED = 512, 128, 4242, 9000


def c1_instance() -> Instance:
    return Instance.uniform(100, 8, side=1000, dmin=1, dmax=10, slack_pct=110, seed=1234)


def c2_instance(seed: int = 1) -> Instance:
    return Instance.uniform(C2_N, C2_K, side=1000, dmin=1, dmax=10, slack_pct=110, seed=seed)


def c2_params(**over) -> Params:
    kw = dict(cand_k=C2_CAND_K, ants=C2_ANTS, rho=C2_RHO)
    kw.update(over)
    return Params(**kw)


def c3_instance() -> Instance:
    return Instance.clustered(5000, 40, side=10000, clusters=25, sigma=300, dmin=1, dmax=100,
                              slack_pct=105, seed=1)


def c3_params(**over) -> Params:
    kw = dict(cand_k=C3_CAND_K, ants=C3_ANTS, rho=0.1)
    kw.update(over)
    return Params(**kw)


def tuned_k_table(path: str = TUNED_K_PATH) -> dict:
    """Candidate length per size class (upper node count of the class)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(c): int(k) for c, k in json.load(f)["k_by_size_class"].items()}


def c4_instances(count: int = C4_COUNT):
    from .prior import training_instances
    return training_instances(count, nmin=50, nmax=2000, seed=C4_SEED)


def c4_cand_k(insts, table: dict | None = None):
    from . import tuner
    return tuner.candidate_lengths(insts, tuned_k_table() if table is None else table,
                                   default=16)


def c4_seeds(count: int = C4_COUNT):
    return [C4_SOLVER_SEED + i for i in range(count)]


def c4_params(**over) -> Params:
    kw = dict(cand_k=16, ants=C4_ANTS, rho=0.1)
    kw.update(over)
    return Params(**kw)


def c5_instance(t: int) -> Instance:
    if t % 2 == 0:
        return Instance.uniform(2000, 16, side=1000, dmin=1, dmax=10, slack_pct=110, seed=500 + t)
    return Instance.clustered(2000, 16, side=1000, clusters=25, sigma=30, dmin=1, dmax=10,
                              slack_pct=110, seed=500 + t)
