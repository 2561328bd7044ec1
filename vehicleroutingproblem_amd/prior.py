"""Learned heuristic prior: Python view of the aco_hip_prior_* calls.

The edge-scoring network (8 features -> 64 -> 64 -> 1, BF16 on the matrix
cores, FP32 Adam) is trained on the GPU on candidate edges labelled by elite
tours the solver itself produces, then installed on an instance as
eta = (1/d)^beta * gain * sigmoid(logit) through the learned-heuristic seam
(aco_hip_set_candidate_eta).  Plain ACO stays the fallback: an instance
without an applied prior keeps (1/d)^beta.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .solver import AcoError, HipSolver

FEATURES = 8
PARAMS = 4801


class EdgePrior:
    def __init__(self, seed: int = 0, device: int = 0):
        self.lib = N.load()
        self._h = C.c_void_p()
        rc = self.lib.aco_hip_prior_create(C.byref(self._h), device, seed & 0xFFFFFFFF)
        if rc:
            raise AcoError(rc, "could not create the edge prior")

    def _check(self, rc: int) -> None:
        if rc:
            msg = self.lib.aco_hip_prior_last_error(self._h)
            raise AcoError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self._h:
            self.lib.aco_hip_prior_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ---------------------------------------------------------
    def params(self) -> np.ndarray:
        out = np.zeros(PARAMS, np.float32)
        self._check(self.lib.aco_hip_prior_get_params(self._h, N.ptr(out, C.c_float)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (PARAMS,)
        self._check(self.lib.aco_hip_prior_set_params(self._h, N.ptr(p, C.c_float)))

    # ---- raw network calls (features as BF16 bits, uint16) --------------------
    def forward(self, xbits) -> np.ndarray:
        x = np.ascontiguousarray(xbits, np.uint16).reshape(-1, FEATURES)
        out = np.zeros(len(x), np.float32)
        self._check(self.lib.aco_hip_prior_forward(self._h, N.ptr(x, C.c_uint16), len(x),
                                                   N.ptr(out, C.c_float)))
        return out

    def gradient(self, xbits, labels):
        x = np.ascontiguousarray(xbits, np.uint16).reshape(-1, FEATURES)
        y = np.ascontiguousarray(labels, np.uint8)
        assert len(y) == len(x)
        g = np.zeros(PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.lib.aco_hip_prior_gradient(self._h, N.ptr(x, C.c_uint16),
                                                    N.ptr(y, C.c_uint8), len(x),
                                                    N.ptr(g, C.c_float), C.byref(loss)))
        return float(loss.value), g

    # ---- training set and training ------------------------------------------
    def add_instance(self, solver: HipSolver, tour, instance: int = 0) -> None:
        t = np.ascontiguousarray(tour, np.int32)
        self._check(self.lib.aco_hip_prior_add_instance(self._h, solver._ctx, instance,
                                                        N.ptr(t, C.c_int32), len(t)))

    def clear(self) -> None:
        self._check(self.lib.aco_hip_prior_clear(self._h))

    def size(self) -> int:
        return int(self.lib.aco_hip_prior_size(self._h))

    def train(self, steps: int, lr: float = 3e-3) -> float:
        loss = C.c_float()
        self._check(self.lib.aco_hip_prior_train(self._h, steps, lr, C.byref(loss)))
        return float(loss.value)

    def apply(self, solver: HipSolver, instance: int = 0, gain: float = 2.0) -> None:
        self._check(self.lib.aco_hip_prior_apply(self._h, solver._ctx, instance, gain))


def edge_features(solver: HipSolver, instance: int = 0) -> np.ndarray:
    """BF16 feature bits [(n+1)*k, 8] of an instance's candidate slots."""
    inst = solver.instances[instance]
    k = solver.cand_k(instance)
    out = np.zeros(((inst.n + 1) * k, FEATURES), np.uint16)
    solver._check(solver.lib.aco_hip_edge_features(solver._ctx, instance,
                                                   N.ptr(out, C.c_uint16)))
    return out


def training_instances(count: int, nmin: int = 200, nmax: int = 2000, seed: int = 0):
    """Seeded training instances: n ~ U[nmin, nmax], even indices from the
    uniform generator, odd from the clustered one (25 Gaussian components,
    sigma 30, in [0,1000]^2), K = ceil(n/125), demands U[1,10],
    Q = ceil(1.1 * sum / K)."""
    import math
    from .solver import Instance
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        K = max(1, math.ceil(n / 125))
        s = (seed * 100003 + i) & 0x7FFFFFFF
        if i % 2 == 0:
            out.append(Instance.uniform(n, K, side=1000, dmin=1, dmax=10, slack_pct=110, seed=s))
        else:
            out.append(Instance.clustered(n, K, side=1000, clusters=25, sigma=30, dmin=1, dmax=10,
                                          slack_pct=110, seed=s))
    return out


def train_prior(instances, epochs: int = 100, steps: int = 600, lr: float = 2e-3,
                cand_k: int = 16, ants: int = 64, batch: int = 64, seed: int = 1,
                device: int = 0) -> "EdgePrior":
    """Harvest elite tours with plain ACO (`batch` instances per context, one
    colony each, `epochs` epochs) and train a prior on them on the GPU."""
    from .solver import Params
    prior = EdgePrior(seed=seed, device=device)
    for b in range(0, len(instances), batch):
        part = instances[b:b + batch]
        with HipSolver(part, Params(cand_k=cand_k, ants=ants, rho=0.1),
                       seeds=[(seed + b + i) & 0xFFFFFFFF for i in range(len(part))],
                       device=device) as s:
            s.prepare()
            s.run(fixed_epochs=epochs, chunk_epochs=min(epochs, 32))
            for i in range(len(part)):
                try:
                    tour, _ = s.best(i)
                except AcoError:  # no feasible tour found: nothing to learn from
                    continue
                prior.add_instance(s, tour, instance=i)
    prior.train(steps, lr=lr)
    return prior
