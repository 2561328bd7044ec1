"""Python view of the MI355X ACO-CVRP solver (a thin layer over the C ABI).

Mirrors the reference's solver interface: ``aco_vrp_with_capacity``
(include/aco.h:127-131) becomes :func:`solve`, the kernel launchers of
include/aco_cuda_kernels.h:151-181 become the methods of :class:`HipSolver`.
Every compute call goes to the gfx950 kernels; nothing here computes a tour.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


class AcoError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status


@dataclass
class Instance:
    """A CVRP instance: node 0 is the depot, integer coordinates in [0, 32767]."""

    x: np.ndarray
    y: np.ndarray
    demand: np.ndarray
    K: int
    capacity: int
    name: str = "instance"

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.int32)
        self.y = np.ascontiguousarray(self.y, dtype=np.int32)
        self.demand = np.ascontiguousarray(self.demand, dtype=np.int32)

    @property
    def n(self) -> int:
        return int(self.x.shape[0]) - 1

    @property
    def total_demand(self) -> int:
        return int(self.demand[1:].sum())

    # ---- construction -------------------------------------------------------
    @staticmethod
    def _from_c(v: N.VrpInstance) -> "Instance":
        n = v.n
        x = np.ctypeslib.as_array(v.x, shape=(n + 1,)).copy()
        y = np.ctypeslib.as_array(v.y, shape=(n + 1,)).copy()
        d = np.ctypeslib.as_array(v.demand, shape=(n + 1,)).copy()
        return Instance(x, y, d, v.K, v.capacity, v.name.decode() or "instance")

    def _to_c(self) -> N.VrpInstance:
        v = N.VrpInstance()
        v.n, v.K, v.capacity = self.n, self.K, self.capacity
        v.x, v.y, v.demand = N.ptr(self.x, C.c_int32), N.ptr(self.y, C.c_int32), N.ptr(
            self.demand, C.c_int32)
        v.unit_demand = int(np.all(self.demand[1:] == 1))
        v.name = self.name.encode()[:127]
        return v

    @staticmethod
    def read(path: str, require_unit_demand: bool = False) -> "Instance":
        lib = N.load()
        v = N.VrpInstance()
        if lib.vrp_read(path.encode(), C.byref(v), int(require_unit_demand)) != 0:
            raise AcoError(N.ACO_HIP_ERR_INVALID_INPUT, f"cannot read {path}")
        try:
            return Instance._from_c(v)
        finally:
            lib.vrp_free(C.byref(v))

    def write(self, path: str) -> None:
        if N.load().vrp_write(path.encode(), C.byref(self._to_c())) != 0:
            raise AcoError(N.ACO_HIP_ERR_INVALID_INPUT, f"cannot write {path}")

    @staticmethod
    def uniform(n: int, K: int, side: int = 1000, dmin: int = 1, dmax: int = 10,
                slack_pct: int = 110, seed: int = 1234) -> "Instance":
        """Depot at the centre, customers U{0..side}^2, demands U{dmin..dmax},
        Q = ceil(slack_pct/100 * sum(d) / K)."""
        lib = N.load()
        v = N.VrpInstance()
        rc = lib.vrp_generate_uniform(C.byref(v), n, K, side, dmin, dmax, slack_pct, seed)
        if rc:
            raise AcoError(N.ACO_HIP_ERR_INVALID_INPUT, "uniform generator rejected its arguments")
        try:
            return Instance._from_c(v)
        finally:
            lib.vrp_free(C.byref(v))

    @staticmethod
    def clustered(n: int, K: int, side: int = 10000, clusters: int = 25, sigma: float = 300.0,
                  dmin: int = 1, dmax: int = 100, slack_pct: int = 105,
                  seed: int = 1234) -> "Instance":
        """Gaussian-mixture customers (centres U[0,side]^2, sigma per axis, clipped),
        depot at the box centre."""
        lib = N.load()
        v = N.VrpInstance()
        rc = lib.vrp_generate_clustered(C.byref(v), n, K, side, clusters, sigma, dmin, dmax,
                                        slack_pct, seed)
        if rc:
            raise AcoError(N.ACO_HIP_ERR_INVALID_INPUT,
                           "clustered generator rejected its arguments")
        try:
            return Instance._from_c(v)
        finally:
            lib.vrp_free(C.byref(v))


@dataclass
class Params:
    """Solver parameters; defaults are the reference CLI's except cand_k/rho."""

    cand_k: int = 20
    ants: int = 256
    colonies: int = 1
    alpha: float = 1.0
    beta: float = 2.0
    rho: float = 0.1
    tau0: float = 1.0
    Q: float = 1.0
    depot_close_weight: float = 2.0
    knn_key: int = N.KNN_EXACT
    ants_per_group: int = 0
    stage_lds: int = -1
    fallback_matrix: int = -1
    dist_mode: int = N.DIST_NINT

    def to_c(self) -> N.AcoHipParams:
        p = N.AcoHipParams()
        for f in ("cand_k", "ants", "colonies", "alpha", "beta", "rho", "tau0", "Q",
                  "depot_close_weight", "knn_key", "ants_per_group", "stage_lds",
                  "fallback_matrix", "dist_mode"):
            setattr(p, f, getattr(self, f))
        return p


def _state_dict(s: N.AcoHipColonyState) -> dict:
    return {f: getattr(s, f) for f, _ in s._fields_}


class HipSolver:
    """One context of the C ABI: a batch of instances, each with `colonies`
    colonies of `ants` ants, resident on one GPU."""

    def __init__(self, instances, params: Params | None = None, seeds=None, device: int = 0,
                 cand_k=None):
        self.lib = N.load()
        if isinstance(instances, Instance):
            instances = [instances]
        self.instances = list(instances)
        self.params = params or Params()
        if seeds is None:
            seeds = [1234] * len(self.instances)
        self.seeds = [int(s) & 0xFFFFFFFF for s in seeds]
        self._cand_k = [int(k) for k in cand_k] if cand_k is not None else None
        self._ctx = C.c_void_p()
        rc = self.lib.aco_hip_create(C.byref(self._ctx), device)
        if rc:
            raise AcoError(rc, "no usable HIP device")
        arr = (N.AcoHipInstance * len(self.instances))()
        for i, (inst, seed) in enumerate(zip(self.instances, self.seeds)):
            arr[i].n, arr[i].K, arr[i].capacity = inst.n, inst.K, inst.capacity
            arr[i].x = N.ptr(inst.x, C.c_int32)
            arr[i].y = N.ptr(inst.y, C.c_int32)
            arr[i].demand = N.ptr(inst.demand, C.c_int32)
            arr[i].seed = seed
            arr[i].cand_k = int(cand_k[i]) if cand_k is not None else 0
        self._check(self.lib.aco_hip_load(self._ctx, arr, len(self.instances),
                                          C.byref(self.params.to_c())))
        self._read_shape()

    def _read_shape(self) -> None:
        """ants, ants_per_group and stage_lds of bucket 0 (the first instance's)."""
        k, m, stride, apg, stage = (C.c_int32() for _ in range(5))
        self._check(self.lib.aco_hip_get_shape(self._ctx, C.byref(k), C.byref(m),
                                               C.byref(stride), C.byref(apg), C.byref(stage)))
        self.ants = m.value
        self.ants_per_group = apg.value
        self.stage_lds = stage.value

    # ---- plumbing -----------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc:
            msg = self.lib.aco_hip_last_error(self._ctx)
            raise AcoError(rc, msg.decode() if msg else "error")

    def close(self) -> None:
        if self._ctx:
            self.lib.aco_hip_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_colonies(self) -> int:
        return self.lib.aco_hip_num_colonies(self._ctx)

    def instance_of(self, colony: int) -> Instance:
        return self.instances[colony // self.params.colonies]

    def cand_k(self, instance: int = 0) -> int:
        k = self._cand_k[instance] if self._cand_k and self._cand_k[instance] else self.params.cand_k
        return min(k, self.instances[instance].n, 64)

    def tour_stride(self, colony: int = 0) -> int:
        inst = self.instance_of(colony)
        return (inst.n + inst.K + 1 + 63) // 64 * 64

    # ---- kernels ------------------------------------------------------------
    def build_candidates(self):
        self._check(self.lib.aco_hip_build_candidates(self._ctx))

    def init_tau(self):
        self._check(self.lib.aco_hip_init_tau(self._ctx))

    def construct(self):
        self._check(self.lib.aco_hip_construct(self._ctx))

    def select_best(self):
        self._check(self.lib.aco_hip_select_best(self._ctx))

    def evaporate(self):
        self._check(self.lib.aco_hip_evaporate(self._ctx))

    def deposit_iter(self):
        self._check(self.lib.aco_hip_deposit_iter(self._ctx))

    def deposit_best(self):
        self._check(self.lib.aco_hip_deposit_best(self._ctx))

    def epochs(self, count: int):
        self._check(self.lib.aco_hip_epochs(self._ctx, count))

    def sync(self):
        self._check(self.lib.aco_hip_sync(self._ctx))

    def prepare(self):
        self.build_candidates()
        self.init_tau()

    def run(self, timeout_seconds=0.0, stagnation_epochs=0, min_rel_improvement=1e-3,
            fixed_epochs=0, max_epochs=0, progress_interval_seconds=0.0, print_progress=False,
            chunk_epochs=8, use_graph=True) -> dict:
        cfg = N.AcoHipRunConfig(timeout_seconds, stagnation_epochs, min_rel_improvement,
                                fixed_epochs, max_epochs, progress_interval_seconds,
                                int(print_progress), chunk_epochs, int(use_graph))
        st = N.AcoHipRunStats()
        self._check(self.lib.aco_hip_run(self._ctx, C.byref(cfg), C.byref(st)))
        return {"epochs": st.epochs, "seconds": st.seconds, "ant_tours_per_s": st.ant_tours_per_s}

    # ---- downloads / uploads ----------------------------------------------
    def candidates(self, instance: int = 0):
        inst = self.instances[instance]
        k = self.cand_k(instance)
        idx = np.zeros((inst.n + 1) * k, np.int32)
        dist = np.zeros_like(idx)
        eta = np.zeros((inst.n + 1) * k, np.float32)
        self._check(self.lib.aco_hip_get_candidates(self._ctx, instance, N.ptr(idx, C.c_int32),
                                                    N.ptr(dist, C.c_int32),
                                                    N.ptr(eta, C.c_float)))
        return idx.reshape(-1, k), dist.reshape(-1, k), eta.reshape(-1, k)

    def tau(self, colony: int = 0):
        inst = self.instance_of(colony)
        k = self.cand_k(colony // self.params.colonies)
        tau = np.zeros((inst.n + 1, inst.n + 1), np.uint8)
        tc = np.zeros((inst.n + 1, k), np.uint8)
        self._check(self.lib.aco_hip_get_tau(self._ctx, colony, N.ptr(tau, C.c_uint8),
                                             N.ptr(tc, C.c_uint8)))
        return tau, tc

    def scores(self, colony: int = 0) -> np.ndarray:
        inst = self.instance_of(colony)
        k = self.cand_k(colony // self.params.colonies)
        out = np.zeros((inst.n + 1, k), np.float32)
        self._check(self.lib.aco_hip_get_scores(self._ctx, colony, N.ptr(out, C.c_float)))
        return out

    def fallback_weights(self, colony: int = 0) -> np.ndarray:
        """(n+1, n+1) all-pairs fallback weights (requires fallback_matrix on)."""
        inst = self.instance_of(colony)
        out = np.zeros((inst.n + 1, inst.n + 1), np.float32)
        self._check(self.lib.aco_hip_get_fallback_weights(self._ctx, colony,
                                                          N.ptr(out, C.c_float)))
        return out

    # ---- size buckets and autotuning -------------------------------------------
    @staticmethod
    def _bucket_dict(b) -> dict:
        return {f: getattr(b, f) for f, _ in b._fields_}

    def buckets(self) -> list:
        out = []
        for i in range(self.lib.aco_hip_num_buckets(self._ctx)):
            b = N.AcoHipBucketInfo()
            self._check(self.lib.aco_hip_get_bucket(self._ctx, i, C.byref(b)))
            out.append(self._bucket_dict(b))
        return out

    def set_bucket_shape(self, bucket: int, ants_per_group: int, stage_lds: int = -1,
                         ants_per_wave: int = 0):
        """ants_per_wave: 0 automatic, 1, or 2 (two ants per wavefront)."""
        self._check(self.lib.aco_hip_set_bucket_layout(self._ctx, bucket, ants_per_group,
                                                       stage_lds, ants_per_wave))
        self._read_shape()

    def autotune(self, launches: int = 3) -> list:
        nb = self.lib.aco_hip_num_buckets(self._ctx)
        arr = (N.AcoHipBucketInfo * max(nb, 1))()
        self._check(self.lib.aco_hip_autotune(self._ctx, launches, arr, nb))
        self._read_shape()
        return [self._bucket_dict(arr[i]) for i in range(nb)]

    def set_candidate_eta(self, eta, instance: int = 0):
        """Learned heuristic per candidate slot ((n+1, k) floats >= 0) replacing
        (1/d)^beta in the candidate roulette; None restores it."""
        if eta is None:
            self._check(self.lib.aco_hip_set_candidate_eta(self._ctx, instance, None))
            return
        e = np.ascontiguousarray(eta, np.float32)
        self._check(self.lib.aco_hip_set_candidate_eta(self._ctx, instance, N.ptr(e, C.c_float)))

    def set_tau(self, tau: np.ndarray, colony: int = 0):
        tau = np.ascontiguousarray(tau, dtype=np.uint8)
        self._check(self.lib.aco_hip_set_tau(self._ctx, colony, N.ptr(tau, C.c_uint8)))

    def tours(self, colony: int = 0):
        stride = self.tour_stride(colony)
        m = self.ants
        tours = np.zeros((m, stride), np.int32)
        ln = np.zeros(m, np.int32)
        cost = np.zeros(m, np.int64)
        unv = np.zeros(m, np.int32)
        self._check(self.lib.aco_hip_get_tours(self._ctx, colony, N.ptr(tours, C.c_int32),
                                               N.ptr(ln, C.c_int32), N.ptr(cost, C.c_int64),
                                               N.ptr(unv, C.c_int32)))
        return tours, ln, cost, unv

    def colony_state(self, colony: int = 0) -> dict:
        s = N.AcoHipColonyState()
        self._check(self.lib.aco_hip_get_colony(self._ctx, colony, C.byref(s)))
        return _state_dict(s)

    def set_epoch(self, epoch: int, colony: int = 0):
        self._check(self.lib.aco_hip_set_epoch(self._ctx, colony, epoch))

    def best(self, colony: int = 0):
        stride = self.tour_stride(colony)
        tour = np.zeros(stride, np.int32)
        ln = C.c_int32()
        cost = C.c_int64()
        self._check(self.lib.aco_hip_get_best(self._ctx, colony, N.ptr(tour, C.c_int32),
                                              C.byref(ln), C.byref(cost)))
        return tour[:ln.value].copy(), int(cost.value)

    def deposit_tour(self, tour, delta: float, colony: int = 0):
        t = np.ascontiguousarray(tour, dtype=np.int32)
        self._check(self.lib.aco_hip_deposit_tour(self._ctx, colony, N.ptr(t, C.c_int32),
                                                  len(t), delta))

    def route_costs(self, tours, lens, instance: int = 0) -> np.ndarray:
        tours = np.ascontiguousarray(tours, dtype=np.int32)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        out = np.zeros(tours.shape[0], np.int64)
        self._check(self.lib.aco_hip_route_costs(self._ctx, instance, N.ptr(tours, C.c_int32),
                                                 N.ptr(lens, C.c_int32), tours.shape[0],
                                                 tours.shape[1], N.ptr(out, C.c_int64)))
        return out

    def route_costs_noalloc(self, tours, lens, instance: int = 0) -> np.ndarray:
        """route_costs staged in the library's fixed device array (no device
        allocation; batches above ACO_HIP_NOALLOC_SCRATCH_BYTES are refused)."""
        tours = np.ascontiguousarray(tours, dtype=np.int32)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        out = np.zeros(tours.shape[0], np.int64)
        self._check(self.lib.aco_hip_route_costs_noalloc(
            self._ctx, instance, N.ptr(tours, C.c_int32), N.ptr(lens, C.c_int32), tours.shape[0],
            tours.shape[1], N.ptr(out, C.c_int64)))
        return out

    def tables(self) -> dict:
        deq = np.zeros(256, np.float32)
        thr = np.zeros(255, np.float32)
        tp = np.zeros(256, np.float32)
        n = C.c_int32()
        self._check(self.lib.aco_hip_get_tables(self._ctx, None, None, None, None, C.byref(n),
                                                None, None))
        eta = np.zeros(n.value, np.float32)
        q0, qe = C.c_uint8(), C.c_uint8()
        self._check(self.lib.aco_hip_get_tables(self._ctx, N.ptr(deq, C.c_float),
                                                N.ptr(thr, C.c_float), N.ptr(tp, C.c_float),
                                                N.ptr(eta, C.c_float), C.byref(n), C.byref(q0),
                                                C.byref(qe)))
        return {"deq": deq, "thr": thr, "taupow": tp, "eta": eta, "q_tau0": q0.value,
                "q_evap": qe.value}

    def set_stats(self, enable: bool):
        self._check(self.lib.aco_hip_set_stats(self._ctx, int(enable)))

    def stats(self, colony: int = 0) -> dict:
        st = N.AcoHipStats()
        self._check(self.lib.aco_hip_get_stats(self._ctx, colony, C.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in st._fields_}

    def set_timing(self, enable: bool):
        self._check(self.lib.aco_hip_set_timing(self._ctx, int(enable)))

    def timing(self) -> dict:
        t = N.AcoHipTiming()
        self._check(self.lib.aco_hip_get_timing(self._ctx, C.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}


def solve(instance: Instance, ants: int = 256, seed: int = 1234, alpha=1.0, beta=2.0, rho=0.5,
          tau0=1.0, Q=1.0, dist_mode: int = N.DIST_NINT):
    """aco_vrp_with_capacity (include/aco.h:127-131) on the GPU: returns
    (best tour, cost in cost units: the nint length, or the real length
    times 2^23 with dist_mode=DIST_REAL).  Stop rules come from ACO_SOLVER_*."""
    lib = N.load()
    tour = np.zeros(instance.n + instance.K + 1 + 64, np.int32)
    ln = C.c_int32()
    cost = C.c_int64()
    rc = lib.aco_vrp_hip_with_capacity(instance.n, instance.K, instance.capacity, ants,
                                       N.ptr(instance.x, C.c_int32), N.ptr(instance.y, C.c_int32),
                                       N.ptr(instance.demand, C.c_int32), alpha, beta, rho, tau0,
                                       Q, seed, dist_mode, N.ptr(tour, C.c_int32), C.byref(ln),
                                       C.byref(cost))
    if rc:
        raise AcoError(rc, lib.aco_status_string(rc).decode())
    return tour[:ln.value].copy(), int(cost.value)


def real_dist(s) -> np.ndarray:
    """The float nearest to sqrt(s) for squared lengths s (uint32), with the
    kernels' arithmetic (host build of the same inline code)."""
    s = np.ascontiguousarray(s, np.uint32)
    out = np.zeros(s.shape, np.float32)
    N.load().aco_real_dist_batch(s.ctypes.data_as(C.POINTER(C.c_uint32)), s.size,
                                 N.ptr(out, C.c_float))
    return out


def split_routes(tour) -> list:
    """Step-interleaved tour -> list of customer lists (one per vehicle)."""
    routes, cur = [], []
    for node in list(tour)[1:]:
        if node == 0:
            routes.append(cur)
            cur = []
        else:
            cur.append(int(node))
    return routes
