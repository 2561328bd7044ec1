"""ctypes loader of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
benchmark's CPU-baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


class OracleState(C.Structure):
    _fields_ = [("epoch", C.c_int32), ("iter_since_best", C.c_int32), ("halted", C.c_int32),
                ("found", C.c_int32), ("best_ant", C.c_int32), ("improved", C.c_int32),
                ("best_len", C.c_int32), ("feasible_ants", C.c_int32),
                ("iter_best_cost", C.c_int64), ("global_best_cost", C.c_int64),
                ("dep_iter", C.c_float), ("dep_best", C.c_float)]


_SIG = {
    "oracle_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, C.c_int,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_int, C.c_uint32, C.c_uint32, C.c_int]),
    "oracle_destroy": (None, [C.c_void_p]),
    "oracle_cand_k": (C.c_int, [C.c_void_p]),
    "oracle_tour_stride": (C.c_int, [C.c_void_p]),
    "oracle_get_candidates": (None, [C.c_void_p, i32p, i32p, f32p]),
    "oracle_get_tables": (None, [C.c_void_p, f32p, f32p, f32p, f32p, i32p, u8p, u8p]),
    "oracle_get_tau": (None, [C.c_void_p, u8p, u8p]),
    "oracle_set_tau": (None, [C.c_void_p, u8p]),
    "oracle_set_epoch": (None, [C.c_void_p, C.c_int]),
    "oracle_set_candidate_eta": (None, [C.c_void_p, C.c_void_p]),
    "oracle_set_stop_rules": (None, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
    "oracle_construct_ant": (C.c_int, [C.c_void_p, C.c_int, C.c_int, i32p, i64p, i32p, i64p]),
    "oracle_construct": (None, [C.c_void_p, C.c_int]),
    "oracle_get_tours": (None, [C.c_void_p, i32p, i32p, i64p, i32p]),
    "oracle_select_best": (None, [C.c_void_p]),
    "oracle_significant": (C.c_int, [C.c_double, C.c_double, C.c_double]),
    "oracle_set_ant_results": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "oracle_evaporate": (None, [C.c_void_p]),
    "oracle_deposit_tour": (None, [C.c_void_p, i32p, C.c_int, C.c_float]),
    "oracle_deposit_routes": (None, [C.c_void_p, i32p, i32p, C.c_int, C.c_int, C.c_longlong,
                                     C.c_int, C.c_float, C.c_int]),
    "oracle_evaporate_levels": (None, [C.c_void_p, C.c_uint8, C.c_uint8]),
    "oracle_deposit_iter": (None, [C.c_void_p]),
    "oracle_deposit_best": (None, [C.c_void_p]),
    "oracle_epochs": (None, [C.c_void_p, C.c_int, C.c_int]),
    "oracle_get_state": (None, [C.c_void_p, C.POINTER(OracleState)]),
    "oracle_get_best": (C.c_int, [C.c_void_p, i32p, i32p, i64p]),
    "oracle_route_cost": (C.c_int64, [C.c_void_p, i32p, C.c_int]),
    "oracle_fallback_count": (C.c_int64, [C.c_void_p]),
    "oracle_philox4x32_10": (None, [u32p, u32p, u32p]),
    "oracle_nint_dist": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "oracle_scan_rows": (None, [f32p, C.c_int, f32p]),
    "oracle_quantize": (C.c_uint8, [C.c_void_p, C.c_float]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
        lib = C.CDLL(LIB)
        for name, (res, args) in _SIG.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """One colony of the CPU restatement (same parameters as the GPU context)."""

    def __init__(self, inst, cand_k=20, ants=256, alpha=1.0, beta=2.0, rho=0.1, tau0=1.0, Q=1.0,
                 depot_w=2.0, knn_key=0, seed=1234, colony=0, dist_mode=0):
        self.lib = load()
        self.inst = inst
        self.x = np.ascontiguousarray(inst.x, np.int32)
        self.y = np.ascontiguousarray(inst.y, np.int32)
        self.d = np.ascontiguousarray(inst.demand, np.int32)
        self.h = self.lib.oracle_create(inst.n, inst.K, inst.capacity, _p(self.x, C.c_int32),
                                        _p(self.y, C.c_int32), _p(self.d, C.c_int32), cand_k,
                                        ants, alpha, beta, rho, tau0, Q, depot_w, knn_key, seed,
                                        colony, dist_mode)
        if not self.h:
            raise ValueError("oracle_create rejected its arguments")
        self.n, self.m = inst.n, ants
        self.k = self.lib.oracle_cand_k(self.h)
        self.stride = self.lib.oracle_tour_stride(self.h)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def candidates(self):
        tot = (self.n + 1) * self.k
        idx = np.zeros(tot, np.int32)
        dist = np.zeros(tot, np.int32)
        eta = np.zeros(tot, np.float32)
        self.lib.oracle_get_candidates(self.h, _p(idx, C.c_int32), _p(dist, C.c_int32),
                                       _p(eta, C.c_float))
        return idx.reshape(-1, self.k), dist.reshape(-1, self.k), eta.reshape(-1, self.k)

    def tables(self):
        deq = np.zeros(256, np.float32)
        thr = np.zeros(255, np.float32)
        tp = np.zeros(256, np.float32)
        n = C.c_int32()
        self.lib.oracle_get_tables(self.h, None, None, None, None, C.byref(n), None, None)
        eta = np.zeros(n.value, np.float32)
        q0, qe = C.c_uint8(), C.c_uint8()
        self.lib.oracle_get_tables(self.h, _p(deq, C.c_float), _p(thr, C.c_float),
                                   _p(tp, C.c_float), _p(eta, C.c_float), C.byref(n),
                                   C.byref(q0), C.byref(qe))
        return {"deq": deq, "thr": thr, "taupow": tp, "eta": eta, "q_tau0": q0.value,
                "q_evap": qe.value}

    def tau(self):
        tau = np.zeros((self.n + 1, self.n + 1), np.uint8)
        tc = np.zeros((self.n + 1, self.k), np.uint8)
        self.lib.oracle_get_tau(self.h, _p(tau, C.c_uint8), _p(tc, C.c_uint8))
        return tau, tc

    def set_tau(self, tau):
        t = np.ascontiguousarray(tau, np.uint8)
        self.lib.oracle_set_tau(self.h, _p(t, C.c_uint8))

    def set_epoch(self, e):
        self.lib.oracle_set_epoch(self.h, e)

    def set_candidate_eta(self, eta):
        """Learned heuristic per candidate slot ((n+1, k) floats); None
        restores (1/d)^beta."""
        if eta is None:
            self.lib.oracle_set_candidate_eta(self.h, None)
            return
        self._ceta = np.ascontiguousarray(eta, np.float32).reshape(-1)
        self.lib.oracle_set_candidate_eta(self.h, self._ceta.ctypes.data_as(C.c_void_p))

    def set_stop_rules(self, stagnation=0, min_rel=1e-3, fixed=0):
        self.lib.oracle_set_stop_rules(self.h, stagnation, min_rel, fixed)

    def construct_ant(self, ant, epoch):
        tour = np.zeros(self.stride, np.int32)
        cost = C.c_int64()
        unv = C.c_int32()
        fb = C.c_int64()
        ln = self.lib.oracle_construct_ant(self.h, ant, epoch, _p(tour, C.c_int32),
                                           C.byref(cost), C.byref(unv), C.byref(fb))
        return tour[:ln], int(cost.value), int(unv.value), int(fb.value)

    def construct(self, threads=0):
        self.lib.oracle_construct(self.h, threads)

    def tours(self):
        tours = np.zeros((self.m, self.stride), np.int32)
        ln = np.zeros(self.m, np.int32)
        cost = np.zeros(self.m, np.int64)
        unv = np.zeros(self.m, np.int32)
        self.lib.oracle_get_tours(self.h, _p(tours, C.c_int32), _p(ln, C.c_int32),
                                  _p(cost, C.c_int64), _p(unv, C.c_int32))
        return tours, ln, cost, unv

    def select_best(self):
        self.lib.oracle_select_best(self.h)

    def set_ant_results(self, cost=None, unvisited=None, lens=None):
        """Replace this epoch's per-ant costs / unvisited counts / lengths."""
        keep = []

        def arr(a, t):
            if a is None:
                return None
            a = np.ascontiguousarray(a, t)
            assert a.shape == (self.m,)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        self.lib.oracle_set_ant_results(self.h, arr(cost, np.int64), arr(unvisited, np.int32),
                                        arr(lens, np.int32))

    def evaporate(self):
        self.lib.oracle_evaporate(self.h)

    def deposit_tour(self, tour, delta):
        t = np.ascontiguousarray(tour, np.int32)
        self.lib.oracle_deposit_tour(self.h, _p(t, C.c_int32), len(t), delta)

    def deposit_solution(self, routes, lens, K, m, ant, delta):
        """kernel_deposit_solution's layout: routes[steps][m], lens[m][K]."""
        r = np.ascontiguousarray(routes, np.int32)
        ln = np.ascontiguousarray(lens, np.int32)
        assert r.ndim == 2 and r.shape[1] == m and ln.shape == (m, K)
        self.lib.oracle_deposit_routes(self.h, _p(r, C.c_int32), _p(ln, C.c_int32), K, m,
                                       r.shape[0], ant, delta, 1)

    def deposit_flat_solution(self, routes, lens, delta):
        """kernel_deposit_flat_solution's layout: routes[K][n+1], lens[K]."""
        r = np.ascontiguousarray(routes, np.int32)
        ln = np.ascontiguousarray(lens, np.int32)
        K = ln.shape[0]
        assert r.shape == (K, self.n + 1)
        self.lib.oracle_deposit_routes(self.h, _p(r, C.c_int32), _p(ln, C.c_int32), K, 1, 0, 0,
                                       delta, 0)

    def evaporate_levels(self, delta, q_min):
        self.lib.oracle_evaporate_levels(self.h, delta, q_min)

    def deposit_iter(self):
        self.lib.oracle_deposit_iter(self.h)

    def deposit_best(self):
        self.lib.oracle_deposit_best(self.h)

    def epochs(self, count, threads=0):
        self.lib.oracle_epochs(self.h, count, threads)

    def state(self):
        s = OracleState()
        self.lib.oracle_get_state(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def best(self):
        tour = np.zeros(self.stride, np.int32)
        ln = C.c_int32()
        cost = C.c_int64()
        if not self.lib.oracle_get_best(self.h, _p(tour, C.c_int32), C.byref(ln), C.byref(cost)):
            return None, None
        return tour[:ln.value].copy(), int(cost.value)

    def route_cost(self, tour):
        t = np.ascontiguousarray(tour, np.int32)
        return int(self.lib.oracle_route_cost(self.h, _p(t, C.c_int32), len(t)))

    def fallback_count(self):
        return int(self.lib.oracle_fallback_count(self.h))

    def quantize(self, tau):
        return int(self.lib.oracle_quantize(self.h, tau))


def philox(ctr, key):
    lib = load()
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib.oracle_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def scan_rows(w, nlanes):
    lib = load()
    w = np.ascontiguousarray(np.pad(np.asarray(w, np.float32), (0, 64 - len(w))), np.float32)
    out = np.zeros(64, np.float32)
    lib.oracle_scan_rows(_p(w, C.c_float), nlanes, _p(out, C.c_float))
    return out


def significant(prev, now, min_rel):
    return bool(load().oracle_significant(prev, now, min_rel))


def nint_dist(x1, y1, x2, y2):
    return int(load().oracle_nint_dist(x1, y1, x2, y2))
