"""CPU checks that pin the oracle: known-answer vectors, closed forms,
brute force on small inputs, the reference's own candidate lists (golden
fixtures from aco_mpi_shared_init) and invariants of the construction and
pheromone update.  No GPU needed."""
import math
import os

import numpy as np
import pytest

from vehicleroutingproblem_amd import Instance, _native as N
import ctypes as C
import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert O.philox(ctr, key).tolist() == want
    lib = N.load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.aco_philox_host(c, k, o)
    assert list(o) == want


def test_nint_distance_exact():
    rng = np.random.default_rng(0)
    pts = rng.integers(0, 32768, size=(20000, 4))
    pts[:50] = [0, 0, 32767, 32767]
    lib = N.load()
    for x1, y1, x2, y2 in pts[:3000]:
        s = (int(x1) - int(x2)) ** 2 + (int(y1) - int(y2)) ** 2
        r = math.isqrt(s)
        want = r + (1 if 4 * s >= (2 * r + 1) ** 2 else 0)  # nint: round half up
        assert O.nint_dist(int(x1), int(y1), int(x2), int(y2)) == want
        assert lib.vrp_nint_dist(int(x1), int(y1), int(x2), int(y2)) == want


def test_scan_rows_is_exact_on_integers():
    # integer-valued floats sum exactly in any order: every prefix is exact
    rng = np.random.default_rng(1)
    for nl in (1, 5, 15, 16, 17, 20, 32, 33, 64):
        w = rng.integers(0, 1000, size=nl).astype(np.float32)
        p = O.scan_rows(w, nl)
        assert np.array_equal(p[:nl], np.cumsum(w).astype(np.float32))


def test_scan_rows_order():
    # a dropped or transposed term shows up as a wrong prefix on values whose
    # rounding depends on the grouping: compare with the documented grouping
    w = np.array([1.0, 1e8, -1e8 + 0.0, 3.0] + [0.5] * 16, np.float32)
    w = np.abs(w)
    p = O.scan_rows(w, 20)
    # row 0 (lanes 0..15) is a Kogge-Stone scan, row 1 adds lane 15's prefix
    ks = w.astype(np.float32).copy()
    for off in (1, 2, 4, 8):
        nxt = ks.copy()
        for l in range(20):
            if (l & 15) >= off:
                nxt[l] = np.float32(ks[l] + ks[l - off])
        ks = nxt
    ks[16:20] = ks[16:20] + ks[15]
    assert np.array_equal(p[:20], ks[:20])


def test_tau_tables_closed_forms():
    o = O.Oracle(Instance.uniform(20, 2, seed=1), cand_k=5, ants=4, rho=0.5)
    t = o.tables()
    assert t["q_tau0"] == 170  # round(4/6 * 255) for tau0 = 1: (log 1 - log 1e-4) / (log 100 - log 1e-4) = 4/6
    assert t["q_evap"] == 13   # rho = 0.5
    o2 = O.Oracle(Instance.uniform(20, 2, seed=1), cand_k=5, ants=4, rho=0.1)
    assert o2.tables()["q_evap"] == 2
    lmin = float(np.float32(math.log(np.float32(1e-4))))
    deq, thr = t["deq"], t["thr"]
    assert np.all(np.diff(deq) > 0) and np.all(np.diff(thr) > 0)
    lmax = np.float32(math.log(np.float32(100.0)))
    step_f = np.float32((lmax - np.float32(lmin)) / np.float32(255.0))
    step = float(step_f)
    for q in range(256):
        # dequantize_tau's argument q*step + log_min as one fused multiply-add
        # (exact in double, rounded once to float), then exp
        arg = np.float32(float(q) * step + lmin)
        want = math.exp(float(arg))
        assert abs(float(deq[q]) - want) <= 0.5 * float(np.spacing(np.float32(want)))
        assert o.quantize(float(deq[q])) == q
    for b in range(255):
        # thr[b] is the boundary of round((log tau - log min)/step) at b + 1/2
        lo = np.nextafter(thr[b], np.float32(0))
        assert o.quantize(float(thr[b])) == b + 1
        assert o.quantize(float(lo)) == b
    rng = np.random.default_rng(3)
    for tau in np.exp(rng.uniform(math.log(1e-5), math.log(200), 5000)).astype(np.float32):
        qf = (math.log(float(tau)) - lmin) / step
        if abs(qf - round(qf)) > 0.499 or abs(qf % 1 - 0.5) < 1e-4:
            continue
        want = 0 if qf <= 0 else 255 if qf >= 255 else int(math.floor(qf + 0.5))
        assert o.quantize(float(tau)) == want


def test_oracle_tables_equal_host_tables():
    lib = N.load()
    for rho in (0.1, 0.5, 0.9):
        tt = N.AcoTauTables()
        lib.aco_tau_tables_build(C.byref(tt), 1.0, rho, 1.0)
        o = O.Oracle(Instance.uniform(30, 2, seed=2), cand_k=5, ants=4, rho=rho)
        t = o.tables()
        assert np.array_equal(np.frombuffer(tt.deq, np.float32).view(np.uint32),
                              t["deq"].view(np.uint32))
        assert np.array_equal(np.frombuffer(tt.thr, np.float32).view(np.uint32),
                              t["thr"].view(np.uint32))
        assert tt.q_evap == t["q_evap"] and tt.q_tau0 == t["q_tau0"]


def _brute_knn(inst, k, ref_key=False):
    n = inst.n
    out = np.zeros((n + 1, k), np.int32)
    for i in range(n + 1):
        js = np.array([j for j in range(1, n + 1) if j != i])
        dx = inst.x[js].astype(np.int64) - inst.x[i]
        dy = inst.y[js].astype(np.int64) - inst.y[i]
        s = dx * dx + dy * dy
        key = np.sqrt(s.astype(np.float64)).astype(np.float32) if ref_key else s
        order = np.lexsort((js, key))[:k]
        out[i, :len(order)] = js[order]
    return out


@pytest.mark.parametrize("ref_key", [False, True])
def test_knn_brute_force_with_ties(ref_key):
    # many coincident and equidistant points force the index tie-break
    rng = np.random.default_rng(4)
    x = rng.integers(0, 8, 120)
    y = rng.integers(0, 8, 120)
    inst = Instance(x, y, np.r_[0, np.ones(119, int)], 4, 40)
    o = O.Oracle(inst, cand_k=12, ants=4, knn_key=int(ref_key))
    assert np.array_equal(o.candidates()[0], _brute_knn(inst, 12, ref_key))


@pytest.mark.parametrize("name", ["n500_k8_s19000", "n1000_k8_s19001"])
@pytest.mark.parametrize("k", [16, 20, 32])
def test_knn_matches_reference_golden(name, k):
    """Reference-key candidate lists equal the reference's own
    aco_mpi_shared_init output (tests/golden, made by make_golden.py)."""
    inst = Instance.read(os.path.join(GOLDEN, f"ref_{name}.vrp"), require_unit_demand=True)
    gold = np.load(os.path.join(GOLDEN, f"ref_knn_{name}.npz"))[f"k{k}"]
    o = O.Oracle(inst, cand_k=k, ants=4, knn_key=1)
    assert np.array_equal(o.candidates()[0], gold)


def _check_tours(inst, tours, lens, cost, unv, o):
    for a in range(len(lens)):
        t = tours[a, :lens[a]]
        assert t[0] == 0 and t[-1] == 0 and (t == 0).sum() == inst.K + 1
        cust = t[t > 0]
        assert len(np.unique(cust)) == len(cust) == inst.n - unv[a]
        load, loads = 0, []
        for v in t[1:]:
            if v == 0:
                loads.append(load)
                load = 0
            else:
                load += inst.demand[v]
        assert max(loads) <= inst.capacity
        assert o.route_cost(t) == cost[a]


@pytest.mark.parametrize("seed", [1, 2])
def test_construct_invariants_and_replay(seed):
    inst = Instance.uniform(150, 8, seed=seed)
    o = O.Oracle(inst, cand_k=15, ants=32, rho=0.1, seed=seed)
    o.construct(2)
    tours, lens, cost, unv = o.tours()
    _check_tours(inst, tours, lens, cost, unv, o)
    assert (unv == 0).mean() > 0.5  # most ants feasible at 10% slack
    o2 = O.Oracle(inst, cand_k=15, ants=32, rho=0.1, seed=seed)
    o2.construct(1)
    assert np.array_equal(o2.tours()[0], tours)  # thread count does not matter
    o3 = O.Oracle(inst, cand_k=15, ants=32, rho=0.1, seed=seed + 100)
    o3.construct(1)
    assert not np.array_equal(o3.tours()[0], tours)


def test_learned_eta_seam():
    """The learned-prior seam: a per-slot heuristic equal to the distance
    term reproduces the distance mode bit for bit; an all-zero one leaves no
    candidate weight, so every decision takes the all-customer fallback."""
    inst = Instance.uniform(120, 6, seed=4)
    base = O.Oracle(inst, cand_k=12, ants=16, rho=0.1, seed=4)
    base.epochs(3)
    o = O.Oracle(inst, cand_k=12, ants=16, rho=0.1, seed=4)
    _, _, eta = o.candidates()
    o.set_candidate_eta(eta)
    o.epochs(3)
    assert np.array_equal(o.tours()[0], base.tours()[0])
    assert np.array_equal(o.tau()[0], base.tau()[0])
    z = O.Oracle(inst, cand_k=12, ants=4, rho=0.1, seed=4)
    z.set_candidate_eta(np.zeros_like(eta))
    z.construct()
    tours, lens, cost, unv = z.tours()
    _check_tours(inst, tours, lens, cost, unv, z)
    decisions = int(sum(lens) - (inst.K + 1) * len(lens))  # customer moves
    assert z.fallback_count() >= decisions


def test_unit_demand_always_feasible():
    # the reference close rule guarantees feasibility for unit demand
    inst = Instance.read(os.path.join(GOLDEN, "ref_n500_k8_s19000.vrp"), require_unit_demand=True)
    o = O.Oracle(inst, cand_k=32, ants=16, rho=0.5)
    o.construct()
    tours, lens, cost, unv = o.tours()
    assert np.all(unv == 0)
    _check_tours(inst, tours, lens, cost, unv, o)


def test_evaporation_closed_form():
    inst = Instance.uniform(40, 3, seed=5)
    o = O.Oracle(inst, cand_k=8, ants=8, rho=0.5)
    rng = np.random.default_rng(2)
    tau = rng.integers(0, 256, size=(41, 41), dtype=np.uint8)
    tau[0, :20] = np.arange(20)
    o.set_tau(tau)
    o.construct()
    o.select_best()
    assert o.state()["found"]
    o.evaporate()
    want = np.where(tau > 13, tau - 13, 0).astype(np.uint8)
    assert np.array_equal(o.tau()[0], want)
    idx = o.candidates()[0]
    assert np.array_equal(o.tau()[1], np.take_along_axis(o.tau()[0], idx, axis=1))


def test_deposit_sequential_semantics():
    inst = Instance.uniform(30, 4, seed=6)
    o = O.Oracle(inst, cand_k=8, ants=4, rho=0.5)
    t = o.tables()
    deq, thr = t["deq"], t["thr"]

    def quant(v):
        return int(np.searchsorted(thr, np.float32(v), side="right")) if v > 0 else 0

    rng = np.random.default_rng(7)
    tau = rng.integers(0, 256, size=(31, 31), dtype=np.uint8)
    o.set_tau(tau)
    tour = [0, 4, 0, 0, 9, 3, 0, 0, 0]
    delta = np.float32(0.7)
    ref = tau.astype(np.int64).copy()
    for u, v in zip(tour[:-1], tour[1:]):
        nq = quant(np.float32(deq[ref[u, v]] + delta))
        ref[u, v] = max(ref[u, v], nq)
        ref[v, u] = max(ref[v, u], nq)
    o.deposit_tour(tour, float(delta))
    assert np.array_equal(o.tau()[0], ref.astype(np.uint8))
    idx = o.candidates()[0]
    assert np.array_equal(o.tau()[1], np.take_along_axis(o.tau()[0], idx, axis=1))


def test_evaporate_levels_closed_form():
    """kernel_evaporate_tau with any delta and q_min (the saturation in int:
    q_min + delta may exceed 255): equals max(q - delta, q_min) clamped at
    0, over every level and a spread of (delta, q_min)."""
    inst = Instance.uniform(15, 2, seed=5)
    o = O.Oracle(inst, cand_k=4, ants=4)
    q = np.resize(np.arange(256, dtype=np.uint8), (16, 16))
    for delta, q_min in [(13, 0), (2, 0), (1, 7), (40, 200), (255, 255), (0, 0), (200, 100),
                         (0, 255)]:
        o.set_tau(q)
        o.evaporate_levels(delta, q_min)
        want = np.maximum(np.maximum(q.astype(np.int64) - delta, 0), q_min).astype(np.uint8)
        assert np.array_equal(o.tau()[0], want), (delta, q_min)


def reference_layouts(tour, n, K, m, ant, steps):
    """The reference's two route layouts of one giant tour: the colony's
    step-major routes[steps][m] / lens[m][K] as kernel_construct_solutions
    writes them (aco_cuda_kernels.cu:520-548: step 0 is the depot, each
    route's nodes follow, its closing depot is the next route's first), and
    the flat routes[K][n+1] / lens[K] of aco_cuda.cu:373-379 (each route's
    nodes with both depots)."""
    routes, cur = [], [0]
    for v in tour[1:]:
        cur.append(int(v))
        if v == 0:
            routes.append(cur)
            cur = [0]
    assert len(routes) == K
    step = np.full((steps, m), -7, np.int32)
    slen = np.zeros((m, K), np.int32)
    g = 0
    step[0, ant] = 0
    for v, r in enumerate(routes):
        for node in r[1:]:
            g += 1
            step[g, ant] = node
        slen[ant, v] = len(r)
    flat = np.full((K, n + 1), -7, np.int32)
    flen = np.zeros(K, np.int32)
    for v, r in enumerate(routes):
        flat[v, :len(r)] = r
        flen[v] = len(r)
    return step, slen, flat, flen


@pytest.mark.parametrize("tour", [
    [0, 4, 0, 0, 9, 3, 0, 0, 0],            # single-customer and empty routes
    [0, 0, 0, 0, 0],                         # all routes empty
    [0, 5, 6, 7, 0, 1, 2, 0, 8, 0, 3, 0],   # ordinary routes
])
def test_route_layout_deposits_equal_tour_deposit(tour):
    """Per-route deposits read in the reference's step-major and flat layouts
    equal the deposit of the same routes as one giant tour (the layouts are
    built independently, by replaying the reference's writers)."""
    inst = Instance.uniform(30, 4, seed=6)
    K = tour.count(0) - 1
    inst = Instance(inst.x, inst.y, inst.demand, K, 30)
    m, ant = 5, 3
    step, slen, flat, flen = reference_layouts(tour, 30, K, m, ant, 40)
    rng = np.random.default_rng(11)
    tau = rng.integers(0, 256, size=(31, 31), dtype=np.uint8)
    want = O.Oracle(inst, cand_k=8, ants=4)
    want.set_tau(tau)
    want.deposit_tour(tour, 0.7)
    for kind in ("solution", "flat"):
        o = O.Oracle(inst, cand_k=8, ants=4)
        o.set_tau(tau)
        if kind == "solution":
            o.deposit_solution(step, slen, K, m, ant, 0.7)
        else:
            o.deposit_flat_solution(flat, flen, 0.7)
        assert np.array_equal(o.tau()[0], want.tau()[0]), kind
        assert np.array_equal(o.tau()[1], want.tau()[1]), kind
    assert not np.array_equal(want.tau()[0], tau)


def test_route_layout_deposits_skip_bad_entries():
    """A node outside 0..n, or a route running past the buffer, deposits
    nothing on the edges that touch it; the other edges are unchanged."""
    inst = Instance.uniform(30, 3, seed=6)
    tour = [0, 5, 6, 7, 0, 1, 2, 0, 8, 0]
    m, ant = 2, 0
    step, slen, flat, flen = reference_layouts(tour, 30, 3, m, ant, 12)
    rng = np.random.default_rng(3)
    tau = rng.integers(0, 256, size=(31, 31), dtype=np.uint8)
    # node 6 of route 0 becomes 99: edges (5,99), (99,7) deposit nothing
    step[2, ant] = 99
    flat[0, 2] = 99
    # route 2 runs 3 steps past the end of the step buffer / is longer than its row
    slen[ant, 2] = 6
    flen[2] = 33
    flat[2, 3:] = 0
    want = O.Oracle(inst, cand_k=8, ants=4)
    want.set_tau(tau)
    want.deposit_tour([0, 5], 0.7)
    want.deposit_tour([7, 0, 1, 2, 0, 8, 0], 0.7)
    o = O.Oracle(inst, cand_k=8, ants=4)
    o.set_tau(tau)
    o.deposit_solution(step, slen, 3, m, ant, 0.7)
    assert np.array_equal(o.tau()[0], want.tau()[0])
    # flat: route 2 is 0 8 0 then zeros to the end of its row: 0->0 edges
    # (29 of them) that the 33-node length runs past the last row do not add
    want_f = O.Oracle(inst, cand_k=8, ants=4)
    want_f.set_tau(tau)
    want_f.deposit_tour([0, 5], 0.7)
    want_f.deposit_tour([7, 0, 1, 2, 0, 8] + [0] * 29, 0.7)
    o = O.Oracle(inst, cand_k=8, ants=4)
    o.set_tau(tau)
    o.deposit_flat_solution(flat, flen, 0.7)
    assert np.array_equal(o.tau()[0], want_f.tau()[0])


def test_epoch_loop_improves_and_stops():
    inst = Instance.uniform(100, 8, seed=1234)
    o = O.Oracle(inst, cand_k=15, ants=64, rho=0.1, seed=1234)
    o.set_stop_rules(stagnation=10, min_rel=1e-3)
    o.epochs(1)
    first = o.state()["global_best_cost"]
    o.epochs(500)
    st = o.state()
    assert st["halted"] and st["iter_since_best"] >= 10
    assert st["global_best_cost"] < first
    tour, cost = o.best()
    assert o.route_cost(tour) == cost


def test_smoke_instance_golden_runs():
    """The 8-customer smoke instance: brute-force optimum (real distances) and
    the reference seq / OpenMP CLI outputs (golden): both are valid route sets
    whose printed cost equals their recomputed real cost, and neither beats
    the optimum."""
    import itertools
    import json
    inst = Instance.read(os.path.join(GOLDEN, "ref_smoke.vrp"), require_unit_demand=True)
    runs = json.load(open(os.path.join(GOLDEN, "ref_smoke_runs.json")))
    lib = N.load()

    def d(a, b):
        return lib.vrp_real_dist(int(inst.x[a]), int(inst.y[a]), int(inst.x[b]), int(inst.y[b]))

    def route_len(r):
        seq = [0] + list(r) + [0]
        return sum(d(a, b) for a, b in zip(seq[:-1], seq[1:]))

    best = float("inf")
    for mask in range(1 << 8):
        a = [c for c in range(1, 9) if mask >> (c - 1) & 1]
        b = [c for c in range(1, 9) if not mask >> (c - 1) & 1]
        if len(a) > 5 or len(b) > 5:
            continue
        la = min((route_len(p) for p in itertools.permutations(a)), default=0.0)
        lb = min((route_len(p) for p in itertools.permutations(b)), default=0.0)
        best = min(best, la + lb)
    assert f"{best:.3f}" == "244.853"  # brute-force optimum
    for tag in ("seq", "omp"):
        lines = runs[tag]
        routes = [list(map(int, l.split(":")[1].split())) for l in lines if l.startswith("Route")]
        assert sorted(sum(routes, [])) == list(range(1, 9))
        assert all(len(r) <= inst.capacity for r in routes)
        cost = sum(route_len(r) for r in routes)
        assert lines[-2] == f"Cost: {cost:.3f}" and lines[-1] == f"best cost: {cost:.3f}"
        assert cost >= best - 1e-9


@pytest.mark.parametrize("name,inst_name", [
    ("n500_k8_s19000_seq_run1", "n500_k8_s19000"),
    ("n500_k8_s19000_cuda_run1", "n500_k8_s19000"),
    ("n1000_k8_s19001_cuda_run1", "n1000_k8_s19001")])
def test_tracked_reference_solutions(name, inst_name):
    """Tracked reference solution files (golden): the parser, the real route
    cost and the validator reproduce them (seq prints double sums, CUDA FP32
    sums, hence the tolerances)."""
    inst = Instance.read(os.path.join(GOLDEN, f"ref_{inst_name}.vrp"), require_unit_demand=True)
    lines = open(os.path.join(GOLDEN, f"ref_{name}.solution.txt")).read().splitlines()
    routes = [list(map(int, l.split(":")[1].split())) for l in lines if l.startswith("Route")]
    printed = float(lines[-1].split(":")[1])
    assert len(routes) == inst.K
    assert sorted(sum(routes, [])) == list(range(1, inst.n + 1))
    assert all(len(r) <= inst.capacity for r in routes)
    lib = N.load()
    cost = 0.0
    for r in routes:
        seq = [0] + r + [0]
        cost += sum(lib.vrp_real_dist(int(inst.x[a]), int(inst.y[a]), int(inst.x[b]),
                                      int(inst.y[b])) for a, b in zip(seq[:-1], seq[1:]))
    tol = 5e-4 if "seq" in name else 5e-3
    assert abs(cost - printed) <= tol
