"""Real-distance mode on the GPU against the oracle, and the drop-in entry
point with the reference's prototype.

* Parity: tours, costs (exact sums of float lengths in 2^-23 units),
  pheromone, candidate scores, all-pairs fallback weights and colony state
  bit for bit, epoch after epoch, on the shapes of the nint parity suite,
  the reference's own grid-100 benchmark instance (co-located customers:
  zero-length edges) and with the weight matrix off.
* The device's float square root with its exact correction, over ten
  million random edges of a full-range instance (route-cost reduction).
* The CLIs: aco_vrp_hip.out optimises and prints the real length by
  default; aco_vrp_cuda_compat.out (a C++ caller declaring the reference's
  prototype, src/cuda/main_vrp.cu:29-35) solves through
  aco_vrp_cuda_with_capacity and prints a validated solution.
"""
import os
import subprocess

import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N
import oracle as O
from test_gpu_parity import CASES, smoke_instance

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COMPAT_CLI = os.path.join(os.path.dirname(N.CLI_PATH), "aco_vrp_cuda_compat.out")


def code_matrix(inst):
    """Heuristic-table code of every pair: bin of the float distance."""
    x = np.asarray(inst.x, np.int64)
    y = np.asarray(inst.y, np.int64)
    s = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    d = np.sqrt(s.astype(np.float64)).astype(np.float32)
    bits = d.view(np.uint32).astype(np.int64)
    return np.where(d >= 1.0, 1 + ((bits - 0x3F800000) >> 13), 0)


def ref_n500():
    return Instance.read(os.path.join(GOLDEN, "ref_n500_k8_s19000.vrp"), require_unit_demand=True)


REAL_CASES = {
    "c1_n100": CASES["c1_n100"],
    "short_routes": CASES["short_routes"],
    "n300_clustered": CASES["n300_clustered"],
    "n1030_two_chunks": CASES["n1030_two_chunks"],
    "wide_coords": CASES["wide_coords"],
    "ref_n500_grid": (ref_n500, dict(cand_k=32, ants=64, rho=0.5)),
}


def make_real_pair(name, **over):
    mk, kw = REAL_CASES[name]
    inst = mk()
    kw = dict(kw)
    kw.update(over)
    p = Params(dist_mode=N.DIST_REAL, **kw)
    s = HipSolver(inst, p, seeds=[4321])
    o = O.Oracle(inst, cand_k=p.cand_k, ants=p.ants, alpha=p.alpha, beta=p.beta, rho=p.rho,
                 tau0=p.tau0, Q=p.Q, depot_w=p.depot_close_weight, knn_key=p.knn_key, seed=4321,
                 dist_mode=N.DIST_REAL)
    return inst, s, o


@pytest.mark.parametrize("name,fm", [(n, 1) for n in REAL_CASES] + [("n1030_two_chunks", 0),
                                                                    ("ref_n500_grid", 0)])
def test_real_mode_epochs_match_oracle(gpu, name, fm):
    inst, s, o = make_real_pair(name, fallback_matrix=fm)
    s.build_candidates()
    tg, to = s.tables(), o.tables()
    assert np.array_equal(tg["eta"].view(np.uint32), to["eta"].view(np.uint32))
    gi, gd, _ = s.candidates()
    oi, od, _ = o.candidates()
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    s.init_tau()
    codes = code_matrix(inst) if fm else None
    for e in range(4):
        s.epochs(1)
        o.epochs(1)
        gt, gl, gc, gu = s.tours()
        ot, ol, oc, ou = o.tours()
        assert np.array_equal(gl, ol) and np.array_equal(gc, oc) and np.array_equal(gu, ou), e
        for a in range(len(gl)):
            assert np.array_equal(gt[a, :gl[a]], ot[a, :ol[a]]), (e, a)
        gs, os_ = s.colony_state(), o.state()
        for key in ("epoch", "iter_since_best", "found", "best_ant", "best_len",
                    "feasible_ants", "iter_best_cost", "global_best_cost", "improved"):
            assert gs[key] == os_[key], (e, key, gs[key], os_[key])
        assert np.float32(gs["dep_iter"]) == np.float32(os_["dep_iter"])
        assert np.float32(gs["dep_best"]) == np.float32(os_["dep_best"])
        tgau, cg = s.tau()
        tor, co = o.tau()
        assert np.array_equal(tgau, tor) and np.array_equal(cg, co), e
        want = (to["taupow"][co] * to["eta"][od]).astype(np.float32)
        assert np.array_equal(s.scores().view(np.uint32), want.view(np.uint32)), e
        if fm:
            want = (to["taupow"][tor] * to["eta"][codes]).astype(np.float32)
            assert np.array_equal(s.fallback_weights().view(np.uint32), want.view(np.uint32)), e
    if o.best()[1] is not None:
        assert s.best()[1] == o.best()[1] and np.array_equal(s.best()[0], o.best()[0])
    s.close()


def test_real_distances_on_device_over_ten_million_edges(gpu):
    inst = Instance.uniform(5000, 8, side=32767, seed=77)
    p = Params(cand_k=8, ants=4, dist_mode=N.DIST_REAL)
    rng = np.random.default_rng(3)
    count = 2000
    tours = np.zeros((count, inst.n + 2), np.int32)
    for t in range(count):
        tours[t, 1:-1] = rng.permutation(np.arange(1, inst.n + 1))
    lens = np.full(count, inst.n + 2, np.int32)
    with HipSolver(inst, p, seeds=[1]) as s:
        s.prepare()
        got = s.route_costs(tours, lens)
        assert np.array_equal(got, s.route_costs_noalloc(tours[:200], lens[:200]).tolist()
                              + got[200:].tolist())
    o = O.Oracle(inst, cand_k=8, ants=4, dist_mode=N.DIST_REAL)
    want = np.array([o.route_cost(tours[t]) for t in range(count)], np.int64)
    assert np.array_equal(got, want)


def test_cli_real_by_default(gpu, tmp_path):
    path = tmp_path / "smoke.vrp"
    inst = smoke_instance()
    inst.write(str(path))
    env = dict(os.environ, ACO_SOLVER_TIMEOUT_SECONDS="10", ACO_SOLVER_STAGNATION_EPOCHS="20")
    env.pop("ACO_HIP_DIST", None)
    costs = {}
    for mode in ("real", "nint"):
        if mode == "nint":
            env["ACO_HIP_DIST"] = "nint"
        out = subprocess.run([N.CLI_PATH, str(path), "2", "0", "1234"], env=env,
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        lines = out.stdout.splitlines()
        routes = [list(map(int, l.split(":")[1].split())) for l in lines if l.startswith("Route ")]
        assert sorted(sum(routes, [])) == list(range(1, 9))
        cost = float([l for l in lines if l.startswith("Cost: ")][0].split()[1])
        x, y = inst.x.astype(np.float64), inst.y.astype(np.float64)
        seq = [0]
        for r in routes:
            seq += r + [0]
        seq = np.array(seq)
        d = np.hypot(np.diff(x[seq]), np.diff(y[seq]))
        want = d.sum() if mode == "real" else np.rint(d).sum()
        assert abs(cost - want) <= 1e-3, (mode, cost, want)
        costs[mode] = cost
    # the brute-force optimum of the smoke instance is 244.853 in real distances
    assert costs["real"] >= 244.853 - 1e-3


def test_drop_in_entry_point_through_reference_prototype(gpu):
    env = dict(os.environ, ACO_SOLVER_TIMEOUT_SECONDS="20", ACO_SOLVER_STAGNATION_EPOCHS="20")
    for name, K, m in (("ref_smoke.vrp", 2, 0), ("ref_n500_k8_s19000.vrp", 8, 64)):
        path = os.path.join(GOLDEN, name)
        out = subprocess.run([COMPAT_CLI, path, str(K), str(m), "1234"], env=env,
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        inst = Instance.read(path, require_unit_demand=True)
        lines = out.stdout.splitlines()
        routes = [list(map(int, l.split(":")[1].split())) for l in lines if l.startswith("Route ")]
        assert len(routes) == K
        assert sorted(sum(routes, [])) == list(range(1, inst.n + 1))
        assert max(len(r) for r in routes) <= inst.capacity
        seq = [0]
        for r in routes:
            seq += r + [0]
        seq = np.array(seq)
        x, y = inst.x.astype(np.float64), inst.y.astype(np.float64)
        want = np.hypot(np.diff(x[seq]), np.diff(y[seq])).sum()
        printed = float([l for l in lines if l.startswith("best cost: ")][0].split()[2])
        assert abs(printed - want) <= 1e-3 + len(seq) * want * 2.0 ** -24, (name, printed, want)
