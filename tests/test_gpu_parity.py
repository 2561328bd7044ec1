"""GPU parity: every kernel of the hot path against the CPU oracle, called
through the C ABI.  Integer, byte and index outputs are compared bit-exactly;
the float tables too (they are shared bit patterns, not recomputed)."""
import os
import subprocess

import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N
from vehicleroutingproblem_amd.solver import AcoError
import oracle as O

pytestmark = pytest.mark.gpu


def smoke_instance():
    """The 8-customer instance of the reference smoke test
    (tools/bash/smoke_test.sh:17-48): unit demand, K=2, Q=5."""
    x = [50, 10, 20, 10, 20, 80, 90, 80, 90]
    y = [50, 10, 10, 20, 20, 80, 80, 90, 90]
    d = [0] + [1] * 8
    return Instance(np.array(x), np.array(y), np.array(d), 2, 5, "smoke")


CASES = {
    "c1_n100": (lambda: Instance.uniform(100, 8, seed=1234), dict(cand_k=15, ants=64, rho=0.1)),
    "n300_clustered": (lambda: Instance.clustered(300, 10, side=10000, sigma=300, dmin=1,
                                                  dmax=100, slack_pct=105, seed=7),
                       dict(cand_k=32, ants=96, rho=0.1)),
    "smoke_unit": (smoke_instance, dict(cand_k=32, ants=32, rho=0.5)),
    "tiny_n5": (lambda: Instance.uniform(5, 2, seed=3), dict(cand_k=20, ants=16, rho=0.5)),
    "n1030_two_chunks": (lambda: Instance.uniform(1030, 9, seed=11),
                         dict(cand_k=20, ants=64, rho=0.1)),
    # full coordinate range: exercises the nint square root near 2*32767^2
    "wide_coords": (lambda: Instance.uniform(200, 4, side=32767, seed=5),
                    dict(cand_k=24, ants=32, rho=0.3)),
    # customers beside (and on) the depot with a capacity of a third of the
    # demand: best tours hold empty (0,0) routes and single-customer 0-c-0
    # routes, the conflict cases of the deposits
    "short_routes": (lambda: Instance(
        np.array([500, 500, 500, 500, 438, 433, 858, 85, 697, 201, 94, 526, 975, 735, 761]),
        np.array([500, 498, 497, 500, 128, 839, 450, 500, 370, 182, 926, 781, 643, 402, 822]),
        np.array([0, 7, 1, 7, 6, 2, 5, 2, 6, 5, 3, 1, 7, 4, 7]), 4, 21, "short_routes"),
        dict(cand_k=6, ants=64, rho=0.5)),
    # short candidate lists: most late decisions take the all-customer
    # fallback, over three 1024-node chunks
    "n2100_k6_fallbacks": (lambda: Instance.uniform(2100, 12, seed=21),
                           dict(cand_k=6, ants=32, rho=0.1)),
    # candidate lists over three and four 16-lane rows (the scan's readlane
    # carries between rows 1-2 and 2-3; k = 64 fills the wavefront)
    "k48_three_rows": (lambda: Instance.uniform(200, 4, seed=8),
                       dict(cand_k=48, ants=32, rho=0.2)),
    "k64_four_rows": (lambda: Instance.clustered(150, 3, side=1000, sigma=60, seed=9),
                      dict(cand_k=64, ants=32, rho=0.2)),
}


def nint_matrix(inst):
    """nint(d(u, v)) for every node pair (sqrt of an integer is never a
    half-integer, so rounding the double square root is exact)."""
    x = np.asarray(inst.x, np.int64)
    y = np.asarray(inst.y, np.int64)
    s = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    return np.rint(np.sqrt(s.astype(np.float64))).astype(np.int64)


def check_fallback_weights(s, o, inst):
    """The all-pairs fallback weights equal taupow[tau[u][v]] * eta[d(u,v)]
    (one fp32 multiply) for the oracle's pheromone matrix."""
    t = o.tables()
    tau, _ = o.tau()
    want = (t["taupow"][tau] * t["eta"][nint_matrix(inst)]).astype(np.float32)
    got = s.fallback_weights()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def make_pair(name, **over):
    mk, kw = CASES[name]
    inst = mk()
    kw = dict(kw)
    kw.update(over)
    seed = kw.pop("seed", 1234)
    p = Params(**kw)
    s = HipSolver(inst, p, seeds=[seed])
    o = O.Oracle(inst, cand_k=p.cand_k, ants=p.ants, alpha=p.alpha, beta=p.beta, rho=p.rho,
                 tau0=p.tau0, Q=p.Q, depot_w=p.depot_close_weight, knn_key=p.knn_key, seed=seed,
                 dist_mode=p.dist_mode)
    return inst, s, o


@pytest.mark.parametrize("name", list(CASES))
def test_tables_and_candidates(gpu, name):
    inst, s, o = make_pair(name)
    s.build_candidates()
    tg, to = s.tables(), o.tables()
    for key in ("deq", "thr", "taupow", "eta"):
        assert np.array_equal(tg[key].view(np.uint32), to[key].view(np.uint32)), key
    assert tg["q_tau0"] == to["q_tau0"] and tg["q_evap"] == to["q_evap"]
    gi, gd, ge = s.candidates()
    oi, od, oe = o.candidates()
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)
    assert np.array_equal(ge.view(np.uint32), oe.view(np.uint32))
    s.close()


@pytest.mark.parametrize("name", ["c1_n100", "n300_clustered"])
def test_candidates_reference_key(gpu, name):
    inst, s, o = make_pair(name, knn_key=N.KNN_REF_F32)
    s.build_candidates()
    assert np.array_equal(s.candidates()[0], o.candidates()[0])
    s.close()


@pytest.mark.parametrize("name", ["n500_k8_s19000", "n1000_k8_s19001"])
@pytest.mark.parametrize("k", [16, 20, 32])
def test_candidates_equal_reference_golden(gpu, name, k):
    """k_knn with the reference key on the reference's own benchmark
    instances gives exactly the lists the reference's aco_mpi_shared_init
    builds (src/openmp-mpi/aco_parallel.c:218-250; golden lists in
    tests/golden, written by make_golden.py from the reference sources)."""
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    inst = Instance.read(os.path.join(golden, f"ref_{name}.vrp"), require_unit_demand=True)
    gold = np.load(os.path.join(golden, f"ref_knn_{name}.npz"))[f"k{k}"]
    with HipSolver(inst, Params(cand_k=k, ants=4, knn_key=N.KNN_REF_F32)) as s:
        s.build_candidates()
        assert np.array_equal(s.candidates()[0], gold)


def _compare_tours(s, o):
    gt, gl, gc, gu = s.tours()
    ot, ol, oc, ou = o.tours()
    assert np.array_equal(gl, ol)
    for a in range(len(gl)):
        assert np.array_equal(gt[a, :gl[a]], ot[a, :ol[a]]), f"ant {a}"
    assert np.array_equal(gc, oc)
    assert np.array_equal(gu, ou)
    return gt, gl, gc, gu


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("stage,fm", [(1, 1), (0, 1), (1, 0)])
def test_construct_matches_oracle(gpu, name, stage, fm):
    inst, s, o = make_pair(name, stage_lds=stage, fallback_matrix=fm)
    s.prepare()
    if fm:
        check_fallback_weights(s, o, inst)
    s.construct()
    o.construct()
    gt, gl, gc, gu = _compare_tours(s, o)
    # every tour is a feasible-or-reported route set with exact integer cost
    for a in range(len(gl)):
        t = gt[a, :gl[a]]
        assert t[0] == 0 and t[-1] == 0 and (t == 0).sum() == inst.K + 1
        cust = t[t > 0]
        assert len(np.unique(cust)) == len(cust) == inst.n - gu[a]
        assert o.route_cost(t) == gc[a]
    s.close()


# two ants per wavefront (k_construct_pair) takes candidate lists of up to 32
PAIR_CASES = [n for n, (_, kw) in CASES.items() if kw["cand_k"] <= 32]


@pytest.mark.parametrize("name", PAIR_CASES)
@pytest.mark.parametrize("stage,fm,dist,ring", [(1, 1, 0, 0), (0, 1, 0, 0), (0, 0, 0, 1),
                                                (0, 1, 1, 1), (1, 1, 0, 1)])
def test_pair_layout_matches_oracle(gpu, name, stage, fm, dist, ring, monkeypatch):
    """Two ants per wavefront: four epochs give the oracle's tours, costs,
    unvisited counts, colony state and pheromone, with rows staged in LDS or
    read from global memory, with and without the all-pairs fallback
    weights, in both distance modes, with the tours built in LDS or in
    registers (stored out in 128-byte chunks)."""
    monkeypatch.setenv("ACO_HIP_TOUR_RING", str(ring))
    inst, s, o = make_pair(name, fallback_matrix=fm, dist_mode=dist)
    try:
        s.set_bucket_shape(0, 4, stage, ants_per_wave=2)
    except AcoError:
        pytest.skip("staged rows do not fit the LDS with two ants per wavefront")
    assert s.buckets()[0]["ants_per_wave"] == 2
    assert s.stage_lds == stage
    s.prepare()
    for e in range(4):
        s.epochs(1)
        o.epochs(1)
        _compare_tours(s, o)
        gs, os_ = s.colony_state(), o.state()
        for key in ("epoch", "best_ant", "best_len", "feasible_ants", "global_best_cost"):
            assert gs[key] == os_[key], (e, key, gs[key], os_[key])
    tg, cg = s.tau()
    to, co = o.tau()
    assert np.array_equal(tg, to) and np.array_equal(cg, co)
    s.close()


@pytest.mark.parametrize("n,K,seed", [(24, 2, 2), (40, 3, 3)])
def test_epochs_without_a_feasible_ant(gpu, n, K, seed):
    """Heavy demands (U[30..70] or U[20..60]) and a capacity with no slack
    leave most epochs without a feasible ant: the reference then neither
    evaporates nor deposits (aco_cuda.cu:355-396).  Tours, colony state and
    pheromone equal the oracle's after every epoch, and both kinds of epoch
    occur."""
    dmin = 30 if n == 24 else 20
    dmax = 70 if n == 24 else 60
    inst = Instance.uniform(n, K, dmin=dmin, dmax=dmax, slack_pct=100, seed=seed)
    p = Params(cand_k=8, ants=4, rho=0.3)
    o = O.Oracle(inst, cand_k=8, ants=4, rho=0.3, seed=7)
    fe = []
    with HipSolver(inst, p, seeds=[7]) as s:
        s.prepare()
        for e in range(10):
            s.epochs(1)
            o.epochs(1)
            _compare_tours(s, o)
            gs, os_ = s.colony_state(), o.state()
            for key in ("epoch", "best_ant", "feasible_ants", "global_best_cost", "iter_since_best"):
                assert gs[key] == os_[key], (e, key, gs[key], os_[key])
            tg, cg = s.tau()
            to, co = o.tau()
            assert np.array_equal(tg, to) and np.array_equal(cg, co), e
            fe.append(gs["feasible_ants"])
    o.close()
    assert 0 in fe and max(fe) > 0, fe  # both kinds of epoch occurred


@pytest.mark.parametrize("ring", [0, 1])
@pytest.mark.parametrize("ants,apg", [(33, 4), (7, 1), (2, 3), (1, 2)])
def test_pair_layout_ragged_ant_counts(gpu, ants, apg, ring, monkeypatch):
    """Odd ant counts leave the last wavefront with one ant; workgroups past
    the colony's ants exit; the tours still equal the one-ant layout's."""
    monkeypatch.setenv("ACO_HIP_TOUR_RING", str(ring))
    mk, kw = CASES["c1_n100"]
    inst = mk()
    p = Params(**dict(kw, ants=ants))
    out = []
    for apw in (1, 2):
        with HipSolver(inst, p, seeds=[99]) as s:
            s.set_bucket_shape(0, apg, -1, ants_per_wave=apw)
            assert s.buckets()[0]["ants_per_wave"] == apw
            s.prepare()
            s.epochs(4)
            out.append(s.tours())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", PAIR_CASES)
@pytest.mark.parametrize("stage,fm,ring", [(0, 1, 0), (0, 1, 1), (1, 1, 1), (0, 0, 1)])
def test_pair_byte_states_match_oracle(gpu, name, stage, fm, ring, monkeypatch):
    """Two ants per wavefront with byte states (every demand of these
    instances is at most 254; the room a state is compared with is capped at
    254, below the visited mark 255): four epochs give the oracle's tours,
    costs, colony state and pheromone, with rows staged or global, tours in
    LDS or in registers, with and without the all-pairs fallback weights.
    n300_clustered and n1030_two_chunks have capacities above 254, so the
    cap is exercised."""
    monkeypatch.setenv("ACO_HIP_STATE8", "1")
    monkeypatch.setenv("ACO_HIP_TOUR_RING", str(ring))
    inst, s, o = make_pair(name, fallback_matrix=fm)
    try:
        s.set_bucket_shape(0, 4, stage, ants_per_wave=2)
    except AcoError:
        pytest.skip("staged rows do not fit the LDS with two ants per wavefront")
    b = s.buckets()[0]
    assert b["ants_per_wave"] == 2 and b["state_bytes"] == 1
    assert s.stage_lds == stage
    s.prepare()
    for e in range(4):
        s.epochs(1)
        o.epochs(1)
        _compare_tours(s, o)
        gs, os_ = s.colony_state(), o.state()
        for key in ("epoch", "best_ant", "best_len", "feasible_ants", "global_best_cost"):
            assert gs[key] == os_[key], (e, key, gs[key], os_[key])
    tg, cg = s.tau()
    to, co = o.tau()
    assert np.array_equal(tg, to) and np.array_equal(cg, co)
    s.close()


def test_pair_byte_states_refused_for_heavy_demands(gpu, monkeypatch):
    """A demand above 254 keeps u16 states even when byte states are asked
    for, and the tours still equal the oracle's."""
    monkeypatch.setenv("ACO_HIP_STATE8", "1")
    inst = Instance.uniform(120, 6, dmin=200, dmax=400, seed=17)
    assert int(np.max(inst.demand)) > 254
    p = Params(cand_k=16, ants=32, rho=0.2)
    with HipSolver(inst, p, seeds=[5]) as s:
        s.set_bucket_shape(0, 2, 0, ants_per_wave=2)
        b = s.buckets()[0]
        assert b["ants_per_wave"] == 2 and b["state_bytes"] == 2
        o = O.Oracle(inst, cand_k=16, ants=32, rho=0.2, seed=5)
        s.prepare()
        s.epochs(3)
        o.epochs(3)
        _compare_tours(s, o)
        assert s.colony_state()["global_best_cost"] == o.state()["global_best_cost"]
        o.close()


@pytest.mark.parametrize("name,fm", [("c1_n100", 1), ("n300_clustered", 1), ("smoke_unit", 1),
                                     ("n1030_two_chunks", 1), ("n1030_two_chunks", 0),
                                     ("n2100_k6_fallbacks", 1), ("short_routes", 1)])
def test_epochs_match_oracle(gpu, name, fm):
    inst, s, o = make_pair(name, fallback_matrix=fm)
    s.prepare()
    empty = single = 0
    for e in range(6):
        s.epochs(1)
        o.epochs(1)
        gt, gl, _, _ = s.tours()
        b = o.state()["best_ant"]
        if b >= 0:
            t = gt[b, :gl[b]]
            empty += int(np.any((t[:-1] == 0) & (t[1:] == 0)))
            single += int(np.any((t[:-2] == 0) & (t[2:] == 0) & (t[1:-1] != 0)))
        gs, os_ = s.colony_state(), o.state()
        for key in ("epoch", "iter_since_best", "found", "best_ant", "best_len",
                    "feasible_ants", "iter_best_cost", "global_best_cost", "improved"):
            assert gs[key] == os_[key], (e, key, gs[key], os_[key])
        assert np.float32(gs["dep_iter"]) == np.float32(os_["dep_iter"])
        tg, cg = s.tau()
        to, co = o.tau()
        assert np.array_equal(tg, to), f"tau differs after epoch {e}"
        assert np.array_equal(cg, co), f"tau_cand differs after epoch {e}"
        # the score cache equals taupow[tau_cand] * eta[d] (one fp32 multiply)
        t = o.tables()
        _, dist, _ = o.candidates()
        want = t["taupow"][co] * t["eta"][dist]
        assert np.array_equal(s.scores().view(np.uint32), want.astype(np.float32).view(np.uint32))
        if fm:
            check_fallback_weights(s, o, inst)
    if o.best()[1] is not None:
        bt, bc = s.best()
        ot, oc = o.best()
        assert bc == oc and np.array_equal(bt, ot)
    if name == "short_routes":
        assert empty > 0 and single > 0, (empty, single)
    s.close()


@pytest.mark.parametrize("name", ["short_routes", "c1_n100"])
def test_fused_update_matches_oracle(gpu, name, monkeypatch):
    """The fused evaporate + single-workgroup update kernel (the path of
    instances too large for the LDS tables of the multi-workgroup deposit)
    gives the oracle's pheromone and state too."""
    monkeypatch.setenv("ACO_HIP_FUSED_UPDATE", "1")
    inst, s, o = make_pair(name)
    s.prepare()
    for e in range(5):
        s.epochs(1)
        o.epochs(1)
        gs, os_ = s.colony_state(), o.state()
        for key in ("epoch", "iter_since_best", "best_ant", "global_best_cost", "best_len"):
            assert gs[key] == os_[key], (e, key)
        assert np.array_equal(s.tau()[0], o.tau()[0]), f"tau differs after epoch {e}"
        check_fallback_weights(s, o, inst)
    s.close()


@pytest.mark.parametrize("name", ["c1_n100", "n300_clustered"])
def test_learned_eta_matches_oracle(gpu, name):
    """The learned-prior seam: a per-slot heuristic replaces (1/d)^beta in
    the candidate roulette of the GPU path and of the oracle alike."""
    inst, s, o = make_pair(name)
    s.prepare()
    k = s.cand_k(0)
    rng = np.random.default_rng(17)
    eta = rng.uniform(1e-6, 1e-3, size=(inst.n + 1, k)).astype(np.float32)
    eta[rng.random(eta.shape) < 0.1] = 0.0  # some slots never drawn
    s.set_candidate_eta(eta)
    o.set_candidate_eta(eta)
    for e in range(4):
        s.epochs(1)
        o.epochs(1)
        _compare_tours(s, o)
        assert np.array_equal(s.tau()[0], o.tau()[0]), f"tau differs after epoch {e}"
        t = o.tables()
        want = (t["taupow"][o.tau()[1]] * eta).astype(np.float32)
        assert np.array_equal(s.scores().view(np.uint32), want.view(np.uint32))
    s.set_candidate_eta(None)  # back to (1/d)^beta
    t = o.tables()
    _, dist, _ = o.candidates()
    want = (t["taupow"][o.tau()[1]] * t["eta"][dist]).astype(np.float32)
    assert np.array_equal(s.scores().view(np.uint32), want.view(np.uint32))
    s.close()


def test_ants_per_group_invariance(gpu):
    ref = None
    for apg in (1, 3, 8):
        inst, s, o = make_pair("c1_n100", ants_per_group=apg)
        s.prepare()
        s.epochs(3)
        t = s.tours()
        if ref is None:
            ref = t
        else:
            for a, b in zip(ref, t):
                assert np.array_equal(a, b)
        s.close()


def test_route_costs(gpu):
    inst, s, o = make_pair("n300_clustered")
    s.prepare()
    s.construct()
    tours, lens, cost, _ = s.tours()
    got = s.route_costs(tours, lens)
    assert np.array_equal(got, cost)
    rng = np.random.default_rng(5)
    perm = np.concatenate([[0], rng.permutation(np.arange(1, inst.n + 1)), [0]]).astype(np.int32)
    got = s.route_costs(perm[None, :], np.array([len(perm)], np.int32))
    assert got[0] == o.route_cost(perm)
    s.close()


def test_deposit_tour_conflicts(gpu):
    """Single-customer routes (0,c,0) and empty routes (0,0) touch the same
    pheromone entries twice: the result must equal the sequential order."""
    inst, s, o = make_pair("c1_n100")
    s.prepare()
    rng = np.random.default_rng(9)
    tau = rng.integers(0, 256, size=(inst.n + 1, inst.n + 1), dtype=np.uint8)
    s.set_tau(tau)
    o.set_tau(tau)
    check_fallback_weights(s, o, inst)
    tour = [0, 5, 0, 0, 7, 9, 11, 0, 0, 13, 0, 2, 3, 0]
    for delta in (0.0013, 0.05, 3.0):
        s.deposit_tour(tour, delta)
        o.deposit_tour(tour, delta)
        assert np.array_equal(s.tau()[0], o.tau()[0])
        assert np.array_equal(s.tau()[1], o.tau()[1])
        check_fallback_weights(s, o, inst)
    s.close()


def test_run_stagnation_matches_oracle(gpu):
    inst, s, o = make_pair("c1_n100")
    s.prepare()
    r = s.run(stagnation_epochs=15, min_rel_improvement=1e-3, max_epochs=400, chunk_epochs=8,
              use_graph=True)
    o.set_stop_rules(stagnation=15, min_rel=1e-3)
    o.epochs(400)
    gs, os_ = s.colony_state(), o.state()
    assert gs["epoch"] == os_["epoch"] == r["epochs"]
    assert gs["global_best_cost"] == os_["global_best_cost"]
    assert np.array_equal(s.best()[0], o.best()[0])
    assert np.array_equal(s.tau()[0], o.tau()[0])
    s.close()


def test_consecutive_runs_continue(gpu):
    """A second run with a later fixed-epoch target continues the colony
    (warmup then timed region, as bench.py does)."""
    inst, s, o = make_pair("c1_n100")
    s.prepare()
    assert s.run(fixed_epochs=5, chunk_epochs=4)["epochs"] == 5
    assert s.run(fixed_epochs=12, chunk_epochs=8)["epochs"] == 7
    assert s.run(fixed_epochs=12, chunk_epochs=8)["epochs"] == 0
    o.set_stop_rules(fixed=12)
    o.epochs(12)
    assert s.colony_state()["epoch"] == 12
    assert s.colony_state()["global_best_cost"] == o.state()["global_best_cost"]
    assert np.array_equal(s.tau()[0], o.tau()[0])
    s.close()


def test_graph_survives_a_change_of_stop_rules(gpu):
    """Runs with one chunk length replay one captured epoch graph while their
    stop rules change (the rules are read from device memory at each epoch's
    end): fixed epochs 4, then an epoch cap of 12, then a stagnation stop of 3
    epochs under a cap of 60.  Epoch counts, halting and the colony equal the
    oracle under the same sequence of rules."""
    inst, s, o = make_pair("c1_n100")
    s.prepare()
    assert s.run(fixed_epochs=4, chunk_epochs=4)["epochs"] == 4
    assert s.run(max_epochs=12, chunk_epochs=4)["epochs"] == 8
    o.set_stop_rules(fixed=12)
    o.epochs(12)
    assert s.colony_state()["epoch"] == 12
    assert s.colony_state()["global_best_cost"] == o.state()["global_best_cost"]
    got = s.run(stagnation_epochs=3, max_epochs=60, chunk_epochs=4)["epochs"]
    o.set_stop_rules(stagnation=3)
    want = 0
    while not o.state()["halted"] and want < 48:
        o.epochs(1)
        want += 1
    gs, os_ = s.colony_state(), o.state()
    assert got == want and gs["epoch"] == os_["epoch"], (got, want)
    for key in ("iter_since_best", "global_best_cost", "best_ant", "halted"):
        assert gs[key] == os_[key], key
    assert np.array_equal(s.tau()[0], o.tau()[0])
    s.close()


def test_batch_matches_single(gpu):
    """Instances batched in one context with two colonies each give the
    tours of each instance run alone (the Philox stream depends only on the
    instance seed, the colony index, the epoch, the ant and the decision)."""
    insts = [Instance.uniform(120, 6, seed=21), Instance.clustered(200, 8, seed=22)]
    p = Params(cand_k=16, ants=32, colonies=2, rho=0.1)
    b = HipSolver(insts, p, seeds=[101, 202])
    b.prepare()
    b.epochs(3)
    for i, inst in enumerate(insts):
        single = HipSolver(inst, Params(cand_k=16, ants=32, colonies=1, rho=0.1),
                           seeds=[[101, 202][i]])
        single.prepare()
        single.epochs(3)
        ts = single.tours(0)
        tb = b.tours(i * 2)
        for x, y in zip(ts, tb):
            assert np.array_equal(x, y)
        assert np.array_equal(single.tau(0)[0], b.tau(i * 2)[0])
        single.close()
    b.close()


def test_cli_smoke_instance(gpu, tmp_path):
    path = tmp_path / "smoke.vrp"
    smoke_instance().write(str(path))
    env = dict(os.environ, ACO_SOLVER_TIMEOUT_SECONDS="10", ACO_SOLVER_STAGNATION_EPOCHS="10")
    out = subprocess.run([N.CLI_PATH, str(path), "2", "0", "1234"], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert sum(1 for l in lines if l.startswith("Route ")) == 2
    assert any(l.startswith("Cost: ") for l in lines)
    assert any(l.startswith("best cost: ") for l in lines)


def test_large_clustered_matches_oracle(gpu):
    """The n=5000 clustered shape (k=32 rows read from global memory, the
    fallback over five 1024-node chunks, a 25M-entry pheromone matrix): two
    epochs of 8 ants equal the oracle's tours, states and pheromone."""
    inst = Instance.clustered(5000, 40, side=10000, clusters=25, sigma=300, dmin=1, dmax=100,
                              slack_pct=105, seed=3)
    p = Params(cand_k=32, ants=8, rho=0.1)
    with HipSolver(inst, p, seeds=[77]) as s:
        assert s.stage_lds == 0
        o = O.Oracle(inst, cand_k=32, ants=8, rho=0.1, seed=77)
        s.prepare()
        for e in range(2):
            s.epochs(1)
            o.epochs(1)
            gt, gl, gc, gu = s.tours()
            ot, ol, oc, ou = o.tours()
            assert np.array_equal(gl, ol) and np.array_equal(gc, oc) and np.array_equal(gu, ou)
            for a in range(len(gl)):
                assert np.array_equal(gt[a, :gl[a]], ot[a, :ol[a]]), (e, a)
            gs, os_ = s.colony_state(), o.state()
            for key in ("epoch", "best_ant", "feasible_ants", "global_best_cost"):
                assert gs[key] == os_[key], (e, key)
            assert np.array_equal(s.tau()[0], o.tau()[0]), e
