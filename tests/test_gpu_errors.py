"""Error behaviour of the C ABI (GPU): every entry point returns a status
mirroring AcoStatus (include/aco.h:82-88) with a message from
aco_hip_last_error, instead of the reference's void launchers whose errors
surface at the next device synchronisation (src/cuda/aco_cuda.cu:19-27).

Bad instances and parameters are refused by aco_hip_load before any device
work; instances that admit no feasible tour run normally and report
ACO_HIP_ERR_NO_SOLUTION for the best tour, as the reference returns
ACO_ERR_NO_SOLUTION when no ant was ever feasible (src/cuda/aco_cuda.cu:453)."""
import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N
from vehicleroutingproblem_amd.solver import AcoError

pytestmark = pytest.mark.gpu

P = dict(cand_k=8, ants=16, rho=0.5)


def _base():
    return Instance.uniform(12, 3, seed=7)


def _with(inst, **kw):
    f = dict(x=inst.x.copy(), y=inst.y.copy(), demand=inst.demand.copy(), K=inst.K,
             capacity=inst.capacity)
    f.update(kw)
    return Instance(f["x"], f["y"], f["demand"], f["K"], f["capacity"])


def _set(a, i, v):
    a = a.copy()
    a[i] = v
    return a


BAD_INSTANCES = {
    "depot_only": (lambda b: Instance([500], [500], [0], 1, 10), "1 <= n <= 65534"),
    "no_vehicles": (lambda b: _with(b, K=0), "1 <= K <= 65535"),
    # (K-1)*Q would overflow the device's int32 remaining-capacity test
    "too_many_vehicles": (lambda b: _with(b, K=70000), "1 <= K <= 65535"),
    "zero_capacity": (lambda b: _with(b, capacity=0), "1 <= capacity <= 32767"),
    "capacity_too_large": (lambda b: _with(b, capacity=40000), "1 <= capacity <= 32767"),
    "x_out_of_range": (lambda b: _with(b, x=_set(b.x, 3, 40000)), "coordinates"),
    "negative_y": (lambda b: _with(b, y=_set(b.y, 5, -1)), "coordinates"),
    "depot_demand": (lambda b: _with(b, demand=_set(b.demand, 0, 2)), "depot demand"),
    "demand_too_large": (lambda b: _with(b, demand=_set(b.demand, 4, 40000)), "demand out of"),
    "negative_demand": (lambda b: _with(b, demand=_set(b.demand, 4, -3)), "demand out of"),
}


@pytest.mark.parametrize("name", sorted(BAD_INSTANCES))
def test_load_rejects_bad_instances(gpu, name):
    make, text = BAD_INSTANCES[name]
    with pytest.raises(AcoError) as e:
        HipSolver(make(_base()), Params(**P))
    assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
    assert text in str(e.value), str(e.value)


BAD_PARAMS = {
    "cand_k_0": (dict(cand_k=0), "cand_k"),
    "cand_k_65": (dict(cand_k=65), "cand_k"),
    "ants_0": (dict(ants=0), "ants"),
    "rho_0": (dict(rho=0.0), "rho"),
    "rho_1": (dict(rho=1.0), "rho"),
    "tau0_0": (dict(tau0=0.0), "tau0"),
    "colonies_0": (dict(colonies=0), "colonies"),
    "colonies_too_many": (dict(colonies=(1 << 20) + 1), "colonies"),
    "dist_mode_2": (dict(dist_mode=2), "dist_mode"),
}


@pytest.mark.parametrize("name", sorted(BAD_PARAMS))
def test_load_rejects_bad_params(gpu, name):
    kw, text = BAD_PARAMS[name]
    p = dict(P)
    p.update(kw)
    with pytest.raises(AcoError) as e:
        HipSolver(_base(), Params(**p))
    assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
    assert text in str(e.value), str(e.value)


def test_per_instance_cand_k_range(gpu):
    with pytest.raises(AcoError) as e:
        HipSolver([_base()], Params(**P), cand_k=[65])
    assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT


@pytest.mark.parametrize("kind", ["heavy_customer", "capacity_short"])
def test_infeasible_instances_report_no_solution(gpu, kind):
    b = _base()
    if kind == "heavy_customer":
        # one customer heavier than a vehicle: it can never be visited
        inst = _with(b, demand=_set(b.demand, 6, b.capacity + 1))
    else:
        # unit demands, K*Q below the demand total
        n = b.n
        inst = _with(b, demand=np.r_[0, np.ones(n, np.int32)], K=2, capacity=n // 2 - 1)
    with HipSolver(inst, Params(**P), seeds=[3]) as s:
        s.prepare()
        s.run(fixed_epochs=6, chunk_epochs=3)
        _, ln, cost, unv = s.tours(0)
        assert (unv >= 1).all()
        assert s.colony_state(0)["global_best_cost"] >= 2 ** 62
        with pytest.raises(AcoError) as e:
            s.best(0)
        assert e.value.status == N.ACO_HIP_ERR_NO_SOLUTION
        # every ant still returns a closed, well-formed partial tour
        tours, ln, _, _ = s.tours(0)
        for a in range(s.ants):
            t = tours[a, :ln[a]]
            assert t[0] == 0 and t[-1] == 0
            cust = t[t != 0]
            assert len(np.unique(cust)) == len(cust)
            assert len(cust) == inst.n - unv[a]


def test_bad_indices_and_tours(gpu):
    inst = _base()
    with HipSolver(inst, Params(**P), seeds=[5]) as s:
        s.prepare()
        s.run(fixed_epochs=2, chunk_epochs=2)
        for call in (lambda: s.colony_state(5), lambda: s.tours(-1), lambda: s.set_epoch(0, colony=7)):
            with pytest.raises(AcoError) as e:
                call()
            assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
        with pytest.raises(AcoError) as e:  # node out of range
            s.deposit_tour([0, 1, inst.n + 1, 0], 0.5)
        assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
        with pytest.raises(AcoError) as e:  # too short
            s.deposit_tour([0], 0.5)
        assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
        tau0 = s.tau()[0].copy()
        for bad in ([0, 3, 5, 3, 0], [0, 2, 4, 2, 0, 0], [1, 2, 0], [0, 2, 4]):
            with pytest.raises(AcoError) as e:  # repeated customer / not depot-delimited
                s.deposit_tour(bad, 0.5)
            assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT, bad
        assert np.array_equal(s.tau()[0], tau0)  # nothing was deposited
        tours = np.zeros((2, 8), np.int32)
        with pytest.raises(AcoError) as e:  # a length past the tour stride
            s.route_costs(tours, np.array([2, 9], np.int32))
        assert e.value.status == N.ACO_HIP_ERR_INVALID_INPUT
        # the context stays usable after refused calls
        s.run(fixed_epochs=4, chunk_epochs=2)
        tour, cost = s.best(0)
        assert cost > 0 and tour[0] == 0
