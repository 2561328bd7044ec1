"""GPU tests of batched solving: size buckets, per-instance candidate
lengths, the launch-shape autotuner and its cache.  The tours of every
instance of a batch equal those of the instance solved alone (the draws
depend only on the instance seed, colony, epoch, ant and decision), so
batching, bucketing and launch shapes are pure scheduling."""
import os

import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import tuner
import oracle as O

pytestmark = pytest.mark.gpu


def _batch():
    insts = [Instance.uniform(60, 2, seed=31), Instance.clustered(300, 4, side=1000, sigma=30,
                                                                  dmin=1, dmax=10, seed=32),
             Instance.uniform(700, 6, seed=33), Instance.clustered(1500, 12, side=1000, sigma=30,
                                                                   dmin=1, dmax=10, seed=34)]
    ks = [10, 16, 20, 12]
    seeds = [301, 302, 303, 304]
    return insts, ks, seeds


def test_buckets_match_single_instance_runs(gpu):
    insts, ks, seeds = _batch()
    p = Params(cand_k=16, ants=32, colonies=2, rho=0.1)
    b = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    bk = b.buckets()
    assert len(bk) == 4
    assert sum(x["instances"] for x in bk) == 4 and sum(x["colonies"] for x in bk) == 8
    assert [b.cand_k(i) for i in range(4)] == ks
    b.prepare()
    b.epochs(3)
    for i, inst in enumerate(insts):
        single = HipSolver(inst, Params(cand_k=ks[i], ants=32, colonies=1, rho=0.1),
                           seeds=[seeds[i]])
        single.prepare()
        single.epochs(3)
        for x, y in zip(single.tours(0), b.tours(2 * i)):
            assert np.array_equal(x, y)
        assert np.array_equal(single.tau(0)[0], b.tau(2 * i)[0])
        assert single.colony_state(0)["global_best_cost"] == b.colony_state(2 * i)["global_best_cost"]
        single.close()
    b.close()


def test_autotune_is_pure_scheduling(gpu, tmp_path):
    insts, ks, seeds = _batch()
    p = Params(cand_k=16, ants=64, rho=0.1)
    ref = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    ref.prepare()
    ref.epochs(2)
    s = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    s.prepare()
    cache = str(tmp_path / "tune.json")
    shapes = tuner.tune_shapes(s, cache=cache, launches=2)
    assert all(x["tuned_ms"] > 0 for x in shapes)
    assert os.path.exists(cache)
    s.epochs(2)
    for c in range(len(insts)):
        for x, y in zip(ref.tours(c), s.tours(c)):
            assert np.array_equal(x, y)
    # a second solver takes the shapes from the cache without timing
    s2 = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    s2.prepare()
    shapes2 = tuner.tune_shapes(s2, cache=cache)
    assert [(x["ants_per_group"], x["stage_lds"]) for x in shapes2] == \
        [(x["ants_per_group"], x["stage_lds"]) for x in shapes]
    assert all(x["tuned_ms"] == 0 for x in shapes2)
    for h in (ref, s, s2):
        h.close()


def test_bucket_shape_requests(gpu):
    insts, ks, seeds = _batch()
    s = HipSolver(insts, Params(cand_k=16, ants=32, rho=0.1), seeds=seeds, cand_k=ks)
    s.set_bucket_shape(0, 2, 1)
    b0 = s.buckets()[0]
    assert b0["ants_per_group"] == 2 and b0["stage_lds"] == 1
    s.set_bucket_shape(3, 8, 0)
    assert s.buckets()[3]["stage_lds"] == 0
    # at most 16 wavefronts per workgroup (8 unless the pairs keep byte states)
    with pytest.raises(Exception):
        s.set_bucket_shape(0, 17, 1)
    s.close()


def test_many_colonies_match_oracle(gpu):
    """More colonies than the flat evaporation stages offsets for in LDS
    (1040 > 1024): 520 small instances, two colonies each.  Sampled colonies
    equal the oracle run of that instance and colony alone (pheromone,
    candidate-aligned copy, best cost) after three epochs, so the update
    kernels address every colony's range correctly on that path too."""
    rng = np.random.default_rng(5)
    insts = [Instance.uniform(int(rng.integers(5, 31)), 2, seed=600 + i) for i in range(520)]
    seeds = [7000 + i for i in range(520)]
    p = Params(cand_k=8, ants=8, colonies=2, rho=0.5)
    b = HipSolver(insts, p, seeds=seeds)
    assert b.num_colonies == 1040
    b.prepare()
    b.epochs(3)
    for i, c in ((0, 0), (1, 1), (257, 0), (518, 1), (519, 0), (519, 1)):
        o = O.Oracle(insts[i], cand_k=8, ants=8, rho=0.5, seed=seeds[i], colony=c)
        o.epochs(3)
        tg, cg = b.tau(2 * i + c)
        to, co = o.tau()
        assert np.array_equal(tg, to), (i, c)
        assert np.array_equal(cg, co), (i, c)
        assert b.colony_state(2 * i + c)["global_best_cost"] == o.state()["global_best_cost"], (i, c)
    b.close()


def test_bucket_chains_match_context_wide_updates(gpu, monkeypatch):
    """One construct/evaporate/deposit chain per size bucket (the default for
    several buckets) gives the same tours, pheromone and states as the
    context-wide update kernels after all buckets (ACO_HIP_SERIAL_UPDATES=1)."""
    insts, ks, seeds = _batch()
    p = Params(cand_k=16, ants=32, colonies=2, rho=0.1)
    monkeypatch.delenv("ACO_HIP_SERIAL_UPDATES", raising=False)
    a = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    monkeypatch.setenv("ACO_HIP_SERIAL_UPDATES", "1")
    b = HipSolver(insts, p, seeds=seeds, cand_k=ks)
    assert len(a.buckets()) == 4
    for s in (a, b):
        s.prepare()
        s.run(fixed_epochs=5, chunk_epochs=5)
    for c in range(a.num_colonies):
        for x, y in zip(a.tours(c), b.tours(c)):
            assert np.array_equal(x, y)
        assert np.array_equal(a.tau(c)[0], b.tau(c)[0]) and np.array_equal(a.tau(c)[1], b.tau(c)[1])
        assert a.colony_state(c) == b.colony_state(c)
    a.close()
    b.close()


def test_lanes_are_pure_scheduling(gpu, monkeypatch):
    """A size class split into two lanes (two chains of launches on two
    streams, the second starting after the first's first construct, forked
    and joined once per graph chunk) gives exactly the tours, pheromone and
    colony states of the same class run as one bucket."""
    insts = [Instance.uniform(150, 3, seed=400 + i) for i in range(12)]
    seeds = [500 + i for i in range(12)]
    p = Params(cand_k=12, ants=256, rho=0.1)
    out = []
    for lanes in (2, 1):
        monkeypatch.setenv("ACO_HIP_LANES", str(lanes))
        with HipSolver(insts, p, seeds=seeds) as s:
            bk = s.buckets()
            assert len(bk) == lanes and all(b["lanes"] == lanes for b in bk)
            s.prepare()
            s.run(fixed_epochs=6, chunk_epochs=3, use_graph=True)
            out.append([(s.tours(c), s.tau(c), s.colony_state(c)) for c in range(len(insts))])
    for c, (a, b) in enumerate(zip(*out)):
        for x, y in zip(a[0], b[0]):
            assert np.array_equal(x, y), c
        assert np.array_equal(a[1][0], b[1][0]) and np.array_equal(a[1][1], b[1][1]), c
        assert a[2] == b[2], c


def test_counter_variants_match(gpu):
    """The counter-collecting construct variants (plain workgroup order, no
    XCD renumbering) build the same tours as the production variants of a
    batched context, epoch after epoch, and report one decision per step of
    every tour."""
    insts, ks, seeds = _batch()
    p = Params(cand_k=16, ants=24, rho=0.1, colonies=2)
    out = []
    for stats in (False, True):
        with HipSolver(insts, p, seeds=seeds, cand_k=ks) as s:
            s.prepare()
            if stats:
                s.set_stats(True)
            s.run(fixed_epochs=4, chunk_epochs=2, use_graph=not stats)
            out.append([s.tours(c) for c in range(len(insts) * 2)])
            if stats:
                st = s.stats(0)
                assert st["ants"] == 4 * p.ants
                assert st["decisions"] > 0
    for c, (a, b) in enumerate(zip(out[0], out[1])):
        for x, y in zip(a, b):
            assert np.array_equal(x, y), c
