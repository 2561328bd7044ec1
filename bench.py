"""Benchmark of the ACO-CVRP hot path on one MI355X.

Workload (plain ACO at the reference's benchmark shape): n=1000 customers,
K=8 vehicles, integer coordinates U{0..1000}^2, depot (500,500), demands
U{1..10}, Q=ceil(1.1*sum(d)/8), 256 ants, candidate list 20, quantized
pheromone, alpha=1, beta=2, rho=0.1, stagnation stop at 80 epochs, over
seeds 1..32 (seed s sets the instance and the solver seed), the 32 seeds
solved together in one context, one colony each.  A step is one epoch of
every colony: every ant builds a tour, then best selection, evaporation and
both deposits.  ant-tours/s = 256 * (epochs summed over the colonies) / wall
seconds of the epoch loop (instances, kNN and pheromone already resident in
HBM).  The single-seed figure (one colony of 256 ants) is the
`single_seed` leg.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vehicleroutingproblem_amd import workloads as W  # noqa: E402

N_CUST, K_VEH, ANTS, CAND_K, RHO = 1000, 8, 256, 20, 0.1
HBM_PEAK_GBS = 8000.0
# HBM bytes the construct phase of one epoch of the n=1000 workload moves, by
# number of seeds (1: one launch; 32: one launch per lane, two lanes), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes over this bench, each scaled by
# what the counter reads for a 256 MiB device copy (fetch x2.0, write x1.0);
# the commands and figures are in profiles/k_construct_n1000_m256_hbm_traffic.md
# (1 seed) and profiles/k_construct_n1000_32seeds_hbm_traffic.md (32 seeds).
CONSTRUCT_TRAFFIC_BYTES = {1: 22631044, 32: 505230354}
# Instructions the construct launch issues per ant-decision at this workload
# (VALU + SALU + LDS + branch SQ counters over decisions;
# profiles/k_construct_n1000_m256_sq_counters.md for one seed, one ant per
# wavefront; profiles/k_construct_n1000_32seeds_sq_counters.md for 32 seeds,
# two ants per wavefront, whose wavefront makes two decisions per pass): a
# lone wavefront issues at most one instruction per 4-cycle slot of its SIMD,
# so 4 x this is the issue bound of one decision in shader cycles.
CONSTRUCT_INSTS_PER_DECISION = {1: 96.5, 32: 50.2}


def tour_bytes(n, K, k):
    """Algorithmic bytes of one ant-tour: S = n+K decisions, each reading k
    candidates (4 B index + 4 B heuristic + 1 B pheromone) and writing one
    4 B route entry."""
    return (n + K) * (9 * k + 4)


def cpu_baseline(inst, seconds_budget=12.0, threads=0):
    """The CPU oracle (a C port of the same algorithm) on the host cores:
    whole epochs of 256 ants, OpenMP over ants (`threads`, default all the
    host's), for about `seconds_budget`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    o = O.Oracle(inst, cand_k=CAND_K, ants=ANTS, rho=RHO, seed=1)
    o.epochs(1, threads)  # warm caches
    t0 = time.perf_counter()
    epochs = 0
    while time.perf_counter() - t0 < seconds_budget and epochs < 5000:
        o.epochs(1, threads)
        epochs += 1
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(ANTS * epochs / dt, 1), "unit": "ant-tours/s", "cores": threads,
            "kind": "port",
            "sample": f"{epochs} epochs x {ANTS} ants of the n=1000 workload (seed 1, "
                      f"{dt:.1f} s), oracle/aco_oracle.c with OpenMP over ants"}


def recorded_reference_baselines(port_value: float, port_single: float):
    """The reference CLIs as timed in the build container, where their
    sources are (tests/golden/ref_cpu_baselines.json, written by
    tests/golden/make_golden.py; no reference binary runs on this host).
    Each carries its speed relative to the port timed the same way on the
    same cores there; `estimate_here` scales this host's port figure by that
    ratio (an estimate: the two programs may scale differently across CPU
    models)."""
    path = os.path.join(ROOT, "tests", "golden", "ref_cpu_baselines.json")
    if not os.path.exists(path):
        return None
    doc = json.load(open(path))
    out = {"measured_on": doc["host"], "port_there": doc["port"]["value"],
           "port_single_thread_there": doc["port_single_thread"]["value"]}
    for key in ("reference_openmp", "reference_openmp_width20"):
        r = dict(doc[key])
        r["estimate_here"] = round(port_value * r["over_port"], 1)
        out[key] = r
    r = dict(doc["reference_sequential"])
    r["estimate_here"] = round(port_single * r["over_port_single_thread"], 1)
    out["reference_sequential"] = r
    return out


# algorithmic flops per edge and training step: forward 2(8*64 + 64*64 + 64),
# backward 2(64*64 [dW2] + 64*64 [dH1] + 64*8 [dW1] + 64 [dW3])
PRIOR_FLOPS_PER_EDGE_STEP = 2 * (8 * 64 + 64 * 64 + 64) + 2 * (64 * 64 * 2 + 64 * 8 + 64)


# the edge-scoring network trains on the kNN-20 graph of its training
# instances; the test solves with candidate lists of 16
PRIOR_GRAPH_K = 20


def learned_prior_leg(budget_s: float, seeds: int, train_count: int):
    """Batched learned ACO against batched plain ACO at equal wall-clock on
    n=2000, K=16, 256 ants, k=16 instances (half uniform, half clustered):
    the prior trains on the kNN-20 edges of `train_count` seeded n in
    [200, 2000] instances, labelled by their elite tours (harvested with
    plain ACO on 20-candidate lists, 100 epochs, 64 ants), then the
    `seeds` test instances run together in one batched context for `budget_s`
    seconds, once plain and once with the prior applied to every instance (the
    learned run's budget includes the prior's feature pass and inference)."""
    from vehicleroutingproblem_amd import HipSolver, Instance, Params
    from vehicleroutingproblem_amd.prior import train_prior, training_instances
    t0 = time.perf_counter()
    train = training_instances(train_count, seed=7)
    prior = train_prior(train, epochs=100, steps=0, batch=512, cand_k=PRIOR_GRAPH_K)
    t_harvest = time.perf_counter() - t0
    edges = prior.size()
    steps = 600
    t0 = time.perf_counter()
    prior.train(steps, lr=2e-3)
    t_train = time.perf_counter() - t0
    flops = edges * steps * PRIOR_FLOPS_PER_EDGE_STEP
    test = [W.c5_instance(t) for t in range(seeds)]
    res = {}
    for use in (False, True):
        with HipSolver(test, Params(cand_k=16, ants=256, rho=0.1),
                       seeds=[77 + t for t in range(seeds)]) as sv:
            sv.prepare()
            sv.sync()
            ts = time.perf_counter()
            if use:
                for i in range(seeds):
                    prior.apply(sv, instance=i)
            sv.run(timeout_seconds=budget_s - (time.perf_counter() - ts), max_epochs=10 ** 6,
                   chunk_epochs=16)
            res[use] = ([sv.best(i)[1] for i in range(seeds)], sv.colony_state(0)["epoch"])
    prior.close()
    plain, learned = res[False][0], res[True][0]
    gains = [(a - b) / a for a, b in zip(plain, learned)]
    return {"test": f"{seeds} instances n=2000 K=16 m=256 k=16 in one batched context, half "
                    "uniform / half clustered (sigma 30)",
            "budget_s": budget_s, "seeds": seeds,
            "plain_best": plain, "learned_best": learned,
            "plain_epochs": res[False][1], "learned_epochs": res[True][1],
            "mean_gain_pct": round(100 * sum(gains) / len(gains), 3),
            "instances_improved": sum(1 for g in gains if g > 0),
            "train_instances": train_count, "train_edges": edges,
            "harvest_s": round(t_harvest, 3), "train_steps": steps,
            "train_s": round(t_train, 3),
            "train_tflops": round(flops / t_train / 1e12, 2),
            "mfma_peak_tflops_bf16": 2500.0}


def c2_to_stagnation_leg(seeds: int = W.C2_SEEDS):
    """The n=1000 workload over seeds 1..`seeds` in one context, every colony
    run to its own 80-epoch stagnation stop (relative improvement 1e-3):
    whole-run ant-tours/s, epochs per colony, best costs and the fraction of
    the last epoch's ants whose tour is feasible."""
    from vehicleroutingproblem_amd import HipSolver
    sd = list(range(1, seeds + 1))
    insts = [W.c2_instance(x) for x in sd]
    with HipSolver(insts, W.c2_params(), seeds=sd) as s:
        s.prepare()
        s.sync()
        t0 = time.perf_counter()
        s.run(stagnation_epochs=80, max_epochs=5000, chunk_epochs=16)
        s.sync()
        dt = time.perf_counter() - t0
        ep = [s.colony_state(i)["epoch"] for i in range(seeds)]
        best = [s.colony_state(i)["global_best_cost"] for i in range(seeds)]
        feas = [float((s.tours(i)[3] == 0).mean()) for i in range(seeds)]
    return {"workload": f"{seeds} seeds of the n=1000 workload in one context, each colony to "
                        "its 80-epoch stagnation stop",
            "ant_tours_per_s": round(W.C2_ANTS * sum(ep) / dt, 1), "seconds": round(dt, 3),
            "epochs_mean": round(sum(ep) / seeds, 1), "epochs_min": min(ep),
            "epochs_max": max(ep), "best_cost_mean": round(sum(best) / seeds, 1),
            "feasible_ant_fraction_last_epoch": round(sum(feas) / seeds, 4)}


def batched_leg(epochs: int, fallback_matrix: int = -1):
    """C4: 512 instances in one context (n ~ U[50, 2000], half uniform,
    half clustered with sigma 30, K = ceil(n/125), demands U[1,10],
    Q = ceil(1.1 sum/K)), 128 ants each, per-instance seeds, candidate
    lengths from the tuned per-size-class table, launch shapes autotuned on
    the device; `epochs` timed epochs in graph replays after 2 warmup."""
    from vehicleroutingproblem_amd import HipSolver
    from vehicleroutingproblem_amd import tuner
    insts = W.c4_instances()
    count = len(insts)
    ks = W.c4_cand_k(insts)
    t0 = time.perf_counter()
    s = HipSolver(insts, W.c4_params(fallback_matrix=fallback_matrix), seeds=W.c4_seeds(),
                  cand_k=ks)
    s.prepare()
    s.sync()
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    shapes = tuner.tune_shapes(s, cache=None, launches=2)
    t_tune = time.perf_counter() - t0
    s.run(fixed_epochs=2, chunk_epochs=2, use_graph=True)
    s.sync()
    chunk = div_chunk(epochs, 64)
    t0 = time.perf_counter()
    s.run(fixed_epochs=2 + epochs, chunk_epochs=chunk, use_graph=True)
    s.sync()
    dt = time.perf_counter() - t0
    done = s.colony_state(0)["epoch"] - 2
    feasible = sum(1 for c in range(count) if s.colony_state(c)["global_best_cost"] < 2 ** 62)
    s.close()
    k_table = W.tuned_k_table()
    return {"workload": f"{count} instances n~U[50,2000] (half uniform, half clustered sigma 30), "
                        f"K=ceil(n/125), {W.C4_ANTS} ants each",
            "value": round(count * W.C4_ANTS * done / dt, 1), "unit": "ant-tours/s",
            "epochs": done, "ms_per_epoch": round(1e3 * dt / max(done, 1), 3),
            "instances_with_feasible_tour": feasible,
            "cand_k_by_size_class": {str(c): k for c, k in sorted(k_table.items())} or "16",
            "buckets": [{"n1": [b["n1_min"], b["n1_max"]], "instances": b["instances"],
                         "ants_per_group": b["ants_per_group"], "stage_lds": b["stage_lds"],
                         "ants_per_wave": b["ants_per_wave"], "state_bytes": b["state_bytes"],
                         "lanes": b["lanes"],
                         "construct_ms": round(b["tuned_ms"], 3)} for b in shapes],
            "prepare_s": round(t_prep, 3), "autotune_s": round(t_tune, 3)}


# HBM bytes one k_evaporate_flat launch of the n=5000 workload moves with the
# all-pairs fallback weights on: rocprofv3 FETCH_SIZE + WRITE_SIZE, each scaled
# by what the counter reads for a 256 MiB device copy (fetch x2.0, write x1.0);
# commands and figures in profiles/k_evaporate_flat_n5000_m1024.md.
C3_EVAP_TRAFFIC_BYTES = 152800000


def large_clustered_leg(epochs: int = 20):
    """C3: n=5000 clustered (25 Gaussian components, centres U[0,10000]^2,
    sigma 300, depot at the centre), demands U[1,100], K=40,
    Q=ceil(1.05 sum/40), 1024 ants, k=32; `epochs` timed epochs after 3, with
    the all-pairs fallback weight matrix on (the library's choice wherever it
    fits) and off.  The evaporation's roofline counts its algorithmic bytes,
    2(n+1)^2 (each pheromone level read and written once); the weight
    matrix the kernel also rewrites is reported beside it as traffic."""
    from vehicleroutingproblem_amd import HipSolver
    from vehicleroutingproblem_amd import tuner
    inst = W.c3_instance()
    runs = {}
    for fm in (1, 0):
        with HipSolver(inst, W.c3_params(fallback_matrix=fm), seeds=[1]) as s:
            s.prepare()
            shapes = tuner.tune_shapes(s, cache=None, launches=4)
            s.run(fixed_epochs=3, chunk_epochs=3)
            s.sync()
            t0 = time.perf_counter()
            s.run(fixed_epochs=3 + epochs, chunk_epochs=div_chunk(epochs, 64))
            s.sync()
            dt = time.perf_counter() - t0
            done = s.colony_state()["epoch"] - 3
            best = s.colony_state()["global_best_cost"]
            # per-kernel durations: 8 more epochs, eager launches with hipEvents
            s.set_timing(True)
            s.run(fixed_epochs=3 + epochs + 8, chunk_epochs=8, use_graph=False)
            s.sync()
            t = s.timing()
            s.set_timing(False)
        ep = max(1, t["construct_launches"])
        runs[fm] = {"value": round(W.C3_ANTS * done / dt, 1), "unit": "ant-tours/s",
                    "ms_per_epoch": round(1e3 * dt / max(done, 1), 3), "epochs": done,
                    "best_cost_after": best,
                    "ants_per_group": shapes[0]["ants_per_group"],
                    "stage_lds": shapes[0]["stage_lds"],
                    "kernel_ms_per_epoch": {
                        "construct": round(t["construct_ms"] / ep, 4),
                        "evaporate": round(t["evaporate_ms"] / ep, 4),
                        "select_and_deposit": round((t["best_ms"] + t["deposit_ms"]) / ep, 4)},
                    "evaporate_avg_ms": t["evaporate_ms"] / max(1, t["evaporate_launches"])}
    n1 = inst.n + 1
    S = (n1 + 15) // 16 * 16
    out = dict(runs[1])
    evap_ms = out.pop("evaporate_avg_ms")
    alg = 2 * n1 * n1
    achieved = alg / (evap_ms * 1e-3) / 1e9
    out["workload"] = ("n=5000 K=40 clustered (sigma 300 in [0,10000]^2), demands U[1,100], "
                       "Q=ceil(1.05 sum/40), 1024 ants, k=32")
    out["evaporation_roofline"] = {
        "bound": "hbm", "kernel": "k_evaporate_flat", "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "algorithmic_bytes_per_launch": alg, "avg_launch_ms": round(evap_ms, 4),
        "traffic": C3_EVAP_TRAFFIC_BYTES,
        "weight_matrix_rewrite_bytes": 4 * n1 * S,
        "note": "with the fallback weight matrix on, the kernel also rewrites taupow[q]*eta "
                "for every pair (4 B each), about 2x the pheromone bytes"}
    off = runs[0]
    off_evap = off.pop("evaporate_avg_ms")
    off["evaporate_achieved_gbs"] = round(alg / (off_evap * 1e-3) / 1e9, 1)
    out["fallback_matrix_off"] = off
    out["fallback_matrix_on_faster"] = runs[1]["ms_per_epoch"] <= runs[0]["ms_per_epoch"]
    return out


def real_tour_cost(inst, tour) -> float:
    """Sum of real (double) Euclidean edge lengths along a depot-delimited
    tour: the cost the reference CLI prints."""
    import numpy as np
    x = np.asarray(inst.x, np.float64)[tour]
    y = np.asarray(inst.y, np.float64)[tour]
    return float(np.hypot(np.diff(x), np.diff(y)).sum())


def tour_is_feasible(inst, tour) -> bool:
    """Starts and ends at the depot, every customer exactly once, every
    route's load <= Q, at most K routes."""
    import numpy as np
    t = np.asarray(tour, np.int64)
    if len(t) < 2 or t[0] != 0 or t[-1] != 0:
        return False
    cust = np.sort(t[t != 0])
    if len(cust) != inst.n or (cust != np.arange(1, inst.n + 1)).any():
        return False
    routes, load = 0, 0
    for a, b in zip(t[:-1], t[1:]):
        if b == 0:
            if a != 0:
                routes += 1
            load = 0
        else:
            load += int(inst.demand[b])
            if load > inst.capacity:
                return False
    return routes <= inst.K


def quality_leg():
    """CPU-baseline leg: solution quality against the reference's sequential
    solver at equal epochs.  The reference sequential CLI ran each of 32
    seeded unit-demand instances (n=100, K=8, Q=ceil(1.2 n/8), 64 ants) to its
    50-epoch stagnation stop in the build container; its best cost and epoch
    count are the fixture tests/golden/ref_seq_quality.json.  This solver
    runs exactly as many epochs with the same parameters (alpha 1, beta 2,
    rho 0.5, tau0 1, Q 1, k=32) on the same metric (real-distance mode) and
    its best tour is costed in real (double) distances like the reference's
    print."""
    from vehicleroutingproblem_amd import HipSolver, Instance, Params
    from vehicleroutingproblem_amd import _native as N
    path = os.path.join(ROOT, "tests", "golden", "ref_seq_quality.json")
    if not os.path.exists(path):
        return None
    doc = json.load(open(path))
    runs = [r for r in doc["runs"] if r["set"] == "bench"]
    ref, ours, epochs = [], [], []
    feasible = True
    for r in runs:
        inst = Instance.uniform(100, 8, side=1000, dmin=1, dmax=1, slack_pct=120,
                                seed=r["instance_seed"])
        with HipSolver(inst, Params(cand_k=32, ants=doc["ants"], alpha=1.0, beta=2.0, rho=0.5,
                                    tau0=1.0, Q=1.0, dist_mode=N.DIST_REAL),
                       seeds=[r["solver_seed"]]) as s:
            s.prepare()
            s.run(fixed_epochs=r["epochs"], chunk_epochs=div_chunk(r["epochs"], 64))
            tour, _ = s.best()
        feasible &= tour_is_feasible(inst, tour)
        ours.append(real_tour_cost(inst, tour))
        ref.append(r["best_cost"])
        epochs.append(r["epochs"])
    diff = [(a - b) / b for a, b in zip(ours, ref)]
    return {"workload": f"{len(runs)} unit-demand uniform instances, n=100, K=8, "
                        f"Q=ceil(1.2n/8), {doc['ants']} ants; reference sequential solver to its "
                        "stagnation stop (50), recorded in the build container; this solver for "
                        "the same number of epochs, both on real (unrounded) distances",
            "mean_best_cost_diff_pct": round(100 * sum(diff) / len(diff), 3),
            "all_tours_feasible": feasible,
            "mean_epochs": round(sum(epochs) / len(epochs), 1),
            "reference_mean_cost": round(sum(ref) / len(ref), 3),
            "mean_cost": round(sum(ours) / len(ours), 3)}


def reference_campaign_leg(budget_s: float = 2.0):
    """The reference's own benchmark instances n500_k8_s19000 and
    n1000_k8_s19001 (copied in tests/golden) with the reference CUDA
    campaign's settings (m=256, solver seeds 1234-1236, k=32), solved by the
    campaign runner: all six runs in one batched context under one wall-clock
    budget (ACO_SOLVER_TIMEOUT_SECONDS), against the best costs and per-run
    elapsed times that campaign recorded
    (tests/golden/ref_cuda_backend_sizes_n500_n1000.csv).  Both sides
    optimise and report real (unrounded) route lengths."""
    import csv
    import tempfile
    from vehicleroutingproblem_amd import campaign
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "ref_cuda_backend_sizes_n500_n1000.csv"), newline="") as f:
        recorded = {(r["name"], int(r["run_id"])): r for r in csv.DictReader(f)}
    keys = ("ACO_SOLVER_TIMEOUT_SECONDS", "ACO_SOLVER_STAGNATION_EPOCHS",
            "ACO_SOLVER_FIXED_EPOCHS", "ACO_HIP_CAND_K", "ACO_HIP_DIST")
    saved = {k: os.environ.get(k) for k in keys}
    try:
        for k in keys:
            os.environ.pop(k, None)
        os.environ["ACO_SOLVER_TIMEOUT_SECONDS"] = str(budget_s)
        with tempfile.TemporaryDirectory() as td:
            manifest = os.path.join(td, "manifest.csv")
            campaign.write_manifest(manifest, [
                {"profile": "generated", "name": name, "instance_path": f"ref_{name}.vrp",
                 "n": n, "K": 8, "m": 256, "solver_seed": 1234, "instance_seed": seed,
                 "layout_id": "grid100", "capacity_formula": "ceil((1+20/100)*n/K)"}
                for name, n, seed in (("n500_k8_s19000", 500, 19000),
                                      ("n1000_k8_s19001", 1000, 19001))])
            rows = campaign.solve(manifest, os.path.join(td, "runs.csv"), os.path.join(td, "sol"),
                                  repeats=3, root=golden)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    runs, diffs = [], []
    for r in rows:
        ref = recorded[(r["name"], r["run_id"])]
        ok = r["status"] == "ok"
        if ok:
            diffs.append((float(r["best_cost"]) - float(ref["best_cost"])) / float(ref["best_cost"]))
        runs.append({"name": r["name"], "seed": r["solver_seed"],
                     "best_cost": float(r["best_cost"]) if ok else None,
                     "reference_cuda_best_cost": float(ref["best_cost"]),
                     "reference_cuda_elapsed_s": float(ref["elapsed_s"])})
    return {"workload": "reference instances n500_k8_s19000 and n1000_k8_s19001 (unit demand, "
                        "K=8), m=256, seeds 1234-1236, k=32, rho=0.5, real distances: the "
                        "reference CUDA campaign's recorded runs against one batched solve of "
                        "all six",
            "budget_s_all_six_runs": budget_s,
            "batch_elapsed_s": float(rows[0]["elapsed_s"]) if rows and rows[0]["elapsed_s"] else None,
            "all_ok": all(r["status"] == "ok" for r in rows),
            "mean_best_cost_diff_pct": round(100 * sum(diffs) / len(diffs), 3) if diffs else None,
            "runs": runs}


def div_chunk(epochs: int, cap: int) -> int:
    """Largest divisor of `epochs` up to `cap`: graph replays of that many
    epochs end exactly on the last epoch."""
    return max([c for c in range(1, min(cap, max(epochs, 1)) + 1) if epochs % c == 0] or [1])


def c2_run(seeds, steps: int, warmup: int, chunk_cap: int, ants_per_group: int = 0) -> dict:
    """The n=1000 workload (256 ants, k=20, rho 0.1, stagnation stop at 80)
    for the given seeds in one context (one colony per seed): `warmup`
    epochs, then exactly `steps` timed epochs of the solver loop in graph
    replays.  Per-kernel durations come from a second context that replays
    the same epochs bit for bit with hipEvents around every launch (eager
    launches on the context stream), device counters from a third."""
    from vehicleroutingproblem_amd import HipSolver
    seeds = list(seeds)
    insts = [W.c2_instance(x) for x in seeds]
    p = W.c2_params(ants_per_group=ants_per_group)
    # the library replays at most 64 epochs per graph (its per-epoch progress ring)
    chunk = div_chunk(steps, min(chunk_cap, 64))

    def fresh():
        r_ = HipSolver(insts, p, seeds=seeds)
        r_.prepare()
        if warmup > 0:
            # the timed run's chunk length, so its epoch graph is captured here: the
            # epochs of a chunk past max_epochs find their colonies halted and return
            r_.run(stagnation_epochs=80, max_epochs=warmup, chunk_epochs=chunk, use_graph=True)
        r_.sync()
        return r_

    s = fresh()
    t0 = time.perf_counter()
    s.run(stagnation_epochs=80, max_epochs=warmup + steps, chunk_epochs=chunk, use_graph=True)
    s.sync()
    dt = time.perf_counter() - t0
    states = [s.colony_state(i) for i in range(len(seeds))]
    epochs = [st["epoch"] - warmup for st in states]
    feasible_last = sum(float((s.tours(i)[3] == 0).mean()) for i in range(len(seeds))) / len(seeds)
    shape = (s.ants_per_group, s.stage_lds, s.buckets()[0]["ants_per_wave"])
    lanes = len(s.buckets())
    s.close()

    def replay(timing: bool):
        r_ = fresh()
        if timing:
            r_.set_timing(True)
        else:
            r_.set_stats(True)
        r_.run(stagnation_epochs=80, max_epochs=warmup + steps, chunk_epochs=chunk,
               use_graph=not timing)
        r_.sync()
        for i, st in enumerate(states):
            assert r_.colony_state(i)["global_best_cost"] == st["global_best_cost"]
        if timing:
            out_ = r_.timing()
        else:
            out_ = {}
            for i in range(len(seeds)):
                for k_, v_ in r_.stats(i).items():
                    out_[k_] = v_ if k_ == "wall_clock_hz" else out_.get(k_, 0) + v_
        r_.close()
        return out_

    t = replay(True)
    assert t["construct_launches"] == max(epochs), t
    dev = replay(False)
    launches = max(1, t["construct_launches"])
    construct_ms = t["construct_ms"] / launches
    bytes_per_launch = len(seeds) * W.C2_ANTS * tour_bytes(W.C2_N, W.C2_K, W.C2_CAND_K)
    achieved = bytes_per_launch / (construct_ms * 1e-3) / 1e9
    cpd = dev["ant_cycles"] / max(1, dev["decisions"])
    return {
        "value": round(W.C2_ANTS * sum(epochs) / dt, 1),
        "ms_per_step": round(1e3 * dt / max(epochs), 4),
        "epochs_per_s": round(max(epochs) / dt, 1),
        "epochs_timed": max(epochs),
        "feasible_ant_fraction_last_epoch": round(feasible_last, 4),
        "best_cost": [st["global_best_cost"] for st in states][:8],
        "ants_per_group": shape[0], "stage_lds": shape[1], "ants_per_wave": shape[2],
        "lanes": lanes,
        "roofline": {"bound": "hbm", "kernel": "k_construct_pair" if shape[2] == 2 else "k_construct",
                     "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": CONSTRUCT_TRAFFIC_BYTES.get(len(seeds)),
                     "avg_launch_ms": round(construct_ms, 5),
                     "bytes_per_launch": bytes_per_launch,
                     "concurrent_launches": lanes,
                     "note": ("the epoch's construct phase: one launch per lane, the lanes' launches "
                              "running concurrently on their own streams; avg_launch_ms spans "
                              "all of them (first start to last end) and bytes_per_launch and "
                              "traffic are their sum") if lanes > 1 else
                             "one construct launch per epoch"},
        "kernel_ms_per_epoch": {
            "construct": round(construct_ms, 5),
            "evaporate": round(t["evaporate_ms"] / launches, 5),
            "select_and_deposit": round((t["best_ms"] + t["deposit_ms"]) / launches, 5)},
        "construct_counters": {
            "decisions_per_tour": round(dev["decisions"] / max(1, dev["ants"]), 1),
            "fallback_rate": round(dev["fallbacks"] / max(1, dev["decisions"]), 4),
            "fallback_cycle_share": round(dev["fallback_cycles"] / max(1, dev["ant_cycles"]), 4),
            "cycles_per_decision": round(cpd, 1),
            "cycles_per_ant": round(dev["ant_cycles"] / max(1, dev["ants"])),
            "insts_per_decision": CONSTRUCT_INSTS_PER_DECISION.get(len(seeds)),
            "issue_bound_cycles_per_decision": (
                4 * CONSTRUCT_INSTS_PER_DECISION[len(seeds)]
                if len(seeds) in CONSTRUCT_INSTS_PER_DECISION else None)},
        "update_phases_us": ({
            k: round(1e6 * dev[k + "_ticks"] / max(1, dev["wall_clock_hz"]) / dev["updates"], 2)
            for k in ("evap", "select", "deposit")} if dev["updates"] else None),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seeds", type=int, default=W.C2_SEEDS,
                    help="seeds of the n=1000 workload solved together (one colony each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ants-per-group", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=64,
                    help="max epochs per graph replay (the largest divisor of --steps up to it, "
                         "at most 64)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-seed n=1000 leg")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the stagnation-stop, batched 512-instance and n=5000 legs")
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the quality comparison with the reference sequential solver "
                         "(part of the CPU baselines, so --no-cpu skips it too)")
    ap.add_argument("--no-campaign", action="store_true",
                    help="skip the reference-instance campaign leg")
    ap.add_argument("--no-prior", action="store_true",
                    help="skip the learned-prior vs plain ACO leg")
    ap.add_argument("--prior-seconds", type=float, default=3.0,
                    help="wall-clock budget per test run of the learned-prior leg")
    ap.add_argument("--prior-train-count", type=int, default=4096,
                    help="seeded n in [200, 2000] instances whose elite tours train the "
                         "learned prior")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device twice (calibrates the HBM "
                         "byte counters of a profiler against a copy of known size)")
    ap.add_argument("--train-probe", type=int, default=0, metavar="STEPS",
                    help="only harvest 16 training instances and train the edge prior "
                         "STEPS steps (the network kernels alone, for counter runs)")
    ap.add_argument("--large-probe", type=int, default=0, metavar="EPOCHS",
                    help="only run the n=5000 clustered workload (1024 ants, k=32, one ant "
                         "per workgroup) for EPOCHS eager epochs (its kernels alone, for "
                         "counter runs)")
    ap.add_argument("--batched-probe", type=int, default=0, metavar="EPOCHS",
                    help="only run the batched 512-instance workload for EPOCHS timed epochs "
                         "(its kernels alone, for profiler runs)")
    ap.add_argument("--headline-probe", type=int, default=0, metavar="EPOCHS",
                    help="only run the n=1000 workload over --seeds seeds for EPOCHS eager "
                         "epochs (its kernels alone, for counter runs)")
    args = ap.parse_args()
    if args.headline_probe:
        from vehicleroutingproblem_amd import HipSolver
        sd = list(range(1, args.seeds + 1))
        with HipSolver([W.c2_instance(x) for x in sd], W.c2_params(), seeds=sd) as s:
            s.prepare()
            s.run(fixed_epochs=args.headline_probe, chunk_epochs=args.headline_probe,
                  use_graph=False)
            s.sync()
            print(json.dumps({"headline_probe_epochs": s.colony_state()["epoch"],
                              "seeds": len(sd)}))
        return
    if args.batched_probe:
        print(json.dumps(batched_leg(args.batched_probe)))
        return
    if args.large_probe:
        from vehicleroutingproblem_amd import HipSolver
        with HipSolver(W.c3_instance(), W.c3_params(ants_per_group=1), seeds=[1]) as s:
            s.prepare()
            s.run(fixed_epochs=args.large_probe, chunk_epochs=args.large_probe, use_graph=False)
            s.sync()
            print(json.dumps({"large_probe_epochs": s.colony_state()["epoch"],
                              "best_cost": s.colony_state()["global_best_cost"]}))
        return
    if args.train_probe:
        from vehicleroutingproblem_amd.prior import train_prior, training_instances
        prior = train_prior(training_instances(16, seed=7), epochs=30, steps=0)
        prior.train(2, lr=2e-3)  # warm-up; train() returns after the device is done
        t0 = time.perf_counter()
        prior.train(args.train_probe, lr=2e-3)
        dt = time.perf_counter() - t0
        flops = prior.size() * args.train_probe * PRIOR_FLOPS_PER_EDGE_STEP
        print(json.dumps({"train_probe_steps": args.train_probe, "edges": prior.size(),
                          "train_s": round(dt, 4), "train_tflops": round(flops / dt / 1e12, 2)}))
        prior.close()
        return
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "copies": 2}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a built path.", file=sys.stderr)
        sys.exit(2)

    seeds = list(range(1, args.seeds + 1))
    head = c2_run(seeds, args.steps, args.warmup, args.chunk, args.ants_per_group)
    out = {
        "metric": "ant-tours/s",
        "value": head.pop("value"),
        "unit": "ant-tours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u8",
        "data": "synthetic",
        "config": {"workload": f"plain ACO n=1000 K=8 uniform [0,1000]^2 demands U[1,10] "
                               f"Q=ceil(1.1*sum/8), 256 ants, k=20, rho 0.1, quantized "
                               f"pheromone, stagnation stop 80; seeds 1..{len(seeds)} (seed s "
                               f"sets instance and solver) solved together, one colony each; "
                               f"a step is one epoch of every colony",
                   "n": W.C2_N, "K": W.C2_K, "ants": W.C2_ANTS, "cand_k": W.C2_CAND_K,
                   "rho": W.C2_RHO, "seeds": len(seeds),
                   "ants_per_group": head.pop("ants_per_group"),
                   "stage_lds": head.pop("stage_lds"),
                   "ants_per_wave": head.pop("ants_per_wave"),
                   "lanes": head.pop("lanes")},
        "roofline": head.pop("roofline"),
    }
    out.update(head)
    if not args.no_single and len(seeds) > 1:
        single = c2_run([1], max(args.steps, 20), args.warmup, args.chunk, args.ants_per_group)
        single["workload"] = "the same n=1000 workload, seed 1 only (one colony of 256 ants)"
        out["single_seed"] = single
    if not args.no_batch:
        out["c2_to_stagnation"] = c2_to_stagnation_leg(len(seeds))
        out["batched"] = batched_leg(20)
        out["large_clustered"] = large_clustered_leg(20)
    if not args.no_prior:
        out["learned_prior"] = learned_prior_leg(args.prior_seconds, 32, args.prior_train_count)
    if not args.no_campaign:
        out["reference_campaign"] = reference_campaign_leg(2.0)
    if not args.no_cpu:
        # the CPU baseline: the port on this host's cores (the reference's
        # own CLIs were timed in the build container, beside the port)
        inst = W.c2_instance(1)
        out["cpu_baseline"] = cpu_baseline(inst)
        single = cpu_baseline(inst, 5.0, threads=1)
        out["cpu_baseline"]["port_single_thread"] = single
        rec = recorded_reference_baselines(out["cpu_baseline"]["value"], single["value"])
        if rec:
            out["cpu_baseline"]["reference_recorded"] = rec
        if not args.no_quality:
            out["cpu_baseline"]["quality_vs_reference_seq"] = quality_leg()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
