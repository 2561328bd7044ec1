"""Benchmark of the ACO-CVRP hot path on one MI355X.

Workload (plain ACO at the reference's benchmark shape): n=1000 customers,
K=8 vehicles, integer coordinates U{0..1000}^2, depot (500,500), demands
U{1..10}, Q=ceil(1.1*sum(d)/8), 256 ants, candidate list 20, quantized
pheromone, alpha=1, beta=2, rho=0.1.  A step is one epoch: every ant builds
a tour, then best selection, evaporation and both deposits.  ant-tours/s =
256 * epochs / wall seconds of the epoch loop (instance, kNN and pheromone
already resident in HBM).

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

N_CUST, K_VEH, ANTS, CAND_K, RHO = 1000, 8, 256, 20, 0.1
HBM_PEAK_GBS = 8000.0
# HBM bytes one k_construct launch of this workload moves, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over this bench, each scaled by what the
# counter reads for a 256 MiB device copy (fetch x2.0, write x1.0); the
# commands and figures are in profiles/k_construct_n1000_m256_hbm_traffic.md.
CONSTRUCT_TRAFFIC_BYTES = 20869168
# Instructions k_construct issues per decision at this workload (VALU + SALU +
# LDS + branch SQ counters over decisions, profiles/k_construct_n1000_m256_sq_counters.md):
# a lone wavefront issues at most one instruction per 4-cycle slot of its SIMD,
# so 4 x this is the issue bound of one decision in shader cycles.
CONSTRUCT_INSTS_PER_DECISION = 103.0


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
    return {"value": round(ANTS * epochs / dt, 1), "unit": "ant-tours/s", "cores": threads,
            "kind": "port",
            "sample": f"{epochs} epochs x {ANTS} ants of the same n=1000 workload "
                      f"({dt:.1f} s), oracle/aco_oracle.c with OpenMP over ants"}


def reference_openmp_baseline(seconds_budget=8.0, width20=False):
    """The reference OpenMP CLI (built from its sources by oracle/ref.mk) on a
    unit-demand n=1000, K=8 instance of the same generator (the reference
    parser rejects weighted demand), 256 ants, ACO_SOLVER_FIXED_EPOCHS; its own
    "Time:" line (epoch loop only).  Its candidate width comes from the host's
    L3 size (k = 0.7 L3 / ((n+1) 8), clamped to [16, 512]); width20 runs the
    same solver with the L3 probe answered "224K" (oracle/ref_omp_l3_driver.c),
    which gives k = 20, this solver's width.  Reported beside the port; None
    when the binary was not built."""
    import re
    import subprocess
    import tempfile
    from vehicleroutingproblem_amd import Instance
    exe = os.path.join(ROOT, "oracle", "_ref", "aco_omp_l3" if width20 else "aco_omp")
    if not os.path.exists(exe):
        return None
    inst = Instance.uniform(N_CUST, K_VEH, side=1000, dmin=1, dmax=1, slack_pct=120, seed=1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "unit_n1000_k8.vrp")
        inst.write(path)

        def run(epochs):
            env = dict(os.environ, ACO_SOLVER_FIXED_EPOCHS=str(epochs),
                       ACO_SOLVER_PROGRESS_INTERVAL_SECONDS="0", OMP_NUM_THREADS=str(threads))
            if width20:
                env["REF_L3_SIZE"] = "224K"
            out = subprocess.run([exe, path, str(K_VEH), str(ANTS), "1"], env=env,
                                 capture_output=True, text=True, timeout=300).stdout
            m = re.search(r"Time: ([0-9.]+)s", out)
            return float(m.group(1)) if m else None

        t = run(4)
        if not t:
            return None
        epochs = max(4, min(400, int(4 * seconds_budget / max(t, 1e-3))))
        t = run(epochs)
    if not t:
        return None
    return {"value": round(ANTS * epochs / t, 1), "unit": "ant-tours/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference src/openmp-mpi CLI, {epochs} fixed epochs x {ANTS} ants, "
                      f"unit-demand n=1000 K=8 (real distances, rho=0.5, "
                      + ("candidate width 20: L3 probe answered 224K" if width20 else
                         "L3-tuned candidate width as shipped")
                      + f"), {t:.2f} s solver time"}


def reference_sequential_baseline(seconds_budget=8.0):
    """The reference sequential CLI (aco_sequential.c, built from its sources
    by oracle/ref.mk; candidate list = all n customers, one core) on the same
    unit-demand n=1000, K=8 instance and 256 ants, stopped by its timeout:
    epochs completed over the elapsed time of its last progress line.  None
    when the binary was not built."""
    import re
    import subprocess
    import tempfile
    from vehicleroutingproblem_amd import Instance
    exe = os.path.join(ROOT, "oracle", "_ref", "aco_seq")
    if not os.path.exists(exe):
        return None
    inst = Instance.uniform(N_CUST, K_VEH, side=1000, dmin=1, dmax=1, slack_pct=120, seed=1)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "unit_n1000_k8.vrp")
        inst.write(path)
        env = dict(os.environ, ACO_SOLVER_TIMEOUT_SECONDS=str(seconds_budget),
                   ACO_SOLVER_PROGRESS_INTERVAL_SECONDS="0.5")
        out = subprocess.run([exe, path, str(K_VEH), str(ANTS), "1"], env=env,
                             capture_output=True, text=True, timeout=300).stdout
    prog = re.findall(r"epoca=([0-9]+) .*?tempo_trascorso=([0-9.]+)s", out)
    if not prog:
        return None
    epochs, t = int(prog[-1][0]), float(prog[-1][1])
    if t <= 0:
        return None
    return {"value": round(ANTS * epochs / t, 1), "unit": "ant-tours/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference src/seq CLI, {epochs} epochs x {ANTS} ants in {t:.1f} s "
                      "(its timeout stop), unit-demand n=1000 K=8, candidate list = all "
                      "customers"}


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
    test = []
    for t in range(seeds):
        if t % 2 == 0:
            test.append(Instance.uniform(2000, 16, side=1000, dmin=1, dmax=10, slack_pct=110,
                                         seed=500 + t))
        else:
            test.append(Instance.clustered(2000, 16, side=1000, clusters=25, sigma=30, dmin=1,
                                           dmax=10, slack_pct=110, seed=500 + t))
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


def c2_seeds_leg(seeds: int = 32):
    """The n=1000 workload over seeds 1..`seeds` (seed s sets the instance
    and the solver seed), all in one context: once to the 80-epoch
    stagnation stop (relative improvement 1e-3), each colony stopping on its
    own, and once for a fixed 100 epochs.  Reports ant-tours/s over the
    whole context, the epochs each colony ran, best costs and the fraction
    of the last epoch's ants whose tour is feasible."""
    from vehicleroutingproblem_amd import HipSolver, Instance, Params
    insts = [Instance.uniform(1000, 8, side=1000, dmin=1, dmax=10, slack_pct=110, seed=s)
             for s in range(1, seeds + 1)]
    out = {"workload": f"{seeds} seeds of plain ACO n=1000 K=8 uniform, 256 ants, k=20, "
                       "rho 0.1, in one context"}
    for mode in ("stagnation_80", "fixed_100"):
        with HipSolver(insts, Params(cand_k=CAND_K, ants=ANTS, rho=RHO),
                       seeds=list(range(1, seeds + 1))) as s:
            s.prepare()
            s.sync()
            t0 = time.perf_counter()
            if mode == "fixed_100":
                s.run(fixed_epochs=100, chunk_epochs=20)
            else:
                s.run(stagnation_epochs=80, max_epochs=5000, chunk_epochs=16)
            s.sync()
            dt = time.perf_counter() - t0
            ep = [s.colony_state(i)["epoch"] for i in range(seeds)]
            best = [s.colony_state(i)["global_best_cost"] for i in range(seeds)]
            feas = []
            for i in range(seeds):
                _, _, _, unv = s.tours(i)
                feas.append(float((unv == 0).mean()))
        out[mode] = {"ant_tours_per_s": round(ANTS * sum(ep) / dt, 1),
                     "seconds": round(dt, 3),
                     "epochs_mean": round(sum(ep) / seeds, 1), "epochs_min": min(ep),
                     "epochs_max": max(ep),
                     "best_cost_mean": round(sum(best) / seeds, 1),
                     "feasible_ant_fraction_last_epoch": round(sum(feas) / seeds, 4)}
    return out


def batched_leg(epochs: int, count: int = 512, ants: int = 128, fallback_matrix: int = -1):
    """C4: `count` instances in one context (n ~ U[50, 2000], half uniform,
    half clustered with sigma 30, K = ceil(n/125), demands U[1,10],
    Q = ceil(1.1 sum/K)), `ants` ants each, per-instance seeds, candidate
    lengths from the tuned per-size-class table, launch shapes autotuned on
    the device; `epochs` timed epochs in graph replays after 2 warmup."""
    import json as _json
    from vehicleroutingproblem_amd import HipSolver, Params
    from vehicleroutingproblem_amd.prior import training_instances
    from vehicleroutingproblem_amd import tuner
    insts = training_instances(count, nmin=50, nmax=2000, seed=4242)
    table_path = os.path.join(ROOT, "vehicleroutingproblem_amd", "tuned_k.json")
    k_table = {}
    if os.path.exists(table_path):
        k_table = {int(c): k for c, k in _json.load(open(table_path))["k_by_size_class"].items()}
    ks = tuner.candidate_lengths(insts, k_table, default=16)
    t0 = time.perf_counter()
    s = HipSolver(insts, Params(cand_k=16, ants=ants, rho=0.1, fallback_matrix=fallback_matrix),
                  seeds=[9000 + i for i in range(count)], cand_k=ks)
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
    return {"workload": f"{count} instances n~U[50,2000] (half uniform, half clustered sigma 30), "
                        f"K=ceil(n/125), {ants} ants each",
            "value": round(count * ants * done / dt, 1), "unit": "ant-tours/s",
            "epochs": done, "ms_per_epoch": round(1e3 * dt / max(done, 1), 3),
            "instances_with_feasible_tour": feasible,
            "cand_k_by_size_class": {str(c): k for c, k in sorted(k_table.items())} or "16",
            "buckets": [{"n1": [b["n1_min"], b["n1_max"]], "instances": b["instances"],
                         "ants_per_group": b["ants_per_group"], "stage_lds": b["stage_lds"],
                         "construct_ms": round(b["tuned_ms"], 3)} for b in shapes],
            "prepare_s": round(t_prep, 3), "autotune_s": round(t_tune, 3)}


def large_clustered_instance():
    """The n=5000 clustered workload: 25 Gaussian components, centres
    U[0,10000]^2, sigma 300, depot at the centre, demands U[1,100], K=40,
    Q=ceil(1.05 sum/40), seed 1."""
    from vehicleroutingproblem_amd import Instance
    return Instance.clustered(5000, 40, side=10000, clusters=25, sigma=300, dmin=1, dmax=100,
                              slack_pct=105, seed=1)


def large_clustered_leg(epochs: int = 20):
    """C3: n=5000 clustered (25 Gaussian components, centres U[0,10000]^2,
    sigma 300, depot at the centre), demands U[1,100], K=40,
    Q=ceil(1.05 sum/40), 1024 ants, k=32; `epochs` timed epochs after 3."""
    from vehicleroutingproblem_amd import HipSolver, Params
    from vehicleroutingproblem_amd import tuner
    inst = large_clustered_instance()
    with HipSolver(inst, Params(cand_k=32, ants=1024, rho=0.1), seeds=[1]) as s:
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
        try:  # the all-pairs fallback weights exist iff the matrix is on
            s.fallback_weights(0)
            fs_on = True
        except Exception:
            fs_on = False
    ep = max(1, t["construct_launches"])
    evap_ms = t["evaporate_ms"] / max(1, t["evaporate_launches"])
    # evaporation of the (n+1) x S pheromone matrix (S = n+1 rounded up to 16)
    # in 16-byte items: each reads and writes its 16 levels and, with the
    # all-pairs fallback weights on, writes their 16 fp32 weights
    n1 = inst.n + 1
    items = n1 * ((n1 + 15) // 16)
    evap_bytes = items * (32 + (64 if fs_on else 0))
    evap_gbs = evap_bytes / (evap_ms * 1e-3) / 1e9
    return {"workload": "n=5000 K=40 clustered (sigma 300 in [0,10000]^2), demands U[1,100], "
                        "Q=ceil(1.05 sum/40), 1024 ants, k=32",
            "value": round(1024 * done / dt, 1), "unit": "ant-tours/s",
            "ms_per_epoch": round(1e3 * dt / max(done, 1), 3), "epochs": done,
            "best_cost_after": best,
            "ants_per_group": shapes[0]["ants_per_group"], "stage_lds": shapes[0]["stage_lds"],
            "kernel_ms_per_epoch": {
                "construct": round(t["construct_ms"] / ep, 4),
                "evaporate": round(t["evaporate_ms"] / ep, 4),
                "select_and_deposit": round((t["best_ms"] + t["deposit_ms"]) / ep, 4)},
            "evaporation_roofline": {
                "bound": "hbm", "kernel": "k_evaporate_flat", "achieved": round(evap_gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(evap_gbs / HBM_PEAK_GBS, 4),
                "bytes_per_launch": evap_bytes, "avg_launch_ms": round(evap_ms, 4),
                "pheromone_entries": n1 * n1, "fallback_matrix": bool(fs_on)}}


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


def quality_leg(seeds: int = 32, n: int = 100, ants: int = 64):
    """CPU-baseline leg: solution quality against the reference's sequential solver at equal
    epochs: unit-demand uniform instances (the reference parser's format),
    n customers, K=8, Q=ceil(1.2 n/8); the reference sequential CLI (built
    from its sources) runs to its stagnation stop (50 epochs) and reports
    its epochs and best cost; this solver then runs exactly as many epochs
    (alpha 1, beta 2, rho 0.5, tau0 1, Q 1, k=32, both CLIs' defaults) and its
    best tour is costed in real (double) distances like the reference's."""
    import re
    import subprocess
    import tempfile
    import numpy as np
    from vehicleroutingproblem_amd import HipSolver, Instance, Params
    exe = os.path.join(ROOT, "oracle", "_ref", "aco_seq")
    if not os.path.exists(exe):
        return None
    ref, ours, epochs = [], [], []
    feasible = True
    with tempfile.TemporaryDirectory() as td:
        for i in range(seeds):
            inst = Instance.uniform(n, 8, side=1000, dmin=1, dmax=1, slack_pct=120, seed=700 + i)
            path = os.path.join(td, f"u{i}.vrp")
            inst.write(path)
            env = dict(os.environ, ACO_SOLVER_STAGNATION_EPOCHS="50",
                       ACO_SOLVER_TIMEOUT_SECONDS="120", ACO_SOLVER_PROGRESS_INTERVAL_SECONDS="0")
            out = subprocess.run([exe, path, "8", str(ants), str(1234 + i)], env=env,
                                 capture_output=True, text=True, timeout=300).stdout
            m = re.search(r"best cost: ([0-9.]+)", out)
            ep = [int(x) for x in re.findall(r"epoca=([0-9]+)", out)]
            if not m or not ep:
                return None
            ref.append(float(m.group(1)))
            epochs.append(max(ep))
            with HipSolver(inst, Params(cand_k=32, ants=ants, alpha=1.0, beta=2.0, rho=0.5,
                                        tau0=1.0, Q=1.0), seeds=[1234 + i]) as s:
                s.prepare()
                s.run(fixed_epochs=max(ep), chunk_epochs=div_chunk(max(ep), 64))
                tour, _ = s.best()
            feasible &= tour_is_feasible(inst, tour)
            ours.append(real_tour_cost(inst, tour))
    diff = [(a - b) / b for a, b in zip(ours, ref)]
    return {"workload": f"{seeds} unit-demand uniform instances, n={n}, K=8, Q=ceil(1.2n/8), "
                        f"{ants} ants; reference sequential solver to its stagnation stop (50), "
                        "this solver for the same number of epochs",
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
    (tests/golden/ref_cuda_backend_sizes_n500_n1000.csv).  Costs are real
    (unrounded) route lengths on both sides."""
    import csv
    import tempfile
    from vehicleroutingproblem_amd import campaign
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "ref_cuda_backend_sizes_n500_n1000.csv"), newline="") as f:
        recorded = {(r["name"], int(r["run_id"])): r for r in csv.DictReader(f)}
    keys = ("ACO_SOLVER_TIMEOUT_SECONDS", "ACO_SOLVER_STAGNATION_EPOCHS",
            "ACO_SOLVER_FIXED_EPOCHS", "ACO_HIP_CAND_K")
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
                        "K=8), m=256, seeds 1234-1236, k=32, rho=0.5: the reference CUDA "
                        "campaign's recorded runs against one batched solve of all six",
            "budget_s_all_six_runs": budget_s,
            "batch_elapsed_s": float(rows[0]["elapsed_s"]) if rows and rows[0]["elapsed_s"] else None,
            "all_ok": all(r["status"] == "ok" for r in rows),
            "mean_best_cost_diff_pct": round(100 * sum(diffs) / len(diffs), 3) if diffs else None,
            "runs": runs}


def div_chunk(epochs: int, cap: int) -> int:
    """Largest divisor of `epochs` up to `cap`: graph replays of that many
    epochs end exactly on the last epoch."""
    return max([c for c in range(1, min(cap, max(epochs, 1)) + 1) if epochs % c == 0] or [1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ants-per-group", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=64,
                    help="max epochs per graph replay (the largest divisor of --steps up to it)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 32-seed n=1000, batched 512-instance and n=5000 legs")
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
    args = ap.parse_args()
    if args.batched_probe:
        print(json.dumps(batched_leg(args.batched_probe)))
        return
    if args.large_probe:
        from vehicleroutingproblem_amd import HipSolver, Params
        inst = large_clustered_instance()
        with HipSolver(inst, Params(cand_k=32, ants=1024, rho=0.1, ants_per_group=1),
                       seeds=[1]) as s:
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

    from vehicleroutingproblem_amd import HipSolver, Instance, Params

    inst = Instance.uniform(N_CUST, K_VEH, side=1000, dmin=1, dmax=10, slack_pct=110, seed=1)
    p = Params(cand_k=CAND_K, ants=ANTS, rho=RHO, ants_per_group=args.ants_per_group)
    s = HipSolver(inst, p, seeds=[1])
    s.prepare()
    s.sync()
    # warmup (also captures the epoch graph)
    if args.warmup > 0:
        s.run(fixed_epochs=args.warmup, chunk_epochs=div_chunk(args.warmup, args.chunk),
              use_graph=True)
    s.sync()
    # timed region: exactly `steps` epochs of the solver loop, in graph
    # replays of `chunk` epochs (a divisor of steps, so no replay runs past
    # the last epoch)
    chunk = div_chunk(args.steps, args.chunk)
    t0 = time.perf_counter()
    r = s.run(fixed_epochs=args.warmup + args.steps, chunk_epochs=chunk, use_graph=True)
    s.sync()
    dt = time.perf_counter() - t0
    st = s.colony_state()
    epochs = st["epoch"] - args.warmup
    assert epochs == args.steps, (epochs, args.steps)
    value = ANTS * epochs / dt
    feasible_last = float((s.tours(0)[3] == 0).mean())

    # Per-kernel durations over the same epochs as the timed region: a second
    # solver with the same instance and seed replays them bit for bit (the
    # path is deterministic), with hipEvents on its stream around every
    # launch (eager launches).  A third replay collects the device counters.
    def replay(timing: bool, stats: bool):
        r_ = HipSolver(inst, p, seeds=[1])
        r_.prepare()
        if args.warmup > 0:
            r_.run(fixed_epochs=args.warmup, chunk_epochs=div_chunk(args.warmup, args.chunk),
                   use_graph=True)
        r_.sync()
        if timing:
            r_.set_timing(True)
        if stats:
            r_.set_stats(True)
        r_.run(fixed_epochs=args.warmup + args.steps, chunk_epochs=chunk,
               use_graph=not timing)
        r_.sync()
        assert r_.colony_state()["global_best_cost"] == st["global_best_cost"]
        out_ = r_.timing() if timing else r_.stats()
        r_.close()
        return out_

    t = replay(True, False)
    assert t["construct_launches"] == args.steps, t
    dev = replay(False, True)
    construct_ms = t["construct_ms"] / max(1, t["construct_launches"])
    evap_ms = t["evaporate_ms"] / max(1, t["evaporate_launches"])
    k = min(CAND_K, inst.n)
    bytes_per_launch = ANTS * tour_bytes(inst.n, inst.K, k)
    achieved = bytes_per_launch / (construct_ms * 1e-3) / 1e9

    out = {
        "metric": "ant-tours/s",
        "value": round(value, 1),
        "unit": "ant-tours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / epochs, 4),
        "epochs_per_s": round(epochs / dt, 1),
        "feasible_ant_fraction_last_epoch": round(feasible_last, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u8",
        "data": "synthetic",
        "config": {"workload": "plain ACO n=1000 K=8 uniform [0,1000]^2 demands U[1,10] "
                               "Q=ceil(1.1*sum/8), 256 ants, k=20, quantized pheromone",
                   "n": inst.n, "K": inst.K, "ants": ANTS, "cand_k": k, "rho": RHO,
                   "ants_per_group": s.ants_per_group, "stage_lds": s.stage_lds,
                   "best_cost": st["global_best_cost"]},
        "roofline": {"bound": "hbm", "kernel": "k_construct", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": CONSTRUCT_TRAFFIC_BYTES,
                     "avg_launch_ms": round(construct_ms, 5),
                     "bytes_per_launch": bytes_per_launch,
                     "evaporate_avg_ms": round(evap_ms, 5)},
        "kernel_ms_per_epoch": {
            "construct": round(t["construct_ms"] / max(1, t["construct_launches"]), 5),
            "evaporate": round(t["evaporate_ms"] / max(1, t["construct_launches"]), 5),
            "select_and_deposit": round((t["best_ms"] + t["deposit_ms"])
                                        / max(1, t["construct_launches"]), 5)},
        "construct_counters": {
            "decisions_per_tour": round(dev["decisions"] / max(1, dev["ants"]), 1),
            "fallback_rate": round(dev["fallbacks"] / max(1, dev["decisions"]), 4),
            "fallback_cycle_share": round(dev["fallback_cycles"] / max(1, dev["ant_cycles"]), 4),
            "cycles_per_decision": round(dev["ant_cycles"] / max(1, dev["decisions"]), 1),
            "cycles_per_ant": round(dev["ant_cycles"] / max(1, dev["ants"])),
            "staging_cycles_per_group": round(dev["staging_cycles"] * s.buckets()[0]["ants_per_group"]
                                              / max(1, dev["ants"])),
            "issue_bound_cycles_per_decision": 4 * CONSTRUCT_INSTS_PER_DECISION,
            "issue_bound_frac": round(4 * CONSTRUCT_INSTS_PER_DECISION
                                      / max(1e-9, dev["ant_cycles"] / max(1, dev["decisions"])),
                                      3)},
        "update_phases_us": ({
            k: round(1e6 * dev[k + "_ticks"] / max(1, dev["wall_clock_hz"]) / dev["updates"], 2)
            for k in ("evap", "select", "deposit")} if dev["updates"] else None),
    }
    s.close()
    if not args.no_batch:
        out["c2_seeds"] = c2_seeds_leg(32)
        out["batched"] = batched_leg(20)
        out["large_clustered"] = large_clustered_leg(20)
    if not args.no_prior:
        out["learned_prior"] = learned_prior_leg(args.prior_seconds, 32, args.prior_train_count)
    if not args.no_campaign:
        out["reference_campaign"] = reference_campaign_leg(2.0)
    if not args.no_cpu:
        # the CPU baselines: the port, the reference's OpenMP binary, and the
        # reference's sequential solver as the quality baseline at equal epochs
        out["cpu_baseline"] = cpu_baseline(inst)
        out["cpu_baseline"]["port_single_thread"] = cpu_baseline(inst, 5.0, threads=1)
        ref = reference_openmp_baseline()
        if ref:
            out["cpu_baseline"]["reference_openmp"] = ref
        ref20 = reference_openmp_baseline(width20=True)
        if ref20:
            out["cpu_baseline"]["reference_openmp_width20"] = ref20
        seq = reference_sequential_baseline()
        if seq:
            out["cpu_baseline"]["reference_sequential"] = seq
        if not args.no_quality:
            out["cpu_baseline"]["quality_vs_reference_seq"] = quality_leg(32)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
