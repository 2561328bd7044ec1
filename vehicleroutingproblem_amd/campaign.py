"""Instance campaigns and the result pipeline, in the reference's formats.

* ``generate``: seeded weighted-demand instances (TSPLIB CVRP ``.vrp``, the
  library's uniform or clustered generator) and a manifest CSV in the schema
  of the reference's ``instances/generated_benchmark/manifest.csv``:
  ``profile,name,instance_path,n,K,m,solver_seed,instance_seed,layout_id,capacity_formula``.
* ``solve``: every manifest row, ``repeats`` runs each, solved on the GPU.
  Run r of a row uses seed ``solver_seed + r - 1`` and ``m`` ants, as the
  reference campaign does (tools/bash/solve_cuda.sh:104-109).  All runs with
  the same ``m`` go into ONE batched context (one colony per run, launched by
  size bucket), instead of one process per run.  The solver settings are the
  CLI's (alpha 1, beta 2, rho 0.5, tau0 1, Q 1; k = min(ACO_HIP_CAND_K or 32,
  n); stop rules from ACO_SOLVER_*, 100 stagnant epochs when none is set),
  optimising the reference's unrounded length (ACO_HIP_DIST=nint optimises
  the integer TSPLIB length instead).
  Each run writes a solution file in the CLI's text format ("Route i: ..."
  per vehicle, then "Cost: %.3f" and "best cost: %.3f", cli_common.c:49-68;
  the real, unrounded cost like the reference's print) and a CSV row in the
  schema of the reference CUDA campaign (solve_cuda.sh:56,161-165):
  ``name,profile,instance_path,n,K,m,solver_seed,instance_seed,layout_id,run_id,status,elapsed_s,best_cost,error``.
  ``elapsed_s`` is the wall time of the batched solve the run belonged to.
* ``validate``: a solution file against its instance with exact arithmetic:
  every customer exactly once, at most K routes, every route's demand within
  the capacity, and the printed cost equal to the recomputed real cost
  within the print's rounding (the check the reference delegates to pyvrp).

    python -m vehicleroutingproblem_amd.campaign generate --out DIR --sizes 200,500 --seeds 3
    python -m vehicleroutingproblem_amd.campaign solve --manifest DIR/manifest.csv --csv runs.csv --solutions sol/
    python -m vehicleroutingproblem_amd.campaign validate INSTANCE.vrp SOLUTION.txt
"""
import csv
import math
import os
import time

import numpy as np

MANIFEST_FIELDS = ["profile", "name", "instance_path", "n", "K", "m", "solver_seed",
                   "instance_seed", "layout_id", "capacity_formula"]
RESULT_FIELDS = ["name", "profile", "instance_path", "n", "K", "m", "solver_seed",
                 "instance_seed", "layout_id", "run_id", "status", "elapsed_s", "best_cost",
                 "error"]


# ---- manifests ----------------------------------------------------------------

def read_manifest(path: str) -> list:
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    for r in rows:
        missing = [k for k in MANIFEST_FIELDS if k not in r]
        if missing:
            raise ValueError(f"{path}: manifest row lacks {missing}")
        for k in ("n", "K", "m", "solver_seed", "instance_seed"):
            r[k] = int(r[k])
    return rows


def write_manifest(path: str, rows: list) -> None:
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=MANIFEST_FIELDS, lineterminator="\n")
        w.writeheader()
        for r in rows:
            w.writerow({k: r[k] for k in MANIFEST_FIELDS})


def generate(out_dir: str, sizes, seeds: int = 1, first_seed: int = 19000, kind: str = "uniform",
             vehicles=None, m: int = 256, solver_seed: int = 1234, dmin: int = 1, dmax: int = 10,
             slack_pct: int = 110, side: int = 1000, clusters: int = 25, sigma: float = 30.0,
             profile: str = "generated", root: str = None) -> str:
    """Writes n{n}_k{K}_s{seed}.vrp for every size and seed, and manifest.csv
    (instance paths relative to `root`, default the output directory).
    K = vehicles(n) if given, else ceil(n/125).  Returns the manifest path."""
    from .solver import Instance
    os.makedirs(out_dir, exist_ok=True)
    root = os.path.abspath(root or out_dir)
    rows = []
    seed = first_seed
    for n in sizes:
        for _ in range(seeds):
            K = int(vehicles(n)) if vehicles else max(1, math.ceil(n / 125))
            if kind == "uniform":
                inst = Instance.uniform(n, K, side=side, dmin=dmin, dmax=dmax,
                                        slack_pct=slack_pct, seed=seed)
                layout = f"uniform{side}"
            elif kind == "clustered":
                inst = Instance.clustered(n, K, side=side, clusters=clusters, sigma=sigma,
                                          dmin=dmin, dmax=dmax, slack_pct=slack_pct, seed=seed)
                layout = f"clustered{side}c{clusters}s{sigma:g}"
            else:
                raise ValueError(f"unknown generator {kind!r}")
            name = f"n{n}_k{K}_s{seed}"
            path = os.path.join(out_dir, name + ".vrp")
            inst.write(path)
            rows.append({"profile": profile, "name": name,
                         "instance_path": os.path.relpath(os.path.abspath(path), root),
                         "n": n, "K": K, "m": m, "solver_seed": solver_seed,
                         "instance_seed": seed, "layout_id": layout,
                         "capacity_formula": f"ceil({slack_pct}/100*sum(d)/K)"})
            seed += 1
    manifest = os.path.join(out_dir, "manifest.csv")
    write_manifest(manifest, rows)
    return manifest


# ---- solution files -------------------------------------------------------------

def format_solution(routes, cost: float) -> str:
    """The CLI's text: one "Route i:" line per vehicle (customers only), then
    the cost twice (cli_print_solution_routes / cli_print_solution_cost)."""
    lines = ["Route %d:%s" % (i + 1, "".join(" %d" % c for c in r)) for i, r in enumerate(routes)]
    lines.append("Cost: %.3f" % cost)
    lines.append("best cost: %.3f" % cost)
    return "\n".join(lines) + "\n"


def read_solution(path: str):
    """(routes, printed best cost or None) from a solution file of the CLI
    (the reference's or this one's)."""
    routes, cost = [], None
    with open(path) as f:
        for line in f:
            if line.startswith("Route "):
                routes.append([int(t) for t in line.split(":", 1)[1].split()])
            elif line.startswith("best cost:"):
                cost = float(line.split(":", 1)[1])
            elif line.startswith("Cost:") and cost is None:
                cost = float(line.split(":", 1)[1])
    return routes, cost


def real_cost(inst, routes) -> float:
    """Sum of real (double) Euclidean edge lengths, depot legs included."""
    x = np.asarray(inst.x, np.float64)
    y = np.asarray(inst.y, np.float64)
    total = 0.0
    for r in routes:
        if not r:
            continue
        p = np.array([0] + list(r) + [0])
        total += float(np.hypot(np.diff(x[p]), np.diff(y[p])).sum())
    return total


def nint_cost(inst, routes) -> int:
    """Exact TSPLIB EUC_2D cost: sum of nint(sqrt(dx^2 + dy^2)) (integers)."""
    x = np.asarray(inst.x, np.int64)
    y = np.asarray(inst.y, np.int64)
    total = 0
    for r in routes:
        if not r:
            continue
        p = np.array([0] + list(r) + [0])
        s = np.diff(x[p]) ** 2 + np.diff(y[p]) ** 2
        q = np.array([math.isqrt(int(v)) for v in s], np.int64)
        q += (s > q * q + q).astype(np.int64)  # round half up; sqrt never lands on .5
        total += int(q.sum())
    return total


def validate(inst, routes, printed=None, cost_rtol: float = 4e-6) -> dict:
    """Exact checks of a solution; `ok` is False with the first failure in
    `error`.  A printed cost must match the recomputed real cost within the
    %.3f print (5e-4) plus cost_rtol relative: the reference CUDA CLI sums
    float32 edge lengths (aco_cuda_kernels.cu:540,550), a few ulps of float
    per edge."""
    out = {"ok": False, "error": "", "routes": len(routes), "customers": 0, "max_load": 0,
           "cost_real": None, "cost_nint": None, "printed": printed}
    seen = np.zeros(inst.n + 1, np.int64)
    dem = np.asarray(inst.demand, np.int64)
    if len(routes) > inst.K:
        out["error"] = f"{len(routes)} routes for {inst.K} vehicles"
        return out
    for i, r in enumerate(routes):
        for c in r:
            if c < 1 or c > inst.n:
                out["error"] = f"route {i + 1}: node {c} out of range"
                return out
            seen[c] += 1
        load = int(dem[list(r)].sum()) if r else 0
        out["max_load"] = max(out["max_load"], load)
        if load > inst.capacity:
            out["error"] = f"route {i + 1}: load {load} > capacity {inst.capacity}"
            return out
    out["customers"] = int((seen[1:] > 0).sum())
    if (seen[1:] != 1).any():
        bad = int(np.nonzero(seen[1:] != 1)[0][0]) + 1
        out["error"] = f"customer {bad} served {int(seen[bad])} times"
        return out
    out["cost_real"] = real_cost(inst, routes)
    out["cost_nint"] = nint_cost(inst, routes)
    if printed is not None and abs(printed - out["cost_real"]) > 5e-4 + cost_rtol * out["cost_real"]:
        out["error"] = f"printed cost {printed:.3f} != recomputed {out['cost_real']:.6f}"
        return out
    out["ok"] = True
    return out


# ---- solving ----------------------------------------------------------------------

def _runtime_config():
    import ctypes as C
    from . import _native as N
    rt = N.AcoRuntimeConfig()
    N.load().aco_runtime_config_load_env(C.byref(rt))
    return rt


def solve(manifest: str, csv_path: str, solutions_dir: str, repeats: int = 1, root: str = None,
          m_override: int = 0, backend: str = "hip", device: int = 0,
          dist_mode: int = None) -> list:
    """Solves every row x repeats; writes the solution files and the CSV
    (rewritten whole).  Returns the CSV rows.  dist_mode: the metric
    optimised (default: real unless ACO_HIP_DIST=nint)."""
    from . import _native as N
    from .solver import AcoError, HipSolver, Instance, Params, split_routes
    if dist_mode is None:
        dist_mode = N.DIST_NINT if os.environ.get("ACO_HIP_DIST") == "nint" else N.DIST_REAL
    rows = read_manifest(manifest)
    root = os.path.abspath(root or os.path.dirname(os.path.abspath(manifest)))
    os.makedirs(solutions_dir, exist_ok=True)
    rt = _runtime_config()
    kcap = int(os.environ.get("ACO_HIP_CAND_K", "32") or 32)
    runs = []  # (row, run_id, seed, m)
    for r in rows:
        for run_id in range(1, repeats + 1):
            m = m_override or r["m"] or 256
            runs.append((r, run_id, r["solver_seed"] + run_id - 1, m))
    results = {}
    insts = {}
    for r in rows:
        path = os.path.join(root, r["instance_path"])
        try:
            inst = Instance.read(path)
            if inst.K != r["K"] or inst.n != r["n"]:
                raise ValueError(f"instance has n={inst.n} K={inst.K}, manifest n={r['n']} "
                                 f"K={r['K']}")
            insts[r["name"]] = inst
        except Exception as e:  # the rows of an unreadable instance fail, the others run
            insts[r["name"]] = e
    for m in sorted({x[3] for x in runs}):
        group = [x for x in runs if x[3] == m and not isinstance(insts[x[0]["name"]], Exception)]
        if not group:
            continue
        g_insts = [insts[x[0]["name"]] for x in group]
        t0 = time.perf_counter()
        status, err = "ok", ""
        try:
            with HipSolver(g_insts, Params(cand_k=32, ants=m, alpha=1.0, beta=2.0, rho=0.5,
                                           tau0=1.0, Q=1.0, dist_mode=dist_mode),
                           seeds=[x[2] for x in group],
                           cand_k=[min(kcap, inst.n) for inst in g_insts], device=device) as s:
                s.prepare()
                stop = dict(timeout_seconds=rt.timeout_seconds,
                            stagnation_epochs=rt.stagnation_epochs,
                            min_rel_improvement=rt.min_rel_improvement,
                            fixed_epochs=rt.fixed_epochs)
                if not (stop["fixed_epochs"] or stop["stagnation_epochs"]
                        or stop["timeout_seconds"] > 0):
                    stop["stagnation_epochs"] = 100
                s.run(**stop)
                s.sync()
                elapsed = time.perf_counter() - t0
                for i, x in enumerate(group):
                    try:
                        tour, _ = s.best(i)
                        results[(x[0]["name"], x[1])] = (split_routes(tour), elapsed, "")
                    except AcoError as e:
                        results[(x[0]["name"], x[1])] = (None, elapsed, f"no solution: {e}")
        except AcoError as e:
            status, err = "error", str(e)
            for x in group:
                results[(x[0]["name"], x[1])] = (None, time.perf_counter() - t0, err)
    out = []
    for r, run_id, seed, m in runs:
        inst = insts[r["name"]]
        row = {"name": r["name"], "profile": r["profile"], "instance_path": r["instance_path"],
               "n": r["n"], "K": r["K"], "m": m, "solver_seed": seed,
               "instance_seed": r["instance_seed"], "layout_id": r["layout_id"],
               "run_id": run_id, "status": "error", "elapsed_s": "", "best_cost": "", "error": ""}
        if isinstance(inst, Exception):
            row["error"] = str(inst)
        else:
            routes, elapsed, err = results[(r["name"], run_id)]
            row["elapsed_s"] = "%.3f" % elapsed
            if routes is None:
                row["error"] = err
            else:
                v = validate(inst, routes)
                if not v["ok"]:
                    row["error"] = v["error"]
                else:
                    text = format_solution(routes, v["cost_real"])
                    sol = os.path.join(solutions_dir,
                                       f"{r['name']}_{backend}_run{run_id}_solution.txt")
                    with open(sol, "w") as f:
                        f.write(text)
                    row["status"] = "ok"
                    row["best_cost"] = "%.3f" % v["cost_real"]
        row["error"] = row["error"].replace(",", ";").replace("\n", " ")
        out.append(row)
    with open(csv_path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=RESULT_FIELDS, lineterminator="\n")
        w.writeheader()
        for row in out:
            w.writerow(row)
    return out


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m vehicleroutingproblem_amd.campaign")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="seeded instances and a manifest CSV")
    g.add_argument("--out", required=True)
    g.add_argument("--sizes", required=True, help="comma-separated customer counts")
    g.add_argument("--seeds", type=int, default=1, help="instances per size")
    g.add_argument("--first-seed", type=int, default=19000)
    g.add_argument("--kind", choices=("uniform", "clustered"), default="uniform")
    g.add_argument("--m", type=int, default=256)
    g.add_argument("--solver-seed", type=int, default=1234)
    g.add_argument("--dmin", type=int, default=1)
    g.add_argument("--dmax", type=int, default=10)
    g.add_argument("--slack-pct", type=int, default=110)
    g.add_argument("--root", default=None, help="instance paths are relative to this")
    s = sub.add_parser("solve", help="solve a manifest; CSV rows and solution files")
    s.add_argument("--manifest", required=True)
    s.add_argument("--csv", required=True)
    s.add_argument("--solutions", required=True)
    s.add_argument("--repeats", type=int, default=1)
    s.add_argument("--m", type=int, default=0, help="ants per run (0: the manifest's m)")
    s.add_argument("--root", default=None, help="instance paths are relative to this")
    v = sub.add_parser("validate", help="check a solution file against its instance")
    v.add_argument("instance")
    v.add_argument("solution")
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        print(generate(a.out, [int(x) for x in a.sizes.split(",")], seeds=a.seeds,
                       first_seed=a.first_seed, kind=a.kind, m=a.m, solver_seed=a.solver_seed,
                       dmin=a.dmin, dmax=a.dmax, slack_pct=a.slack_pct, root=a.root))
        return 0
    if a.cmd == "solve":
        rows = solve(a.manifest, a.csv, a.solutions, repeats=a.repeats, root=a.root,
                     m_override=a.m)
        bad = [r for r in rows if r["status"] != "ok"]
        print(f"{len(rows) - len(bad)} ok, {len(bad)} errors -> {a.csv}")
        return 1 if bad else 0
    from .solver import Instance
    inst = Instance.read(a.instance)
    routes, printed = read_solution(a.solution)
    res = validate(inst, routes, printed)
    print(json.dumps(res))
    return 0 if res["ok"] else 1


if __name__ == "__main__":
    raise SystemExit(main())
