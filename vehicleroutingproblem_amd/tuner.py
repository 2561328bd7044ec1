"""Autotuning for batched solves: launch shapes per size bucket and the
candidate-list length per size class, cached on disk.

* Shapes (ants per workgroup, LDS staging) are timed on the device by
  aco_hip_autotune: every bucket's construct launch for each shape that fits.
  Results are cached per (bucket node range, candidate length, ants).
* The candidate length changes the search itself, so it is tuned on solution
  quality: a sample instance of each size class is solved for a fixed
  wall-clock budget with each candidate length, and the length with the best
  cost wins (ties to the shorter list).  Cached per (size class, ants).
"""
import json
import math
import os
import time

from .solver import HipSolver, Params

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "aco_hip_tune.json")


def size_class(n1: int) -> int:
    """Upper node-count bound of the bucket an instance of n1 nodes joins
    (the same classes as the library's size buckets)."""
    c = 256
    while c < n1:
        c <<= 1
    return c


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _shape_key(bucket: dict, k: int, ants: int) -> str:
    return f"shape|n1<={size_class(bucket['n1_max'])}|k{k}|m{ants}"


def tune_shapes(solver: HipSolver, cache: str | None = DEFAULT_CACHE, launches: int = 3) -> list:
    """Apply cached shapes to every bucket of `solver`; time the buckets with
    no cache entry (the device tuner times all buckets) and store them."""
    table = _load(cache) if cache else {}
    ants = solver.ants

    def bucket_k(info):
        ks = [solver.cand_k(i) for i, inst in enumerate(solver.instances)
              if info["n1_min"] <= inst.n + 1 <= info["n1_max"]]
        return max(ks) if ks else solver.params.cand_k

    buckets = solver.buckets()
    missing = False
    for b, info in enumerate(buckets):
        hit = table.get(_shape_key(info, bucket_k(info), ants))
        if hit:
            solver.set_bucket_shape(b, hit["ants_per_group"], hit["stage_lds"],
                                    hit.get("ants_per_wave", 0))
        else:
            missing = True
    if missing:
        buckets = solver.autotune(launches)
        for info in buckets:
            table[_shape_key(info, bucket_k(info), ants)] = {"ants_per_group": info["ants_per_group"],
                                                "stage_lds": info["stage_lds"],
                                                "ants_per_wave": info["ants_per_wave"],
                                                "ms": info["tuned_ms"]}
        if cache:
            _store(cache, table)
    return solver.buckets()


def tune_candidate_length(samples: dict, ants: int, ks=(12, 16, 20, 24), seconds: float = 1.0,
                          cache: str | None = DEFAULT_CACHE, seed: int = 1) -> dict:
    """samples: {size class: [Instance, ...]}.  Each instance is solved for
    `seconds` with each candidate length; a length's score is its cost
    relative to the best length on that instance, averaged over the class's
    instances; the lowest score wins (ties to the shorter list).  Returns
    {size class: best k}."""
    table = _load(cache) if cache else {}
    out = {}
    for cls, insts in sorted(samples.items()):
        key = f"k|n1<={cls}|m{ants}"
        if key in table:
            out[cls] = table[key]["k"]
            continue
        usable = [k for k in ks if all(k <= inst.n for inst in insts)]
        costs = {k: [] for k in usable}
        for inst in insts:
            for k in usable:
                with HipSolver(inst, Params(cand_k=k, ants=ants, rho=0.1), seeds=[seed]) as s:
                    s.prepare()
                    s.run(timeout_seconds=seconds, max_epochs=10 ** 6, chunk_epochs=16)
                    try:
                        costs[k].append(float(s.best()[1]))
                    except Exception:
                        costs[k].append(math.inf)
        score = {}
        for k in usable:
            rel = [c / min(costs[q][i] for q in usable) for i, c in enumerate(costs[k])]
            score[k] = sum(rel) / len(rel)
        best = min(usable, key=lambda k: (score[k], k))
        out[cls] = best
        table[key] = {"k": best, "score": {str(k): score[k] for k in usable}, "seconds": seconds}
    if cache:
        _store(cache, table)
    return out


def candidate_lengths(instances, k_by_class: dict, default: int = 16) -> list:
    """Per-instance candidate lengths from a {size class: k} table."""
    return [int(k_by_class.get(size_class(inst.n + 1), default)) for inst in instances]


def main(argv=None):
    """python -m vehicleroutingproblem_amd.tuner --k-table OUT.json: tune the
    candidate length per size class on sample instances of the batched
    generator (n ~ middle of each class) and write the table."""
    import argparse
    from .prior import training_instances
    ap = argparse.ArgumentParser()
    ap.add_argument("--k-table", required=True)
    ap.add_argument("--ants", type=int, default=128)
    ap.add_argument("--seconds", type=float, default=1.5)
    ap.add_argument("--ks", default="10,12,16,20,24")
    a = ap.parse_args(argv)
    ks = tuple(int(x) for x in a.ks.split(","))
    samples = {}
    for cls, n in ((256, 200), (512, 400), (1024, 900), (2048, 1800)):
        samples[cls] = training_instances(2, nmin=n, nmax=n, seed=cls)  # uniform + clustered
    table = tune_candidate_length(samples, a.ants, ks=ks, seconds=a.seconds, cache=None)
    out = {"k_by_size_class": {str(c): k for c, k in sorted(table.items())},
           "method": f"best cost after {a.seconds} s of plain ACO ({a.ants} ants) per candidate "
                     f"length in {list(ks)}, mean relative cost over a uniform and a clustered "
                     "sample instance per size class",
           "command": "python -m vehicleroutingproblem_amd.tuner --k-table " + a.k_table}
    with open(a.k_table, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
