"""Campaign manifests, solution files and the exact validator (CPU).

The validator is pinned by the reference's own tracked solution files (the
sequential and CUDA runs of n500_k8_s19000 and the CUDA run of
n1000_k8_s19001, with their instances, in tests/golden): each validates and
its printed cost equals the recomputed real cost within the print's rounding;
planted mistakes in those files are caught."""
import os

import numpy as np
import pytest

from vehicleroutingproblem_amd import Instance
from vehicleroutingproblem_amd import campaign as CP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

# header lines of the reference's instances/generated_benchmark/manifest.csv:1
# and of the CSV its CUDA campaign writes (tools/bash/solve_cuda.sh:56)
REF_MANIFEST_HEADER = "profile,name,instance_path,n,K,m,solver_seed,instance_seed,layout_id,capacity_formula"
REF_RESULT_HEADER = ("name,profile,instance_path,n,K,m,solver_seed,instance_seed,layout_id,"
                     "run_id,status,elapsed_s,best_cost,error")

REF_RUNS = [("n500_k8_s19000", "seq_run1"), ("n500_k8_s19000", "cuda_run1"),
            ("n1000_k8_s19001", "cuda_run1")]


def test_headers_match_reference():
    assert ",".join(CP.MANIFEST_FIELDS) == REF_MANIFEST_HEADER
    assert ",".join(CP.RESULT_FIELDS) == REF_RESULT_HEADER


def test_generate_and_read_manifest(tmp_path):
    m = CP.generate(str(tmp_path / "inst"), [30, 60], seeds=2, first_seed=500, m=16,
                    solver_seed=77)
    assert open(m).readline().strip() == REF_MANIFEST_HEADER
    rows = CP.read_manifest(m)
    assert [r["n"] for r in rows] == [30, 30, 60, 60]
    assert [r["instance_seed"] for r in rows] == [500, 501, 502, 503]
    for r in rows:
        inst = Instance.read(os.path.join(os.path.dirname(m), r["instance_path"]))
        assert inst.n == r["n"] and inst.K == r["K"] == max(1, -(-r["n"] // 125))
        assert r["m"] == 16 and r["solver_seed"] == 77
        assert r["name"] == f"n{r['n']}_k{r['K']}_s{r['instance_seed']}"
        again = Instance.uniform(r["n"], r["K"], seed=r["instance_seed"])
        assert np.array_equal(inst.x, again.x) and np.array_equal(inst.demand, again.demand)


@pytest.mark.parametrize("name,run", REF_RUNS)
def test_reference_solution_files_validate(name, run):
    inst = Instance.read(os.path.join(GOLDEN, f"ref_{name}.vrp"))
    routes, printed = CP.read_solution(os.path.join(GOLDEN, f"ref_{name}_{run}.solution.txt"))
    assert len(routes) == inst.K
    res = CP.validate(inst, routes, printed)
    assert res["ok"], res
    assert res["customers"] == inst.n
    # the sequential CLI prints the double sum; the CUDA one sums float32
    assert abs(res["cost_real"] - printed) <= (5e-4 if "seq" in run else 5e-3)
    # nint per edge differs from the real length by at most 1/2
    edges = sum(len(r) + 1 for r in routes if r)
    assert abs(res["cost_nint"] - res["cost_real"]) <= 0.5 * edges


def test_validator_catches_planted_mistakes():
    name, run = REF_RUNS[0]
    inst = Instance.read(os.path.join(GOLDEN, f"ref_{name}.vrp"))
    routes, printed = CP.read_solution(os.path.join(GOLDEN, f"ref_{name}_{run}.solution.txt"))
    full = [r for r in routes if r]
    dup = [list(r) for r in routes]
    light = min(range(len(dup)), key=lambda t: len(dup[t]))
    dup[light] = dup[light] + [full[0][0]]  # a customer served twice (capacity to spare)
    assert "served 2 times" in CP.validate(inst, dup)["error"]
    miss = [list(r) for r in routes]
    miss[-1] = miss[-1][:-1]  # a customer left out
    assert "served 0 times" in CP.validate(inst, miss)["error"]
    heavy = [list(r) for r in routes]
    i = max(range(len(heavy)), key=lambda t: len(heavy[t]))
    j = min((t for t in range(len(heavy)) if t != i and heavy[t]), key=lambda t: len(heavy[t]))
    heavy[i] = heavy[i] + heavy[j]  # merge two routes: over capacity
    heavy[j] = []
    assert "capacity" in CP.validate(inst, heavy)["error"]
    assert "routes for" in CP.validate(inst, routes + [[]])["error"]
    assert CP.validate(inst, routes, printed + 0.0004)["ok"]
    assert "printed cost" in CP.validate(inst, routes, printed + 0.5)["error"]
    assert "out of range" in CP.validate(inst, [[inst.n + 1]] + routes[1:])["error"]


def test_solution_text_round_trip(tmp_path):
    routes = [[], [3, 1], [2, 5, 4]]
    p = tmp_path / "s.txt"
    p.write_text(CP.format_solution(routes, 123.4567))
    assert p.read_text().splitlines() == ["Route 1:", "Route 2: 3 1", "Route 3: 2 5 4",
                                          "Cost: 123.457", "best cost: 123.457"]
    r, c = CP.read_solution(str(p))
    assert r == routes and c == 123.457


def test_nint_cost_exact():
    # 3-4-5 triangles and a half-way case: nint of sqrt(s) by integer arithmetic
    inst = Instance(np.array([0, 3, 3, 1]), np.array([0, 4, 0, 1]), np.array([0, 1, 1, 1]), 3,
                    10, "t")
    assert CP.nint_cost(inst, [[1], [2], [3]]) == 10 + 6 + 2 * round(2 ** 0.5)
    assert abs(CP.real_cost(inst, [[1], [2], [3]]) - (10 + 6 + 2 * 2 ** 0.5)) < 1e-12
