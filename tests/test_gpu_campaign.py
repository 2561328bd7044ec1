"""GPU test of the campaign runner: a generated manifest solved in one
batched context, with CSV rows and solution files in the reference's formats.
Every run's solution file validates exactly, the CSV cost is the file's cost,
and each run equals the same instance solved alone with the run's seed (the
batch only schedules)."""
import csv
import os

import numpy as np
import pytest

from vehicleroutingproblem_amd import HipSolver, Instance, Params
from vehicleroutingproblem_amd import _native as N
from vehicleroutingproblem_amd import campaign as CP
from vehicleroutingproblem_amd.solver import split_routes

pytestmark = pytest.mark.gpu


def test_campaign_solve(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("ACO_SOLVER_FIXED_EPOCHS", "6")
    for k in ("ACO_SOLVER_STAGNATION_EPOCHS", "ACO_SOLVER_TIMEOUT_SECONDS", "ACO_HIP_CAND_K",
              "ACO_HIP_DIST"):
        monkeypatch.delenv(k, raising=False)
    manifest = CP.generate(str(tmp_path / "inst"), [25, 70, 140], seeds=1, first_seed=900, m=24,
                           solver_seed=55)
    out_csv = str(tmp_path / "runs.csv")
    sol_dir = str(tmp_path / "sol")
    rows = CP.solve(manifest, out_csv, sol_dir, repeats=2)
    assert len(rows) == 6 and all(r["status"] == "ok" for r in rows), rows
    with open(out_csv, newline="") as f:
        assert f.readline().strip() == ",".join(CP.RESULT_FIELDS)
        back = list(csv.DictReader(f, fieldnames=CP.RESULT_FIELDS))
    assert [(b["name"], b["run_id"], b["solver_seed"]) for b in back] == \
        [(r["name"], str(r["run_id"]), str(r["solver_seed"])) for r in rows]
    mrows = {r["name"]: r for r in CP.read_manifest(manifest)}
    for r in rows:
        inst = Instance.read(os.path.join(os.path.dirname(manifest), r["instance_path"]))
        routes, printed = CP.read_solution(os.path.join(sol_dir,
                                                        f"{r['name']}_hip_run{r['run_id']}_solution.txt"))
        res = CP.validate(inst, routes, printed)
        assert res["ok"], res
        assert float(r["best_cost"]) == printed
        assert r["solver_seed"] == mrows[r["name"]]["solver_seed"] + r["run_id"] - 1
        # the same run alone: same seed, settings, metric (real, the campaign's
        # default) and epochs give the same best tour
        with HipSolver(inst, Params(cand_k=min(32, inst.n), ants=24, alpha=1.0, beta=2.0,
                                    rho=0.5, tau0=1.0, Q=1.0, dist_mode=N.DIST_REAL),
                       seeds=[r["solver_seed"]]) as s:
            s.prepare()
            s.run(fixed_epochs=6)
            tour, _ = s.best()
        assert split_routes(tour) == routes
