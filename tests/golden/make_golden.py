"""Writes the golden fixtures of tests/golden from the reference checkout.

Runs in the build container only (it reads /root/reference and the binaries
oracle/ref.mk builds from its sources).  Fixtures written:
  ref_<name>.vrp            reference instances (data files, copied as-is)
  ref_knn_<name>.npz        candidate lists of the reference's
                            aco_mpi_shared_init (aco_parallel.c:218-250), k=16,20,32
  ref_<run>.solution.txt    tracked reference solution files (routes + cost lines)
  ref_smoke_runs.json       routes/costs of the reference seq and OpenMP CLIs on
                            the smoke-test instance (tools/bash/smoke_test.sh:17-48)
  ref_cuda_backend_sizes_n500_n1000.csv
                            the reference CUDA campaign's result rows (data) for
                            the two instances above, m=256, seeds 1234-1236
usage: python tests/golden/make_golden.py
"""
import json
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref")

INSTANCES = ["n500_k8_s19000", "n1000_k8_s19001"]
SOLUTIONS = [
    "results/manual_campaign/backend_sizes/seq/n500/solutions/seq/n500_k8_s19000_seq_run1_solution.txt",
    "results/manual_campaign/backend_sizes/cuda/n1000/solutions/cuda/n1000_k8_s19001_cuda_run1_solution.txt",
    "results/manual_campaign/backend_sizes/cuda/n500/solutions/cuda/n500_k8_s19000_cuda_run1_solution.txt",
]

SMOKE = """NAME : smoke
TYPE : CVRP
DIMENSION : 9
VEHICLES : 2
EDGE_WEIGHT_TYPE : EUC_2D
CAPACITY : 5
NODE_COORD_SECTION
1 50 50
2 10 10
3 20 10
4 10 20
5 20 20
6 80 80
7 90 80
8 80 90
9 90 90
DEMAND_SECTION
1 0
2 1
3 1
4 1
5 1
6 1
7 1
8 1
9 1
DEPOT_SECTION
1
-1
EOF
"""


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-f", "ref.mk"], check=True)
    for name in INSTANCES:
        src = os.path.join(REF, "instances", "generated_benchmark", name + ".vrp")
        dst = os.path.join(HERE, f"ref_{name}.vrp")
        shutil.copyfile(src, dst)
        out = {}
        for k in (16, 20, 32):
            txt = subprocess.run([os.path.join(REF_BIN, "ref_knn"), src, str(k)], check=True,
                                 capture_output=True, text=True).stdout.split("\n")
            rows = [list(map(int, l.split())) for l in txt[1:] if l.strip()]
            out[f"k{k}"] = np.array(rows, dtype=np.int32)
        np.savez_compressed(os.path.join(HERE, f"ref_knn_{name}.npz"), **out)
    for rel in SOLUTIONS:
        lines = open(os.path.join(REF, rel)).read().splitlines()
        keep = [l for l in lines if l.startswith(("Route ", "Cost:", "best cost:"))]
        base = os.path.basename(rel).replace("_solution.txt", "")
        with open(os.path.join(HERE, f"ref_{base}.solution.txt"), "w") as f:
            f.write("\n".join(keep) + "\n")
    rows = []
    for n in (500, 1000):
        lines = open(os.path.join(REF, "results/manual_campaign/backend_sizes/cuda", f"n{n}",
                                  "csv/manifest_cuda_per_instance_results.csv")).read().splitlines()
        header = lines[0]
        rows += [l for l in lines[1:] if l.strip()]
    with open(os.path.join(HERE, "ref_cuda_backend_sizes_n500_n1000.csv"), "w") as f:
        f.write("\n".join([header] + rows) + "\n")
    smoke = os.path.join(HERE, "ref_smoke.vrp")
    with open(smoke, "w") as f:
        f.write(SMOKE)
    env = dict(os.environ, ACO_SOLVER_TIMEOUT_SECONDS="10", ACO_SOLVER_STAGNATION_EPOCHS="10",
               OMP_NUM_THREADS="2")
    runs = {}
    for tag, exe in (("seq", "aco_seq"), ("omp", "aco_omp")):
        r = subprocess.run([os.path.join(REF_BIN, exe), smoke, "2", "0", "1234"], env=env,
                           capture_output=True, text=True, check=True)
        keep = [l for l in r.stdout.splitlines() if l.startswith(("Route ", "Cost:", "best cost:"))]
        runs[tag] = keep
    with open(os.path.join(HERE, "ref_smoke_runs.json"), "w") as f:
        json.dump(runs, f, indent=1)


if __name__ == "__main__":
    main()
