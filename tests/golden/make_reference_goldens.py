"""Writes tests/golden/reference_goldens.json from the reference's own result
files (run once in the build container, where the reference is mounted).

Sources (values copied as data, not code):
  cluster_bringup/accel_regression_golden.txt:4     ap_cpu RESULT_VM of
      `hh1952_ap_verify.g 8 200` (chanmode 4, CPU)
  experiments/data/campaign_divergence.csv:2-8      cpu_vm of the same model
      after N+10 steps, N = 100..6400
  cluster_bringup/accel_regression_golden.txt:7-8, 11-12   chanmode-1 CPU
      results of the multi-compartment benchmark scripts (recorded for
      reference; a different table arithmetic, not used as a bit-level pin)
"""
import csv
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = {"ap_verify": {"model": "genesis/Scripts/benchmark/hh1952_ap_verify.g", "n_neurons": 8,
                     "points": []}}
gold = open(os.path.join(REF, "cluster_bringup/accel_regression_golden.txt")).read().splitlines()
for line in gold:
    if line.startswith("ap_cpu|RESULT_VM="):
        out["ap_verify"]["result_vm_210"] = line.split("=", 1)[1].strip()
    if line.startswith("multi_cpu|") or line.startswith("branch_cpu|"):
        k, v = line.split("=", 1)
        out.setdefault("chanmode1_cpu", {})[k] = v.strip()
with open(os.path.join(REF, "experiments/data/campaign_divergence.csv")) as f:
    for row in csv.DictReader(f):
        out["ap_verify"]["points"].append({"steps": int(row["steps"]) + 10, "cpu_vm": row["cpu_vm"]})
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")
with open(dst, "w") as f:
    json.dump(out, f, indent=1)
print("wrote", dst)
