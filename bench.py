#!/usr/bin/env python3
"""bench.py -- hsolve chanmode-4 step throughput on one MI355X.

Workload: 50,000 neurons x 16-compartment random Hines trees (800,000
compartments), soma Na/K + dendrite leak/Kdr, fp64, dt = 20 us, BEULER.  A
"step" is one hsolve PROCESS step of the whole population (channel-opcode
interpreter, Hines elimination, voltage update).  The state is resident in
HBM before the timed region starts.

Measured, in one run:
  * batched path  -- persistent fused kernel, K steps per launch, replayed from
                     a captured hipGraph  (``value``)
  * plain path    -- per-step drop-in: channel kernel + elimination kernel
  * parity        -- both paths vs the CPU oracle after the CPU sample's steps
  * CPU baseline  -- the oracle (a restatement of the reference CPU hsolve)
                     on a bounded number of steps of the same population,
                     single-threaded and on all host cores

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "genesis-2.4_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
NEURONS, NCOMP = 50000, 16


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10000)
    p.add_argument("--warmup", type=int, default=256)
    p.add_argument("--neurons", type=int, default=NEURONS)
    p.add_argument("--ncomp", type=int, default=NCOMP)
    p.add_argument("--batch-k", type=int, default=0, help="steps per batched launch (0 = autotuned)")
    p.add_argument("--cpu-steps", type=int, default=0,
                   help="steps of the CPU sample (0 = sized to ~10 s single-thread)")
    p.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-cores CPU leg (0 = nproc)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle parity")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--flags", type=int, default=0, help="hsh_options.flags for the batched handle")
    p.add_argument("--copy-calibrate", type=int, default=0, metavar="MB",
                   help="only run a device-to-device copy of MB megabytes (counter calibration "
                        "for FETCH_SIZE/WRITE_SIZE) and exit")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(neurons shard with no per-step exchange; see DESIGN.md)"}))
        sys.exit(2)

    import torch
    from hsolve_hip import device as D
    from hsolve_hip import model as M

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if a.copy_calibrate:
        nel = a.copy_calibrate * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibrate_bytes": nel * 8, "launches": 4}))
        return

    t_e2e = time.perf_counter()
    m = M.population(a.neurons, (a.ncomp,), (1.0,), seed=a.seed)
    t_build = time.perf_counter() - t_e2e
    n = m.ncompts
    print(f"[bench] model: {a.neurons} neurons x {a.ncomp} = {n} compartments, "
          f"{m.chip.size} chip values, {m.ops.size} ops (built in {t_build:.2f}s)", file=sys.stderr, flush=True)

    # ---- batched path (value)
    gb = D.HipSolver(m, batch_k=a.batch_k, flags=a.flags)
    info = gb.info()
    K = info["batch_k"]
    t_ready = time.perf_counter() - t_e2e
    gb.step(a.warmup, D.BATCHED)
    gb.sync()
    torch.cuda.synchronize()
    gb.set_timing(True)
    t0 = time.perf_counter()
    r = gb.step(a.steps, D.BATCHED)
    gb.sync()
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    dev_b_ms = gb.last_time(D.T_TOTAL)
    gb.set_timing(False)
    launches_b = -(-a.steps // K)
    t_e2e_total = time.perf_counter() - t_e2e
    print(f"[bench] batched K={K}: {wall_b*1e3:.1f} ms wall, {dev_b_ms:.1f} ms device, "
          f"{r['launches']} graph/direct launches", file=sys.stderr, flush=True)

    # ---- plain path
    gp = D.HipSolver(m)
    gp.step(a.warmup, D.PLAIN)
    gp.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gp.step(a.steps, D.PLAIN)
    gp.sync()
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - t0
    # per-kernel split of the plain step (events around every launch)
    nprof = min(a.steps, 500)
    gp.set_timing(True)
    gp.step(nprof, D.PLAIN)
    gp.sync()
    ch_ms = gp.last_time(D.T_CHANNEL) / nprof
    el_ms = gp.last_time(D.T_ELIM) / nprof
    gp.set_timing(False)
    print(f"[bench] plain: {wall_p*1e3:.1f} ms wall; channel {ch_ms*1e3:.1f} us, elim {el_ms*1e3:.1f} us "
          f"per step", file=sys.stderr, flush=True)

    bytes_step = info["bytes_per_step"]
    value = a.steps * n / wall_b
    plain_value = a.steps * n / wall_p
    launch_s = dev_b_ms * 1e-3 / launches_b
    achieved = bytes_step * K / launch_s / 1e9
    out = {
        "metric": "compartment-steps/s + %HBM peak, 50k x 16 Hines trees, plain vs batched @1e-7 V",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall_b * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.neurons} random Hines trees x {a.ncomp} compartments "
                               "(HH soma Na/K, dendrite leak+Kdr), fp64, dt 20 us, BEULER",
                   "neurons": a.neurons, "compartments_per_neuron": a.ncomp, "compartments": n,
                   "batch_k": K, "batch_waves": info["batch_waves"], "tiles": info["ntiles"],
                   "path": "batched persistent kernel in a captured hipGraph",
                   "specialised_kernels": bool(info["jit"]), "jit_compile_s": info["jit_compile_s"]},
        "plain": {"value": plain_value, "ms_per_step": wall_p * 1e3 / a.steps,
                  "channel_kernel_us": ch_ms * 1e3, "elim_kernel_us": el_ms * 1e3,
                  "batched_over_plain": value / plain_value},
        "roofline": {"bound": "hbm",
                     "kernel": ("hsh_jit_batched_r4" if info["jit"] else "k_batched_r<4>"),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_compartment_step": bytes_step / n,
                     "launch_us": launch_s * 1e6, "steps_per_launch": K},
        "end_to_end_s": {"build_create_upload": t_ready, "total_batched_incl_warmup": t_e2e_total},
    }

    # ---- CPU baseline + parity (oracle)
    if not a.no_cpu:
        from oracle import Oracle
        from oracle.oracle import step_parallel
        o = Oracle(m, givals=False)
        t0 = time.perf_counter()
        o.step(2)
        per = (time.perf_counter() - t0) / 2
        S = a.cpu_steps or max(8, min(2000, int(10.0 / per)))
        S = max(S, 3)
        rest = S - 2
        t0 = time.perf_counter()
        o.step(rest)
        cpu1 = (time.perf_counter() - t0)
        cpu1_rate = rest * n / cpu1
        vref = o.vm
        # parity: fresh device handles, same number of steps from RESET
        par = {}
        for name, mode in (("batched", D.BATCHED), ("plain", D.PLAIN)):
            with D.HipSolver(m, batch_k=K) as g:
                g.step(S, mode)
                v = g.vm
            par[name] = float(np.max(np.abs(v - vref)))
            par[name + "_bitexact"] = bool(np.array_equal(v, vref))
        # all host cores: neurons split into independent blocks
        nth = a.cpu_threads or (os.cpu_count() or 1)
        nblk = nth
        bounds = np.linspace(0, a.neurons, nblk + 1).astype(int)
        subs = [Oracle(M.population(int(bounds[i + 1] - bounds[i]), (a.ncomp,), (1.0,), seed=a.seed,
                                    first_neuron=int(bounds[i])), givals=False)
                for i in range(nblk) if bounds[i + 1] > bounds[i]]
        t0 = time.perf_counter()
        step_parallel(subs, S, nth)
        cpuN = time.perf_counter() - t0
        cpuN_rate = S * n / cpuN
        out["cpu_baseline"] = {"value": cpu1_rate, "unit": "compartment-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{rest} steps of the full {n}-compartment population "
                                         f"({cpu1:.1f} s), oracle compiled -O2 -ffp-contract=off"}
        out["cpu_baseline_all_cores"] = {"value": cpuN_rate, "unit": "compartment-steps/s",
                                         "cores": nth, "kind": "port",
                                         "sample": f"{S} steps, neurons split into {len(subs)} "
                                                   f"independent blocks ({cpuN:.1f} s)"}
        out["parity"] = {"steps": S, "max_abs_dv_batched_V": par["batched"],
                         "max_abs_dv_plain_V": par["plain"],
                         "bitexact_batched": par["batched_bitexact"],
                         "bitexact_plain": par["plain_bitexact"], "tolerance_V": 1e-7}
        out["speedup_vs_cpu_1core"] = value / cpu1_rate
    print(json.dumps(out))


if __name__ == "__main__":
    main()
