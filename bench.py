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
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "genesis-2.4_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
NEURONS, NCOMP = 50000, 16
CAL_MB = 256                   # device copies used to scale FETCH_SIZE / WRITE_SIZE


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
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 --pmc child runs that measure HBM bytes per launch")
    p.add_argument("--traffic-dir", default="",
                   help="keep the rocprofv3 databases of the traffic passes under this directory")
    p.add_argument("--no-hetero", action="store_true",
                   help="skip the heterogeneous-population section (100k neurons, 1..1024 compartments)")
    p.add_argument("--hetero-neurons", type=int, default=100000)
    p.add_argument("--hetero-steps", type=int, default=5000)
    p.add_argument("--no-network", action="store_true",
                   help="skip the network section (4000 HH cells, synchans, device event delivery)")
    p.add_argument("--net-cells", type=int, default=4000)
    p.add_argument("--net-steps", type=int, default=50000)
    p.add_argument("--no-calcium", action="store_true",
                   help="skip the calcium section (Ca pools, Ca-gated K, GHK, table currents)")
    p.add_argument("--ca-neurons", type=int, default=50000)
    p.add_argument("--ca-steps", type=int, default=2000)
    p.add_argument("--no-shells", action="store_true",
                   help="skip the shells section (diffusion shells, buffers, pumps, concpools)")
    p.add_argument("--shell-neurons", type=int, default=50000)
    p.add_argument("--shell-steps", type=int, default=2000)
    p.add_argument("--no-percall", action="store_true",
                   help="skip the one-step-per-call leg (keeps a profile of the headline to K-step launches)")
    p.add_argument("--no-adjoint", action="store_true",
                   help="skip the adjoint section (1024 x 64-compartment fit, 2000 steps)")
    p.add_argument("--adj-neurons", type=int, default=1024)
    p.add_argument("--adj-steps", type=int, default=2000)
    p.add_argument("--no-iso", action="store_true",
                   help="skip the isopotential section (1000 single-compartment HH cells, 200 steps)")
    p.add_argument("--iso-cells", type=int, default=1000)
    p.add_argument("--iso-steps", type=int, default=200)
    p.add_argument("--only", choices=("isopotential", "heterogeneous", "calcium", "shells", "network", "adjoint"),
                   default=None,
                   help="run one section alone and print its JSON (for profiling it)")
    p.add_argument("--probe-workload", choices=("population", "shells"), default="population",
                   help="workload of --traffic-probe")
    p.add_argument("--traffic-probe", action="store_true",
                   help="(child of the traffic measurement) copy calibration, then a few batched "
                        "launches, and exit")
    p.add_argument("--probe-layout", default="16,1024,4,1",
                   help="(traffic probe) tile_max_rows,pack_cap,pack_waves,pack_stage of the measured run")
    return p.parse_args()


def cpu_model():
    """The host CPU's model name (the first "model name" of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_calibration(torch, mb):
    """Copies of ``mb`` MiB whose counter readings scale FETCH_SIZE and
    WRITE_SIZE: four by an elementwise kernel (``torch.mul`` by 1.0, one
    8-byte load and store per element: the access kind of this library's
    kernels) and four by the runtime's blit (``copy_``), kept for
    comparison -- the blit's reads and writes count at half the rate of
    ordinary kernels' on this GPU."""
    nel = mb * (1 << 20) // 8
    src = torch.ones(nel, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        torch.mul(src, 1.0, out=dst)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    return nel * 8


def read_counter(db, counter, kernel):
    """Mean per-dispatch value of ``counter`` for the batched kernel, for the
    elementwise calibration copy and for the blit copy, from a rocprofv3
    --pmc database."""
    # a built-in kernel's name ends in its first template arguments
    # ("k_batched_r<4>"); the trace also lists the rest ("k_batched_r<4, false>")
    pat = kernel[:-1] if kernel.endswith(">") else kernel
    con = sqlite3.connect(db)
    kern, elem, blit = [], [], []
    for name, cn, val in con.execute("select kernel_name, counter_name, value from counters_collection"):
        if cn != counter:
            continue
        if pat in name:
            kern.append(val)
        elif "MulFunctor" in name:
            elem.append(val)
        elif "copyBuffer" in name:
            blit.append(val)
    con.close()
    if not kern or not elem or not blit:
        raise RuntimeError(f"{counter}: no dispatches of the batched kernel or a calibration copy in {db}")
    return sum(kern) / len(kern), sum(elem) / len(elem), sum(blit) / len(blit)


def measure_traffic(a, K, info, kernel, workload="population"):
    """HBM bytes per batched launch: two child runs of this script under
    ``rocprofv3 --pmc`` (FETCH_SIZE, then WRITE_SIZE) on the same workload
    ("population": the headline trees, "shells": the shells section's).  Each
    counter is scaled by what it reads for an elementwise copy of CAL_MB MiB
    in the same run (true bytes / counter), so the unit of the counter does
    not matter."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    res = {}
    env = dict(os.environ, TMPDIR="/tmp")
    keep = os.path.abspath(a.traffic_dir) if a.traffic_dir else None
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(keep or tmp, "pmc_" + counter)
            # the child runs the same layout, kernel shape and K without tuning
            lay = f"{info['tile_max_rows'] or -1},{info['pack_cap']},{info['pack_waves']},{info['pack_stage']}"
            cmd = [exe, "--pmc", counter, "-d", d, "-o", "p", "--", sys.executable, os.path.abspath(__file__),
                   "--traffic-probe", "--neurons", str(a.neurons), "--ncomp", str(a.ncomp),
                   "--seed", str(a.seed), "--batch-k", str(K), "--flags", str(a.flags | 1),
                   "--probe-layout", lay, "--probe-workload", workload,
                   "--shell-neurons", str(a.shell_neurons)]
            print(f"[bench] traffic pass {counter}: {' '.join(cmd)}", file=sys.stderr, flush=True)
            try:
                p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                   timeout=180)
            except subprocess.TimeoutExpired:
                return None, f"{counter} pass timed out"
            if p.returncode != 0:
                return None, f"{counter} pass exited {p.returncode}: {p.stdout.decode(errors='replace')[-300:]}"
            dbs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("_results.db")]
            if not dbs:
                return None, f"{counter}: no results database"
            res[counter] = read_counter(dbs[0], counter, kernel)
    cal_kib = CAL_MB * 1024.0
    f_fetch = cal_kib / res["FETCH_SIZE"][1]
    f_write = cal_kib / res["WRITE_SIZE"][1]
    t = {"bytes_per_launch": (res["FETCH_SIZE"][0] * f_fetch + res["WRITE_SIZE"][0] * f_write) * 1024.0,
         "fetch_kib_per_launch": res["FETCH_SIZE"][0], "write_kib_per_launch": res["WRITE_SIZE"][0],
         "fetch_scale": f_fetch, "write_scale": f_write,
         "calibration": f"{CAL_MB} MiB elementwise copy read as {res['FETCH_SIZE'][1]:.0f} KiB FETCH_SIZE, "
                        f"{res['WRITE_SIZE'][1]:.0f} KiB WRITE_SIZE",
         "blit_calibration": f"{CAL_MB} MiB runtime blit copy read as {res['FETCH_SIZE'][2]:.0f} KiB "
                             f"FETCH_SIZE, {res['WRITE_SIZE'][2]:.0f} KiB WRITE_SIZE",
         "bytes_per_launch_blit_scale": (res["FETCH_SIZE"][0] * cal_kib / res["FETCH_SIZE"][2] +
                                         res["WRITE_SIZE"][0] * cal_kib / res["WRITE_SIZE"][2]) * 1024.0}
    return t, None


def launch_plan(steps, K):
    """The kernel launches ``hsh_step(h, steps, HSH_BATCHED)`` issues at K
    steps per launch: ``steps // K`` launches of K steps (replayed from
    captured graphs of up to 64 launches) and one direct launch of the
    remaining ``steps % K`` steps."""
    if K < 1:
        raise ValueError("K must be at least 1")
    full, rem = divmod(int(steps), int(K))
    launches = full + (1 if rem else 0)
    return {"batch_k": int(K), "full_launches": full, "remainder_steps": rem, "kernel_launches": launches,
            "steps_per_launch": steps / launches if launches else 0.0,
            # the launch shape the traffic passes repeat: a full launch when
            # there is one, else the single short launch
            "probe_k": int(K) if full > 0 else rem,
            "graph_replayed": full > 0}


def roofline_fields(bytes_step, steps, dev_ms, lp):
    """Achieved algorithmic GB/s, its fraction of the HBM peak and the mean
    launch duration, from the device time ``dev_ms`` of ``steps`` batched
    steps issued as ``lp`` (launch_plan) describes."""
    achieved = bytes_step * steps / (dev_ms * 1e-3) / 1e9
    return {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
            "launch_us": dev_ms * 1e3 / max(1, lp["kernel_launches"]),
            "steps_per_launch": lp["steps_per_launch"], "device_us_per_step": dev_ms * 1e3 / steps}


def batched_kernel_name(info):
    """Name of the kernel a single-class batched launch runs."""
    jit = bool(info["jit"])
    if info["ntiles"] == 0:
        return f"hsh_jit_forest{max(0, info['pack_stage'] - 1)}" if jit else "k_forest"
    r = info["max_rows"]
    if r == 1:
        return "hsh_jit_iso" if jit else "k_iso"
    if r <= 16:
        rpw = 1 if r <= 4 else (2 if r <= 8 else 4)
        return f"hsh_jit_batched_r{rpw}" if jit else f"k_batched_r<{rpw}>"
    return "hsh_jit_batched" if jit else "k_batched"


def hetero_section(a, D, M, Oracle):
    """Heterogeneous population: tree sizes {1, 8, 16, 64, 256, 1024} with
    P proportional to size^-0.67 (mean ~17 compartments), layout, pack shape
    and K chosen by the autotuner on the first batched step; a K sweep on the
    chosen layout; parity against the oracle on a bounded number of steps."""
    sizes = (1, 8, 16, 64, 256, 1024)
    p = np.array(sizes, float) ** -0.67
    p /= p.sum()
    t0 = time.perf_counter()
    m = M.population(a.hetero_neurons, sizes, tuple(p), seed=a.seed)
    t_build = time.perf_counter() - t0
    n = m.ncompts
    print(f"[bench] hetero: {a.hetero_neurons} neurons, {n} compartments (built {t_build:.1f}s)",
          file=sys.stderr, flush=True)
    res = {"workload": f"{a.hetero_neurons} random Hines trees, sizes {list(sizes)} with P ~ size^-0.67",
           "compartments": n, "steps": a.hetero_steps}
    with D.HipSolver(m) as g:
        g.step(64, D.BATCHED)              # the autotuner runs here
        g.sync()
        info = g.info()
        g.set_timing(True)
        t0 = time.perf_counter()
        g.step(a.hetero_steps, D.BATCHED)
        g.sync()
        wall = time.perf_counter() - t0
        dev_ms = g.last_time(D.T_TOTAL)
        g.set_timing(False)
    K = info["batch_k"]
    lp = launch_plan(a.hetero_steps, K)
    rf = roofline_fields(info["bytes_per_step"], a.hetero_steps, dev_ms, lp)
    packs = info["tile_max_rows"] == 0
    res["layout"] = {"autotuned": bool(info["autotuned"]), "autotune_s": info["autotune_s"],
                     "tile_max_rows": info["tile_max_rows"], "pack_cap": info["pack_cap"],
                     "pack_waves": info["pack_waves"], "schedule_in_lds": {2: "all", 3: "indices"}.get(info["pack_stage"], "none"),
                     "packs": info["npacks"], "tiles": info["ntiles"], "batch_k": K,
                     "kernel": batched_kernel_name(info), "lane_fill": info["lane_fill"]}
    res["batched"] = {"value": a.hetero_steps * n / wall, "unit": "compartment-steps/s",
                      "ms_per_step": wall * 1e3 / a.hetero_steps, "device_us_per_step": rf["device_us_per_step"],
                      "hbm_frac_algorithmic": rf["frac"], "launch_us": rf["launch_us"], "launches": lp}
    opts = dict(flags=D.F_NO_AUTOTUNE, tile_max_rows=info["tile_max_rows"] if not packs else -1,
                pack_cap=info["pack_cap"], pack_waves=info["pack_waves"], pack_stage=info["pack_stage"])
    # steps per launch on the chosen layout: K = 1, 8, 32, 128 (1024 steps each)
    sweep = []
    for Ks in (1, 8, 32, 128):
        with D.HipSolver(m, **dict(opts, batch_k=Ks)) as g:
            g.step(2 * Ks, D.BATCHED)
            g.sync()
            t0 = time.perf_counter()
            g.step(1024, D.BATCHED)
            g.sync()
            wall = time.perf_counter() - t0
        sweep.append({"batch_k": Ks, "ms_per_step": wall * 1e3 / 1024, "value": 1024 * n / wall})
        print(f"[bench] hetero K={Ks}: {wall * 1e6 / 1024:.1f} us/step", file=sys.stderr, flush=True)
    res["batch_k_sweep"] = sweep
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        g.step(16, D.PLAIN)
        g.sync()
        t0 = time.perf_counter()
        g.step(200, D.PLAIN)
        g.sync()
        wall = time.perf_counter() - t0
        res["plain"] = {"value": 200 * n / wall, "unit": "compartment-steps/s", "ms_per_step": wall * 1e3 / 200}
    if not a.no_cpu:
        S = 40
        o = Oracle(m, givals=False)
        t0 = time.perf_counter()
        o.step(S)
        cpu = time.perf_counter() - t0
        ref = o.vm
        with D.HipSolver(m, **dict(opts, batch_k=K)) as g:
            g.step(S, D.BATCHED)
            v = g.vm
        res["parity"] = {"steps": S, "max_abs_dv_V": float(np.max(np.abs(v - ref))),
                         "bitexact": bool(np.array_equal(v, ref)),
                         "full_horizon": "5,000 steps bit-exact in tests/test_gpu_horizon.py"}
        res["cpu_baseline"] = {"value": S * n / cpu, "unit": "compartment-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{S} steps of the population ({cpu:.1f} s)"}
    return res


def iso_section(a, D, M, Oracle):
    """Isopotential cells: 1000 single-compartment HH cells (30x30 um soma,
    Na/K 1200/360 S/m^2, inject U(0.1, 0.5) nA from seed 0), dt 20 us, 200
    steps, fp64 -- the persistent multi-step kernel with V in a register for
    K steps.  Step phase (device time), end to end (build, create, upload,
    steps, download), roofline by algorithmic bytes, parity against the
    oracle over the whole run, and the oracle's time for the same run."""
    n, T = a.iso_cells, a.iso_steps
    t_e2e = time.perf_counter()
    m = M.isopotential_hh(n, seed=0)
    with D.HipSolver(m) as g:
        g.step(T, D.BATCHED)               # autotunes, then runs: end to end as a caller sees it
        v = g.vm
        t_e2e = time.perf_counter() - t_e2e
        info = g.info()
        K = info["batch_k"]
        # step phase on the same handle, from the state it reached
        g.step(2 * T, D.BATCHED)
        g.sync()
        g.set_timing(True)
        t0 = time.perf_counter()
        g.step(T, D.BATCHED)
        g.sync()
        wall = time.perf_counter() - t0
        dev_ms = g.last_time(D.T_TOTAL)
        g.set_timing(False)
    lp = launch_plan(T, K)
    rf = roofline_fields(info["bytes_per_step"], T, dev_ms, lp)
    res = {"workload": f"{n} isopotential HH cells, {T} steps, dt 20 us, fp64", "compartments": n, "steps": T,
           "batched": {"value": T * n / wall, "unit": "compartment-steps/s", "us_per_step": wall * 1e6 / T,
                       "device_us_per_step": dev_ms * 1e3 / T, "batch_k": K, "tiles": info["ntiles"],
                       "kernel": batched_kernel_name(info), "launches": lp},
           "end_to_end": {"value": T * n / t_e2e, "unit": "compartment-steps/s", "seconds": t_e2e,
                          "includes": "model build, hsh_create (JIT), autotune, steps, vm download"},
           "roofline": dict(rf, bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
                            algorithmic_bytes_per_compartment_step=info["bytes_per_step"] / n,
                            note="16 tiles of 64 cells: a few waves on a 256-CU GPU, latency-bound")}
    if not a.no_cpu:
        o = Oracle(m, givals=False)
        t0 = time.perf_counter()
        o.step(T)
        cpu = time.perf_counter() - t0
        res["parity"] = {"steps": T, "bitexact": bool(np.array_equal(o.vm, v)),
                         "max_abs_dv_V": float(np.max(np.abs(o.vm - v)))}
        res["cpu_baseline"] = {"value": T * n / cpu, "unit": "compartment-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the whole run ({cpu:.3f} s)"}
    return res


def calcium_section(a, D, M, Oracle):
    """The 16-compartment population with calcium in every compartment: a
    GHK Ca channel feeding a Ca_concen pool, a Ca-gated K channel (NEWCONC1 +
    IPOL1C), a DOMAIN-gated K channel, table currents on Vm and on the pool.
    Batched (autotuned) and plain throughput; parity against the oracle on a
    bounded number of steps (GHK calls the device exp, so the bar there is
    the 1e-7 V tolerance)."""
    t0 = time.perf_counter()
    m = M.ca_population(a.ca_neurons, (16,), (1.0,), features=M.CA_FEATURES[:6] + ("ghk",), seed=a.seed)
    t_build = time.perf_counter() - t0
    n = m.ncompts
    print(f"[bench] calcium: {a.ca_neurons} neurons, {n} compartments (built {t_build:.1f}s)",
          file=sys.stderr, flush=True)
    res = {"workload": f"{a.ca_neurons} random Hines trees x 16 compartments, HH soma, Kdr dendrites; per "
                       "compartment a GHK Ca channel, one Ca_concen pool, Ca- and DOMAIN-gated K, table currents",
           "compartments": n, "steps": a.ca_steps}
    with D.HipSolver(m) as g:
        g.step(64, D.BATCHED)          # the autotuner runs here
        g.sync()
        info = g.info()
        g.set_timing(True)
        t0 = time.perf_counter()
        g.step(a.ca_steps, D.BATCHED)
        g.sync()
        wall = time.perf_counter() - t0
        dev_ms = g.last_time(D.T_TOTAL)
        g.set_timing(False)
        res["batched"] = {"value": a.ca_steps * n / wall, "unit": "compartment-steps/s",
                          "ms_per_step": wall * 1e3 / a.ca_steps, "batch_k": info["batch_k"],
                          "tile_max_rows": info["tile_max_rows"], "bytes_per_compartment_step":
                          info["bytes_per_step"] / n,
                          "hbm_frac_algorithmic": info["bytes_per_step"] * a.ca_steps / (dev_ms * 1e-3) / 1e9
                          / HBM_PEAK_GBS}
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        g.step(16, D.PLAIN)
        g.sync()
        t0 = time.perf_counter()
        g.step(200, D.PLAIN)
        g.sync()
        wall = time.perf_counter() - t0
        res["plain"] = {"value": 200 * n / wall, "unit": "compartment-steps/s", "ms_per_step": wall * 1e3 / 200}
    if not a.no_cpu:
        S = 40
        o = Oracle(m, givals=False)
        t0 = time.perf_counter()
        o.step(S)
        cpu = time.perf_counter() - t0
        with D.HipSolver(m, batch_k=8, flags=D.F_NO_AUTOTUNE) as g:
            g.step(S, D.BATCHED)
            v, c = g.vm, g.conc()
        ref, cref = o.vm, o.conc()
        res["parity"] = {"steps": S, "max_abs_dv_V": float(np.max(np.abs(v - ref))),
                         "max_rel_dconc": float(np.max(np.abs(c - cref) / np.abs(cref))),
                         "bitexact": bool(np.array_equal(v, ref) and np.array_equal(c, cref)),
                         "tolerance_V": 1e-7}
        res["cpu_baseline"] = {"value": S * n / cpu, "unit": "compartment-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{S} steps of the calcium population ({cpu:.1f} s)"}
    return res


SHELL_FEATURES = ("fixbuf", "difbuf", "ctaupump", "mmpump", "hillpump", "elmmpump", "pool", "kc", "leak")


def shells_section(a, D, M, Oracle):
    """The 16-compartment population with a diffusion system of 4 Ca shells
    in every compartment (a fixed and a diffusible buffer per shell,
    Crank-Nicolson), fed by a Ca channel, drained by constant-rate,
    Michaelis-Menten and Hill pumps and an electrogenic pump, with a concpool
    and its buffer filled by a Hill pump from the outermost shell and a
    Ca-gated K channel on that shell.  Each compartment's concentration
    program and solve run at the end of its own program.  Batched
    (autotuned) and plain throughput; bit-exact parity against the oracle on
    a bounded number of steps."""
    t0 = time.perf_counter()
    m = M.shell_population(a.shell_neurons, (16,), (1.0,), shells=4, features=SHELL_FEATURES, seed=a.seed,
                           method=M.CRANK)
    t_build = time.perf_counter() - t0
    n = m.ncompts
    print(f"[bench] shells: {a.shell_neurons} neurons, {n} compartments, {m.ndiffs} diffusion systems "
          f"(built {t_build:.1f}s)", file=sys.stderr, flush=True)
    res = {"workload": f"{a.shell_neurons} random Hines trees x 16 compartments, HH soma, Kdr dendrites, "
                       "Crank-Nicolson; per compartment 4 Ca shells each with a fixed and a diffusible buffer, "
                       "constant-rate / Michaelis-Menten / Hill / electrogenic pumps, a concpool with its "
                       "buffer, a Ca channel and a Ca-gated K channel",
           "compartments": n, "diffusion_systems": int(m.ndiffs), "steps": a.shell_steps,
           "concentrations": int(m.conc.size)}
    with D.HipSolver(m) as g:
        g.step(64, D.BATCHED)          # the autotuner runs here
        g.sync()
        info = g.info()
        g.set_timing(True)
        t0 = time.perf_counter()
        g.step(a.shell_steps, D.BATCHED)
        g.sync()
        wall = time.perf_counter() - t0
        dev_ms = g.last_time(D.T_TOTAL)
        g.set_timing(False)
        bps = info["bytes_per_step"]
        K = info["batch_k"]
        if not a.no_traffic:
            tr, why = measure_traffic(a, K, info, batched_kernel_name(info), workload="shells")
            if tr is not None:
                res["traffic"] = {"bytes_per_compartment_step": tr["bytes_per_launch"] / (n * K),
                                  "traffic_over_algorithmic": tr["bytes_per_launch"] / (bps * K),
                                  "detail": tr}
            else:
                res["traffic"] = {"error": why}
        res["batched"] = {"value": a.shell_steps * n / wall, "unit": "compartment-steps/s",
                          "ms_per_step": wall * 1e3 / a.shell_steps,
                          "device_us_per_step": dev_ms * 1e3 / a.shell_steps, "batch_k": info["batch_k"],
                          "tile_max_rows": info["tile_max_rows"], "kernel": batched_kernel_name(info),
                          "bytes_per_compartment_step": bps / n,
                          "roofline": {"bound": "hbm", "achieved": bps * a.shell_steps / (dev_ms * 1e-3) / 1e9,
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": bps * a.shell_steps / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        g.step(16, D.PLAIN)
        g.sync()
        t0 = time.perf_counter()
        g.step(200, D.PLAIN)
        g.sync()
        wall = time.perf_counter() - t0
        res["plain"] = {"value": 200 * n / wall, "unit": "compartment-steps/s", "ms_per_step": wall * 1e3 / 200}
    if not a.no_cpu:
        S = 40
        o = Oracle(m, givals=False)
        t0 = time.perf_counter()
        o.step(S)
        cpu = time.perf_counter() - t0
        with D.HipSolver(m, batch_k=8, flags=D.F_NO_AUTOTUNE) as g:
            g.step(S, D.BATCHED)
            v, c = g.vm, g.conc()
        ref, cref = o.vm, o.conc()
        res["parity"] = {"steps": S, "max_abs_dv_V": float(np.max(np.abs(v - ref))),
                         "max_rel_dconc": float(np.max(np.abs(c - cref) / np.abs(cref))),
                         "bitexact": bool(np.array_equal(v, ref) and np.array_equal(c, cref))}
        res["cpu_baseline"] = {"value": S * n / cpu, "unit": "compartment-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{S} steps of the shells population ({cpu:.1f} s)"}
    return res


def network_section(a, D, M, Oracle):
    """Event-driven network: isopotential HH cells, 80% excitatory, 2%
    random connectivity, dual-exponential synchans (exc tau 5 ms, Erev 0;
    inh tau 10 ms, Erev -80 mV; gmax 6 / 67 nS), delay 1.5 ms, threshold 0 mV,
    Vm0 uniform -70..-50 mV, constant inject uniform 0.3..0.8 nA, dt 0.1 ms.
    Spikes are delivered on the device; the raster of the whole run is
    compared step-exact with the oracle's event mode."""
    t0 = time.perf_counter()
    m, table = M.network(a.net_cells, seed=a.seed, inject=(0.3e-9, 0.8e-9))
    print(f"[bench] network: {a.net_cells} cells, {table[0].size} connections, {a.net_steps} steps "
          f"(built {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    res = {"workload": f"{a.net_cells} isopotential HH cells (80% exc), 2% connectivity, dual-exponential "
                       "synchans (5/10 ms, 6/67 nS), delay 1.5 ms, dt 0.1 ms, constant inject U(0.3, 0.8) nA",
           "connections": int(table[0].size), "steps": a.net_steps,
           "simulated_s": a.net_steps * m.dt}
    for name, mode, kw in (("batched", D.BATCHED, dict()), ("plain", D.PLAIN, dict(flags=D.F_NO_AUTOTUNE))):
        with D.HipSolver(m, **kw) as g:
            g.set_synapses(*table)
            if mode == D.BATCHED:
                g.autotune()         # keeps the state: the timed run starts at step 0
            g.sync()
            t1 = time.perf_counter()
            g.step(a.net_steps, mode)
            g.sync()
            wall = time.perf_counter() - t1
            ops_, steps_ = g.drain_spikes()
            info = g.info()
            res[name] = {"value": a.net_steps * m.ncompts / wall, "unit": "compartment-steps/s",
                         "ms_per_step": wall * 1e3 / a.net_steps, "spikes": int(ops_.size),
                         "rate_hz": ops_.size / a.net_cells / (a.net_steps * m.dt), "batch_k": info["batch_k"]}
            if name == "batched":
                raster = (ops_, steps_)
                vm_dev = g.vm
    if not a.no_cpu:
        o = Oracle(m, givals=False)
        o.set_synapses(*table)
        t1 = time.perf_counter()
        o.step(a.net_steps)
        cpu = time.perf_counter() - t1
        w_op, w_step = o.spikes()
        res["parity"] = {"steps": a.net_steps, "raster_identical": bool(np.array_equal(w_op, raster[0]) and
                                                                        np.array_equal(w_step, raster[1])),
                         "spikes_oracle": int(w_op.size),
                         "max_abs_dv_V": float(np.max(np.abs(o.vm - vm_dev)))}
        res["cpu_baseline"] = {"value": a.net_steps * m.ncompts / cpu, "unit": "compartment-steps/s", "cores": 1,
                               "kind": "port", "sample": f"the whole run ({cpu:.1f} s)"}
    # the same kind of network in the reference's VAnet2 layout: one solver per
    # cell, all of them in one group handle, events between solvers
    t0 = time.perf_counter()
    cells, merged, ctable = M.network_solvers(a.net_cells, seed=a.seed, inject=(0.3e-9, 0.8e-9))
    print(f"[bench] network: {a.net_cells} one-cell solvers as a group (built {time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    with D.HipSolver(cells) as g:
        g.set_synapses(*ctable)
        g.autotune()
        g.sync()
        t1 = time.perf_counter()
        g.step(a.net_steps, D.BATCHED)
        g.sync()
        wall = time.perf_counter() - t1
        ops_, steps_ = g.drain_spikes()
        vm_dev = g.vm
        res["one_solver_per_cell"] = {"solvers": a.net_cells, "connections": int(ctable[0].size),
                                      "value": a.net_steps * a.net_cells / wall, "unit": "compartment-steps/s",
                                      "ms_per_step": wall * 1e3 / a.net_steps, "spikes": int(ops_.size),
                                      "batch_k": g.info()["batch_k"]}
    if not a.no_cpu:
        o = Oracle(merged, givals=False)
        o.set_synapses(*ctable)
        o.step(a.net_steps)
        w_op, w_step = o.spikes()
        res["one_solver_per_cell"]["raster_identical"] = bool(np.array_equal(w_op, ops_) and
                                                              np.array_equal(w_step, steps_))
        res["one_solver_per_cell"]["max_abs_dv_V"] = float(np.max(np.abs(o.vm - vm_dev)))
    return res


def row_crossings(run, n, j, h):
    """Steps at which the runs theta[n, j] +- h interpolate a different gate
    table row in some compartment of neuron n (rows from the device
    voltages, which are the oracle's bit for bit)."""
    from hsolve_hip import fit as F
    N = run.pmap.n_neurons
    rows = np.nonzero(F.tree_of_row(run.pmap.base) == n)[0]
    tp = np.ones((N, 5))
    tp[n, j] += h
    tm = np.ones((N, 5))
    tm[n, j] -= h
    return int(np.sum(np.any(run.table_rows(tp, rows) != run.table_rows(tm, rows), axis=1)))


def fd_sweep(run, tgt, g, j, h, chunk=128):
    """Central differences of every neuron's part of the loss at once
    (theta[:, j] +- h for all neurons: neurons are independent) against the
    adjoint g[:, j], split by whether the neuron's +-h runs interpolate the
    same table rows at every step.  The quotient's rounding noise is measured
    per neuron as |FD(h) - FD(1.1 h)| (the secant itself moves by O(h^2)
    between the two on a crossing-free neuron): its bar is 1e-5 |g| plus
    three times that noise, or the summation floor 10 eps L_n / h."""
    import torch
    from hsolve_hip import fit as F
    N = run.pmap.n_neurons
    nrn = F.tree_of_row(run.pmap.base)

    def quotient(hh):
        tp = np.ones((N, 5))
        tp[:, j] += hh
        tm = np.ones((N, 5))
        tm[:, j] -= hh
        with torch.no_grad():
            trp = run(torch.from_numpy(tp)).numpy()
            trm = run(torch.from_numpy(tm)).numpy()
        q = (((trp - tgt) ** 2).sum(axis=0) - ((trm - tgt) ** 2).sum(axis=0)) / trp.size / (2 * hh)
        cross = np.zeros(N, np.int64)
        for n0 in range(0, N, chunk):
            rows = np.nonzero((nrn >= n0) & (nrn < n0 + chunk))[0]
            d = run.table_rows(tp, rows) != run.table_rows(tm, rows)
            for n in range(n0, min(N, n0 + chunk)):
                cross[n] = int(np.sum(np.any(d[:, nrn[rows] == n], axis=1)))
        return q, cross

    with torch.no_grad():
        tr0 = run(torch.ones(N, 5, dtype=torch.float64)).numpy()
    L_n = ((tr0 - tgt) ** 2).sum(axis=0) / tr0.size
    fd, cross = quotient(h)
    fd2, cross2 = quotient(1.1 * h)
    gj = g[:, j]
    rel = np.abs(gj - fd) / np.abs(gj)
    floor = np.maximum(10 * np.finfo(float).eps * L_n / h, 3 * np.abs(fd - fd2)) / np.abs(gj)
    free = (cross == 0) & (cross2 == 0)
    ok = rel <= 1e-5 + floor
    return {"param": F.PARAMS[j], "neurons": int(N), "crossing_free": int(free.sum()),
            "crossing_free_within_bar": int((ok & free).sum()),
            "crossing_free_within_1e-5": int(((rel <= 1e-5) & free).sum()),
            "median_rel_crossing_free": float(np.median(rel[free])) if free.any() else None,
            "max_rel_crossing_free_quiet": float(np.max(rel[free & (floor < 1e-5)], initial=0.0)),
            "with_crossings": int((~free).sum()),
            "max_rel_with_crossings": float(np.max(rel[~free], initial=0.0))}


def adjoint_section(a, torch, Oracle, TangentLinear):
    """Adjoint parameter fit: 64-compartment random trees, soma-Vm MSE over
    the run against CPU-solver targets at hidden per-neuron gbar/Ra/Cm
    (+-30%), checkpoint every 100 steps; gradient checked against the CPU
    tangent-linear model and central differences."""
    from hsolve_hip import fit as F
    N, T = a.adj_neurons, a.adj_steps
    t0 = time.perf_counter()
    pm = F.ParamMap(N, (64,), (1.0,), seed=a.seed)
    rng = np.random.default_rng(a.seed + 1)
    th_star = 1 + 0.3 * rng.uniform(-1, 1, (N, 5))
    tgt_np, _ = TangentLinear(pm.model_at(th_star)).run(T, pm.soma)
    tgt = torch.from_numpy(tgt_np)
    print(f"[bench] adjoint: {N} x 64 compartments, {T} steps (setup + targets {time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    n = pm.base.ncompts

    def timed(layout, cap):
        """(runner, forward s, backward s, trace, loss, theta gradient) of one
        fit step on a fresh Runner, after a warm-up pass"""
        r = F.Runner(pm, T, ckpt=100, layout=layout, pack_cap=cap)
        r(torch.ones(N, 5, dtype=torch.float64))        # warm (JIT, layout)
        th = torch.ones(N, 5, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        tr = r(th)
        t_f = time.perf_counter() - t0
        ls = torch.mean((tr - tgt) ** 2)
        t0 = time.perf_counter()
        ls.backward()
        t_b = time.perf_counter() - t0
        return r, t_f, t_b, tr, ls, th.grad.numpy().copy()

    # pack sizes for the pack layout: forward + backward of one fit step
    sweep_caps = []
    for cap in (64, 128, 256, 512):
        r, t_f, t_b, _, _, _ = timed("packs", cap)
        inf = r.solver.info()
        sweep_caps.append({"pack_cap": cap, "packs": inf["npacks"], "forward_s": t_f, "backward_s": t_b})
        r.close()
    cap_best = min(sweep_caps, key=lambda x: x["forward_s"] + x["backward_s"])["pack_cap"]
    rt, tf_t, tb_t, tr_t, _, g_t = timed("tiles", 0)
    rt.close()
    run, t_fwd, t_bwd, tr, loss, g = timed("packs", cap_best)
    info = run.solver.info()
    same_as_tiles = bool(np.array_equal(g, g_t) and np.array_equal(tr.detach().numpy(), tr_t.detach().numpy()))
    print(f"[bench] adjoint packs of {cap_best}: forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms; "
          f"tiles {tf_t * 1e3:.1f} / {tb_t * 1e3:.1f} ms; same gradients {same_as_tiles}", file=sys.stderr, flush=True)
    w = (2 * (tr.detach() - tgt) / tr.numel()).numpy()
    checks = []
    # per parameter type: the neuron with the largest gradient component (the
    # most spike-sensitive, the hardest case for difference quotients) and one
    # drawn at random
    pick_rng = np.random.default_rng(a.seed + 2)
    picks = [(int(np.argmax(np.abs(g[:, j]))), j) for j in range(5)]
    picks += [(int(pick_rng.integers(N)), j) for j in range(5)]
    for (nn, j) in picks:
        e = torch.zeros(N, 5, dtype=torch.float64)
        e[nn, j] = 1
        _, (dc, dr) = torch.func.jvp(pm.arrays, (torch.ones(N, 5, dtype=torch.float64),), (e,))
        _, dtr = TangentLinear(pm.base, dchip=dc.numpy(), dravals=dr.numpy()).run(T, pm.soma)
        g_tl = float(np.sum(w * dtr))
        fd = {}
        for h in (1e-6, 1e-8):
            with torch.no_grad():
                tp = torch.ones(N, 5, dtype=torch.float64)
                tp[nn, j] += h
                tm = torch.ones(N, 5, dtype=torch.float64)
                tm[nn, j] -= h
                fd[h] = (torch.mean((run(tp) - tgt) ** 2) - torch.mean((run(tm) - tgt) ** 2)).item() / (2 * h)
        L0 = loss.item()
        floor = {h: 10 * np.finfo(float).eps * L0 / h for h in fd}
        checks.append({"param": f"{F.PARAMS[j]}[{nn}]", "adjoint": float(g[nn, j]), "tangent_linear": g_tl,
                       "rel_vs_tl": abs(g[nn, j] - g_tl) / abs(g_tl),
                       "rel_vs_fd_h1e-6": abs(g[nn, j] - fd[1e-6]) / abs(g[nn, j]),
                       "rel_vs_fd_h1e-8": abs(g[nn, j] - fd[1e-8]) / abs(g[nn, j]),
                       "row_crossings_h1e-6": row_crossings(run, nn, j, 1e-6),
                       "row_crossings_h1e-8": row_crossings(run, nn, j, 1e-8),
                       "fd_rounding_floor_rel": {f"h{h:.0e}": floor[h] / abs(g[nn, j]) for h in fd}})
    # every neuron at once: neurons are independent, so one pair of runs with
    # parameter j of all neurons at +-h gives each neuron's own difference
    # quotient of its part of the loss, and its row crossings
    sweep = [fd_sweep(run, tgt.numpy(), g, j, 1e-6) for j in range(5)]
    run.close()
    return {"workload": f"{N} random Hines trees x 64 compartments, {T} steps, dt 20 us, fp64, checkpoint "
                        f"every 100 steps; loss = soma-Vm MSE vs CPU targets at hidden +-30% gNa/gK/gKdr/Ra/Cm",
            "parameters": N * 5, "loss": loss.item(),
            "layout": {"kind": "packs", "pack_cap": cap_best, "packs": info["npacks"],
                       "pack_waves": info["pack_waves"]},
            "forward_s": t_fwd, "backward_s": t_bwd,
            "forward_cs_per_s": T * n / t_fwd, "backward_cs_per_s": T * n / t_bwd,
            "pack_cap_sweep": sweep_caps,
            "tile_layout": {"forward_s": tf_t, "backward_s": tb_t, "forward_cs_per_s": T * n / tf_t,
                            "backward_cs_per_s": T * n / tb_t, "gradients_bitexact_vs_packs": same_as_tiles},
            "gradient_checks": checks,
            "max_rel_vs_tl": max(c["rel_vs_tl"] for c in checks),
            "fd_sweep_h1e-6": sweep,
            "note": "gate tables make the model piecewise linear in Vm: a central difference measures its "
                    "derivative only when the +h and -h runs interpolate the same table rows at every step "
                    "(row_crossings = steps where some compartment of the neuron does not).  fd_sweep: every "
                    "neuron's quotient at h=1e-6 per parameter; the 1e-5 bar (plus the quotient's rounding "
                    "floor 10 eps L_n/h) holds on the neurons without crossings"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(neurons shard across GPUs with no per-step exchange, one handle per GPU)"}))
        sys.exit(2)

    import torch
    from hsolve_hip import device as D
    from hsolve_hip import model as M

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if a.copy_calibrate:
        nbytes = copy_calibration(torch, a.copy_calibrate)
        print(json.dumps({"copy_calibrate_bytes": nbytes, "launches": 4}))
        return
    if a.traffic_probe:
        copy_calibration(torch, CAL_MB)
        if a.probe_workload == "shells":
            m = M.shell_population(a.shell_neurons, (16,), (1.0,), shells=4, features=SHELL_FEATURES,
                                   seed=a.seed, method=M.CRANK)
        else:
            m = M.population(a.neurons, (a.ncomp,), (1.0,), seed=a.seed)
        tmr, cap, wv, st = (int(x) for x in a.probe_layout.split(","))
        with D.HipSolver(m, batch_k=a.batch_k, flags=a.flags, tile_max_rows=tmr, pack_cap=cap, pack_waves=wv,
                         pack_stage=st) as g:
            g.step(4 * g.info()["batch_k"], D.BATCHED)
            g.sync()
        print(json.dumps({"traffic_probe": "done"}))
        return

    if a.only:
        from oracle import Oracle as _O
        if a.only == "isopotential":
            r = iso_section(a, D, M, _O)
        elif a.only == "heterogeneous":
            r = hetero_section(a, D, M, _O)
        elif a.only == "calcium":
            r = calcium_section(a, D, M, _O)
        elif a.only == "shells":
            r = shells_section(a, D, M, _O)
        elif a.only == "network":
            r = network_section(a, D, M, _O)
        else:
            from oracle import TangentLinear as _TL
            r = adjoint_section(a, torch, _O, _TL)
        print(json.dumps({a.only: r}))
        return

    t_e2e = time.perf_counter()
    m = M.population(a.neurons, (a.ncomp,), (1.0,), seed=a.seed)
    t_build = time.perf_counter() - t_e2e
    n = m.ncompts
    print(f"[bench] model: {a.neurons} neurons x {a.ncomp} = {n} compartments, "
          f"{m.chip.size} chip values, {m.ops.size} ops (built in {t_build:.2f}s)", file=sys.stderr, flush=True)

    # ---- batched path (value)
    gb = D.HipSolver(m, batch_k=a.batch_k, flags=a.flags)
    t_ready = time.perf_counter() - t_e2e
    gb.step(a.warmup, D.BATCHED)          # the first batched step runs the autotuner
    gb.sync()
    torch.cuda.synchronize()
    info = gb.info()
    K = info["batch_k"]
    gb.set_timing(True)
    t0 = time.perf_counter()
    r = gb.step(a.steps, D.BATCHED)
    gb.sync()
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    dev_b_ms = gb.last_time(D.T_TOTAL)
    gb.set_timing(False)
    lp = launch_plan(a.steps, K)
    t_vm = time.perf_counter()
    _ = gb.vm                             # the caller's view: state back on the host
    t_e2e_total = time.perf_counter() - t_e2e
    t_vm = time.perf_counter() - t_vm
    print(f"[bench] batched K={K}: {wall_b*1e3:.1f} ms wall, {dev_b_ms:.1f} ms device, "
          f"{r['launches']} graph/direct launches", file=sys.stderr, flush=True)

    # ---- one step per call, as a per-step caller (the GENESIS glue) makes
    # them: each hsh_step(h, 1, HSH_BATCHED) is one fused launch
    percall = None
    if not a.no_percall:
        ncall = min(a.steps, 2000)
        gb.sync()
        t0 = time.perf_counter()
        for _ in range(ncall):
            gb.step(1, D.BATCHED)
        gb.sync()
        wall_c = time.perf_counter() - t0
        print(f"[bench] one step per call: {wall_c * 1e6 / ncall:.1f} us per call", file=sys.stderr, flush=True)
        percall = {"value": ncall * n / wall_c, "unit": "compartment-steps/s",
                   "us_per_call": wall_c * 1e6 / ncall, "calls": ncall,
                   "path": "hsh_step(h, 1, HSH_BATCHED) from Python: one fused launch per call"}

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
    # algorithmic bytes of every step the timed launches ran over their
    # device time (HIP events on the handle's stream around the timed call)
    rf = roofline_fields(bytes_step, a.steps, dev_b_ms, lp)
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
                   "packs": info["npacks"], "autotuned": bool(info["autotuned"]),
                   "autotune_s": info["autotune_s"],
                   "path": ("batched persistent kernel, K steps per launch, replayed from a captured hipGraph"
                            if lp["graph_replayed"] else
                            "batched persistent kernel, one direct launch of all timed steps (fewer than K)"),
                   "launches": lp,
                   "specialised_kernels": bool(info["jit"]), "jit_compile_s": info["jit_compile_s"]},
        "plain": {"value": plain_value, "ms_per_step": wall_p * 1e3 / a.steps,
                  "channel_kernel_us": ch_ms * 1e3, "elim_kernel_us": el_ms * 1e3,
                  "batched_over_plain": value / plain_value},
        "one_step_per_call": percall,
        "roofline": {"bound": "hbm",
                     "kernel": batched_kernel_name(info),
                     "achieved": rf["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rf["frac"], "traffic": None,
                     "algorithmic_bytes_per_compartment_step": bytes_step / n,
                     "launch_us": rf["launch_us"], "steps_per_launch": rf["steps_per_launch"],
                     "device_us_per_step": rf["device_us_per_step"]},
        "end_to_end_s": {"build_create_upload": t_ready, "total_batched_incl_warmup": t_e2e_total,
                         "vm_download": t_vm},
        "end_to_end": {"value": (a.warmup + a.steps) * n / t_e2e_total, "unit": "compartment-steps/s",
                       "includes": "model build, hsh_create (JIT), autotune, warmup and timed steps, vm download"},
    }

    # ---- HBM traffic of the batched kernel (two rocprofv3 --pmc child runs)
    gb.close()
    gp.close()
    if not a.no_traffic:
        Kp = lp["probe_k"]
        t, why = measure_traffic(a, Kp, info, out["roofline"]["kernel"])
        if t is None:
            print(f"[bench] traffic not measured: {why}", file=sys.stderr, flush=True)
            out["roofline"]["traffic_note"] = why
        else:
            # the child runs repeat launches of Kp steps with the same layout
            out["roofline"]["traffic"] = t["bytes_per_launch"]
            out["roofline"]["traffic_detail"] = dict(t, steps_per_probe_launch=Kp)
            out["roofline"]["traffic_over_algorithmic"] = t["bytes_per_launch"] / (bytes_step * Kp)
            # measured bytes per step over the timed device time per step
            out["roofline"]["traffic_gbs"] = t["bytes_per_launch"] / Kp * a.steps / (dev_b_ms * 1e-3) / 1e9

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
        host = cpu_model()
        out["cpu_baseline"] = {"value": cpu1_rate, "unit": "compartment-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{rest} steps of the full {n}-compartment population "
                                         f"({cpu1:.1f} s), oracle compiled -O2 -ffp-contract=off",
                               "cpu": host}
        out["cpu_baseline_all_cores"] = {"value": cpuN_rate, "unit": "compartment-steps/s",
                                         "cores": nth, "kind": "port",
                                         "sample": f"{S} steps, neurons split into {len(subs)} "
                                                   f"independent blocks ({cpuN:.1f} s)",
                                         "cpu": host, "nproc": os.cpu_count()}
        out["parity"] = {"steps": S, "max_abs_dv_batched_V": par["batched"],
                         "max_abs_dv_plain_V": par["plain"],
                         "bitexact_batched": par["batched_bitexact"],
                         "bitexact_plain": par["plain_bitexact"], "tolerance_V": 1e-7}
        out["speedup_vs_cpu_1core"] = value / cpu1_rate
    if not a.no_iso:
        from oracle import Oracle as _O
        out["isopotential"] = iso_section(a, D, M, _O)
    if not a.no_hetero:
        from oracle import Oracle as _O
        out["heterogeneous"] = hetero_section(a, D, M, _O)
    if not a.no_calcium:
        from oracle import Oracle as _O
        out["calcium"] = calcium_section(a, D, M, _O)
    if not a.no_shells:
        from oracle import Oracle as _O
        out["shells"] = shells_section(a, D, M, _O)
    if not a.no_network:
        from oracle import Oracle as _O
        out["network"] = network_section(a, D, M, _O)
    if not a.no_adjoint:
        from oracle import Oracle as _O, TangentLinear as _TL
        out["adjoint"] = adjoint_section(a, torch, _O, _TL)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
