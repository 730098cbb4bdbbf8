"""The multi-compartment workloads at their stated sizes and horizons
against the oracle: 50,000 neurons x 16 compartments for 10,000 steps
(plain and batched), and 100,000 neurons of 1..1024 compartments for 5,000
steps (batched, packs).  The oracle runs on the host's cores with the
neurons split into independent blocks (each block the same neurons as the
whole population: one generator per neuron; the whole population's rows
hold the blocks in reverse order).  The bar: bit-exact vm."""
import os
import time

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle
from oracle.oracle import step_parallel

pytestmark = pytest.mark.gpu


def oracle_vm(n, sizes, probs, steps, seed=0):
    nth = max(1, min(16, os.cpu_count() or 1))
    bounds = np.linspace(0, n, nth + 1).astype(int)
    subs = [Oracle(M.population(int(bounds[i + 1] - bounds[i]), sizes, probs, seed=seed,
                                first_neuron=int(bounds[i])), givals=False)
            for i in range(nth) if bounds[i + 1] > bounds[i]]
    t0 = time.perf_counter()
    step_parallel(subs, steps, nth)
    print(f"[case] oracle {n} neurons x {steps} steps on {nth} threads: {time.perf_counter() - t0:.1f} s",
          flush=True)
    return np.concatenate([s.vm for s in subs[::-1]])


def test_c2_50k_by_16_for_10000_steps(gpu):
    n, steps = 50000, 10000
    m = M.population(n, (16,), (1.0,), seed=0)
    print(f"[case] 50k x 16, {steps} steps, plain and batched", flush=True)
    got = {}
    for name, mode, kw in (("batched", D.BATCHED, dict()), ("plain", D.PLAIN, dict(flags=D.F_NO_AUTOTUNE))):
        with D.HipSolver(m, **kw) as g:
            g.step(steps, mode)
            got[name] = g.vm
    want = oracle_vm(n, (16,), (1.0,), steps)
    for name, v in got.items():
        assert np.array_equal(v, want), (name, float(np.max(np.abs(v - want))))


def test_c3_heterogeneous_100k_for_5000_steps(gpu):
    sizes = (1, 8, 16, 64, 256, 1024)
    p = np.array(sizes, float) ** -0.67
    p /= p.sum()
    n, steps = 100000, 5000
    m = M.population(n, sizes, tuple(p), seed=0)
    print(f"[case] 100k heterogeneous ({m.ncompts} compartments), {steps} steps, batched", flush=True)
    with D.HipSolver(m, batch_k=32, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=1024) as g:
        g.step(steps, D.BATCHED)
        v = g.vm
    want = oracle_vm(n, sizes, tuple(p), steps)
    assert np.array_equal(v, want), float(np.max(np.abs(v - want)))
