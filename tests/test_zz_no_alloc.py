"""hsh_set_inputs_static (allocates nothing: static device staging) against
hsh_set_inputs, through the C ABI: two handles on one model take the same
inputs, one through each entry point; after each call and a run of steps
the states are bit-identical.  Two calls in a row, the second smaller; an
oversized call is refused with HSH_ERR_LIMIT and changes nothing."""
import ctypes as C

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M

pytestmark = pytest.mark.gpu

HSH_ERR_LIMIT = -5
HSH_INPUTS_STATIC_MAX = 16384


def inputs(m, rng, n_inj, n_act):
    inj = [int(m.compchips[r]) + 1 for r in rng.choice(m.ncompts, n_inj, replace=False)]
    out = [(D.IN_INJECT, i, 0, 0, float(rng.uniform(0.0, 2e-9))) for i in inj]
    for j in rng.choice(len(m.syn_op), n_act, replace=False):
        op = int(m.syn_op[j])
        out.append((D.IN_ACTIVATION, int(m.syn_chip[j]), int(m.ops[op + 1]), int(rng.integers(0, 2)),
                    float(rng.uniform(0.1, 3.0))))
    # a second overwrite of the first inject and a second activation of the
    # first synchan in the same call: order within a slot matters
    out.append((out[0][0], out[0][1], 0, 0, 0.25e-9))
    a = out[n_inj]
    out.append((a[0], a[1], a[2], 1, 0.5))
    return out


def same(a, b):
    sa, sb = a.state(), b.state()
    return all(np.array_equal(sa[k], sb[k]) for k in ("vm", "chip", "results", "ops"))


def test_static_inputs_equal_allocating_inputs(gpu):
    m, table = M.network(300, seed=11, inject=(0.2e-9, 0.6e-9))
    rng = np.random.default_rng(5)
    L = D.lib()
    print("[case] hsh_set_inputs_static vs hsh_set_inputs, 300-cell network", flush=True)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as a, D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as b:
        a.set_synapses(*table)
        b.set_synapses(*table)
        a.step(50, D.BATCHED)
        b.step(50, D.BATCHED)
        for n_inj, n_act in ((120, 200), (7, 3)):          # the second call smaller
            arr = D.input_array(inputs(m, rng, n_inj, n_act))
            assert L.hsh_set_inputs(a.h, arr.size, arr.ctypes.data) == 0
            assert L.hsh_set_inputs_static(b.h, arr.size, arr.ctypes.data) == 0
            assert same(a, b)
            a.step(40, D.BATCHED)
            b.step(40, D.BATCHED)
            assert same(a, b)
        # more entries than the static staging holds: refused, nothing applied
        before = b.state()
        big = D.input_array([(D.IN_INJECT, int(m.compchips[0]) + 1, 0, 0, 1e-9)] * (HSH_INPUTS_STATIC_MAX + 1))
        assert L.hsh_set_inputs_static(b.h, big.size, big.ctypes.data) == HSH_ERR_LIMIT
        after = b.state()
        assert all(np.array_equal(before[k], after[k]) for k in ("vm", "chip", "results", "ops"))
