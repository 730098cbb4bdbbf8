"""Event-driven networks on the device: SPIKE_OP, device-side delivery of
synaptic events (hsh_set_synapses) and SYN2_OP, against the oracle's event
mode on the same connection table.

The bar: the spike raster (ops index, step) is identical and vm, chip,
results and ops are bit-exact -- events of one step reaching one synchan
all carry the same weight here, so the order in which the device adds them
cannot change the sum.
"""
import os

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["jit", "interp"], autouse=True)
def kernels(request, monkeypatch):
    if request.param == "interp":
        monkeypatch.setenv("HSH_NO_JIT", "1")
    else:
        monkeypatch.delenv("HSH_NO_JIT", raising=False)
    return request.param


def check_same(o, g):
    os_, gs = o.state(), g.state()
    for k in ("vm", "chip", "results", "ops"):
        assert np.array_equal(os_[k], gs[k]), k


@pytest.mark.parametrize("mode,kw", [
    (D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=15, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=256)),
    (D.BATCHED, dict()),
], ids=["plain", "iso-K15", "packs-K8", "autotuned"])
def test_network_raster_and_state(gpu, mode, kw):
    m, table = M.network(400, seed=3, inject=(0.3e-9, 0.8e-9))
    T = 1500
    print(f"[case] network 400 cells, {table[0].size} connections, mode {mode} {kw}", flush=True)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    o.step(T)
    want_op, want_step = o.spikes()
    assert want_op.size > 100           # the network is active
    with D.HipSolver(m, **kw) as g:
        g.set_synapses(*table)
        assert g.info()["batch_k"] <= 15
        g.step(T, mode)
        got_op, got_step = g.drain_spikes()
        check_same(o, g)
    assert np.array_equal(got_step, want_step)
    assert np.array_equal(got_op, want_op)


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
def test_event_arrival_both_op_orders(gpu, mode):
    """A held-above-threshold cell drives one synchan; the arrival step
    follows the op order (one step later when the synchan's op comes first)."""
    for src_first in (True, False):
        src, dst = (0, 1) if src_first else (1, 0)
        Em = [0.0, 0.0]
        Em[src], Em[dst] = 0.01, -0.065
        comps = dict(parent=[-1, -1], Cm=[1e-11] * 2, Rm=[1e8] * 2, Ra=[1e6] * 2, Em=Em,
                     initVm=list(Em), inject=[0.0, 0.0])
        m = M.build(comps, [], [], dt=1e-4, spikes=[(src, 0.0, 4e-3)],
                    syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(dst, 0, 1e-9)])
        table = ([int(m.spike_op[0])], [int(m.syn_op[0])], [1.5], [1.5e-3])
        o = Oracle(m, givals=False)
        o.set_synapses(*table)
        o.step(200)
        with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
            g.set_synapses(*table)
            g.step(200, mode)
            check_same(o, g)
            got = g.drain_spikes()
        want = o.spikes()
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[0], want[0])


@pytest.mark.parametrize("mode,kw", [
    (D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=15, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=256)),
], ids=["plain", "iso-K15", "packs-K8"])
def test_network_with_mg_block(gpu, mode, kw):
    """Excitatory synchans with a Mg block (NMDA-like): the block calls the
    device's exp, so vm is held to 1e-7 V; the raster must still be the
    oracle's."""
    from helpers import with_mgblock
    m0, table0 = M.network(400, seed=3, inject=(0.3e-9, 0.8e-9))
    m, remap = with_mgblock(m0, range(0, m0.syn_op.size, 2))
    table = (remap(table0[0]), remap(table0[1]), table0[2], table0[3])
    T = 1500
    print(f"[case] network + Mg block, mode {mode} {kw}", flush=True)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    o.step(T)
    want_op, want_step = o.spikes()
    assert want_op.size > 100
    with D.HipSolver(m, **kw) as g:
        g.set_synapses(*table)
        g.step(T, mode)
        got_op, got_step = g.drain_spikes()
        gs = g.state()
    os_ = o.state()
    assert float(np.max(np.abs(os_["vm"] - gs["vm"]))) <= 1e-7
    assert np.array_equal(os_["ops"], gs["ops"])
    assert np.array_equal(got_step, want_step) and np.array_equal(got_op, want_op)
