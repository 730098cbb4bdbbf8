"""Event-driven networks on the device: SPIKE_OP, the synchans' pending
event lists kept by the device's event pass (hsh_set_synapses,
hsh_post_events) and SYN2_OP, against the oracle's restatement of the
reference's lists (newconn/synchan.c:161-358) on the same connections.

The bar: the spike raster (ops index, step) is identical and vm, chip,
results and ops (the SYN2_OP counters included) are bit-exact -- with
unequal weights and delays too: the device sums a step's events in the
list's order, in single precision, as the reference's float activation does.
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
    (D.BATCHED, dict(batch_k=14, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=256)),
    (D.BATCHED, dict()),
], ids=["plain", "iso", "packs", "autotuned"])
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
        assert g.info()["batch_k"] == 1           # one step per launch with pending lists
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
    (D.BATCHED, dict(batch_k=14, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=256)),
], ids=["plain", "iso", "packs"])
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


def unequal_table(table, seed):
    """The same connections with unequal weights (0.5..2) and delays
    (1.5..4 ms): events queue behind and ahead of one another, and a step's
    events carry different weights."""
    rng = np.random.default_rng(seed)
    n = table[0].size
    return (table[0], table[1], rng.uniform(0.5, 2.0, n), rng.uniform(1.5e-3, 4e-3, n))


@pytest.mark.parametrize("mode,kw", [
    (D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=14, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=5, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=256)),
], ids=["plain", "iso", "packs"])
def test_network_unequal_weights_and_delays(gpu, mode, kw):
    m, table0 = M.network(400, seed=5, inject=(0.3e-9, 0.8e-9))
    table = unequal_table(table0, 9)
    T = 2000
    print(f"[case] network unequal weights/delays, mode {mode} {kw}", flush=True)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    o.step(T)
    want_op, want_step = o.spikes()
    assert want_op.size > 100
    with D.HipSolver(m, **kw) as g:
        g.set_synapses(*table)
        g.step(T // 2, mode)
        g.step(T - T // 2, mode)              # a call boundary inside the run
        got_op, got_step = g.drain_spikes()
        check_same(o, g)
    assert np.array_equal(got_step, want_step)
    assert np.array_equal(got_op, want_op)


def one_synchan_model():
    comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[0.0])
    return M.build(comps, [], [], dt=1e-4, syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)],
                   syns=[(0, 0, 1e-9)], float_fields=False)


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
def test_posted_events_queue_like_the_reference(gpu, mode):
    """hsh_post_events against the oracle: events before and after a step's
    ops, queued behind the head, ahead of it, several in one step, and the
    re-timing interval (t_23 - t_8)/dt that truncates to 14."""
    m = one_synchan_model()
    syn = int(m.syn_op[0])
    # (steps already run when posting, steps to the hit, weight, after)
    plan = [(0, 15, 2e-9, False), (3, 15, 5e-9, False), (3, 4, 1.3e-9, True), (8, 30, 1e-9, False),
            (23, 20, 2e-9, False), (23, 20, 7.53e-9, False), (23, 20, 1.02e-9, False), (30, 0, 3e-9, False),
            (41, 6, 8.48e-9, True)]
    o = Oracle(m, givals=False)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE, batch_k=7) as g:
        done = 0
        for at, T, w, after in plan:
            if at > done:
                o.step(at - done)
                g.step(at - done, mode)
                done = at
            o.post_event(syn, T, w, after=after)
            g.post_events([syn], [T], [w], after=after)
            check_same(o, g)                 # HSEVENT counters right after posting
        for _ in range(6):
            o.step(9)
            g.step(9, mode)
            check_same(o, g)


def test_event_state_survives_autotune_and_put(gpu):
    """The autotuner's trial launches fire and add events; it puts the lists
    back.  hsh_put_state with ops sets the counters the event pass counts."""
    m, table0 = M.network(200, seed=7, inject=(0.3e-9, 0.8e-9))
    table = unequal_table(table0, 3)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    o.step(300)
    with D.HipSolver(m) as g:
        g.set_synapses(*table)
        g.step(300, D.PLAIN)
        g.autotune()
        check_same(o, g)
        o.step(400)
        g.step(400, D.BATCHED)
        check_same(o, g)
        st = o.state()
        g.put_state(ops=st["ops"])
        o.step(100)
        g.step(100, D.BATCHED)
        check_same(o, g)


def test_fused_and_per_step_spike_passes_at_4000_cells(gpu):
    """4000 isopotential cells (63 one-row tiles, ~80 connections per
    source, more than a wave scatters at once): the batched step kernel
    runs the next step's spike pass for its own blocks (arrivals in the
    buffers of the next step's parity), the plain loop launches the pass
    before every step, whose last wave holds lanes past the last source.
    Over call boundaries -- each batched call starts with a spike launch and
    its last step leaves the next pass alone -- both equal the oracle: raster
    identical, vm, chip, results and ops bit for bit."""
    m, table0 = M.network(4000, seed=2, inject=(0.3e-9, 0.8e-9))
    table = unequal_table(table0, 4)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    o.step(2000)
    want_op, want_step = o.spikes()
    assert want_op.size > 1000
    for mode in (D.BATCHED, D.PLAIN):
        print(f"[case] 4000-cell network mode {mode}", flush=True)
        with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
            g.set_synapses(*table)
            for n in (1500, 1, 17, 482):
                g.step(n, mode)
            got_op, got_step = g.drain_spikes()
            check_same(o, g)
        assert np.array_equal(got_step, want_step)
        assert np.array_equal(got_op, want_op)
