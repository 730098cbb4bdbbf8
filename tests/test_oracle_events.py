"""Oracle pins for spikes and synaptic events, from closed forms.

* SPIKE_OP: a compartment held above threshold spikes at steps 0, r, 2r, ...
  with r = (int)(abs_refract/dt + 1.9) (chip_put_spike, hines_data.c:127-147;
  hines_chip.c:266-276);
* arrival step of an event: a spike at step s reaches a synchan at
  s + (int)(delay/dt + 0.9) (newconn/synchan.c:193-194), one step later when
  the synchan's op precedes the spike's op in ops[];
* SYN2 impulse response: one event of weight w at step t0 gives
  X_t = a x2^n and Y_t = a y1 (x2^(n+1) - y2^(n+1)) / (x2 - y2), n = t - t0,
  a = dt * (w/dt) * stablist[k+3]: the geometric sums of the two first-order
  filters (not the recursion the oracle runs);
* X and Y stay exactly 0 before the arrival step;
* the pending list of an hsolved synchan (newconn/synchan.c:161-358,
  hines.c:352-357, hines_conn.c:60-89), hand-worked: events queued behind
  another are timed relative to it, the head is re-timed by the whole steps
  elapsed since time_last_event as the floating simulation times give them
  (one step short for most intervals once the accumulated time carries
  rounding), an earlier event becomes the head and re-arms the counter,
  events of one step fire together in reverse arrival order, and the
  activation is summed in single precision.
"""
import numpy as np

from hsolve_hip import model as M
from oracle import Oracle

DT = 1e-4


def two_cells(order_src_first: bool, refract=0.5, delay=1.5e-3, weight=2.0):
    """Cell 'src' held above threshold (Em = +10 mV) with a spikegen; cell
    'dst' passive with one synchan."""
    # compartment order decides the op order; the Hines numbering keeps
    # single-compartment trees in element order
    src, dst = (0, 1) if order_src_first else (1, 0)
    Em = [0.0, 0.0]
    Em[src] = 0.01
    Em[dst] = -0.065
    comps = dict(parent=[-1, -1], Cm=[1e-11] * 2, Rm=[1e8] * 2, Ra=[1e6] * 2, Em=Em,
                 initVm=list(Em), inject=[0.0, 0.0])
    m = M.build(comps, [], [], dt=DT, spikes=[(src, 0.0, refract)],
                syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(dst, 0, 1e-9)], float_fields=False)
    table = (np.array([m.spike_op[0]], np.int32), np.array([m.syn_op[0]], np.int32), np.array([weight]),
             np.array([delay]))
    return m, table


def test_spike_refractory_spacing():
    refract = 5e-3
    m, table = two_cells(True, refract=refract)
    o = Oracle(m, givals=False)
    o.step(400)
    _, steps = o.spikes()
    r = int(float(np.float32(refract)) / DT + 1.9)
    assert r == 51
    assert list(steps) == list(range(0, 400, r))


def test_event_arrival_step_and_impulse_response():
    orders = set()
    for src_first in (True, False):
        m, table = two_cells(src_first)
        src_op, dst_op = int(table[0][0]), int(table[1][0])
        orders.add(dst_op < src_op)
        o = Oracle(m, givals=False)
        assert o.set_synapses(*table) >= 1
        xslot = int(m.syn_chip[0])
        T = 60
        X = np.empty(T)
        Y = np.empty(T)
        for t in range(T):
            o.step(1)
            st = o.state()
            X[t], Y[t] = st["chip"][xslot], st["chip"][xslot + 1]
        _, spk = o.spikes()
        assert spk[0] == 0                        # the held cell spikes at step 0, once
        t0 = int(float(np.float32(1.5e-3)) / DT + 0.9) + (1 if dst_op < src_op else 0)
        assert t0 == (16 if dst_op < src_op else 15)
        assert np.all(X[:t0] == 0.0) and np.all(Y[:t0] == 0.0)
        k = int(m.ops[dst_op + 1])
        x2, y1, y2, c3 = m.stablist[k:k + 4]
        # the synchan's activation is a float field (struct_defs.h:112)
        a = DT * float(np.float32(2.0 / DT)) * c3
        n = np.arange(T - t0)
        wantX = a * x2 ** n
        wantY = a * y1 * (x2 ** (n + 1) - y2 ** (n + 1)) / (x2 - y2)
        assert np.allclose(X[t0:], wantX, rtol=1e-12, atol=0)
        assert np.allclose(Y[t0:], wantY, rtol=1e-11, atol=0)
    assert orders == {True, False}


def test_mgblock_is_jahr_stevens():
    """MGBLOCK_OP: g = g_syn / (1 + [Mg]/A exp(-V/B)) (Jahr & Stevens 1990),
    with KMg_A = A, 1/KMg_B = 1/B in the chip values."""
    from helpers import with_mgblock
    A, invB, Mg = 3.57, 62.0, 1.2
    for V in (-0.08, -0.03, 0.0, 0.02):
        comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[V], inject=[0.0])
        m0 = M.build(comps, [], [], dt=DT, syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(0, 0, 1e-9)],
                     float_fields=False)
        m, _ = with_mgblock(m0, [0], A, invB, Mg)
        xs = int(m.syn_chip[0])
        chip = m.chip.copy()
        chip[xs], chip[xs + 1] = 0.7, 0.3                      # X, Y
        m.chip = chip
        o = Oracle(m, givals=False)
        o.chip_update()
        k = int(m.ops[m.syn_op[0] + 1])
        x2, y1, y2, _ = m.stablist[k:k + 4]
        Y1 = 0.7 * x2 * y1 + 0.3 * y2
        g_syn = chip[xs - 1] * Y1
        want = g_syn / (1 + Mg / A * np.exp(-V * invB))
        st = o.state()
        tbyc, dterm = m.chip[-2], m.chip[-1]
        G = (st["results"][1] - dterm) / tbyc
        assert abs(G - want) <= 1e-9 * want, V


def one_synchan(tau1=5e-3, tau2=1e-3):
    """One passive cell with one synchan, no spikegen: events come from
    ohs_post_event."""
    comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[0.0])
    m = M.build(comps, [], [], dt=DT, syntypes=[dict(Ek=0.0, tau1=tau1, tau2=tau2)], syns=[(0, 0, 1e-9)],
                float_fields=False)
    syn = int(m.syn_op[0])
    k = int(m.ops[syn + 1])
    return m, syn, int(m.syn_chip[0]), m.stablist[k:k + 4]


def sim_times(n):
    """SimulationTime() of steps 0..n-1: simulation_time += dt after each."""
    t = np.zeros(n)
    for i in range(1, n):
        t[i] = t[i - 1] + DT
    return t


def run_x(o, m, xslot, nsteps):
    X = np.empty(nsteps)
    for i in range(nsteps):
        o.step(1)
        X[i] = o.state()["chip"][xslot]
    return X


def want_x(fires, stab, n0, nsteps):
    """X of steps n0..n0+nsteps-1 for events firing at (step, activation):
    each adds dt*activation*c3 after the decay, X_t = sum a_i x2^(t-f_i)."""
    x2, _, _, c3 = stab
    t = np.arange(n0, n0 + nsteps)
    out = np.zeros(nsteps)
    for f, act in fires:
        out += np.where(t >= f, DT * act * c3 * x2 ** np.maximum(t - f, 0), 0.0)
    return out


def f32(x):
    return float(np.float32(x))


def test_pending_queued_behind_head():
    """E1 (15 steps) posted before step 0; E2 (15 steps) before step 3.
    (t_3 - t_0)/dt is 3 exactly here, so the head is re-timed to 12 and E2
    queues 3 steps behind it: fires at 15 and 18."""
    m, syn, xs, stab = one_synchan()
    t = sim_times(40)
    assert int((t[3] - t[0]) / DT) == 3
    o = Oracle(m, givals=False)
    o.post_event(syn, 15, 2e-9)
    assert o.state()["ops"][syn + 2] == 15                 # HSEVENT armed the counter
    X = run_x(o, m, xs, 3)
    assert o.state()["ops"][syn + 2] == 12
    o.post_event(syn, 15, 5e-9)
    times, wdt, tle = o.pending(syn)
    assert list(times) == [12, 3] and tle == t[3]
    assert np.array_equal(wdt, [f32(2e-9) / DT, f32(5e-9) / DT])
    X = np.concatenate([X, run_x(o, m, xs, 30)])
    fires = [(15, f32(f32(2e-9) / DT)), (18, f32(f32(5e-9) / DT))]
    assert np.allclose(X, want_x(fires, stab, 0, 33), rtol=1e-12, atol=0)
    assert X[14] == 0.0 and X[15] != 0.0
    # after the last fire the counter is -1, decremented that step and after
    assert o.pending(syn)[0].size == 0 and o.state()["ops"][syn + 2] == -2 - (32 - 18)


def test_pending_new_head_rearms_counter():
    """E1 (20 steps) before step 0, E2 (5 steps) before step 2: the head
    re-timed to 18 >= 5, so E2 becomes the head (counter 5) and E1 waits 13
    steps behind it: fires at 7 and 20."""
    m, syn, xs, stab = one_synchan()
    o = Oracle(m, givals=False)
    o.post_event(syn, 20, 1e-9)
    X = run_x(o, m, xs, 2)
    o.post_event(syn, 5, 3e-9)
    assert list(o.pending(syn)[0]) == [5, 13]
    assert o.state()["ops"][syn + 2] == 5
    X = np.concatenate([X, run_x(o, m, xs, 30)])
    fires = [(7, f32(f32(3e-9) / DT)), (20, f32(f32(1e-9) / DT))]
    assert np.allclose(X, want_x(fires, stab, 0, 32), rtol=1e-12, atol=0)


def test_pending_retime_truncates_floating_interval():
    """(t_23 - t_8)/dt falls just below 15 with accumulated simulation
    times, so the head posted before step 8 (30 steps, fires at 38) is
    re-timed to 30 - 14 = 16 before step 23, one more than its 15 remaining
    steps; E2 (20 steps) queues 4 behind it and fires at 42, one step before
    its nominal 23 + 20."""
    m, syn, xs, stab = one_synchan()
    t = sim_times(60)
    assert (t[23] - t[8]) / DT < 15 and int((t[23] - t[8]) / DT) == 14
    o = Oracle(m, givals=False)
    o.step(8)
    assert o.time == t[8]
    o.post_event(syn, 30, 1e-9)
    o.step(15)
    o.post_event(syn, 20, 2e-9)
    times, _, tle = o.pending(syn)
    assert list(times) == [16, 4] and tle == t[23]
    X = run_x(o, m, xs, 30)
    fires = [(38, f32(f32(1e-9) / DT)), (42, f32(f32(2e-9) / DT))]
    assert np.allclose(X, want_x(fires, stab, 23, 30), rtol=1e-12, atol=0)


def test_pending_same_step_reverse_order_float_sum():
    """Three events of one step with equal times: each new one goes before
    the head (time <= head's), so they fire together at step 15 with the
    activation summed in reverse arrival order, in single precision."""
    m, syn, xs, stab = one_synchan()
    ws = [8.48e-9, 7.53e-9, 1.02e-9]        # float sums that depend on the order
    o = Oracle(m, givals=False)
    for w in ws:
        o.post_event(syn, 15, w)
    assert list(o.pending(syn)[0]) == [15, 0, 0]
    X = run_x(o, m, xs, 16)
    act = np.float32(0.0)
    for w in reversed(ws):
        act = np.float32(float(act) + f32(w) / DT)
    x2, _, _, c3 = stab
    assert X[14] == 0.0
    assert X[15] == 0.0 * x2 + DT * float(act) * c3
    # the double-precision sum and the arrival-order sum give other bits
    dsum = sum(f32(w) / DT for w in reversed(ws))
    assert DT * dsum * c3 != X[15]
    fwd = np.float32(0.0)
    for w in ws:
        fwd = np.float32(float(fwd) + f32(w) / DT)
    assert DT * float(fwd) * c3 != X[15]


def test_pending_fire_restarts_retime_clock():
    """E1 (5 steps) and E2 (10 steps) before step 0: [5, 5].  E1 fires at
    step 5 and leaves E2 at the head, so time_last_event becomes t_5.  E3
    (3 steps) before step 8: the head re-times to 5 - (t_8 - t_5)/dt = 2,
    E3 queues 1 behind it: E2 fires at 10, E3 at 11."""
    m, syn, xs, stab = one_synchan()
    t = sim_times(20)
    assert int((t[8] - t[5]) / DT) == 3
    o = Oracle(m, givals=False)
    o.post_event(syn, 5, 1e-9)
    o.post_event(syn, 10, 2e-9)
    assert list(o.pending(syn)[0]) == [5, 5]
    X = run_x(o, m, xs, 8)
    times, _, tle = o.pending(syn)
    assert list(times) == [5] and tle == t[5]
    o.post_event(syn, 3, 4e-9)
    assert list(o.pending(syn)[0]) == [2, 1]
    X = np.concatenate([X, run_x(o, m, xs, 10)])
    fires = [(5, f32(f32(1e-9) / DT)), (10, f32(f32(2e-9) / DT)), (11, f32(f32(4e-9) / DT))]
    assert np.allclose(X, want_x(fires, stab, 0, 18), rtol=1e-12, atol=0)


def test_pending_spike_driven_both_op_orders():
    """A spikegen firing every 5 steps (reload 5) into one synchan 15 steps
    away: the events of steps 5 and 10 queue 5 behind the one before when
    the head's re-timing intervals are whole; the synchan's op before the
    spike's in ops[] moves every fire one step later (its counter was
    decremented before the event came)."""
    t = sim_times(30)
    # the re-timings that place the 2nd and 3rd events (at steps 5 and 10)
    assert [int((t[i + 5] - t[i]) / DT) for i in (0, 5)] == [5, 5]
    for src_first in (True, False):
        m, table = two_cells(src_first, refract=3.2e-4)
        src_op, dst_op = int(table[0][0]), int(table[1][0])
        o = Oracle(m, givals=False)
        o.set_synapses(*table)
        r = int(float(np.float32(3.2e-4)) / DT + 1.9)
        assert r == 5
        late = 1 if dst_op < src_op else 0
        xs = int(m.syn_chip[0])
        k = int(m.ops[dst_op + 1])
        stab = m.stablist[k:k + 4]
        X = run_x(o, m, xs, 30)
        _, spk = o.spikes()
        assert list(spk[:6]) == [0, 5, 10, 15, 20, 25]
        a = f32(f32(2.0) / DT)
        fires = [(15 + late + 5 * i, a) for i in range(3)]
        assert np.allclose(X[:30], want_x(fires, stab, 0, 30), rtol=1e-12, atol=0)
