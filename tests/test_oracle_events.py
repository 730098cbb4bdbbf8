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
* X and Y stay exactly 0 before the arrival step.
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
        a = DT * (2.0 / DT) * c3
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
