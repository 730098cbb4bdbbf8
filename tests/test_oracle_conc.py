"""Oracle pins for concentrations, from closed forms and textbook formulas.

* Ca_concen (CACONC_OP, hines_conc_chip.c:196-207): under a constant influx
  I the pool follows u_n = u* + (u0 - u*) r^n, u = C - base,
  r = (2 - i)/i, u* = I B'/(2 (i - 1)), which with i = 1 + dt/(2 tau) and
  B' = dt B is the continuous steady state B I tau; the flux is reset by
  CONC_VAL_OP every step (otherwise it would grow);
* NERNST_OP: Vm settles at the leak/channel equilibrium with
  E = (RT/zF) ln(Cout/Cin);
* GHK_OP: the current Gk (Ek - Vm) is the GHK current
  -P zF X (Cin - Cout e^-X)/(1 - e^-X), X = zF V/RT (inward positive), and
  Gk is minus its derivative in V (slope conductance), in both branches
  (|X| < 1e-6 takes the series form);
* NEWCONC1_OP + IPOL1C_OP: a concentration-gated channel's gate settles at
  A/(2(B - 1)) of the table interpolated at (C - ymin) invdy, clamped at the
  table ends;
* DOMAIN_OP: the gate row comes from scale * Ik of the channel before it;
* power-0 IPOL (table currents): Gk = gmax G(x), Ek = E(x) on the voltage and
  on the concentration table.
Gk and Ek are read back from results[] (Vm + ichan tbyc, sumgchan tbyc +
diagterm)."""
import numpy as np
import pytest

from helpers import spliced
from hsolve_hip import model as M
from oracle import Oracle

DT = 2e-5
CONC_VAL, NERNST, GHK, CHAN, CHAN_EK, ADD_CURR, FLUX = 1000, 2000, 2010, 3000, 3001, 3200, 3300
IPOL1V, IPOL1C, NEWCONC1, DOMAIN, CACONC, LCACONC = 4001, 4002, 5110, 5112, 10200, 10202
F, R, T = 96485.3329, 8.314462618, 308.15
ZFRT = 2 * F / (R * T)


def passive(n, *, Em=-0.065, Rm=1e8, Cm=1e-11, V0=None):
    V0 = [Em] * n if V0 is None else list(V0)
    comps = dict(parent=[-1] * n, Cm=[Cm] * n, Rm=[Rm] * n, Ra=[1e6] * n, Em=[Em] * n, initVm=V0,
                 inject=[0.0] * n)
    return M.build(comps, [], [], dt=DT, float_fields=False)


def seg(m, r):
    """chip values of compartment r: Em/Rm, inject, ..., tbyc, diagterm."""
    lo = int(m.compchips[r])
    hi = int(m.compchips[r + 1]) if r + 1 < m.ncompts else m.chip.size
    return lo, hi


def gk_ek(m, o, r, vm):
    """Gk and Ek*Gk of the channels of compartment r from results[]."""
    st = o.state()
    lo, hi = seg(m, r)
    tbyc, dterm = m.chip[hi - 2], m.chip[hi - 1]
    res = st["results"]
    G = (res[2 * r + 1] - dterm) / tbyc
    EG = (res[2 * r] - vm) / tbyc - m.chip[lo] - m.chip[lo + 1]
    return G, EG


def test_caconc_closed_form():
    Ek, g, Em, Rm = 0.12, 2e-9, -0.065, 1e8
    V = (Em / Rm + g * Ek) / (1 / Rm + g)        # leak + channel equilibrium
    base = passive(1, Em=Em, Rm=Rm, V0=[V])
    tau, B, cbase, C0 = 4e-4, 3e9, 5e-5, 1e-3
    itau, Bp = 1 + DT / (2 * tau), DT * B
    m = spliced(base, lambda r: ([CHAN_EK, ADD_CURR, CONC_VAL, 0, FLUX], [Ek, g]),
                conc=[C0], concops=[CACONC, 0, LCACONC], concchip=[cbase, itau, Bp])
    o = Oracle(m, givals=False)
    I = g * (Ek - V)
    ustar = I * Bp / (2 * (itau - 1))
    assert abs(ustar - B * I * tau) <= 1e-12 * ustar
    rr = (2 - itau) / itau
    n_done = 0
    for n in (1, 2, 10, 100, 1000, 5000):
        o.step(n - n_done)
        n_done = n
        want = cbase + ustar + (C0 - cbase - ustar) * rr ** n
        assert abs(o.conc()[0] - want) <= 1e-12 * abs(want), n
        assert abs(o.vm[0] - V) <= 1e-15


def test_nernst_equilibrium():
    Cin, Cout, g, Em, Rm = 7e-5, 2.0, 1e-8, -0.065, 1e8
    base = passive(1, Em=Em, Rm=Rm)
    m = spliced(base, lambda r: ([CONC_VAL, 0, NERNST, ADD_CURR], [1 / ZFRT, g]), conc=[Cin, Cout])
    o = Oracle(m, givals=False)
    o.step(5000)
    E = np.log(Cout / Cin) / ZFRT
    want = (Em / Rm + g * E) / (1 / Rm + g)
    assert abs(o.vm[0] - want) <= 1e-13


def ghk_current(V, P, Cin, Cout):
    X = ZFRT * V
    if X == 0.0:
        return -P * 2 * F * (Cin - Cout)
    return -P * 2 * F * X * (Cin - Cout * np.exp(-X)) / -np.expm1(-X)


@pytest.mark.parametrize("V", [-0.07, -0.02, 0.03, 2e-5, 1e-9, 0.0, -1e-9])
def test_ghk_current_and_slope(V):
    Cin, Cout, P = 1e-4, 2.0, 3e-15
    base = passive(1)
    m = spliced(base, lambda r: ([CONC_VAL, 0, GHK, ADD_CURR], [ZFRT, 2 * F, P]), conc=[Cin, Cout])
    o = Oracle(m, givals=False)
    o.put(vm=np.array([V]))
    o.chip_update()
    G, EG = gk_ek(m, o, 0, V)
    I = EG - G * V
    if ZFRT * abs(V) < 1e-6:
        # the reference's series form is exact to O(X^2) ~ 1e-15 here
        want_I = ghk_current(V, P, Cin, Cout)
        want_G = P * 2 * F * ZFRT * (Cin + Cout) / 2              # slope at X -> 0
        assert abs(I - want_I) <= 1e-12 * abs(want_I)
        assert abs(G - want_G) <= 1e-6 * want_G
    else:
        h = 1e-6 * max(abs(V), 1e-3)
        want_G = -(ghk_current(V + h, P, Cin, Cout) - ghk_current(V - h, P, Cin, Cout)) / (2 * h)
        assert abs(I - ghk_current(V, P, Cin, Cout)) <= 1e-9 * abs(ghk_current(V, P, Cin, Cout))
        assert abs(G - want_G) <= 1e-6 * abs(want_G)


YDIVS, YMAX = 100, 2e-3
INVDY = YDIVS / YMAX


def conc_table():
    """Two column pairs on the concentration axis: a fast gate (B', A') and a
    table current (G, E)."""
    C = np.arange(YDIVS + 2) / INVDY
    C[-1] = C[-2]
    al = 3000 * C / (C + 5e-4) + 200
    be = 4000.0
    cols = np.stack([1 + DT / 2 * (al + be), DT * al, C / (C + 3e-4), -0.08 + 20 * C], axis=1)
    return cols


def interp(col, c):
    v = min(max((c - 0.0) * INVDY, 0.0), YDIVS)
    r = int(v)
    f = v - r
    tab = conc_table()
    return tab[r, col] + f * (tab[r + 1, col] - tab[r, col])


def test_conc_gate_settles_on_interpolated_table():
    # mid-interval, the same interval again (the reference's row cache,
    # hines_chip.c:248-250), on a row, below and above the table
    Cs = [7.3e-4, 7.35e-4, 4 * 2e-5, -1e-4, 5e-3]
    base = passive(len(Cs))
    m = spliced(base, lambda r: ([CONC_VAL, r, NEWCONC1, CHAN_EK, IPOL1C, 0, 1, ADD_CURR], [-0.09, 1e-10, 0.5]),
                conc=Cs, extra_cols=conc_table(), ydivs=YDIVS, ymin=0.0, invdy=INVDY)
    o = Oracle(m, givals=False)
    o.step(600)
    chip = o.state()["chip"]
    for r, c in enumerate(Cs):
        lo, _ = seg(m, r)
        B, A = interp(0, c), interp(1, c)
        want = A / (2 * (B - 1))
        assert abs(chip[lo + 4] - want) <= 1e-12 * want, (r, c)


def test_domain_gate_reads_previous_current():
    V, g1, E1, scale, X0 = -0.06, 3e-9, 0.1, 2.5, 0.4
    base = passive(1, V0=[V])
    m = spliced(base, lambda r: ([CHAN_EK, ADD_CURR, DOMAIN, CHAN_EK, IPOL1C, 0, 1, ADD_CURR],
                                 [E1, g1, scale, -0.09, 1e-10, X0]),
                extra_cols=conc_table(), ydivs=YDIVS, ymin=0.0, invdy=INVDY)
    o = Oracle(m, givals=False)
    o.chip_update()
    c = scale * g1 * (E1 - V)
    B, A = interp(0, c), interp(1, c)
    got = o.state()["chip"][seg(m, 0)[0] + 7]
    want = (X0 * (2 - B) + A) / B
    assert abs(got - want) <= 1e-14 * want


def test_table_currents():
    C, V, gv, gc = 9.1e-4, -0.052, 2e-9, 5e-10
    base = passive(2, V0=[V, V])
    # voltage table current on a voltage column pair placed in the same table
    xdivs, xmin, invdx = YDIVS, -0.1, YDIVS / 0.15
    Vr = xmin + np.arange(xdivs + 2) / invdx
    vcols = np.stack([1 / (1 + np.exp(-(Vr + 0.04) / 0.005)), -0.03 + 0.1 * Vr], axis=1)
    cols = np.concatenate([conc_table(), vcols], axis=1)
    progs = {0: ([NEWVOLT_OP, CHAN, IPOL1V, 4, 0], [gv]),
             1: ([CONC_VAL, 0, NEWCONC1, CHAN, IPOL1C, 2, 0], [gc])}
    m = spliced(base, lambda r: progs[r], conc=[C], extra_cols=cols, ydivs=YDIVS, ymin=0.0, invdy=INVDY)
    assert m.xdivs == xdivs and m.invdx == invdx
    o = Oracle(m, givals=False)
    o.chip_update()
    v = (V - xmin) * invdx
    rv, fv = int(v), v - int(v)
    Gt = vcols[rv, 0] + fv * (vcols[rv + 1, 0] - vcols[rv, 0])
    Et = vcols[rv, 1] + fv * (vcols[rv + 1, 1] - vcols[rv, 1])
    G, EG = gk_ek(m, o, 0, V)
    assert abs(G - gv * Gt) <= 1e-9 * G and abs(EG - gv * Gt * Et) <= 1e-9 * abs(EG)
    G, EG = gk_ek(m, o, 1, V)
    Gc, Ec = interp(2, C), interp(3, C)
    assert abs(G - gc * Gc) <= 1e-9 * G and abs(EG - gc * Gc * Ec) <= 1e-9 * abs(EG)


NEWVOLT_OP = 5100


IPOL2V, IPOL2C, NEWCONC2, SAME_GHK, DO_TABGHK = 4003, 4004, 5111, 3002, 3003
NY2 = 8                 # concentration rows of the 2-D test tables


def table2d(rows, seed):
    """A 2-D gate block: B' at 2j, A' at 2j + 1 for concentration row j, rows
    along the first axis; values differ along both axes (a transposed read
    gives another number)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(50, 500, (rows, NY2 + 1)) + 30 * np.arange(rows)[:, None]
    blk = np.empty((rows, 2 * (NY2 + 1) + 2))
    blk[:, 0:2 * (NY2 + 1):2] = 1 + DT / 2 * (a + 4000.0)
    blk[:, 1:2 * (NY2 + 1):2] = DT * a
    blk[:, -2:] = blk[:, -4:-2]
    return blk


def bilinear(blk, r, f, j, g, off):
    t0, t1 = blk[r], blk[r + 1]
    B = (1 - f) * ((1 - g) * t0[2 * j] + g * t0[2 * j + 2]) + f * ((1 - g) * t1[2 * j] + g * t1[2 * j + 2])
    A = (1 - f) * ((1 - g) * t0[2 * j + 1] + g * t0[2 * j + 3]) + f * ((1 - g) * t1[2 * j + 1] + g * t1[2 * j + 3])
    return A / (2 * (B - 1))


def test_2d_gate_on_voltage_and_concentration():
    """IPOL2V: the gate settles on the table interpolated bilinearly at the
    voltage row/fraction and the concentration row/fraction."""
    inv_y = NY2 / 1e-3
    V, C = -0.0533, 6.1e-4
    rows = 152
    base = passive(1, Em=V, V0=[V])
    blk = table2d(rows, 1)
    m = spliced(base, lambda r: ([NEWVOLT_OP, CONC_VAL, 0, NEWCONC1, CHAN_EK, IPOL2V, 0, 1, ADD_CURR],
                                 [V, 1e-12, 0.5]),
                conc=[C], extra_cols=blk, ydivs=NY2, ymin=0.0, invdy=inv_y)
    o = Oracle(m, givals=False)
    o.step(400)
    v = (V - m.xmin) * m.invdx
    c = C * inv_y
    want = bilinear(blk, int(v), v - int(v), int(c), c - int(c), 0)
    got = o.state()["chip"][seg(m, 0)[0] + 4]
    assert abs(got - want) <= 1e-12 * want
    assert abs(o.vm[0] - V) <= 1e-15


def test_2d_gate_on_two_concentrations():
    """NEWCONC2 + IPOL2C: rows from the first concentration (the table
    register NEWCONC2 saves), columns from the second."""
    inv_y = NY2 / 1e-3
    C1, C2 = 3.3e-4, 8.6e-4
    blk = table2d(NY2 + 2, 2)
    base = passive(1)
    m = spliced(base, lambda r: ([CONC_VAL, 0, NEWCONC1, NEWCONC2, 1, CHAN_EK, IPOL2C, 0, 1, ADD_CURR],
                                 [-0.065, 1e-12, 0.5]),
                conc=[C1, C2], extra_cols=blk, ydivs=NY2, ymin=0.0, invdy=inv_y)
    o = Oracle(m, givals=False)
    o.step(400)
    a, b = C1 * inv_y, C2 * inv_y
    want = bilinear(blk, int(a), a - int(a), int(b), b - int(b), 0)
    got = o.state()["chip"][seg(m, 0)[0] + 4]
    assert abs(got - want) <= 1e-12 * want
    # swapping the two concentrations gives another value (the test sees the axes)
    assert abs(bilinear(blk, int(b), b - int(b), int(a), a - int(a), 0) - want) > 1e-6 * want


def test_table_ghk_and_same_ghk():
    """DO_TABGHK_OP + power-0 IPOL + SAME_GHK_OP: Ek = E(V), Gk = gmax G(V);
    SAME_GHK_OP after GHK_OP: a second gmax on the same slope."""
    V, g = -0.047, 2e-9
    xdivs, xmin, invdx = 100, -0.1, 100 / 0.15
    Vr = xmin + np.arange(xdivs + 2) / invdx
    cols = np.stack([1 / (1 + np.exp(-(Vr + 0.04) / 0.005)), -0.03 + 0.1 * Vr], axis=1)
    base = passive(1, V0=[V])
    m = spliced(base, lambda r: ([NEWVOLT_OP, DO_TABGHK, IPOL1V, 0, 0, SAME_GHK, ADD_CURR], [g]), extra_cols=cols)
    o = Oracle(m, givals=False)
    o.chip_update()
    v = (V - xmin) * invdx
    r, f = int(v), v - int(v)
    Gt = cols[r, 0] + f * (cols[r + 1, 0] - cols[r, 0])
    Et = cols[r, 1] + f * (cols[r + 1, 1] - cols[r, 1])
    G, EG = gk_ek(m, o, 0, V)
    assert abs(G - g * Gt) <= 1e-9 * G and abs(EG - g * Gt * Et) <= 1e-9 * abs(EG)
    # GHK channel P1 then SAME_GHK with P2: the current is (P1 + P2) times the GHK current per P
    Cin, Cout, P1, P2 = 1e-4, 2.0, 3e-15, 1e-15
    m = spliced(passive(1, V0=[V]), lambda r: ([CONC_VAL, 0, GHK, ADD_CURR, SAME_GHK, ADD_CURR],
                                               [ZFRT, 2 * F, P1, P2]), conc=[Cin, Cout])
    o = Oracle(m, givals=False)
    o.chip_update()
    G, EG = gk_ek(m, o, 0, V)
    want_I = ghk_current(V, P1 + P2, Cin, Cout)
    assert abs((EG - G * V) - want_I) <= 1e-9 * abs(want_I)
