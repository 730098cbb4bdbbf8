"""Oracle pins from the mathematics (independent of the oracle's own code).

* passive isopotential compartment: backward Euler has the closed-form
  solution V_n = Vinf + (V0 - Vinf) q^n, q = 1/(1 + dt/tau);
* Crank-Nicolson: V_n = Vinf + (V0 - Vinf) q^n with q = (1 - dt/2tau)/(1 + dt/2tau);
* elimination: one step on random passive trees equals numpy.linalg.solve of
  the dense Hines matrix assembled independently from the element fields
  (off-diagonal -dt/(Cm_i Ra_link), Ra_link = the child's Ra);
* rest invariance: a passive tree at V = Em with no inject stays at Em;
* gate tables: the composite table equals 1 + dt/2 (alpha+beta) and dt*alpha
  of the closed-form HH rates at the breakpoints.
"""
import math

import numpy as np
import pytest

from hsolve_hip import model as M
from oracle import Oracle


def passive(parent, Cm, Rm, Ra, Em, V0, inject, dt, method=M.BEULER):
    n = len(parent)
    comps = dict(parent=parent, Cm=Cm, Rm=Rm, Ra=Ra, Em=Em, initVm=V0, inject=inject)
    return M.build(comps, [], [], dt=dt, method=method, float_fields=False)


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_passive_isopotential_closed_form(method):
    Cm, Rm, Em, I, V0, dt = 2.8e-11, 1.2e8, -0.065, 0.3e-9, -0.065, 20e-6
    m = passive([-1], [Cm], [Rm], [1e6], [Em], [V0], [I], dt, method)
    o = Oracle(m)
    tau = Rm * Cm
    vinf = Em + I * Rm
    if method == M.BEULER:
        q = 1.0 / (1.0 + dt / tau)
    else:
        q = (1.0 - dt / (2 * tau)) / (1.0 + dt / (2 * tau))
    for n in (1, 7, 200, 1000):
        o2 = Oracle(m)
        o2.step(n)
        want = vinf + (V0 - vinf) * q ** n
        assert abs(o2.vm[0] - want) < 1e-14 + 1e-12 * abs(want)


def _random_tree(rng, n, max_kids=3):
    parent = [-1]
    kids = [0]
    for i in range(1, n):
        while True:
            p = int(rng.integers(0, i))
            if kids[p] < max_kids:
                break
        parent.append(p)
        kids[p] += 1
        kids.append(0)
    # shuffle element order but keep a valid parent relation
    return parent


def _dense_step(parent, Cm, Rm, Ra, Em, V, inj, dt):
    n = len(parent)
    A = np.zeros((n, n))
    b = np.zeros(n)
    for i in range(n):
        A[i, i] = 1.0 + dt / (Cm[i] * Rm[i])
        b[i] = V[i] + dt / Cm[i] * (Em[i] / Rm[i] + inj[i])
    for k in range(n):
        p = parent[k]
        if p < 0:
            continue
        # axial resistance of the link = the child's Ra
        A[p, k] -= dt / (Cm[p] * Ra[k])
        A[p, p] += dt / (Cm[p] * Ra[k])
        A[k, p] -= dt / (Cm[k] * Ra[k])
        A[k, k] += dt / (Cm[k] * Ra[k])
    return np.linalg.solve(A, b)


@pytest.mark.parametrize("seed", range(12))
def test_elimination_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    ntree = int(rng.integers(1, 4))
    parent, Cm, Rm, Ra, Em, V, inj = [], [], [], [], [], [], []
    for t in range(ntree):
        n = int(rng.integers(1, 40))
        base = len(parent)
        par = _random_tree(rng, n, max_kids=int(rng.integers(1, 4)))
        for i in range(n):
            parent.append(-1 if par[i] < 0 else base + par[i])
            L, d = rng.uniform(10e-6, 100e-6), rng.uniform(0.5e-6, 4e-6)
            area = math.pi * d * L
            Cm.append(0.01 * area)
            Rm.append(0.33 / area)
            Ra.append(1.0 * L / (math.pi * (d / 2) ** 2))
            Em.append(-0.065 + rng.uniform(-0.01, 0.01))
            V.append(-0.065 + rng.uniform(-0.02, 0.02))
            inj.append(rng.uniform(0, 1e-10))
    dt = 20e-6
    m = passive(parent, Cm, Rm, Ra, Em, V, inj, dt)
    o = Oracle(m)
    o.step(1)
    got = o.vm[m.hnum]                 # element order
    want = _dense_step(parent, Cm, Rm, Ra, Em, V, inj, dt)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)


def test_crank_nicolson_dense():
    rng = np.random.default_rng(5)
    par = _random_tree(rng, 25)
    n = len(par)
    Cm = rng.uniform(1e-12, 1e-11, n)
    Rm = rng.uniform(1e8, 1e9, n)
    Ra = rng.uniform(1e6, 1e8, n)
    Em = np.full(n, -0.065)
    V = rng.uniform(-0.08, -0.05, n)
    inj = rng.uniform(0, 1e-10, n)
    dt = 50e-6
    m = passive(par, Cm, Rm, Ra, Em, V, inj, dt, M.CRANK)
    o = Oracle(m)
    o.step(1)
    x = _dense_step(par, Cm, Rm, Ra, Em, V, inj, dt / 2)
    assert np.allclose(o.vm[m.hnum], 2 * x - V, rtol=1e-12, atol=1e-15)


def test_rest_is_invariant():
    rng = np.random.default_rng(1)
    par = _random_tree(rng, 30)
    n = len(par)
    m = passive(par, np.full(n, 1e-11), np.full(n, 3e8), rng.uniform(1e6, 1e8, n),
                np.full(n, -0.065), np.full(n, -0.065), np.zeros(n), 20e-6)
    o = Oracle(m)
    o.step(50)
    assert np.max(np.abs(o.vm + 0.065)) < 1e-15


def test_gate_tables_closed_form():
    E = -0.065
    m = M.isopotential_hh(2, dt=20e-6)
    tab = m.tablist.reshape(m.xdivs + 2, m.ncols)
    dt = m.dt
    x = m.xvals[:m.xdivs + 1]

    def rate(a, xx):
        return (a[0] + a[1] * xx) / (a[2] + np.exp((xx + a[3]) / a[4]))
    na, k = M.hh_chantypes(E)
    # columns are registered K-X first (channels are ordered by ascending Ek)
    gates = [k["gates"][0], na["gates"][0], na["gates"][1]]
    for g, (pw, inst, mode, args) in enumerate(gates):
        alpha = rate(args[:5], x)
        beta = rate(args[5:], x)
        ok = np.isfinite(alpha) & (np.abs(args[2] + np.exp((x + args[3]) / args[4])) > 1e-6)
        Bp = 1.0 + dt / 2 * (alpha + beta)
        Ap = dt * alpha
        assert np.allclose(tab[:m.xdivs + 1, 2 * g][ok], Bp[ok], rtol=1e-12)
        assert np.allclose(tab[:m.xdivs + 1, 2 * g + 1][ok], Ap[ok], rtol=1e-12)
    # guard row repeats the top row
    assert np.array_equal(tab[m.xdivs + 1], tab[m.xdivs])


def test_malformed_funcs_stops_inside_the_rows():
    """A funcs[] program whose passes leave the row range (here: no SET_DIAG
    to reach the last of 5 rows) is reported, not run past results[]."""
    import dataclasses
    import pytest
    from hsolve_hip import model as M
    from oracle import Oracle
    m = M.population(5, (1,), (1.0,), seed=1)
    assert m.funcs.tolist() == [1, 1, 1, 7, 6, 6, 6, 6, 7]
    bad = dataclasses.replace(m, funcs=np.array([7, 6, 6, 6, 6, 7], np.int32))
    o = Oracle(bad, givals=False)
    with pytest.raises(RuntimeError):
        o.step(1)
    bad2 = dataclasses.replace(m, funcs=np.array([1, 1, 1, 7, 6, 6, 6, 6, 6, 6, 7], np.int32))
    with pytest.raises(RuntimeError):
        Oracle(bad2, givals=False).step(1)
