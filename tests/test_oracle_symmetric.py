"""Symmetric compartments: the setup compiler's elimination program for
symcompartment trees (hines_init.c:361-1160, symflag) and the oracle's
interpreter of it (hines_solve.c:140-296: COPY_ARRAY, SIBARRAY_ELIM,
FASTSIBARRAY_ELIM), pinned by dense solves.

The matrix is assembled here from circuit theory alone: each compartment is
two half-resistances Ra/2; at a junction (a parent's tail and its kids'
heads) eliminating the junction node couples every pair m, n of its members
with conductance g_m g_n / sum g (g = 2/Ra), i.e. resistance
Ra_n (1 + sum_{l != m} Ra_m/Ra_l)/2.  One step of the oracle on a passive
tree equals numpy.linalg.solve of that matrix; a planted mistake in the
oracle's interpreter or in the program fails it.
"""
import math

import numpy as np
import pytest

from hsolve_hip import model as M
from oracle import Oracle


def random_tree(rng, n, max_kids):
    parent, kids = [-1], [0]
    for i in range(1, n):
        while True:
            p = int(rng.integers(0, i))
            if kids[p] < max_kids:
                break
        parent.append(p)
        kids[p] += 1
        kids.append(0)
    return parent


def population(rng, ntree, nmax, max_kids):
    parent, Cm, Rm, Ra, Em, V, inj = [], [], [], [], [], [], []
    for t in range(ntree):
        n = int(rng.integers(1, nmax + 1))
        base = len(parent)
        par = random_tree(rng, n, max_kids)
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
    return [np.asarray(x, dtype=float) if i else x for i, x in enumerate((parent, Cm, Rm, Ra, Em, V, inj))]


def sym_matrix(parent, Cm, Rm, Ra, dt):
    n = len(parent)
    A = np.diag(1.0 + dt / (Cm * Rm))
    kids = [[] for _ in range(n)]
    for k, p in enumerate(parent):
        if p >= 0:
            kids[p].append(k)
    for p in range(n):
        if not kids[p]:
            continue
        J = [p] + kids[p]
        for m_ in J:
            coeff = (1.0 + sum(Ra[m_] / Ra[l] for l in J if l != m_)) / 2.0
            for n_ in J:
                if n_ == m_:
                    continue
                g = dt / (Cm[m_] * Ra[n_] * coeff)
                A[m_, n_] -= g
                A[m_, m_] += g
    return A


def build(parent, Cm, Rm, Ra, Em, V, inj, dt, method=M.BEULER):
    comps = dict(parent=parent, Cm=Cm, Rm=Rm, Ra=Ra, Em=Em, initVm=V, inject=inj)
    return M.build(comps, [], [], dt=dt, method=method, float_fields=False, symmetric=True)


def test_three_kid_junction_stops_like_the_reference():
    """The reference's program builder stops at its check #8
    (hines_init.c:956-960) when the middle one of three kids of a junction
    eliminates the first: the first kid's off-diagonal to the third is only
    recorded when the third kid's row comes, after the middle one's.  The
    restatement stops at the same check (GENESIS prints the error and keeps
    a partial program)."""
    parent = [-1, 0, 0, 0]            # a soma with three leaf kids
    n = len(parent)
    with pytest.raises(ValueError, match="check 8"):
        build(parent, np.full(n, 1e-11), np.full(n, 3e8), np.full(n, 1e7), np.full(n, -0.065),
              np.full(n, -0.065), np.zeros(n), 20e-6)
    # two kids per junction build
    build([-1, 0, 0, 1, 1], np.full(5, 1e-11), np.full(5, 3e8), np.full(5, 1e7), np.full(5, -0.065),
          np.full(5, -0.065), np.zeros(5), 20e-6)


def test_junction_matrix_is_symmetric_in_conductance():
    """Sanity of the assembly itself: Cm_m A_mn = Cm_n A_nm (a conductance
    matrix), rows sum to the leak term."""
    rng = np.random.default_rng(0)
    parent, Cm, Rm, Ra, Em, V, inj = population(rng, 1, 30, 4)
    dt = 20e-6
    A = sym_matrix(parent, Cm, Rm, Ra, dt)
    G = A * Cm[:, None]
    assert np.allclose(G, G.T, rtol=1e-13, atol=0)
    assert np.allclose(A.sum(axis=1), 1.0 + dt / (Cm * Rm), rtol=1e-12)


@pytest.mark.parametrize("seed", range(10))
def test_symmetric_elimination_matches_dense_solve(seed):
    rng = np.random.default_rng(100 + seed)
    parent, Cm, Rm, Ra, Em, V, inj = population(rng, int(rng.integers(1, 4)), 40, 2)
    dt = 20e-6
    m = build(parent, Cm, Rm, Ra, Em, V, inj, dt)
    o = Oracle(m)
    o.step(1)
    A = sym_matrix(parent, Cm, Rm, Ra, dt)
    b = V + dt / Cm * (Em / Rm + inj)
    want = np.linalg.solve(A, b)
    assert np.allclose(o.vm[m.hnum], want, rtol=1e-12, atol=1e-15)


def test_symmetric_crank_nicolson_dense():
    rng = np.random.default_rng(7)
    parent, Cm, Rm, Ra, Em, V, inj = population(rng, 2, 30, 2)
    dt = 50e-6
    m = build(parent, Cm, Rm, Ra, Em, V, inj, dt, M.CRANK)
    o = Oracle(m)
    o.step(1)
    A = sym_matrix(parent, Cm, Rm, Ra, dt / 2)
    b = V + dt / 2 / Cm * (Em / Rm + inj)
    x = np.linalg.solve(A, b)
    assert np.allclose(o.vm[m.hnum], 2 * x - V, rtol=1e-12, atol=1e-15)


def program_ops(funcs):
    """(forward?, opcode) of every op of an elimination program, parsed with
    each op's operand count (hines_solve.c:153-296)."""
    f = list(funcs)
    out, i, fwd = [], 0, True
    while i < len(f):
        op = f[i]
        out.append((fwd, op))
        if op == 7:
            if not fwd:
                break
            fwd = False
            i += 1
        elif op in (0, 3, 4):
            i += 2
        elif op == 5:
            i += 3
        else:
            i += 1
    return out


def test_programs_use_every_sibling_op():
    """Over branched trees the programs hold every symmetric op (forward
    COPY_ARRAY, FASTSIBARRAY_ELIM and SIBARRAY_ELIM, backward SIBARRAY_ELIM);
    an unbranched cable holds none of them."""
    seen = set()
    rng = np.random.default_rng(3)
    for t in range(6):
        parent, Cm, Rm, Ra, Em, V, inj = population(rng, 1, 40, 2)
        seen |= set(program_ops(build(parent, Cm, Rm, Ra, Em, V, inj, 20e-6).funcs))
    assert {(True, 3), (True, 4), (True, 5), (False, 5)} <= seen
    cable = build(list(range(-1, 9)), np.full(10, 1e-11), np.full(10, 3e8), np.full(10, 1e7),
                  np.full(10, -0.065), np.full(10, -0.06), np.zeros(10), 20e-6)
    assert not ({3, 4, 5} & {op for _, op in program_ops(cable.funcs)})


def test_rest_is_invariant_symmetric():
    rng = np.random.default_rng(11)
    parent, Cm, Rm, Ra, Em, V, inj = population(rng, 2, 30, 2)
    n = len(parent)
    m = build(parent, Cm, Rm, Ra, np.full(n, -0.065), np.full(n, -0.065), np.zeros(n), 20e-6)
    o = Oracle(m)
    o.step(50)
    assert np.max(np.abs(o.vm + 0.065)) < 1e-15
