"""Solvers of symmetric compartments on the device: the funcs[] interpreter
(k_funcs, one lane per solver) against the oracle, bit for bit -- vm,
results[], chip and ops -- for Euler and Crank-Nicolson, passive and HH
trees, requested plain or batched (a handle with units steps plain), after
hsh_set_ravals, and in a group with ordinary solvers."""
import os

import numpy as np
import pytest

from helpers import random_tree_model, symmetric_model
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


def same(os_, gs):
    for k in ("vm", "results", "chip", "ops"):
        assert np.array_equal(os_[k], gs[k]), (k, float(np.max(np.abs(os_[k] - gs[k]))))


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
@pytest.mark.parametrize("trees,nmax,seed", [(1, 40, 1), (6, 25, 2), (1, 200, 3)],
                         ids=["one-cell", "six-trees", "200-rows"])
def test_symmetric_solver_matches_oracle(method, mode, trees, nmax, seed):
    m = symmetric_model(trees, nmax, seed=seed, method=method)
    print(f"[case] symmetric ncompts={m.ncompts} method={method} mode={mode}", flush=True)
    o = Oracle(m, givals=False)
    o.step(300)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        g.step(150, mode)
        g.step(150, mode)
        same(o.state(), g.state())


def test_symmetric_set_ravals():
    """hsh_set_ravals rewrites a unit's ravals (constants; working slots are
    refilled by COPY_ARRAY each step): equal to the oracle on the new array."""
    m = symmetric_model(3, 30, seed=5)
    rav = m.ravals * np.linspace(0.8, 1.2, m.ravals.size)
    m2 = M.Model(**{**m.__dict__, "ravals": rav})
    o = Oracle(m2, givals=False)
    o.step(200)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        g.set_ravals(rav)
        g.step(200, D.PLAIN)
        same(o.state(), g.state())


def test_group_of_symmetric_and_ordinary_solvers():
    """Symmetric one-cell solvers and ordinary trees in one group handle:
    each member equal to its own oracle run."""
    ms = [symmetric_model(1, 30, seed=10 + i) for i in range(5)] + [random_tree_model(20, 16, seed=3)]
    print(f"[case] group of {len(ms)} solvers", flush=True)
    with D.HipSolver(ms, flags=D.F_NO_AUTOTUNE) as g:
        g.step(120, D.PLAIN)
        st = g.state()
        offs = [g.member(i) for i in range(len(ms))]
    for i, m in enumerate(ms):
        o = Oracle(m, givals=False)
        o.step(120)
        w, e = o.state(), offs[i]
        assert np.array_equal(st["vm"][e["row0"]:e["row0"] + e["nrows"]], w["vm"]), i
        assert np.array_equal(st["chip"][e["chip0"]:e["chip0"] + e["nchips"]], w["chip"]), i


def test_adjoint_refuses_symmetric():
    m = symmetric_model(1, 20, seed=7)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        with pytest.raises(D.Refused, match="symmetric compartments"):
            g.adj_forward(10, 5, np.array([m.ncompts - 1], np.int32))
