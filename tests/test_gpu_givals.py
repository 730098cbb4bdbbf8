"""The chanmode-4/5 output stream givals[] written on the device (leak
current, {Gk, Ik, Ek} of every current, total membrane current;
hines_chip.c:164-165, :536-564, :598) against the oracle's, bit for bit,
after plain and batched runs."""
import os

import numpy as np
import pytest

from helpers import ap_verify_model, random_tree_model
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


def run(m, T, mode, table=None, **kw):
    print(f"[case] givals ncompts={m.ncompts} chanmode={m.chanmode} mode={mode} {kw}", flush=True)
    o = Oracle(m, givals=True)
    if table is not None:
        o.set_synapses(*table)
    o.step(T)
    with D.HipSolver(m, givals=True, **kw) as g:
        if table is not None:
            g.set_synapses(*table)
        g.step(T, mode)
        gg, gs = g.givals(), g.state()
    return o.state(), gs, gg


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
@pytest.mark.parametrize("chanmode", [4, 5])
def test_givals_hh_cells(gpu, mode, chanmode):
    m = ap_verify_model(8)
    if chanmode == 5:
        m = M.Model(**{**m.__dict__, "chanmode": 5})
    os_, gs, gg = run(m, 210, mode, batch_k=32, flags=D.F_NO_AUTOTUNE)
    assert np.array_equal(os_["vm"], gs["vm"])
    assert np.array_equal(os_["givals"], gg)
    assert np.any(gg != m.givals)                # written


@pytest.mark.parametrize("kw", [dict(tile_max_rows=16), dict(tile_max_rows=-1, pack_cap=512),
                                dict(tile_max_rows=-1, pack_cap=512, pack_stage=3)],
                         ids=["tiles", "packs", "packs-idx"])
def test_givals_trees_batched(gpu, kw):
    m = random_tree_model(150, 16, seed=3)
    os_, gs, gg = run(m, 100, D.BATCHED, batch_k=16, flags=D.F_NO_AUTOTUNE, **kw)
    assert np.array_equal(os_["vm"], gs["vm"]) and np.array_equal(os_["givals"], gg)


def test_givals_survive_autotuning(gpu):
    m = random_tree_model(150, 16, seed=4)
    os_, gs, gg = run(m, 80, D.BATCHED)
    assert np.array_equal(os_["givals"], gg)


def test_givals_network(gpu):
    """Synchans: {Gk, Ik, Ek} of each SYN2 current."""
    m, table = M.network(300, seed=5, inject=(0.3e-9, 0.8e-9))
    os_, gs, gg = run(m, 600, D.PLAIN, table=table, flags=D.F_NO_AUTOTUNE)
    assert np.array_equal(os_["vm"], gs["vm"]) and np.array_equal(os_["givals"], gg)
