"""Several solvers in one handle (hsh_create_group) against each solver run
on its own by the oracle: vm, chip, ops and conc of every member bit-exact,
in the plain path and in batched layouts; and a network of one-cell solvers
(the reference's VAnet2 layout) with synaptic events between members."""
import os

import numpy as np
import pytest

from helpers import random_tree_model
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


def members():
    return [M.isopotential_hh(50, seed=1),
            random_tree_model(40, 16, seed=2),
            M.ca_population(30, (1, 8, 16), (0.4, 0.3, 0.3), seed=3, ydivs=50,
                            features=("flux", "kc", "kc2d", "c1c2", "tabghk", "domain", "tabcurr", "ctabcurr")),
            M.population(12, (64,), (1.0,), seed=4),
            M.isopotential_hh(10, seed=5)]


@pytest.mark.parametrize("mode,kw", [
    (D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=16, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=512)),
], ids=["plain", "tiles", "packs"])
def test_group_equals_separate_solvers(gpu, mode, kw):
    ms = members()
    T = 300
    print(f"[case] group of {len(ms)} solvers mode={mode} {kw}", flush=True)
    with D.HipSolver(ms, **kw) as g:
        g.step(T, mode)
        st = g.state()
        conc = g.conc()
        offs = [g.member(i) for i in range(len(ms))]
    for i, m in enumerate(ms):
        o = Oracle(m, givals=False)
        o.step(T)
        w = o.state()
        e = offs[i]
        assert np.array_equal(st["vm"][e["row0"]:e["row0"] + e["nrows"]], w["vm"]), i
        assert np.array_equal(st["chip"][e["chip0"]:e["chip0"] + e["nchips"]], w["chip"]), i
        assert np.array_equal(st["ops"][e["op0"]:e["op0"] + e["nops"]], w["ops"][:-1]), i
        if m.conc.size:
            assert np.array_equal(conc[e["conc0"]:e["conc0"] + e["nconcs"]], o.conc()), i


def test_network_of_one_cell_solvers(gpu):
    """400 one-cell solvers, 2% connectivity across solvers: the group's raster
    and state are those of the same cells held by one solver."""
    n = 400
    ms, big, table = M.network_solvers(n, seed=8, inject=(0.3e-9, 0.8e-9))
    op0 = np.cumsum([0] + [m.ops.size - 1 for m in ms[:-1]])
    src = table[0]
    T = 1500
    print(f"[case] {n} one-cell solvers, {len(src)} connections", flush=True)
    o = Oracle(big, givals=False)
    o.set_synapses(*table)
    o.step(T)
    want = o.spikes()
    assert want[0].size > 100
    for mode, kw in [(D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)), (D.BATCHED, dict(batch_k=15, flags=D.F_NO_AUTOTUNE))]:
        with D.HipSolver(ms, **kw) as g:
            assert g.member(5)["op0"] == op0[5]
            g.set_synapses(*table)
            g.step(T, mode)
            got = g.drain_spikes()
            st = g.state()
        w = o.state()
        assert np.array_equal(st["vm"], w["vm"]) and np.array_equal(st["chip"], w["chip"])
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
