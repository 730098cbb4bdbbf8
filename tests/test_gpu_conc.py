"""Concentrations on the device against the oracle: Ca_concen pools updated in
the program of the compartment that feeds them, Ca-gated channels
(NEWCONC1/IPOL1C), DOMAIN_OP, table currents, NERNST_OP and GHK_OP, in every
kernel family (plain, tiles, packs; specialised and interpreter kernels).

Without exp/log (Nernst, GHK) the bar is bit-exact vm, chip, results, ops
and conc[].  NERNST_OP and GHK_OP call the device's log/exp, which are not
the host libm's bit for bit, so there the bar is the solver tolerance:
|vm - oracle| <= 1e-7 V and pool concentrations within 1e-6 relative.
The concentration-row lookup finds the row afresh where the reference keeps a
row cache (hines_chip.c:236-252); the two agree except within rounding of a
table breakpoint.
"""
import os

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle

pytestmark = pytest.mark.gpu

EXACT = ("flux", "scale", "kc", "domain", "tabcurr", "ctabcurr")


@pytest.fixture(params=["jit", "interp"], autouse=True)
def kernels(request, monkeypatch):
    if request.param == "interp":
        monkeypatch.setenv("HSH_NO_JIT", "1")
    else:
        monkeypatch.delenv("HSH_NO_JIT", raising=False)
    return request.param


def run_pair(m, nsteps, mode, **kw):
    print(f"[case] ca ncompts={m.ncompts} steps={nsteps} mode={mode} {kw}", flush=True)
    o = Oracle(m, givals=False)
    o.step(nsteps)
    with D.HipSolver(m, **kw) as g:
        want_jit = 0 if os.environ.get("HSH_NO_JIT") == "1" else 1
        assert g.info()["jit"] == want_jit
        g.step(nsteps, mode)
        st = g.state()
        st["conc"] = g.conc()
    want = o.state()
    want["conc"] = o.conc()
    return want, st


def assert_exact(w, g):
    for k in ("vm", "chip", "results", "ops", "conc"):
        assert np.array_equal(w[k], g[k]), (k, float(np.max(np.abs(w[k] - g[k]))))


LAYOUTS = [
    (D.PLAIN, dict()),
    (D.BATCHED, dict(batch_k=16, flags=D.F_NO_AUTOTUNE)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=64)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=512, pack_stage=2)),
    (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=1024, pack_waves=8,
                     pack_stage=3)),
]
IDS = ["plain", "tiles16", "tiles64", "packs", "packs-idx"]


@pytest.mark.parametrize("mode,kw", LAYOUTS, ids=IDS)
@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_calcium_exact(gpu, mode, kw, method):
    m = M.ca_population(200, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), features=EXACT, seed=3, method=method)
    w, g = run_pair(m, 400, mode, **kw)
    assert np.ptp(w["conc"][0::2]) > 1e-4        # the pools move
    assert_exact(w, g)


@pytest.mark.parametrize("mode,kw", LAYOUTS, ids=IDS)
@pytest.mark.parametrize("rev", ["nernst", "ghk"])
def test_calcium_nernst_ghk(gpu, mode, kw, rev):
    m = M.ca_population(200, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), features=EXACT + (rev,), seed=4)
    w, g = run_pair(m, 1000, mode, **kw)
    assert float(np.max(np.abs(w["vm"] - g["vm"]))) <= 1e-7
    c0, c1 = w["conc"], g["conc"]
    assert float(np.max(np.abs(c0 - c1) / np.abs(c0))) <= 1e-6
    assert np.array_equal(w["ops"], g["ops"])


def test_put_conc_and_continue(gpu):
    """hsh_put_conc replaces the pools mid-run; the run continues as the
    oracle does from the same concentrations."""
    m = M.ca_population(100, (1, 8), (0.5, 0.5), features=EXACT, seed=5)
    print("[case] put_conc", flush=True)
    o = Oracle(m, givals=False)
    with D.HipSolver(m, batch_k=8, flags=D.F_NO_AUTOTUNE) as g:
        o.step(100)
        g.step(100, D.BATCHED)
        assert np.array_equal(o.conc(), g.conc())
        c = o.conc()
        c[0::2] *= 1.5
        o.put_conc(c)
        g.put_conc(c)
        assert np.array_equal(g.conc(), c)
        o.step(200)
        g.step(200, D.BATCHED)
        gs = g.state()
        gs["conc"] = g.conc()
    ws = o.state()
    ws["conc"] = o.conc()
    assert_exact(ws, gs)


def test_adjoint_refuses_concentrations(gpu):
    m = M.ca_population(4, (1,), (1.0,), features=("flux", "kc"), seed=1)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        with pytest.raises(D.Refused, match="concentration"):
            g.adj_forward(10, 5, [0])


EXACT2D = ("flux", "kc", "kc2d", "c1c2", "tabghk", "domain", "tabcurr", "ctabcurr")


@pytest.mark.parametrize("mode,kw", LAYOUTS, ids=IDS)
def test_calcium_2d_tables_exact(gpu, mode, kw):
    """2-D gates over (Vm, Ca) and (Ca, second concentration through the
    table register), table GHK: bit-exact."""
    m = M.ca_population(150, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), features=EXACT2D, seed=6, ydivs=50)
    w, g = run_pair(m, 400, mode, **kw)
    assert_exact(w, g)


@pytest.mark.parametrize("mode,kw", [LAYOUTS[0], LAYOUTS[3]], ids=[IDS[0], IDS[3]])
def test_calcium_same_ghk(gpu, mode, kw):
    m = M.ca_population(150, (1, 8, 16), (0.4, 0.3, 0.3), features=("flux", "kc", "ghk", "sameghk", "kc2d"),
                        seed=7, ydivs=50)
    w, g = run_pair(m, 1000, mode, **kw)
    assert float(np.max(np.abs(w["vm"] - g["vm"]))) <= 1e-7
    assert float(np.max(np.abs(w["conc"] - g["conc"]) / np.abs(w["conc"]))) <= 1e-6
