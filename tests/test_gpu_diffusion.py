"""Diffusion shells, buffers, pumps and concpools on the device against the
oracle (do_h4_conc_chip_update, hines_conc_chip.c:59-222; do_h_conc_solve,
hines_conc.c:39-396): each compartment's concentration program and diffusion
systems run at the end of its own program, in every kernel family (plain,
tiles of <= 16 and <= 64 rows, packs; specialised and interpreter kernels).

Without TAUPUMP_OP (the device's exp) the bar is bit-exact vm, chip, results,
ops and conc[].  With it: |vm - oracle| <= 1e-7 V and concentrations within
1e-9 relative.
"""
import os

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle

pytestmark = pytest.mark.gpu

PURE = ("ctaupump", "mmpump", "hillpump", "kc", "leak")
BUFFERED = ("fixbuf", "difbuf", "ctaupump", "mmpump", "hillpump", "elmmpump", "pool", "kc", "leak")


@pytest.fixture(params=["jit", "interp"], autouse=True)
def kernels(request, monkeypatch):
    if request.param == "interp":
        monkeypatch.setenv("HSH_NO_JIT", "1")
    else:
        monkeypatch.delenv("HSH_NO_JIT", raising=False)
    return request.param


def run_pair(m, nsteps, mode, **kw):
    print(f"[case] shells ncompts={m.ncompts} systems={m.ndiffs} steps={nsteps} mode={mode} {kw}", flush=True)
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
def test_shells_alone_exact(gpu, mode, kw, method):
    m = M.shell_population(150, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), shells=4, features=PURE, seed=3,
                           method=method)
    w, g = run_pair(m, 300, mode, **kw)
    outer = m.dconc + 3
    assert np.ptp(w["conc"][outer]) > 1e-6          # the shells move
    assert_exact(w, g)


@pytest.mark.parametrize("mode,kw", LAYOUTS, ids=IDS)
@pytest.mark.parametrize("shells", [2, 5])
def test_buffers_pools_and_electrogenic_pump_exact(gpu, mode, kw, shells):
    m = M.shell_population(120, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), shells=shells, features=BUFFERED,
                           seed=5, method=M.CRANK)
    w, g = run_pair(m, 300, mode, **kw)
    assert_exact(w, g)


@pytest.mark.parametrize("mode,kw", LAYOUTS[:2], ids=IDS[:2])
def test_taupump_within_tolerance(gpu, mode, kw):
    m = M.shell_population(120, (1, 8, 16), (0.4, 0.3, 0.3), shells=3, features=PURE + ("taupump",), seed=7,
                           method=M.BEULER)
    w, g = run_pair(m, 1000, mode, **kw)
    assert float(np.max(np.abs(w["vm"] - g["vm"]))) <= 1e-7
    rel = np.abs(w["conc"] - g["conc"]) / np.abs(w["conc"])
    assert float(np.max(rel)) <= 1e-9


def test_put_conc_mid_run(gpu):
    """hsh_put_conc between calls: the device continues from the new
    concentrations as the oracle does."""
    m = M.shell_population(64, (1, 8), (0.5, 0.5), shells=3, features=("fixbuf",) + PURE, seed=9, method=M.CRANK)
    o = Oracle(m, givals=False)
    with D.HipSolver(m, batch_k=8, flags=D.F_NO_AUTOTUNE) as g:
        o.step(50)
        g.step(50, D.BATCHED)
        c = o.conc() * 1.3
        o.put_conc(c)
        g.put_conc(c)
        o.step(70)
        g.step(70, D.BATCHED)
        assert np.array_equal(o.conc(), g.conc())
        assert np.array_equal(o.vm, g.vm)


def test_autotune_keeps_shell_state(gpu):
    """The autotuner times layouts on the model and puts its state back:
    afterwards the run equals the oracle's."""
    m = M.shell_population(200, (1, 8), (0.5, 0.5), shells=3, features=PURE, seed=11, method=M.BEULER)
    o = Oracle(m, givals=False)
    o.step(40)
    with D.HipSolver(m) as g:
        g.step(40, D.BATCHED)
        assert np.array_equal(o.conc(), g.conc())
        assert np.array_equal(o.vm, g.vm)
