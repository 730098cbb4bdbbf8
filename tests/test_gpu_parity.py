"""GPU parity: libhsolve_hip against the CPU oracle on the same arrays.

The device path rounds every expression as the reference does, so the bar is
bit-exact equality of vm[], chip[] (gate states), results[] and the mutable
ops[] counters.  The one documented exception is the NEWVOLT row cache across
tile boundaries in the batched kernel, which can differ only when a voltage
lies within rounding of a table breakpoint; the 1e-7 V bound the workload
states is asserted beside the bitwise check there.
"""
import os

import numpy as np
import pytest

from helpers import ap_verify_model, c3_model, max_abs, mixed_model, random_tree_model
from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["jit", "interp"], autouse=True)
def kernels(request, monkeypatch):
    """Every GPU test runs with the specialised kernels and with the built-in
    interpreter kernels."""
    if request.param == "interp":
        monkeypatch.setenv("HSH_NO_JIT", "1")
    else:
        monkeypatch.delenv("HSH_NO_JIT", raising=False)
    return request.param


def run_pair(m, nsteps, mode, **kw):
    print(f"[case] ncompts={m.ncompts} steps={nsteps} mode={mode} {kw}", flush=True)
    o = Oracle(m, givals=False)
    o.step(nsteps)
    with D.HipSolver(m, **kw) as g:
        want_jit = 0 if os.environ.get("HSH_NO_JIT") == "1" else 1
        assert g.info()["jit"] == want_jit
        r = g.step(nsteps, mode)
        assert r["vm_ready"] == 1
        st = g.state()
    return o.state(), st


def assert_same(os_, gs, *, exact=True, tol=1e-7):
    if exact:
        assert np.array_equal(os_["vm"], gs["vm"]), max_abs(os_["vm"], gs["vm"])
        assert np.array_equal(os_["chip"], gs["chip"]), max_abs(os_["chip"], gs["chip"])
        assert np.array_equal(os_["results"], gs["results"]), max_abs(os_["results"], gs["results"])
        assert np.array_equal(os_["ops"], gs["ops"])
    else:
        assert max_abs(os_["vm"], gs["vm"]) <= tol


def test_ap_verify_plain(gpu):
    m = ap_verify_model(8)
    os_, gs = run_pair(m, 210, D.PLAIN)
    assert_same(os_, gs)
    # the reference's CPU golden (cluster_bringup/accel_regression_golden.txt:4)
    assert "%0.10g" % gs["vm"][m.hnum[0]] == "-0.02143485882"


def test_ap_verify_batched_iso(gpu):
    m = ap_verify_model(8)
    os_, gs = run_pair(m, 210, D.BATCHED, batch_k=32)
    assert_same(os_, gs)


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
def test_isopotential_population(gpu, mode):
    m = M.isopotential_hh(1000, seed=0)
    os_, gs = run_pair(m, 200, mode, batch_k=16)
    assert_same(os_, gs)


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_random_trees(gpu, mode, method):
    m = random_tree_model(300, 16, seed=7, method=method)
    os_, gs = run_pair(m, 200, mode, batch_k=8)
    assert_same(os_, gs)


@pytest.mark.parametrize("ncomp", [3, 8, 16, 40, 64])
@pytest.mark.parametrize("waves", [0, 4])
def test_batched_kernel_variants(gpu, ncomp, waves):
    """Tiles only (tile_max_rows 64): rows <= 16 run the four-wave tiles
    (RPW 1/2/4) with register-held coefficients and the row-owner
    elimination; 17..64 rows the LDS kernel with ``waves`` waves (0 = rows/4)."""
    m = random_tree_model(200, ncomp, seed=ncomp)
    os_, gs = run_pair(m, 64, D.BATCHED, batch_k=16, batch_waves=waves, tile_max_rows=64,
                       flags=D.F_NO_AUTOTUNE)
    assert_same(os_, gs)


@pytest.mark.parametrize("ncomp", [17, 40, 64, 128, 150])
@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_plain_tile_elimination(gpu, ncomp, method):
    """Plain steps on tiles of larger trees (tile_max_rows 256): 17..128
    rows eliminate with the tile's rhs/diag in LDS, 150 rows in global
    memory."""
    m = random_tree_model(100, ncomp, seed=ncomp + 1, method=method)
    os_, gs = run_pair(m, 50, D.PLAIN, flags=D.F_NO_AUTOTUNE, tile_max_rows=256)
    assert_same(os_, gs)


@pytest.mark.parametrize("mode", [D.PLAIN, D.BATCHED])
@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_heterogeneous_population(gpu, mode, method):
    """Tree sizes 1..1024 in one solver: small trees in tiles, the 64-, 256-
    and 1024-row trees in packs with level-scheduled elimination."""
    m = c3_model(600, seed=5, method=method)
    os_, gs = run_pair(m, 60, mode, batch_k=8, flags=D.F_NO_AUTOTUNE)
    assert_same(os_, gs)


@pytest.mark.parametrize("tmr,cap,waves,stage", [(-1, 512, 4, 1), (-1, 1024, 8, 2), (16, 2048, 4, 2),
                                                 (64, 1024, 8, 1), (16, 100, 2, 1), (-1, 1024, 8, 3),
                                                 (-1, 256, 4, 3), (16, 2048, 2, 3)])
def test_pack_variants(gpu, tmr, cap, waves, stage):
    """Pack layouts and workgroup shapes: every tree in packs or only the
    large ones, pack capacity, waves per pack, schedule read from memory,
    staged in LDS, or its indices staged in LDS."""
    m = c3_model(300, seed=6)
    os_, gs = run_pair(m, 40, D.BATCHED, batch_k=16, flags=D.F_NO_AUTOTUNE, tile_max_rows=tmr,
                       pack_cap=cap, pack_waves=waves, pack_stage=stage)
    assert_same(os_, gs)


def test_autotune_keeps_state(gpu):
    """The autotuner runs candidate kernels on the handle's own state and
    puts it back: a tuned run continues exactly where it left off.  On this
    heterogeneous model the device build sweeps the layouts one kernel runs
    (all trees in packs); layouts mixing tiles and packs are swept only in
    the host-sanitizer build (tests/test_host_sanitizer.py)."""
    m = c3_model(300, seed=8)
    o = Oracle(m, givals=False)
    o.step(50)
    with D.HipSolver(m) as g:
        g.step(20, D.PLAIN)
        info = g.autotune()
        assert info["autotuned"] == 1 and info["batch_k"] in (1, 8, 32, 128)
        g.step(30, D.BATCHED)
        gs = g.state()
    assert_same(o.state(), gs)


def test_mixed_sizes_plain(gpu):
    m = mixed_model(500)
    os_, gs = run_pair(m, 100, D.PLAIN)
    assert_same(os_, gs)


def test_mixed_sizes_batched(gpu):
    m = mixed_model(500)
    os_, gs = run_pair(m, 100, D.BATCHED, batch_k=10, flags=D.F_NO_AUTOTUNE)
    assert_same(os_, gs)


def test_steps_split_across_calls(gpu):
    """K-step launches and graph replays compose: 3 calls == one run."""
    m = random_tree_model(130, 16, seed=11)
    o = Oracle(m, givals=False)
    o.step(77)
    with D.HipSolver(m, batch_k=8) as g:
        g.step(30, D.BATCHED)
        g.step(7, D.PLAIN)
        g.step(40, D.BATCHED)
        gs = g.state()
    assert_same(o.state(), gs)


@pytest.mark.parametrize("layout", [
    dict(tile_max_rows=16), dict(tile_max_rows=64), dict(tile_max_rows=-1, pack_cap=512, pack_waves=4, pack_stage=1),
    dict(tile_max_rows=-1, pack_cap=1024, pack_waves=8, pack_stage=2)], ids=["tiles16", "tiles64", "packs", "packs-staged"])
@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_one_step_per_call(gpu, layout, method):
    """The glue's per-PROCESS path: hsh_step(h, 1, HSH_BATCHED) once per step
    (each call one K = 1 launch of the fused kernel) equals the oracle."""
    m = M.population(120, (1, 8, 16, 64, 256), (0.3, 0.3, 0.2, 0.15, 0.05), seed=21, method=method)
    T = 60
    print(f"[case] one step per call {layout} method={method}", flush=True)
    o = Oracle(m, givals=False)
    o.step(T)
    with D.HipSolver(m, batch_k=128, flags=D.F_NO_AUTOTUNE, **layout) as g:
        for _ in range(T):
            g.step(1, D.BATCHED)
        gs = g.state()
    assert_same(o.state(), gs)
