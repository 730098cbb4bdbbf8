"""Host-code scenarios run against the sanitizer build of libhsolve_hip
(tests/hostsan/_build: the library's host code under AddressSanitizer and
UBSan, linked with a mock HIP runtime whose launches check their pointers).
No kernel runs, so nothing here checks numbers: what is checked is that the
host code -- analysis, layouts, relayout and state restore in the autotuner,
graph capture and replay, event tables, inputs, groups, adjoint buffers --
reads and writes only memory it owns, hands kernels only live buffers of the
extent their grids use, and frees every device buffer it allocates.

usage: python scenarios.py NAME [NAME ...]   (run under LD_PRELOAD=libasan)
"""
from __future__ import annotations

import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "genesis-2.4_amd"))
sys.path.insert(0, os.path.join(HERE, ".."))

import numpy as np  # noqa: E402

from hsolve_hip import _paths  # noqa: E402

_paths.LIB_DIR = os.path.join(HERE, "_build")

from hsolve_hip import device as D  # noqa: E402
from hsolve_hip import model as M  # noqa: E402
from helpers import c3_model, random_tree_model, symmetric_model  # noqa: E402

INJECT, ACT = 1, 4100
SHELL_ALL = ("fixbuf", "difbuf", "ctaupump", "taupump", "mmpump", "hillpump", "elmmpump", "pool", "kc", "leak")
CA_2D = ("flux", "kc", "kc2d", "c1c2", "tabghk", "domain", "tabcurr", "ctabcurr")


def hetero():
    """The autotuner on trees of 1..1024 rows: every layout candidate, mixed
    classes included, relaid out and timed on the handle's state, the state
    put back, then plain and batched steps on the layout it kept."""
    m = c3_model(300, seed=8)
    with D.HipSolver(m) as g:
        g.step(20, D.PLAIN)
        g.autotune()
        g.step(30, D.BATCHED)
        g.step(7, D.PLAIN)
        st = g.state()
        g.put_state(vm=st["vm"])
        g.step(40, D.BATCHED)
    with D.HipSolver(m, pack_waves=8) as g:          # tuned on the first batched step
        g.step(12, D.BATCHED)


def layouts():
    m = c3_model(200, seed=6)
    for tmr in (16, 64, -1):
        for cap, waves, stage in ((512, 4, 1), (1024, 8, 2), (2048, 4, 2), (2048, 8, 3), (256, 2, 3)):
            with D.HipSolver(m, batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=tmr, pack_cap=cap,
                             pack_waves=waves, pack_stage=stage) as g:
                g.step(20, D.BATCHED)
                g.step(3, D.PLAIN)
                g.step(9, D.BATCHED)
                g.step(2, D.CHANNELS)


def network():
    """Event tables, posted events, message inputs (allocating and static),
    the simulation clock, spike drains, and the autotuner with lists on."""
    m, table = M.network(120, seed=4, inject=(0.2e-9, 0.6e-9))
    rng = np.random.default_rng(3)
    inj = [int(m.compchips[r]) + 1 for r in range(m.ncompts)]
    acts = [(ACT, int(m.syn_chip[j]), int(m.ops[int(m.syn_op[j]) + 1]), 1, 0.5) for j in range(4)]
    for mode, kw in ((D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)), (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE)),
                     (D.BATCHED, dict())):
        with D.HipSolver(m, **kw) as g:
            g.set_synapses(*table)
            g.step(25, mode)
            g.post_events(m.syn_op[:3], [2, 5, 9], [0.4, 0.1, 0.2])
            g.post_events(m.syn_op[3:5], [1, 3], [0.3, 0.3], after=True)
            g.set_inputs([(INJECT, inj[c], 0, 0, 1e-9) for c in rng.choice(m.ncompts, 6, replace=False)] + acts)
            arr = D.input_array(acts)
            assert D.lib().hsh_set_inputs_static(g.h, arr.size, arr.ctypes.data) == 0
            g.step(33, mode)
            g.set_time(0.5)
            g.step(10, mode)
            g.drain_spikes()
            g.state()


def shells():
    m = M.shell_population(40, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), shells=4, features=SHELL_ALL, seed=5,
                           method=M.CRANK)
    for mode, kw in ((D.PLAIN, dict()), (D.BATCHED, dict(batch_k=16, flags=D.F_NO_AUTOTUNE)),
                     (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=64)),
                     (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=512,
                                      pack_stage=2)),
                     (D.BATCHED, dict())):
        with D.HipSolver(m, **kw) as g:
            g.step(11, mode)
            c = g.conc()
            g.put_conc(c)
            g.step(5, mode)


def calcium():
    m = M.ca_population(60, (1, 8, 16, 40), (0.4, 0.3, 0.2, 0.1), features=CA_2D, seed=6, ydivs=50)
    for kw in (dict(), dict(tile_max_rows=-1, flags=D.F_NO_AUTOTUNE)):
        with D.HipSolver(m, **kw) as g:
            g.step(9, D.BATCHED)
            g.step(4, D.PLAIN)
            g.conc()


def givals():
    m = random_tree_model(40, 16, seed=3)
    for kw in (dict(), dict(tile_max_rows=-1, pack_cap=512, flags=D.F_NO_AUTOTUNE)):
        with D.HipSolver(m, givals=True, **kw) as g:
            g.step(9, D.BATCHED)
            g.step(4, D.PLAIN)
            g.givals()


def groups():
    ms, _, table = M.network_solvers(60, seed=8, inject=(0.3e-9, 0.8e-9))
    for mode, kw in ((D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)), (D.BATCHED, dict(batch_k=15, flags=D.F_NO_AUTOTUNE))):
        with D.HipSolver(ms, **kw) as g:
            g.member(5)
            g.set_synapses(*table)
            g.step(40, mode)
            g.drain_spikes()
            g.state()


def symmetric():
    for method in (M.BEULER, M.CRANK):
        m = symmetric_model(12, 40, seed=2, method=method)
        for mode, kw in ((D.PLAIN, dict()), (D.BATCHED, dict(batch_k=8, flags=D.F_NO_AUTOTUNE))):
            with D.HipSolver(m, **kw) as g:
                g.step(10, mode)
                g.state()


def adjoint():
    for method in (M.BEULER, M.CRANK):
        m = M.population(6, (8, 16, 100), (0.4, 0.4, 0.2), seed=5, method=method, float_fields=False)
        obs = np.array([0, 9, m.ncompts - 1], np.int32)
        for kw in (dict(tile_max_rows=16), dict(tile_max_rows=128), dict(tile_max_rows=-1, pack_cap=256),
                   dict(tile_max_rows=-1, pack_cap=256, flags=D.F_NO_JIT)):
            kw = dict(kw, flags=kw.get("flags", 0) | D.F_NO_AUTOTUNE)
            with D.HipSolver(m, **kw) as s:
                tr = s.adj_forward(60, 16, obs)
                s.adj_backward(np.ones_like(tr))
                s.set_ravals(m.ravals * 1.01)
                tr = s.adj_forward(30, 7, obs)
                s.adj_backward(np.ones_like(tr))


def iso():
    m = M.isopotential_hh(100, seed=2)
    with D.HipSolver(m) as g:
        g.step(50, D.BATCHED)
        g.step(3, D.PLAIN)
        g.set_chip([int(m.compchips[0]) + 1], [1e-9])
        g.step(5, D.BATCHED)


SCENARIOS = dict(hetero=hetero, layouts=layouts, network=network, shells=shells, calcium=calcium, givals=givals,
                 groups=groups, symmetric=symmetric, adjoint=adjoint, iso=iso)


def main(names):
    L = D.lib()
    for fn in L.mockhip_launches, L.mockhip_replays, L.mockhip_live, L.mockhip_modules:
        fn.restype = C.c_long
    assert L.mockhip_selfcheck() == 0
    for name in names:
        SCENARIOS[name]()
        live, mods = L.mockhip_live(), L.mockhip_modules()
        print(f"ok {name} launches={L.mockhip_launches()} replays={L.mockhip_replays()} live={live} modules={mods}",
              flush=True)
        assert live == 0 and mods == 0, (name, live, mods)


if __name__ == "__main__":
    main(sys.argv[1:] or list(SCENARIOS))
