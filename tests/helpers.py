"""Shared builders for the tests: the reference's own validation model and the
synthetic workloads at test sizes."""
from __future__ import annotations

import numpy as np

from hsolve_hip import model as M


def g10(x: float) -> float:
    """A GENESIS script float passed as a command argument: formatted with
    %0.10g (genesis/src/sim/sim_main.c:372-373, ss/eval.c:1075) and parsed
    back (Atod)."""
    return float("%0.10g" % x)


def ap_verify_model(n_neurons: int = 8) -> M.Model:
    """genesis/Scripts/benchmark/hh1952_ap_verify.g: N identical HH somas,
    2 nA inject, dt 10 us, chanmode 4, calcmode 1."""
    R = g10
    PI = 3.141592653589793
    E = -0.070
    ENA, EK, ELEAK = E + 0.115, E - 0.012, E + 0.010613
    L = D = 20e-6
    area = L * PI * D
    xarea = PI * (D / 2) * (D / 2)
    na, k = M.hh_chantypes(E, R(ENA), R(EK))
    for t in (na, k):
        t["gates"] = [(p, i, mo, [R(a) for a in args]) for (p, i, mo, args) in t["gates"]]
    n = n_neurons
    comps = dict(parent=[-1] * n, Cm=[R(0.01 * area)] * n, Rm=[R(0.333333 / area)] * n,
                 Ra=[R(0.3 * L / xarea)] * n, Em=[R(ELEAK)] * n, initVm=[R(E)] * n,
                 inject=[R(2e-9)] * n)
    chans = []
    for c in range(n):
        chans.append((c, 0, R(1200.0 * area)))
        chans.append((c, 1, R(360.0 * area)))
    return M.build(comps, [na, k], chans, dt=R(10e-6))


def random_tree_model(n_neurons: int, ncomp: int, *, seed: int = 1, method: int = M.BEULER,
                      dt: float = 20e-6) -> M.Model:
    """The multi-compartment workload at a test size (random Hines trees,
    HH soma, Kdr dendrites)."""
    return M.population(n_neurons, (ncomp,), (1.0,), seed=seed, method=method, dt=dt)


def mixed_model(n_neurons: int, *, seed: int = 3) -> M.Model:
    """Heterogeneous tree sizes in one solver."""
    return M.population(n_neurons, (1, 8, 16, 64), (0.4, 0.3, 0.2, 0.1), seed=seed)


def c3_model(n_neurons: int, *, seed: int = 4, sizes=(1, 8, 16, 64, 256, 1024), method: int = M.BEULER,
             dt: float = 2e-5) -> M.Model:
    """Tree sizes from {1, 8, 16, 64, 256, 1024} with P proportional to
    size^-0.67 (mean about 17 compartments per neuron)."""
    p = np.array(sizes, float) ** -0.67
    p /= p.sum()
    return M.population(n_neurons, tuple(sizes), tuple(p), seed=seed, method=method, dt=dt)


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) if np.size(a) else 0.0


def spliced(base: M.Model, add, *, conc=(), concops=(), concchip=(), extra_cols=None, ydivs=0,
            ymin=0.0, invdy=0.0) -> M.Model:
    """A copy of ``base`` (NEWVOLT / CHAN_EK / CHAN / IPOL1V / ADD_CURR
    programs) with ``add(row) -> (ops, chips)`` appended to each compartment's
    program, before its tbyc/diagterm; the concentration arrays as given and
    ``extra_cols`` (rows x 2k) appended to the gate table."""
    import dataclasses
    segs = M._segments(base)
    ops, chip, compops, compchips = [], [], [], []
    for r, (head, body, pre, post) in enumerate(segs):
        compops.append(len(ops)); compchips.append(len(chip))
        ops.append(head); chip += pre
        for (o, opnd, vals) in body:
            ops += [o] + opnd
            chip += vals
        a_ops, a_chip = add(r)
        ops += list(a_ops); chip += list(a_chip)
        chip += post
    ops.append(M._LCOMPT)
    tab, ncols = base.tablist, base.ncols
    if extra_cols is not None:
        t0 = base.tablist.reshape(base.xdivs + 2, base.ncols) if base.ncols else np.zeros((0, 0))
        rows = max(t0.shape[0], extra_cols.shape[0])
        t = np.zeros((rows, base.ncols + extra_cols.shape[1]))
        if base.ncols:
            t[:t0.shape[0], :base.ncols] = t0
            t[t0.shape[0]:, :base.ncols] = t0[-1]
        t[:extra_cols.shape[0], base.ncols:] = extra_cols
        t[extra_cols.shape[0]:, base.ncols:] = extra_cols[-1]
        tab, ncols = np.ascontiguousarray(t.reshape(-1)), t.shape[1]
    xvals, xdivs, xmin, invdx = base.xvals, base.xdivs, base.xmin, base.invdx
    if not base.ncols and extra_cols is not None:
        # a table with only concentration columns still needs a voltage axis
        xdivs, xmin = extra_cols.shape[0] - 2, -0.1
        invdx = xdivs / 0.15
        xvals = xmin + np.arange(xdivs + 2) / invdx
    return dataclasses.replace(
        base, ops=np.array(ops, np.int32), chip=np.array(chip, np.float64), tablist=tab, ncols=ncols,
        xdivs=xdivs, xvals=xvals, xmin=xmin, invdx=invdx, compops=np.array(compops, np.int32),
        compchips=np.array(compchips, np.int32), conc=np.array(conc, np.float64),
        concops=np.array(concops, np.int32), concchip=np.array(concchip, np.float64), ydivs=ydivs,
        ymin=ymin, invdy=invdy, givals=np.zeros(0))


MGBLOCK_OP = 4200


def with_mgblock(m: M.Model, which, A: float = 3.57, invB: float = 62.0, Mg: float = 1.2):
    """A copy of ``m`` whose synchans ``which`` (indices into m.syn_op) carry
    a Mg block: MGBLOCK_OP after the four SYN2_OP words and the chip values
    KMg_A, 1/KMg_B, [Mg] after X and Y, as SETUP writes them
    (hines_chip_init.c:1180-1187).  Returns (model, remap) where remap maps
    old ops indices to new ones (for connection tables)."""
    import dataclasses
    which = sorted(set(int(w) for w in which))
    op_at = [int(m.syn_op[w]) + 4 for w in which]          # insertion points in ops[]
    ch_at = [int(m.syn_chip[w]) + 2 for w in which]        # insertion points in chip[]
    ops = np.insert(np.asarray(m.ops), op_at, MGBLOCK_OP).astype(np.int32)
    chip = np.asarray(m.chip, np.float64)
    chip = np.insert(chip, np.repeat(ch_at, 3), np.tile([A, invB, Mg], len(ch_at)))
    op_at_a, ch_at_a = np.sort(op_at), np.sort(ch_at)

    def remap(q):
        q = np.asarray(q)
        return q + np.searchsorted(op_at_a, q, side="right")

    def remap_chip(q):
        q = np.asarray(q)
        return q + 3 * np.searchsorted(ch_at_a, q, side="right")
    out = dataclasses.replace(
        m, ops=ops, chip=chip, compops=remap(m.compops).astype(np.int32),
        compchips=remap_chip(m.compchips).astype(np.int32),
        spike_op=remap(m.spike_op).astype(np.int32), syn_op=remap(m.syn_op).astype(np.int32),
        syn_chip=remap_chip(m.syn_chip).astype(np.int32))
    return out, remap


def symmetric_model(n_trees: int, nmax: int, *, seed: int = 1, method: int = M.BEULER,
                    dt: float = 20e-6) -> M.Model:
    """Random trees of symmetric compartments (two kids per junction at most:
    the reference's program builder stops at three), HH Na and K in every
    root, K in the others, one solver."""
    rng = np.random.default_rng(seed)
    na, k = M.hh_chantypes(-0.065)
    parent, Cm, Rm, Ra, Em, V0, inj, chans = [], [], [], [], [], [], [], []
    for t in range(n_trees):
        n = int(rng.integers(1, nmax + 1))
        base = len(parent)
        nk = [0]
        for i in range(n):
            if i == 0:
                parent.append(-1)
            else:
                while True:
                    p = int(rng.integers(0, i))
                    if nk[p] < 2:
                        break
                nk[p] += 1
                parent.append(base + p)
            nk.append(0)
            L, d = (30e-6, 30e-6) if i == 0 else (rng.uniform(10e-6, 100e-6), rng.uniform(0.5e-6, 4e-6))
            area = np.pi * d * L
            Cm.append(0.01 * area)
            Rm.append(0.33 / area)
            Ra.append(1.0 * L / (np.pi * (d / 2) ** 2))
            Em.append(-0.065)
            V0.append(-0.065)
            inj.append(rng.uniform(0.2e-9, 1.0e-9) if i == 0 else 0.0)
            if i == 0:
                chans.append((base + i, 0, 1200.0 * area))
            chans.append((base + i, 1, (360.0 if i == 0 else 50.0) * area))
    comps = dict(parent=parent, Cm=Cm, Rm=Rm, Ra=Ra, Em=Em, initVm=V0, inject=inj)
    return M.build(comps, [na, k], chans, dt=dt, method=method, symmetric=True)
