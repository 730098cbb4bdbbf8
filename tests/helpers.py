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
