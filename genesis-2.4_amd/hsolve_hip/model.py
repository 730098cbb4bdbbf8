"""Synthetic hsolve populations: Python view of libhsolve_model.

The setup compiler (``csrc/model/hsm_model.cpp``) emits the arrays GENESIS's
hsolve builds at SETUP/RESET (``genesis/src/hines/hines_init.c:40-1273``,
``hines_chip_init.c:465-1552``).  This module wraps it with ctypes and returns
numpy copies, so tests, the benchmark and the device library all consume the
same arrays.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from ._paths import lib_path

BEULER = 10
CRANK = 11

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class GateSpec(C.Structure):
    _fields_ = [("power", C.c_int32), ("instant", C.c_int32), ("mode", C.c_int32),
                ("pad", C.c_int32), ("args", C.c_double * 10)]


class ChanType(C.Structure):
    _fields_ = [("Ek", C.c_double), ("gate", GateSpec * 3)]


class SynType(C.Structure):
    _fields_ = [("Ek", C.c_double), ("tau1", C.c_double), ("tau2", C.c_double)]


class Spec(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("method", C.c_int32), ("chanmode", C.c_int32),
        ("xdivs", C.c_int32), ("float_fields", C.c_int32), ("xmin", C.c_double), ("xmax", C.c_double),
        ("ncomp", C.c_int32), ("pad1", C.c_int32), ("parent", _i32p),
        ("Cm", _f64p), ("Rm", _f64p), ("Ra", _f64p), ("Em", _f64p), ("initVm", _f64p),
        ("inject", _f64p),
        ("nchantypes", C.c_int32), ("nchan", C.c_int32), ("chantypes", C.POINTER(ChanType)),
        ("chan_comp", _i32p), ("chan_type", _i32p), ("chan_gbar", _f64p),
        ("nspike", C.c_int32), ("pad2", C.c_int32), ("spike_comp", _i32p),
        ("spike_thresh", _f64p), ("spike_refract", _f64p),
        ("nsyntypes", C.c_int32), ("nsyn", C.c_int32), ("syntypes", C.POINTER(SynType)),
        ("syn_comp", _i32p), ("syn_type", _i32p), ("syn_gmax", _f64p),
        ("symmetric", C.c_int32), ("pad3", C.c_int32),
    ]


class Out(C.Structure):
    _fields_ = [
        ("ncompts", C.c_int32), ("nchips", C.c_int32), ("nops", C.c_int32),
        ("nfuncs", C.c_int32), ("nravals", C.c_int32), ("ncols", C.c_int32),
        ("xdivs", C.c_int32), ("ngivals", C.c_int32), ("nstab", C.c_int32),
        ("method", C.c_int32), ("chanmode", C.c_int32), ("nchildren", C.c_int32),
        ("xmin", C.c_double), ("invdx", C.c_double), ("dt", C.c_double),
        ("ops", _i32p), ("chip", _f64p), ("tablist", _f64p), ("xvals", _f64p),
        ("funcs", _i32p), ("ravals", _f64p), ("vm", _f64p), ("givals", _f64p),
        ("stablist", _f64p), ("compops", _i32p), ("compchips", _i32p),
        ("elmnum", _i32p), ("hnum", _i32p), ("parent_row", _i32p),
        ("spike_row", _i32p), ("spike_op", _i32p), ("syn_op", _i32p), ("syn_chip", _i32p),
    ]


class Pop(C.Structure):
    _fields_ = [
        ("n_neurons", C.c_int32), ("nsizes", C.c_int32), ("sizes", _i32p), ("probs", _f64p),
        ("seed", C.c_uint64), ("dt", C.c_double), ("method", C.c_int32), ("max_kids", C.c_int32),
        ("soma_len", C.c_double), ("soma_dia", C.c_double),
        ("dend_len_lo", C.c_double), ("dend_len_hi", C.c_double),
        ("dend_dia_lo", C.c_double), ("dend_dia_hi", C.c_double),
        ("CM", C.c_double), ("RM", C.c_double), ("RA", C.c_double), ("Em", C.c_double),
        ("gNa", C.c_double), ("gK", C.c_double), ("gKdr", C.c_double),
        ("inject_lo", C.c_double), ("inject_hi", C.c_double), ("erest_act", C.c_double),
        ("float_fields", C.c_int32), ("first_neuron", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        p = lib_path("libhsolve_model.so")
        _lib = C.CDLL(p)
        _lib.hsm_build.argtypes = [C.POINTER(Spec), C.POINTER(C.POINTER(Out))]
        _lib.hsm_build.restype = C.c_int32
        _lib.hsm_build_population.argtypes = [C.POINTER(Pop), C.POINTER(C.POINTER(Out))]
        _lib.hsm_build_population.restype = C.c_int32
        _lib.hsm_free.argtypes = [C.POINTER(Out)]
        _lib.hsm_last_error.restype = C.c_char_p
        _lib.hsm_tabinterp.argtypes = [_f64p, C.c_int32, C.c_double, C.c_double, C.c_double]
        _lib.hsm_tabinterp.restype = C.c_double
    return _lib


@dataclass
class Model:
    """Hsolve-format arrays of one solver (row index = Hines number)."""
    ncompts: int
    dt: float
    method: int
    chanmode: int
    xdivs: int
    xmin: float
    invdx: float
    ncols: int
    ops: np.ndarray
    chip: np.ndarray
    tablist: np.ndarray
    xvals: np.ndarray
    funcs: np.ndarray
    ravals: np.ndarray
    vm: np.ndarray
    givals: np.ndarray
    stablist: np.ndarray
    compops: np.ndarray
    compchips: np.ndarray
    elmnum: np.ndarray
    hnum: np.ndarray
    parent_row: np.ndarray
    spike_row: np.ndarray
    spike_op: np.ndarray
    syn_op: np.ndarray
    syn_chip: np.ndarray
    extra: dict = field(default_factory=dict)
    # concentrations (hines_struct.h:315-376): pool values, the Ca_concen
    # program (CACONC_OP k ... LCACONC_OP) and its three values per pool
    conc: np.ndarray = field(default_factory=lambda: np.zeros(0))
    concops: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    concchip: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ydivs: int = 0
    ymin: float = 0.0
    invdy: float = 0.0
    # diffusion systems of shells and buffers (h_init_conc_solve,
    # hines_conc_init.c:59-291): per system its shells, rows per shell, fixed
    # buffers, first concentration and first buffer-rate value; per
    # compartment (Hines row) its first system (ncompts + 1 entries); the
    # coupling values in the order the solve consumes them; buffer rates
    nshells: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nshrows: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nshfix: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dconc: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    fbufval: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    fdiffs: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    covals0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    bufvals: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def ndiffs(self):
        return int(self.nshells.size)

    @property
    def nchips(self):
        return int(self.chip.size)

    @property
    def nops(self):
        return int(self.ops.size)

    @property
    def nstab(self):
        return int(self.stablist.size // 4)


def _arr(ptr, n, dt):
    if n <= 0:
        return np.zeros(0, dtype=dt)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


def _take(outp, nspike, nsyn) -> Model:
    o = outp.contents
    n = o.ncompts
    m = Model(
        ncompts=n, dt=o.dt, method=o.method, chanmode=o.chanmode, xdivs=o.xdivs,
        xmin=o.xmin, invdx=o.invdx, ncols=o.ncols,
        ops=_arr(o.ops, o.nops, np.int32), chip=_arr(o.chip, o.nchips, np.float64),
        tablist=_arr(o.tablist, (o.xdivs + 2) * o.ncols if o.ncols else 0, np.float64),
        xvals=_arr(o.xvals, o.xdivs + 2 if o.ncols else 0, np.float64),
        funcs=_arr(o.funcs, o.nfuncs, np.int32), ravals=_arr(o.ravals, o.nravals, np.float64),
        vm=_arr(o.vm, n, np.float64), givals=_arr(o.givals, o.ngivals, np.float64),
        stablist=_arr(o.stablist, 4 * o.nstab, np.float64),
        compops=_arr(o.compops, n, np.int32), compchips=_arr(o.compchips, n, np.int32),
        elmnum=_arr(o.elmnum, n, np.int32), hnum=_arr(o.hnum, n, np.int32),
        parent_row=_arr(o.parent_row, n, np.int32),
        spike_row=_arr(o.spike_row, nspike, np.int32), spike_op=_arr(o.spike_op, nspike, np.int32),
        syn_op=_arr(o.syn_op, nsyn, np.int32), syn_chip=_arr(o.syn_chip, nsyn, np.int32),
    )
    lib().hsm_free(outp)
    return m


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def hh_chantypes(erest: float, ena: float | None = None, ek: float | None = None):
    """Na (X^3 Y) and K (X^4) HH squid-axon gate specs as in
    genesis/Scripts/benchmark/hh1952_ap_verify.g:49-62."""
    E = erest
    ena = E + 0.115 if ena is None else ena
    ek = E - 0.012 if ek is None else ek
    na = dict(Ek=ena, gates=[
        (3, 0, 0, [100000.0 * (0.025 + E), -100000.0, -1.0, -1.0 * (0.025 + E), -0.01,
                   4000.0, 0.0, 0.0, 0.0 - E, 0.018]),
        (1, 0, 0, [70.0, 0.0, 0.0, 0.0 - E, 0.02, 1000.0, 0.0, 1.0, -0.03 - E, -0.01]),
    ])
    k = dict(Ek=ek, gates=[
        (4, 0, 0, [10000.0 * (0.010 + E), -10000.0, -1.0, -1.0 * (0.010 + E), -0.01,
                   125.0, 0.0, 0.0, 0.0 - E, 0.08]),
    ])
    return na, k


def build(comps: dict, chantypes: list, chans: list, *, dt: float, method: int = BEULER,
          chanmode: int = 4, xdivs: int = 3000, xmin: float = -0.1, xmax: float = 0.05,
          spikes: list | None = None, syntypes: list | None = None,
          syns: list | None = None, float_fields: bool = True, symmetric: bool = False) -> Model:
    """Build a solver from element-level descriptions.

    comps: dict of equal-length sequences parent, Cm, Rm, Ra, Em, initVm, inject
           (element-list order; parent is an element index or -1).
    chantypes: list of dict(Ek=..., gates=[(power, instant, mode, args10), ...]) (X, Y, Z)
    chans: list of (comp, type, gbar) in creation order.
    spikes: list of (comp, thresh, abs_refract).
    syntypes: list of dict(Ek, tau1, tau2); syns: list of (comp, type, gmax).
    symmetric: the compartments are symcompartments (siblings couple through
    their junction; the elimination program then uses COPY_ARRAY /
    SIBARRAY_ELIM / FASTSIBARRAY_ELIM).
    """
    spikes = spikes or []
    syntypes = syntypes or []
    syns = syns or []
    n = len(comps["parent"])
    arrs = {k: np.ascontiguousarray(comps[k], dtype=np.float64)
            for k in ("Cm", "Rm", "Ra", "Em", "initVm", "inject")}
    parent = np.ascontiguousarray(comps["parent"], dtype=np.int32)
    ct = (ChanType * max(1, len(chantypes)))()
    for i, t in enumerate(chantypes):
        ct[i].Ek = t["Ek"]
        for g, (pw, inst, mode, args) in enumerate(t["gates"]):
            ct[i].gate[g].power = pw
            ct[i].gate[g].instant = inst
            ct[i].gate[g].mode = mode
            for j in range(10):
                ct[i].gate[g].args[j] = args[j]
    st = (SynType * max(1, len(syntypes)))()
    for i, t in enumerate(syntypes):
        st[i].Ek, st[i].tau1, st[i].tau2 = t["Ek"], t["tau1"], t["tau2"]
    chan_comp = np.array([c[0] for c in chans] or [0], dtype=np.int32)
    chan_type = np.array([c[1] for c in chans] or [0], dtype=np.int32)
    chan_gbar = np.array([c[2] for c in chans] or [0.0], dtype=np.float64)
    sp_comp = np.array([s[0] for s in spikes] or [0], dtype=np.int32)
    sp_th = np.array([s[1] for s in spikes] or [0.0], dtype=np.float64)
    sp_rf = np.array([s[2] for s in spikes] or [0.0], dtype=np.float64)
    sy_comp = np.array([s[0] for s in syns] or [0], dtype=np.int32)
    sy_type = np.array([s[1] for s in syns] or [0], dtype=np.int32)
    sy_g = np.array([s[2] for s in syns] or [0.0], dtype=np.float64)
    s = Spec()
    s.dt, s.method, s.chanmode, s.xdivs, s.xmin, s.xmax = dt, method, chanmode, xdivs, xmin, xmax
    s.float_fields = 1 if float_fields else 0
    s.symmetric = 1 if symmetric else 0
    s.ncomp = n
    s.parent = _p(parent, C.c_int32)
    for k in ("Cm", "Rm", "Ra", "Em", "initVm", "inject"):
        setattr(s, k, _p(arrs[k], C.c_double))
    s.nchantypes, s.nchan = len(chantypes), len(chans)
    s.chantypes = ct
    s.chan_comp, s.chan_type, s.chan_gbar = (_p(chan_comp, C.c_int32), _p(chan_type, C.c_int32),
                                             _p(chan_gbar, C.c_double))
    s.nspike = len(spikes)
    s.spike_comp, s.spike_thresh, s.spike_refract = (_p(sp_comp, C.c_int32), _p(sp_th, C.c_double),
                                                     _p(sp_rf, C.c_double))
    s.nsyntypes, s.nsyn = len(syntypes), len(syns)
    s.syntypes = st
    s.syn_comp, s.syn_type, s.syn_gmax = (_p(sy_comp, C.c_int32), _p(sy_type, C.c_int32),
                                          _p(sy_g, C.c_double))
    outp = C.POINTER(Out)()
    rc = lib().hsm_build(C.byref(s), C.byref(outp))
    if rc != 0:
        raise ValueError(lib().hsm_last_error().decode())
    return _take(outp, len(spikes), len(syns))


# ---------------------------------------------------------------------------
# Synthetic workloads

def population(n_neurons: int, sizes=(16,), probs=(1.0,), *, seed: int = 0, dt: float = 20e-6,
               method: int = BEULER, soma=(30e-6, 30e-6), dend_len=(10e-6, 100e-6),
               dend_dia=(0.5e-6, 4e-6), CM=0.01, RM=0.33, RA=1.0, Em=-0.065,
               gNa=1200.0, gK=360.0, gKdr=50.0, inject=(0.2e-9, 1.0e-9),
               max_kids: int = 3, float_fields: bool = True, first_neuron: int = 0) -> Model:
    """Random Hines trees: soma (root, Na/K) plus dendrites (leak + Kdr).

    Compartment i>0 takes a parent uniform over compartments 0..i-1, resampled
    while that parent already has ``max_kids`` children.  Sizes are drawn from
    ``sizes`` with weights ``probs``.  All values SI.
    """
    sz = np.ascontiguousarray(sizes, dtype=np.int32)
    pr = np.ascontiguousarray(probs, dtype=np.float64)
    p = Pop()
    p.n_neurons, p.nsizes = n_neurons, len(sz)
    p.sizes, p.probs = _p(sz, C.c_int32), _p(pr, C.c_double)
    p.seed, p.dt, p.method, p.max_kids = seed, dt, method, max_kids
    p.soma_len, p.soma_dia = soma
    p.dend_len_lo, p.dend_len_hi = dend_len
    p.dend_dia_lo, p.dend_dia_hi = dend_dia
    p.CM, p.RM, p.RA, p.Em = CM, RM, RA, Em
    p.gNa, p.gK, p.gKdr = gNa, gK, gKdr
    p.inject_lo, p.inject_hi = inject
    p.erest_act = Em
    p.float_fields = 1 if float_fields else 0
    p.first_neuron = first_neuron
    outp = C.POINTER(Out)()
    rc = lib().hsm_build_population(C.byref(p), C.byref(outp))
    if rc != 0:
        raise ValueError(lib().hsm_last_error().decode())
    return _take(outp, 0, 0)


def isopotential_hh(n: int, *, seed: int = 0, dt: float = 20e-6, inject=(0.1e-9, 0.5e-9)) -> Model:
    """N single-compartment HH cells: 30x30 um soma, Cm 1 uF/cm^2,
    Rm 3.3 kOhm cm^2, Em -65 mV, Na/K 1200/360 S/m^2, constant inject."""
    return population(n, (1,), (1.0,), seed=seed, dt=dt, inject=inject)


def network(n_cells: int = 4000, *, seed: int = 0, dt: float = 1e-4, frac_exc: float = 0.8,
            p_conn: float = 0.02, g_exc: float = 6e-9, g_inh: float = 67e-9, tau_exc=(5e-3, 1e-3),
            tau_inh=(10e-3, 1e-3), E_exc: float = 0.0, E_inh: float = -0.08, delay: float = 1.5e-3,
            vm0=(-0.070, -0.050), inject=(0.0, 0.0), thresh: float = 0.0, abs_refract: float = 5e-3):
    """Event-driven network of isopotential HH cells (30x30 um soma, Na/K
    1200/360 S/m^2, Em -65 mV), one spikegen and two synchans (excitatory,
    inhibitory; dual exponential, as the reference's VAnet2 synchans) per cell.
    The first frac_exc of the cells are excitatory; each ordered pair (i, j),
    i != j, is connected with probability p_conn to j's synchan of i's kind,
    weight 1 (gmax carries the strength), delay ``delay``.  Initial Vm uniform
    in vm0, inject uniform in ``inject``, from ``seed``.

    Returns (model, (src_op, dst_op, weight, delay)) -- the connection table in
    the form hsh_set_synapses takes."""
    rng = np.random.default_rng(seed)
    L = D = 30e-6
    area = np.pi * D * L
    n = n_cells
    comps = dict(parent=[-1] * n, Cm=[0.01 * area] * n, Rm=[0.33 / area] * n, Ra=[1e6] * n,
                 Em=[-0.065] * n, initVm=list(rng.uniform(vm0[0], vm0[1], n)),
                 inject=list(rng.uniform(inject[0], inject[1], n)))
    chans = []
    for i in range(n):
        chans.append((i, 0, 1200.0 * area))
        chans.append((i, 1, 360.0 * area))
    spikes = [(i, thresh, abs_refract) for i in range(n)]
    syntypes = [dict(Ek=E_exc, tau1=tau_exc[0], tau2=tau_exc[1]),
                dict(Ek=E_inh, tau1=tau_inh[0], tau2=tau_inh[1])]
    syns = []
    for i in range(n):
        syns.append((i, 0, g_exc))
        syns.append((i, 1, g_inh))
    m = build(comps, hh_chantypes(-0.065), chans, dt=dt, spikes=spikes, syntypes=syntypes, syns=syns)
    n_exc = int(round(frac_exc * n))
    src, dst = [], []
    conn = rng.random((n, n)) < p_conn
    np.fill_diagonal(conn, False)
    for i in range(n):
        js = np.nonzero(conn[i])[0]
        kind = 0 if i < n_exc else 1
        src.extend([int(m.spike_op[i])] * js.size)
        dst.extend(int(m.syn_op[2 * j + kind]) for j in js)
    k = len(src)
    table = (np.array(src, np.int32), np.array(dst, np.int32), np.ones(k), np.full(k, delay))
    return m, table


def network_solvers(n_cells: int = 4000, *, seed: int = 0, dt: float = 1e-4, frac_exc: float = 0.8,
                    p_conn: float = 0.02, g_exc: float = 6e-9, g_inh: float = 67e-9, tau_exc=(5e-3, 1e-3),
                    tau_inh=(10e-3, 1e-3), delay: float = 1.5e-3, vm0=(-0.070, -0.050), inject=(0.0, 0.0),
                    thresh: float = 0.0, abs_refract: float = 5e-3):
    """The network of :func:`network` in the reference's VAnet2 layout: one
    solver per cell.  Returns (models, merged, table): the per-cell solvers,
    the same cells as one solver (arrays concatenated; the funcs[] program of
    n isolated rows, SET_DIAG to the last row, then CALC_RESULTS back), and the
    connection table in group ops indices (which are also the merged solver's
    ops indices)."""
    import dataclasses
    rng = np.random.default_rng(seed)
    L = D = 30e-6
    area = np.pi * D * L
    na, k = hh_chantypes(-0.065)
    syntypes = [dict(Ek=0.0, tau1=tau_exc[0], tau2=tau_exc[1]), dict(Ek=-0.08, tau1=tau_inh[0], tau2=tau_inh[1])]
    models = []
    for i in range(n_cells):
        comps = dict(parent=[-1], Cm=[0.01 * area], Rm=[0.33 / area], Ra=[1e6], Em=[-0.065],
                     initVm=[rng.uniform(vm0[0], vm0[1])], inject=[rng.uniform(inject[0], inject[1])])
        models.append(build(comps, [na, k], [(0, 0, 1200.0 * area), (0, 1, 360.0 * area)], dt=dt,
                            spikes=[(0, thresh, abs_refract)], syntypes=syntypes,
                            syns=[(0, 0, g_exc), (0, 1, g_inh)]))
    n = n_cells
    ops = np.concatenate([m.ops[:-1] for m in models] + [np.array([102], np.int32)]).astype(np.int32)
    funcs = np.array([1] * (n - 2) + [7] + [6] * (n - 1) + [7] if n > 1 else [7, 7], np.int32)
    merged = dataclasses.replace(models[0], ncompts=n, ops=ops, chip=np.concatenate([m.chip for m in models]),
                                 vm=np.concatenate([m.vm for m in models]), funcs=funcs, ravals=np.zeros(0),
                                 givals=np.zeros(0))
    op0 = np.cumsum([0] + [m.ops.size - 1 for m in models[:-1]])
    n_exc = int(round(frac_exc * n))
    conn = rng.random((n, n)) < p_conn
    np.fill_diagonal(conn, False)
    src, dst = [], []
    for i in range(n):
        js = np.nonzero(conn[i])[0]
        kind = 0 if i < n_exc else 1
        src.extend([int(op0[i] + models[i].spike_op[0])] * js.size)
        dst.extend(int(op0[j] + models[j].syn_op[kind]) for j in js)
    kk = len(src)
    table = (np.array(src, np.int32), np.array(dst, np.int32), np.ones(kk), np.full(kk, delay))
    return models, merged, table


def tabinterp(table: np.ndarray, xdivs: int, xmin: float, invdx: float, x: float) -> float:
    t = np.ascontiguousarray(table, dtype=np.float64)
    return lib().hsm_tabinterp(_p(t, C.c_double), xdivs, xmin, invdx, x)


# ---------------------------------------------------------------------------
# Calcium: pools, Ca-dependent channels, Nernst / GHK reversal potentials

# opcodes (genesis/src/hines/hines_defs.h)
_COMPT, _FCOMPT, _LCOMPT = 100, 101, 102
_CONC_VAL, _NERNST, _GHK = 1000, 2000, 2010
_CHAN, _CHAN_EK, _ADD_CURR, _FLUX, _SCALE_FLUX = 3000, 3001, 3200, 3300, 3301
_IPOL1V, _IPOL1C, _NEWVOLT, _NEWCONC1, _DOMAIN = 4001, 4002, 5100, 5110, 5112
_CACONC, _LCACONC = 10200, 10202

CA_FEATURES = ("flux", "scale", "kc", "domain", "tabcurr", "ctabcurr", "nernst", "ghk", "kc2d", "c1c2",
               "tabghk", "sameghk")
_IPOL2V, _IPOL2C, _NEWCONC2, _SAME_GHK, _DO_TABGHK = 4003, 4004, 5111, 3002, 3003


def _segments(m: Model):
    """Per compartment (ops order): (head op, [(op, operands, chip values)],
    Em/Rm and inject, tbyc and diagterm) of a model whose programs hold only
    NEWVOLT / CHAN_EK / CHAN / IPOL1V / ADD_CURR ops."""
    ops, chip = m.ops, m.chip
    segs = []
    p = c = 0
    while True:
        head = int(ops[p]); p += 1
        if head == _LCOMPT:
            break
        pre = list(chip[c:c + 2]); c += 2
        body = []
        while int(ops[p]) not in (_COMPT, _FCOMPT, _LCOMPT):
            o = int(ops[p]); p += 1
            if o == _CHAN_EK:
                body.append((o, [], list(chip[c:c + 2]))); c += 2
            elif o == _CHAN:
                body.append((o, [], list(chip[c:c + 1]))); c += 1
            elif o == _IPOL1V:
                nc = 1 if ops[p + 1] != 0 else 0
                body.append((o, [int(ops[p]), int(ops[p + 1])], list(chip[c:c + nc])))
                c += nc
                p += 2
            elif o in (_NEWVOLT, _ADD_CURR):
                body.append((o, [], []))
            else:
                raise ValueError(f"_segments: op {o} not in the supported subset")
        segs.append((head, body, pre, list(chip[c:c + 2])))
        c += 2
    return segs


def ca_population(n_neurons: int, sizes=(1,), probs=(1.0,), *, features=CA_FEATURES, seed: int = 0,
                  dt: float = 20e-6, method: int = BEULER, ydivs: int = 1000, ymax: float = 0.05,
                  ca_base: float = 5e-5, tau: float = 10e-3, cout: float = 2.0,
                  inject=(0.2e-9, 1.0e-9), **kw) -> Model:
    """The population workload (random Hines trees, HH soma, Kdr dendrites)
    with calcium in every compartment, emitted in the order the reference's
    SETUP writes it (hines_chip_init.c:806-1235, hines_conc_init.c:779-795):

    * a Ca channel (HH m gate, power 2): Ek fixed ("flux"), from NERNST_OP
      ("nernst") or from the GHK slope conductance ("ghk"), its current fed
      into the compartment's Ca_concen pool with FLUX_OP (SCALE_FLUX_OP with
      "scale" or "ghk");
    * "kc": a Ca-dependent K channel, Z gate on the pool concentration
      (NEWCONC1_OP, IPOL1C_OP);
    * "domain": a second one whose Z gate reads the Ca current of the channel
      before it (DOMAIN_OP);
    * "tabcurr" / "ctabcurr": table currents on Vm / on the pool (power-0 IPOL);
    * "kc2d": a K channel whose gate is a 2-D table over Vm and the pool
      (IPOL2V); "c1c2": one over the pool and a second, constant
      concentration (NEWCONC2 + IPOL2C through the table register);
    * "tabghk": a channel whose Ek and GHK slope come from a voltage table
      (DO_TABGHK_OP, power-0 IPOL, SAME_GHK_OP); "sameghk" (with "ghk"): a
      second channel on the Ca channel's GHK slope (SAME_GHK_OP);
    * one Ca_concen pool per compartment (CACONC_OP: base, 1 + dt/(2 tau),
      dt B), a constant outside concentration after it in conc[] and a third,
      constant concentration (1 mM) for "c1c2".

    Concentrations in mM (mol/m^3), the conc table on [0, ymax] with ydivs
    intervals, sharing tablist with the voltage columns.
    """
    feats = set(features)
    bad = feats - set(CA_FEATURES)
    if bad:
        raise ValueError(f"unknown features {sorted(bad)}")
    if "nernst" in feats and "ghk" in feats:
        raise ValueError("nernst and ghk both set the Ca channel's Ek; choose one")
    if "sameghk" in feats and "ghk" not in feats:
        raise ValueError("sameghk reuses the Ca channel's GHK slope; add ghk")
    base = population(n_neurons, sizes, probs, seed=seed, dt=dt, method=method, inject=inject, **kw)
    rng = np.random.default_rng(seed + 7919)
    xdivs, ncols0 = base.xdivs, base.ncols
    rows = max(xdivs, ydivs) + 2
    tab0 = base.tablist.reshape(xdivs + 2, ncols0)
    w2 = 2 * (ydivs + 1)                          # columns of one 2-D gate
    n2d = ("kc2d" in feats) + ("c1c2" in feats)
    tab = np.zeros((rows, ncols0 + 6 + n2d * w2))
    tab[:xdivs + 2, :ncols0] = tab0
    tab[xdivs + 2:, :ncols0] = tab0[xdivs + 1]
    invdy = ydivs / ymax
    C_r = np.arange(ydivs + 1) / invdy
    V_r = base.xmin + np.arange(xdivs + 1) / base.invdx
    col_z, col_ct, col_t = ncols0, ncols0 + 2, ncols0 + 4
    # K_C Z gate: alpha = 500 C/(C + 1e-3), beta = 100 (1/s), as the
    # non-instant columns 1 + dt/2 (alpha + beta), dt alpha
    al = 500.0 * C_r / (C_r + 1e-3)
    be = 100.0
    zB, zA = 1.0 + dt / 2.0 * (al + be), dt * al
    def put(col, x, n):
        tab[:n + 1, col] = x
        tab[n + 1:, col] = x[-1]
    put(col_z, zB, ydivs)
    put(col_z + 1, zA, ydivs)
    # table currents: G_tab and E_tab
    put(col_ct, C_r / (C_r + 5e-4), ydivs)
    put(col_ct + 1, -0.08 + 10.0 * C_r, ydivs)
    put(col_t, 1.0 / (1.0 + np.exp(-(V_r + 0.04) / 0.005)), xdivs)
    put(col_t + 1, -0.03 + 0.1 * V_r, xdivs)
    # 2-D gates: B' at col + 2j, A' at col + 2j + 1 for concentration row j
    nxt = ncols0 + 6
    col_2v = col_cc = -1
    if "kc2d" in feats:
        col_2v, nxt = nxt, nxt + w2
        a2 = (300.0 * C_r / (C_r + 1e-3))[None, :] * np.exp(V_r / 0.03)[:, None]
        tab[:xdivs + 1, col_2v:col_2v + w2:2] = 1.0 + dt / 2.0 * (a2 + 80.0)
        tab[:xdivs + 1, col_2v + 1:col_2v + w2:2] = dt * a2
        tab[xdivs + 1:, col_2v:col_2v + w2] = tab[xdivs, col_2v:col_2v + w2]
    if "c1c2" in feats:
        col_cc, nxt = nxt, nxt + w2
        a2 = (400.0 * C_r / (C_r + 5e-4))[:, None] * (1.0 / (1.0 + C_r / 1e-3))[None, :]
        tab[:ydivs + 1, col_cc:col_cc + w2:2] = 1.0 + dt / 2.0 * (a2 + 150.0)
        tab[:ydivs + 1, col_cc + 1:col_cc + w2:2] = dt * a2
        tab[ydivs + 1:, col_cc:col_cc + w2] = tab[ydivs, col_cc:col_cc + w2]

    def ss2(col, r_a, f_a, r_b, f_b):
        """Steady state of a non-instant 2-D gate (rows r_a, fraction f_a;
        concentration row r_b, fraction f_b)."""
        t0, t1 = tab[r_a, col + 2 * r_b:], tab[r_a + 1, col + 2 * r_b:]
        B = (1 - f_a) * ((1 - f_b) * t0[0] + f_b * t0[2]) + f_a * ((1 - f_b) * t1[0] + f_b * t1[2])
        A = (1 - f_a) * ((1 - f_b) * t0[1] + f_b * t0[3]) + f_a * ((1 - f_b) * t1[1] + f_b * t1[3])
        return A / (2.0 * (B - 1.0))

    def ss(col, v):
        """Steady state A/B of a non-instant column at row coordinate v."""
        r = min(max(int(v), 0), tab.shape[0] - 2)
        f = v - r
        B = tab[r, col] + f * (tab[r + 1, col] - tab[r, col])
        A = tab[r, col + 1] + f * (tab[r + 1, col + 1] - tab[r, col + 1])
        return A / (2.0 * (B - 1.0))

    RM, Em = kw.get("RM", 0.33), kw.get("Em", -0.065)
    # the Na m gate column (the soma's power-3 gate) drives the Ca channel
    col_m = next(int(base.ops[i + 1]) for i in range(base.ops.size - 2)
                 if base.ops[i] == _IPOL1V and base.ops[i + 2] == 3)
    segs = _segments(base)
    ops, chip, conc, concops, concchip = [], [], [], [], []
    compops, compchips = [], []
    F, R, T = 96485.3329, 8.314462618, 308.15
    zFRT = 2.0 * F / (R * T)
    # The Ek register carries across compartments: CHAN_OP inherits the last
    # Ek, and the channels added here change it, so a base CHAN_OP whose
    # inherited value no longer holds becomes a CHAN_EK_OP, as SETUP writes it
    # (hines_chip_init.c:1029-1049).
    reg = base_reg = 0.0
    for r, (head, body, pre, post) in enumerate(segs):
        compops.append(len(ops)); compchips.append(len(chip))
        ops.append(head); chip += pre
        for (o, opnd, vals) in body:
            if o == _CHAN_EK:
                base_reg = reg = vals[0]
            elif o == _CHAN and reg != base_reg:
                o, vals = _CHAN_EK, [base_reg] + vals
                reg = base_reg
            ops += [o] + opnd
            chip += vals
        v0 = (base.vm[r] - base.xmin) * base.invdx
        k = len(conc)
        area = float(pre[0]) * RM / Em       # Em/Rm = Em area / RM
        vol = area * 0.2e-6
        B = 1.0 / (2.0 * F * vol)
        conc += [ca_base, cout] + ([1e-3] if "c1c2" in feats else [])
        g_ca = 5.0 * area * rng.uniform(0.5, 1.5)
        m0 = ss(col_m, v0)
        if "nernst" in feats:
            ops += [_CONC_VAL, k, _NERNST, _IPOL1V, col_m, 2, _ADD_CURR, _FLUX]
            chip += [1.0 / zFRT, g_ca, m0]
            reg = None
        elif "ghk" in feats:
            perm = g_ca / (2.0 * F * zFRT * cout)
            ops += [_CONC_VAL, k, _GHK, _IPOL1V, col_m, 2, _ADD_CURR, _SCALE_FLUX]
            chip += [zFRT, 2.0 * F, perm, m0, 1.0]
            reg = None
        else:
            ops += [_CHAN_EK, _IPOL1V, col_m, 2, _ADD_CURR, _CONC_VAL, k]
            chip += [0.12, g_ca, m0]
            reg = 0.12
            if "scale" in feats:
                ops += [_SCALE_FLUX]
                chip += [0.8]
            else:
                ops += [_FLUX]
        if "domain" in feats:
            ops += [_DOMAIN, _CHAN_EK, _IPOL1C, col_z, 1, _ADD_CURR]
            chip += [2.0e5, -0.09, 5.0 * area, ss(col_z, 0.0)]
            reg = -0.09
        if "kc" in feats:
            ops += [_NEWCONC1, _CHAN_EK, _IPOL1C, col_z, 1, _ADD_CURR]
            chip += [-0.09, 10.0 * area, ss(col_z, ca_base * invdy)]
            reg = -0.09
        cb = ca_base * invdy
        if "kc2d" in feats:
            ops += [_NEWCONC1, _CHAN_EK, _IPOL2V, col_2v, 1, _ADD_CURR]
            chip += [-0.09, 2.0 * area, ss2(col_2v, int(v0), v0 - int(v0), int(cb), cb - int(cb))]
            reg = -0.09
        if "c1c2" in feats:
            c2 = 1e-3 * invdy
            ops += [_NEWCONC1, _NEWCONC2, k + 2, _CHAN_EK, _IPOL2C, col_cc, 2, _ADD_CURR]
            chip += [-0.085, 1.5 * area, ss2(col_cc, int(cb), cb - int(cb), int(c2), c2 - int(c2))]
            reg = -0.085
        if "sameghk" in feats:
            ops += [_SAME_GHK, _IPOL1V, col_m, 1, _ADD_CURR]
            chip += [0.3 * perm, m0]
            reg = None
        if "tabghk" in feats:
            ops += [_DO_TABGHK, _IPOL1V, col_t, 0, _SAME_GHK, _ADD_CURR]
            chip += [1.5 * area]
            reg = None
        if "tabcurr" in feats:
            ops += [_CHAN, _IPOL1V, col_t, 0]
            chip += [2.0 * area]
            reg = None
        if "ctabcurr" in feats:
            ops += [_NEWCONC1, _CHAN, _IPOL1C, col_ct, 0]
            chip += [1.0 * area]
            reg = None
        chip += post
        concops += [_CACONC, k]
        concchip += [ca_base, 1.0 + dt / (2.0 * tau), dt * B]
    ops.append(_LCOMPT)
    concops.append(_LCACONC)
    import dataclasses
    return dataclasses.replace(
        base, ops=np.array(ops, np.int32), chip=np.array(chip, np.float64), ncols=tab.shape[1],
        tablist=np.ascontiguousarray(tab.reshape(-1)), compops=np.array(compops, np.int32),
        compchips=np.array(compchips, np.int32), conc=np.array(conc, np.float64),
        concops=np.array(concops, np.int32), concchip=np.array(concchip, np.float64),
        ydivs=ydivs, ymin=0.0, invdy=invdy, givals=np.zeros(0))


# ---------------------------------------------------------------------------
# Diffusion shells, buffers and pumps (the concentration solve)

_CONC, _FCONC, _LCONC, _FLUXCONC = 10100, 10101, 10102, 10103
_TAUPUMP, _CTAUPUMP, _MMPUMP, _HILLPUMP, _PHILLPUMP = 11000, 11001, 11010, 11011, 11012
_BUFFER, _POOL, _ELMMPUMP = 12000, 12100, 500
SHELL_FEATURES = ("fixbuf", "difbuf", "ctaupump", "taupump", "mmpump", "hillpump", "elmmpump", "pool",
                  "kc", "leak")
_FARADAY = 96485.3329


def shell_population(n_neurons: int, sizes=(1,), probs=(1.0,), *, shells: int = 4,
                     features=("fixbuf", "difbuf", "ctaupump", "mmpump", "hillpump", "kc", "leak"),
                     seed: int = 0, dt: float = 20e-6, method: int = CRANK, hill: int = 2,
                     ydivs: int = 1000, ymax: float = 0.5, ca0: float = 5e-5,
                     inject=(0.2e-9, 1.0e-9), **kw) -> Model:
    """The population workload with a diffusion system of ``shells``
    concentric Ca shells in every compartment, emitted as the reference's
    SETUP writes it: the system arrays of h_init_conc_solve /
    h_setup_conc_solve (hines_conc_init.c:59-411), the concentration program
    of h_init_conc_chip (:413-805) and the values of the chip_put_* functions
    (hines_data.c:189-400).

    Per compartment (Hines row r) the children are, in order: the system's
    rows shell by shell from the innermost (its buffers -- "fixbuf" a fixed
    buffer, "difbuf" a diffusible one -- then the shell), the pumps on the
    outermost shell ("ctaupump", "taupump" with a voltage-dependent rate,
    "mmpump", "hillpump" with Hill coefficient ``hill``), and with "pool" a
    concpool fed by a Hill pump from the outermost shell, holding a pool
    buffer.  The outermost shell takes the Ca channel's current (FLUX_OP,
    then FLUXCONC_OP with a "leak"); "kc" adds a K channel gated by that
    shell (NEWCONC1_OP, IPOL1C_OP), "elmmpump" an electrogenic pump on it
    (ELMMPUMP_OP in the channel program).  Shell m (0 innermost) of a
    cylinder of diameter ~U(1, 4) um and length ~U(10, 100) um is
    ``f * 2**(shells - 1 - m)`` of the radius thick, the shells outside the
    innermost together half of it (the outermost thinnest), the innermost
    the rest.  Concentrations in mM; buffers start at equilibrium.  Buffers
    need Crank-Nicolson (the reference refuses them under backward Euler,
    hines_conc_init.c:160-164).
    """
    feats = set(features)
    bad = feats - set(SHELL_FEATURES)
    if bad:
        raise ValueError(f"unknown features {sorted(bad)}")
    if shells < 2:
        raise ValueError("a diffusion system needs at least two shells")
    if ("fixbuf" in feats or "difbuf" in feats) and method != CRANK:
        raise ValueError("buffers need Crank-Nicolson (hines_conc_init.c:160-164)")
    base = population(n_neurons, sizes, probs, seed=seed, dt=dt, method=method, inject=inject, **kw)
    rng = np.random.default_rng(seed + 104729)
    crank = method == CRANK
    nd = 1 if "fixbuf" in feats else 0
    ndif = 1 if "difbuf" in feats else 0
    dw = 1 + nd + ndif
    # chip_put_flux / _taupump / _mmpump / _hillpump: dt/2 for a system of
    # shells alone under Crank-Nicolson (hines_data.c:224-227); the solve's
    # own dt2 is dt/2 under Crank-Nicolson (hines_conc_init.c:312-315)
    dtp = dt / 2.0 if (crank and dw == 1) else dt
    dt2 = dt / 2.0 if crank else dt
    xdivs, ncols0 = base.xdivs, base.ncols
    rows = max(xdivs, ydivs) + 2
    tab0 = base.tablist.reshape(xdivs + 2, ncols0)
    tab = np.zeros((rows, ncols0 + 2))
    tab[:xdivs + 2, :ncols0] = tab0
    tab[xdivs + 2:, :ncols0] = tab0[xdivs + 1]
    invdy = ydivs / ymax
    C_r = np.arange(ydivs + 1) / invdy
    col_z = ncols0
    al = 500.0 * C_r / (C_r + 1e-3)
    zB, zA = 1.0 + dt / 2.0 * (al + 100.0), dt * al
    tab[:ydivs + 1, col_z], tab[ydivs + 1:, col_z] = zB, zB[-1]
    tab[:ydivs + 1, col_z + 1], tab[ydivs + 1:, col_z + 1] = zA, zA[-1]
    col_m = next(int(base.ops[i + 1]) for i in range(base.ops.size - 2)
                 if base.ops[i] == _IPOL1V and base.ops[i + 2] == 3)
    n = base.ncompts
    kf_f, kb_f, bt_f = 1.0e5, 100.0, 0.05          # fixed buffer: /mM/s, /s, mM
    kf_d, kb_d, bt_d, D_b = 2.0e5, 300.0, 0.02, 2.0e-11   # diffusible buffer
    D_ca = 6.0e-10

    # --- per compartment geometry (difshell.c:205-229, shape_mode SHELL)
    dia = rng.uniform(1e-6, 4e-6, n)
    ln = rng.uniform(10e-6, 100e-6, n)
    g_scale = rng.uniform(0.5, 1.5, n)
    rad = dia / 2.0
    f = 0.5 / (2 ** (shells - 1) - 1)
    th = rad[:, None] * f * 2.0 ** (shells - 1 - np.arange(shells))[None, :]
    th[:, 0] = rad - th[:, 1:].sum(axis=1)
    r_out = np.empty((n, shells))
    acc = rad.copy()
    for m_ in range(shells - 1, -1, -1):
        r_out[:, m_] = acc
        acc = acc - th[:, m_]
    r_in = r_out - th
    r_in[:, 0] = 0.0
    vol = np.pi * ln[:, None] * (r_out ** 2 - r_in ** 2)
    s_up = 2.0 * np.pi * r_out * ln[:, None]
    s_dn = 2.0 * np.pi * r_in * ln[:, None]

    # --- h_setup_conc_solve (hines_conc_init.c:294-411): covals0 per system
    def coupling(D, m_):
        Diff = 2.0 * dt2 * D / vol[:, m_]
        if m_ < shells - 1:
            up = Diff * s_up[:, m_] / ((2 * th[:, m_] if m_ == 0 else th[:, m_]) + th[:, m_ + 1])
            if m_ == 0:
                return [up]
            den = (2 * th[:, m_ - 1] + th[:, m_]) if m_ == 1 else (th[:, m_ - 1] + th[:, m_])
            return [Diff * s_dn[:, m_] / den, up]
        return [Diff * s_dn[:, m_] / (th[:, m_ - 1] + th[:, m_])]
    cols = []
    for m_ in range(shells):
        if nd:
            cols.append(np.full(n, dt * kb_f * bt_f))
        if ndif:
            cols.extend(coupling(D_b, m_))
        cols.extend(coupling(D_ca, m_))
    covals0 = np.stack(cols, axis=1).reshape(-1)
    buf1 = ([dt2 * kf_f, dt2 * kb_f, dt * kb_f * bt_f] if nd else []) + \
           ([dt2 * kf_d, dt2 * kb_d, dt * kb_d * bt_d] if ndif else [])
    bufvals = np.tile(np.array(buf1, np.float64), n)

    # --- concentrations: per compartment its system (buffers, then the shell,
    # shell by shell), then with "pool" the concpool and its buffer
    bf_f = bt_f * kb_f / (kb_f + kf_f * ca0)
    bf_d = bt_d * kb_d / (kb_d + kf_d * ca0)
    row = ([bf_f] if nd else []) + ([bf_d] if ndif else []) + [ca0]
    kb_p, kf_p, bt_p, cp0 = 50.0, 5e4, 0.1, 1e-3
    per = row * shells + ([cp0, bt_p * kb_p / (kb_p + kf_p * cp0)] if "pool" in feats else [])
    npc = len(per)
    conc = np.tile(np.array(per, np.float64), n)
    k0 = np.arange(n) * npc
    kout = k0 + (shells - 1) * dw + dw - 1            # the outermost shell
    kp = k0 + shells * dw                             # the concpool

    # --- the concentration program (h_init_conc_chip, hines_conc_init.c:
    # 413-805): per compartment the outermost shell's flux and leak, its
    # pumps, then the Hill pump into the concpool, the pool buffer and the
    # pool; the first switch is FCONC_OP, later ones CONC_OP (after a pool a
    # shell's switch still stores, hines_conc_init.c:642-687)
    leak = 1e-6 if "leak" in feats else 0.0
    v_out = vol[:, -1]
    ops_cols, chip_cols = [], []
    def oc(v): ops_cols.append(np.broadcast_to(np.asarray(v, np.int64), (n,)))
    def cc(v): chip_cols.append(np.broadcast_to(np.asarray(v, np.float64), (n,)))
    sw = np.full(n, _CONC, np.int64)
    sw[0] = _FCONC
    oc(sw); oc(kout); oc(_FLUXCONC); cc(dtp * leak); cc(dtp / (2.0 * _FARADAY * v_out))
    if "ctaupump" in feats:
        T_C, Ceq = 2e-3, 4e-5
        oc(_CTAUPUMP); cc(Ceq * dtp / T_C); cc(dtp / T_C)
    if "taupump" in feats:
        oc(_TAUPUMP); oc(np.arange(n))
        for v in (3e-3, 1e-3, -0.04, 1.0 / 0.01, 5e-5):
            cc(v)
    if "mmpump" in feats:
        oc(_MMPUMP); cc(dtp * 2e-4 / (v_out * 1e12)); cc(2e-4)
    if "hillpump" in feats:
        oc(_HILLPUMP); oc(hill); cc(3e-4 ** hill); cc(dtp * 5e-4 / (v_out * 1e12))
    if "pool" in feats:
        pool_vol = 0.05 * vol.sum(axis=1)
        vmax = 3e-4 / 1e12
        oc(_CONC); oc(kp); oc(_PHILLPUMP); oc(hill); oc(kout)
        cc(2e-4 ** hill); cc(dtp * vmax / v_out); cc(dt * vmax / pool_vol)
        oc(_BUFFER); oc(kp + 1)
        cc(dt * kb_p / 2.0); cc(dt * kb_p * bt_p); cc(dt * kf_p)
        oc(_POOL)
    concops = np.concatenate([np.stack(ops_cols, axis=1).reshape(-1),
                              [_LCACONC if "pool" in feats else _LCONC]]).astype(np.int32)
    concchip = np.stack(chip_cols, axis=1).reshape(-1)

    # --- channel programs: the base compartment's ops, then the Ca channel
    # into the outermost shell, optional electrogenic pump and Ca-gated K
    # channel on it.  The Ek register carries across compartments, and the
    # channels added here change it, so a base CHAN_OP that inherited Ek
    # becomes a CHAN_EK_OP with the base's value, as SETUP writes it
    # (hines_chip_init.c:1029-1049).
    bops, bchip = base.ops, base.chip
    cops = np.append(base.compops.astype(np.int64), bops.size - 1)      # LCOMPT_OP ends the last
    cchp = np.append(base.compchips.astype(np.int64), bchip.size)
    sig_info = {}
    first_chan = np.full(n, -1, np.int64)      # ops index of a CHAN_OP to rewrite
    first_chan_c = np.full(n, -1, np.int64)    # chip index of its gmax
    last_ek_c = np.full(n, -1, np.int64)       # chip index of the compartment's last CHAN_EK Ek
    for r in range(n):
        a, b = cops[r], cops[r + 1]
        key = bops[a:b].tobytes()
        info = sig_info.get(key)
        if info is None:
            p, c = 1, 2
            fc = fcc = lek = -1
            seen = False
            seg = bops[a:b]
            while p < seg.size:
                o = int(seg[p]); p += 1
                if o == _CHAN_EK:
                    lek = c
                    seen = True
                    c += 2
                elif o == _CHAN:
                    if not seen:
                        fc, fcc = p - 1, c
                    seen = True
                    c += 1
                elif o == _IPOL1V:
                    c += 1 if int(seg[p + 1]) != 0 else 0
                    p += 2
                elif o not in (_NEWVOLT, _ADD_CURR):
                    raise ValueError(f"shell_population: base op {o} not in the supported subset")
            info = (fc, fcc, lek)
            sig_info[key] = info
        fc, fcc, lek = info
        if fc >= 0:
            first_chan[r], first_chan_c[r] = a + fc, cchp[r] + fcc
        if lek >= 0:
            last_ek_c[r] = cchp[r] + lek
    # the register each compartment starts with in the base model
    idx = np.where(last_ek_c >= 0, np.arange(n), -1)
    idx = np.maximum.accumulate(idx)
    reg_after = np.where(idx >= 0, bchip[np.where(idx >= 0, last_ek_c[np.maximum(idx, 0)], 0)], 0.0)
    reg_before = np.concatenate([[0.0], reg_after[:-1]])
    rw = first_chan >= 0
    area = bchip[cchp[:-1]] * kw.get("RM", 0.33) / kw.get("Em", -0.065)
    v0 = (base.vm - base.xmin) * base.invdx
    rr = np.clip(v0.astype(np.int64), 0, xdivs)
    fr = v0 - rr
    fB = tab[rr, col_m] + fr * (tab[rr + 1, col_m] - tab[rr, col_m])
    fA = tab[rr, col_m + 1] + fr * (tab[rr + 1, col_m + 1] - tab[rr, col_m + 1])
    m0 = fA / (2.0 * (fB - 1.0))
    g_ca = 2.0 * area * g_scale
    eo, ec = [], []
    def eop(v): eo.append(np.broadcast_to(np.asarray(v, np.int64), (n,)))
    def ech(v): ec.append(np.broadcast_to(np.asarray(v, np.float64), (n,)))
    for v in (_CHAN_EK, _IPOL1V, col_m, 2, _ADD_CURR, _CONC_VAL):
        eop(v)
    eop(kout); eop(_FLUX)
    ech(0.12); ech(g_ca); ech(m0)
    if "elmmpump" in feats:
        vm_ = 2e-16 * s_up[:, -1] / 1e-10
        eop(_ELMMPUMP)
        ech(1e-3); ech(vm_ * 2.0 * _FARADAY); ech(dtp * vm_ / v_out)
    if "kc" in feats:
        for v in (_NEWCONC1, _CHAN_EK, _IPOL1C, col_z, 1, _ADD_CURR):
            eop(v)
        c = ca0 * invdy
        z0 = (tab[0, col_z + 1] + c * (tab[1, col_z + 1] - tab[0, col_z + 1])) / (
            2.0 * (tab[0, col_z] + c * (tab[1, col_z] - tab[0, col_z]) - 1.0))
        ech(-0.09); ech(10.0 * area); ech(z0)
    E_o, E_c = len(eo), len(ec)
    ops2 = bops.astype(np.int64).copy()
    ops2[first_chan[rw]] = _CHAN_EK
    ops2 = np.insert(ops2, np.repeat(cops[1:], E_o), np.stack(eo, axis=1).reshape(-1)).astype(np.int32)
    pos = np.concatenate([first_chan_c[rw], np.repeat(cchp[1:] - 2, E_c)])
    vals = np.concatenate([reg_before[rw], np.stack(ec, axis=1).reshape(-1)])
    order = np.argsort(pos, kind="stable")
    chip2 = np.insert(bchip, pos[order], vals[order])
    nrw = np.concatenate([[0], np.cumsum(rw)[:-1]])
    compops = (cops[:-1] + np.arange(n) * E_o).astype(np.int32)
    compchips = (cchp[:-1] + np.arange(n) * E_c + nrw).astype(np.int32)

    import dataclasses
    extra = dict(base.extra)
    extra.update(shell_vol=vol, shell_params=dict(
        fixbuf=(kf_f, kb_f, bt_f), difbuf=(kf_d, kb_d, bt_d, D_b), D_ca=D_ca, ca0=ca0, dt2=dt2, dtp=dtp))
    return dataclasses.replace(
        base, ops=ops2, chip=chip2, ncols=tab.shape[1],
        tablist=np.ascontiguousarray(tab.reshape(-1)), compops=compops, compchips=compchips,
        conc=conc, concops=concops, concchip=concchip,
        ydivs=ydivs, ymin=0.0, invdy=invdy, givals=np.zeros(0), extra=extra,
        nshells=np.full(n, shells, np.int32), nshrows=np.full(n, dw, np.int32),
        nshfix=np.full(n, nd, np.int32), dconc=k0.astype(np.int32),
        fbufval=(np.arange(n) * len(buf1)).astype(np.int32), fdiffs=np.arange(n + 1, dtype=np.int32),
        covals0=covals0, bufvals=bufvals)
