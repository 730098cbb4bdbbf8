"""ctypes binding of libhsolve_oracle.so -- TEST INFRASTRUCTURE (see __init__)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class Arrays(C.Structure):
    _fields_ = [
        ("ncompts", C.c_int32), ("nchips", C.c_int32), ("nops", C.c_int32),
        ("nfuncs", C.c_int32), ("nravals", C.c_int32),
        ("ncols", C.c_int32), ("xdivs", C.c_int32), ("ydivs", C.c_int32),
        ("nconcs", C.c_int32), ("nstab", C.c_int32), ("ngivals", C.c_int32),
        ("nelm_names", C.c_int32),
        ("method", C.c_int32), ("chanmode", C.c_int32), ("storemode", C.c_int32),
        ("pad", C.c_int32),
        ("xmin", C.c_double), ("invdx", C.c_double), ("ymin", C.c_double),
        ("invdy", C.c_double), ("dt", C.c_double),
        ("ops", _i32p), ("chip", _f64p), ("tablist", _f64p), ("xvals", _f64p),
        ("yvals", _f64p), ("stablist", _f64p), ("funcs", _i32p), ("ravals", _f64p),
        ("vm", _f64p), ("givals", _f64p), ("conc", _f64p),
        ("concops", _i32p), ("concchip", _f64p), ("nconcops", C.c_int32), ("nconcchip", C.c_int32),
        ("ndiffs", C.c_int32), ("ncovals0", C.c_int32), ("nbufvals", C.c_int32), ("pad2", C.c_int32),
        ("nshells", _i32p), ("nshrows", _i32p), ("nshfix", _i32p), ("dconc", _i32p), ("fbufval", _i32p),
        ("covals0", _f64p), ("bufvals", _f64p),
    ]


SYNFN = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32)
RANDFN = C.CFUNCTYPE(C.c_double, C.c_void_p)

_lib = None

INPUT_DTYPE = np.dtype([("op", np.int32), ("index", np.int32), ("stab", np.int32), ("per_dt", np.int32),
                        ("value", np.float64)])


def input_array(inputs):
    """h_in_msgs inputs (op, index, stab, per_dt, value) as the C struct
    array ohs_in_msgs and hsh_set_inputs take."""
    a = np.zeros(len(inputs), INPUT_DTYPE)
    for i, (op, index, stab, per_dt, value) in enumerate(inputs):
        a[i] = (op, index, stab, per_dt, value)
    return a


def oracle_lib():
    global _lib
    if _lib is None:
        p = os.path.join(_HERE, "libhsolve_oracle.so")
        if not os.path.exists(p):
            raise OSError(f"{p} missing: build with `make -C oracle`")
        L = C.CDLL(p)
        L.ohs_create.argtypes = [C.POINTER(Arrays)]
        L.ohs_create.restype = C.c_void_p
        L.ohs_destroy.argtypes = [C.c_void_p]
        L.ohs_set_hooks.argtypes = [C.c_void_p, SYNFN, RANDFN, C.c_void_p]
        L.ohs_chip_update.argtypes = [C.c_void_p]
        L.ohs_chip_update.restype = C.c_int32
        L.ohs_euler_solve.argtypes = [C.c_void_p]
        L.ohs_crank_solve.argtypes = [C.c_void_p]
        L.ohs_step.argtypes = [C.c_void_p, C.c_int32]
        L.ohs_step.restype = C.c_int32
        L.ohs_conc_chip_update.argtypes = [C.c_void_p]
        L.ohs_conc_chip_update.restype = C.c_int32
        L.ohs_conc_solve.argtypes = [C.c_void_p]
        L.ohs_conc_solve.restype = C.c_int32
        L.ohs_get_conc.argtypes = [C.c_void_p, _f64p]
        L.ohs_conc_regs.argtypes = [C.c_void_p, _f64p, _f64p]
        L.ohs_put_conc.argtypes = [C.c_void_p, _f64p]
        L.ohs_get.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _i32p, _f64p]
        L.ohs_put.argtypes = [C.c_void_p, _f64p, _f64p, _i32p]
        L.ohs_nspikes.argtypes = [C.c_void_p]
        L.ohs_nspikes.restype = C.c_int32
        L.ohs_spikes.argtypes = [C.c_void_p, _i32p, _i32p]
        L.ohs_clear_spikes.argtypes = [C.c_void_p]
        L.ohs_stepno.argtypes = [C.c_void_p]
        L.ohs_stepno.restype = C.c_int32
        L.ohs_trace_enable.argtypes = [C.c_void_p, C.c_int32]
        L.ohs_trace.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.ohs_rows_enable.argtypes = [C.c_void_p, C.c_int32]
        L.ohs_rows.argtypes = [C.c_void_p, _i32p]
        L.ohs_step_parallel.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32]
        L.ohs_step_parallel.restype = C.c_int32
        L.ohs_set_synapses.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _f64p, _f64p]
        L.ohs_set_synapses.restype = C.c_int32
        L.ohs_post_event.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32]
        L.ohs_post_event.restype = C.c_int32
        L.ohs_in_msgs.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.ohs_in_msgs.restype = C.c_int32
        L.ohs_set_time.argtypes = [C.c_void_p, C.c_double]
        L.ohs_time.argtypes = [C.c_void_p]
        L.ohs_time.restype = C.c_double
        L.ohs_pending.argtypes = [C.c_void_p, C.c_int32, _i32p, _f64p, C.c_int32, _f64p]
        L.ohs_pending.restype = C.c_int32
        L.ohs_tl_create.argtypes = [C.POINTER(Arrays)]
        L.ohs_tl_create.restype = C.c_void_p
        L.ohs_tl_destroy.argtypes = [C.c_void_p]
        L.ohs_tl_set_tangent.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.ohs_tl_step.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int32, _f64p, _f64p]
        L.ohs_tl_step.restype = C.c_int32
        L.ohs_tl_get.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class Oracle:
    """CPU reference solver over one set of Hsolve-format arrays."""

    def __init__(self, m, *, givals: bool = True, conc=None, yvals=None,
                 ydivs: int = 0, ymin: float = 0.0, invdy: float = 0.0, storemode: int = 0,
                 nelm_names: int = 0):
        L = oracle_lib()
        self.m = m
        # keep every array alive for the lifetime of the handle
        self._keep = dict(
            ops=np.ascontiguousarray(m.ops, np.int32), chip=np.ascontiguousarray(m.chip, np.float64),
            tablist=np.ascontiguousarray(m.tablist, np.float64),
            xvals=np.ascontiguousarray(m.xvals, np.float64),
            stablist=np.ascontiguousarray(m.stablist, np.float64),
            funcs=np.ascontiguousarray(m.funcs, np.int32),
            ravals=np.ascontiguousarray(m.ravals, np.float64),
            vm=np.ascontiguousarray(m.vm, np.float64),
            givals=np.ascontiguousarray(m.givals, np.float64),
        )
        k = self._keep
        a = Arrays()
        a.ncompts, a.nchips, a.nops = m.ncompts, k["chip"].size, k["ops"].size
        a.nfuncs, a.nravals = k["funcs"].size, k["ravals"].size
        # concentration arrays: the model's unless given here
        if conc is None and getattr(m, "conc", None) is not None and np.size(m.conc):
            conc = m.conc
        if not ydivs and getattr(m, "ydivs", 0):
            ydivs, ymin, invdy = m.ydivs, m.ymin, m.invdy
        a.ncols, a.xdivs, a.ydivs = m.ncols, m.xdivs, ydivs
        a.nstab, a.ngivals, a.nelm_names = k["stablist"].size // 4, k["givals"].size, nelm_names
        a.method, a.chanmode, a.storemode = m.method, m.chanmode, storemode
        a.xmin, a.invdx, a.ymin, a.invdy, a.dt = m.xmin, m.invdx, ymin, invdy, m.dt
        a.ops, a.chip = _p(k["ops"], C.c_int32), _p(k["chip"], C.c_double)
        a.tablist = _p(k["tablist"], C.c_double) if k["tablist"].size else None
        a.xvals = _p(k["xvals"], C.c_double) if k["xvals"].size else None
        if yvals is None and ydivs > 0:
            # accumulated as SETUP fills them (hines_chip_init.c:1291-1297),
            # plus one guard entry for the row-cache bound at the top row
            yvals = np.empty(ydivs + 2)
            x, dx = ymin, 1.0 / invdy
            for j in range(ydivs + 2):
                yvals[j] = x
                x += dx
        if yvals is not None:
            k["yvals"] = np.ascontiguousarray(yvals, np.float64)
            a.yvals = _p(k["yvals"], C.c_double)
        a.stablist = _p(k["stablist"], C.c_double) if k["stablist"].size else None
        a.funcs, a.ravals = _p(k["funcs"], C.c_int32), _p(k["ravals"], C.c_double)
        a.vm = _p(k["vm"], C.c_double)
        a.givals = _p(k["givals"], C.c_double) if givals else None
        if conc is not None:
            k["conc"] = np.ascontiguousarray(conc, np.float64)
            a.nconcs = k["conc"].size
            a.conc = _p(k["conc"], C.c_double)
        if getattr(m, "concops", None) is not None and np.size(m.concops):
            k["concops"] = np.ascontiguousarray(m.concops, np.int32)
            k["concchip"] = np.ascontiguousarray(m.concchip, np.float64)
            a.concops, a.nconcops = _p(k["concops"], C.c_int32), k["concops"].size
            a.concchip, a.nconcchip = _p(k["concchip"], C.c_double), k["concchip"].size
        if getattr(m, "nshells", None) is not None and np.size(m.nshells):
            for f in ("nshells", "nshrows", "nshfix", "dconc", "fbufval"):
                k[f] = np.ascontiguousarray(getattr(m, f), np.int32)
                setattr(a, f, _p(k[f], C.c_int32))
            for f in ("covals0", "bufvals"):
                k[f] = np.ascontiguousarray(getattr(m, f), np.float64)
                setattr(a, f, _p(k[f], C.c_double) if k[f].size else None)
            a.ndiffs, a.ncovals0, a.nbufvals = k["nshells"].size, k["covals0"].size, k["bufvals"].size
        self._a = a
        self.h = L.ohs_create(C.byref(a))
        self._hooks = None

    def close(self):
        if getattr(self, "h", None):
            oracle_lib().ohs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_hooks(self, syn=None, rnd=None):
        f1 = SYNFN(syn) if syn else SYNFN(0)
        f2 = RANDFN(rnd) if rnd else RANDFN(0)
        self._hooks = (f1, f2)
        oracle_lib().ohs_set_hooks(self.h, f1, f2, None)

    def step(self, n: int = 1):
        if oracle_lib().ohs_step(self.h, n) != 0:
            raise RuntimeError("oracle: undefined chip operation or a funcs[] program outside the rows")

    def chip_update(self):
        if oracle_lib().ohs_chip_update(self.h) != 0:
            raise RuntimeError("oracle: undefined chip operation")

    def euler_solve(self):
        oracle_lib().ohs_euler_solve(self.h)

    def crank_solve(self):
        oracle_lib().ohs_crank_solve(self.h)

    def conc_chip_update(self):
        if oracle_lib().ohs_conc_chip_update(self.h) != 0:
            raise RuntimeError("oracle: undefined concentration operation")

    def conc_solve(self):
        if oracle_lib().ohs_conc_solve(self.h) != 0:
            raise RuntimeError("oracle: a diffusion system of fewer than two shells")

    @property
    def vm(self):
        v = np.empty(self.m.ncompts)
        oracle_lib().ohs_get(self.h, _p(v, C.c_double), None, None, None, None)
        return v

    def state(self):
        m = self.m
        vm = np.empty(m.ncompts)
        chip = np.empty(self._keep["chip"].size)
        res = np.empty(2 * m.ncompts)
        ops = np.empty(self._keep["ops"].size, np.int32)
        giv = np.empty(self._keep["givals"].size)
        oracle_lib().ohs_get(self.h, _p(vm, C.c_double), _p(chip, C.c_double), _p(res, C.c_double),
                             _p(ops, C.c_int32), _p(giv, C.c_double))
        return dict(vm=vm, chip=chip, results=res, ops=ops, givals=giv)

    def conc(self) -> np.ndarray:
        out = np.empty(self._a.nconcs)
        if out.size:
            oracle_lib().ohs_get_conc(self.h, _p(out, C.c_double))
        return out

    def conc_regs(self):
        """(codiag, coresult) as the concops pass leaves them."""
        d, r = np.empty(self._a.nconcs), np.empty(self._a.nconcs)
        if d.size:
            oracle_lib().ohs_conc_regs(self.h, _p(d, C.c_double), _p(r, C.c_double))
        return d, r

    def put_conc(self, conc):
        a = np.ascontiguousarray(conc, np.float64)
        assert a.size == self._a.nconcs
        if a.size:
            oracle_lib().ohs_put_conc(self.h, _p(a, C.c_double))

    def put(self, vm=None, chip=None, ops=None):
        keep = []
        def pp(a, dt, ct):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _p(a, ct)
        oracle_lib().ohs_put(self.h, pp(vm, np.float64, C.c_double), pp(chip, np.float64, C.c_double),
                             pp(ops, np.int32, C.c_int32))

    def set_synapses(self, src_op, dst_op, weight, delay):
        """Connection table (see ohs_set_synapses); returns the shortest
        connection in steps."""
        a = [np.ascontiguousarray(x, t) for x, t in ((src_op, np.int32), (dst_op, np.int32),
                                                       (weight, np.float64), (delay, np.float64))]
        self._syn = a
        r = oracle_lib().ohs_set_synapses(self.h, a[0].size, _p(a[0], C.c_int32), _p(a[1], C.c_int32),
                                          _p(a[2], C.c_double), _p(a[3], C.c_double))
        if r < 0:
            raise ValueError("oracle: bad synaptic connection")
        return r

    def post_event(self, syn_op, time, weight, after=False):
        """An event from outside the solver (see ohs_post_event)."""
        if oracle_lib().ohs_post_event(self.h, int(syn_op), int(time), float(weight), 1 if after else 0) != 0:
            raise ValueError("oracle: bad posted event")

    def in_msgs(self, inputs):
        """Apply h_in_msgs inputs: a sequence of (op, index, stab, per_dt,
        value) (see ohs_in_msgs)."""
        a = input_array(inputs)
        if oracle_lib().ohs_in_msgs(self.h, a.size, a.ctypes.data) != 0:
            raise ValueError("oracle: bad input")

    def set_time(self, t):
        oracle_lib().ohs_set_time(self.h, float(t))

    @property
    def time(self):
        return oracle_lib().ohs_time(self.h)

    def pending(self, syn_op, cap=4096):
        """(time fields, weight/dt, time_last_event) of a synchan's pending
        list, or None when it has none."""
        t = np.empty(cap, np.int32)
        w = np.empty(cap, np.float64)
        tle = C.c_double(0.0)
        n = oracle_lib().ohs_pending(self.h, int(syn_op), _p(t, C.c_int32), _p(w, C.c_double), cap,
                                     C.byref(tle))
        if n < 0:
            return None
        n = min(n, cap)
        return t[:n].copy(), w[:n].copy(), tle.value

    def spikes(self):
        L = oracle_lib()
        n = L.ohs_nspikes(self.h)
        a = np.empty(n, np.int32)
        b = np.empty(n, np.int32)
        if n:
            L.ohs_spikes(self.h, _p(a, C.c_int32), _p(b, C.c_int32))
        return a, b

    def trace(self):
        t = np.empty(self.m.ncompts, np.uint64)
        oracle_lib().ohs_trace(self.h, t.ctypes.data_as(C.POINTER(C.c_uint64)))
        return t

    def row_trace(self, nsteps, rows=None):
        """Step nsteps times and return the NEWVOLT table row of each
        compartment (or of ``rows``) at every step, shape (nsteps, n)."""
        L = oracle_lib()
        L.ohs_rows_enable(self.h, 1)
        buf = np.empty(self.m.ncompts, np.int32)
        sel = np.arange(self.m.ncompts) if rows is None else np.asarray(rows)
        out = np.empty((nsteps, sel.size), np.int32)
        for t in range(nsteps):
            self.step(1)
            L.ohs_rows(self.h, _p(buf, C.c_int32))
            out[t] = buf[sel]
        L.ohs_rows_enable(self.h, 0)
        return out

    def trace_enable(self, on=True):
        oracle_lib().ohs_trace_enable(self.h, 1 if on else 0)


class TangentLinear:
    """Tangent-linear model of the step (hsolve_tl.c): the oracle's step with
    the derivative of every value along one direction of (vm0, chip, ravals)."""

    def __init__(self, m, dvm=None, dchip=None, dravals=None):
        o = Oracle(m, givals=False)          # reuse the array marshalling
        self._o = o
        self.m = m
        self.h = oracle_lib().ohs_tl_create(C.byref(o._a))
        if not self.h:
            raise ValueError("tangent-linear model: ops outside the tabchannel subset")
        keep = [np.ascontiguousarray(x, np.float64) if x is not None else None
                for x in (dvm, dchip, dravals)]
        self._t = keep
        oracle_lib().ohs_tl_set_tangent(self.h, *[_p(x, C.c_double) for x in keep])

    def run(self, nsteps: int, obs):
        """(trace, dtrace): vm[obs] and its tangent after each of nsteps steps."""
        obs = np.ascontiguousarray(obs, np.int32)
        tr = np.empty((nsteps, obs.size))
        dtr = np.empty((nsteps, obs.size))
        oracle_lib().ohs_tl_step(self.h, nsteps, _p(obs, C.c_int32), obs.size, _p(tr, C.c_double),
                                 _p(dtr, C.c_double))
        return tr, dtr

    def close(self):
        if getattr(self, "h", None):
            oracle_lib().ohs_tl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_parallel(oracles, nsteps: int, nthreads: int):
    L = oracle_lib()
    arr = (C.c_void_p * len(oracles))(*[o.h for o in oracles])
    if L.ohs_step_parallel(arr, len(oracles), nsteps, nthreads) != 0:
        raise RuntimeError("oracle: undefined chip operation")
