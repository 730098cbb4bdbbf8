"""ctypes binding of libhsolve_hip (include/hsolve_hip.h).

``HipSolver`` mirrors the reference accelerator's per-hsolve life cycle
(genesis/src/hines/cuda/cuda_hsolve.c:289-517): create from the hsolve's
arrays (cuda_init / cuda_backend_init), step (cuda_chip_update with its
"vm ready" return protocol), read state back (cuda_sync_chip), destroy
(cuda_backend_cleanup).  There is no CPU fallback inside: if the library is
missing or the device refuses the model, the caller sees an exception or the
refusal list, exactly as GENESIS glue would see HSH_REFUSED and keep its own
CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._paths import lib_path

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

HSH_OK, HSH_REFUSED = 0, 1
# hsh_input (message inputs of hsh_set_inputs) and its op codes
IN_INJECT, IN_BASE, IN_THRESH, IN_FLUX, IN_ACTIVATION = 1, 2, 3, 3300, 4100
INPUT_DTYPE = np.dtype([("op", np.int32), ("index", np.int32), ("stab", np.int32), ("per_dt", np.int32),
                        ("value", np.float64)])


def input_array(inputs):
    """(op, index, stab, per_dt, value) tuples as an hsh_input array."""
    a = np.zeros(len(inputs), INPUT_DTYPE)
    for i, (op, index, stab, per_dt, value) in enumerate(inputs):
        a[i] = (op, index, stab, per_dt, value)
    return a

PLAIN, BATCHED, CHANNELS = 0, 1, 2
T_TOTAL, T_CHANNEL, T_ELIM = 0, 1, 2
F_NO_AUTOTUNE = 1
F_HOST_ONLY = 2
F_NO_JIT = 4

REASONS = {1: "opcode", 2: "Ek live-in", 3: "table live-in", 4: "register live-in",
           5: "funcs (elimination program the device cannot run)", 6: "tree layout", 7: "calcmode", 8: "pending synaptic event",
           9: "shared pool", 10: "diffusion system or concentration op the device does not run"}


class HshModel(C.Structure):
    _fields_ = [
        ("ncompts", C.c_int32), ("nchips", C.c_int32), ("nops", C.c_int32),
        ("nfuncs", C.c_int32), ("nravals", C.c_int32), ("ncols", C.c_int32),
        ("xdivs", C.c_int32), ("nstab", C.c_int32), ("method", C.c_int32),
        ("chanmode", C.c_int32), ("calcmode", C.c_int32), ("nconcs", C.c_int32),
        ("dt", C.c_double), ("xmin", C.c_double), ("invdx", C.c_double),
        ("ops", _i32p), ("chip", _f64p), ("tablist", _f64p), ("xvals", _f64p),
        ("stablist", _f64p), ("funcs", _i32p), ("ravals", _f64p), ("vm", _f64p),
        ("ydivs", C.c_int32), ("nconcops", C.c_int32), ("nconcchip", C.c_int32),
        ("ymin", C.c_double), ("invdy", C.c_double),
        ("conc", _f64p), ("concops", _i32p), ("concchip", _f64p),
        ("ngivals", C.c_int32), ("givals", _f64p),
        ("ndiffs", C.c_int32), ("ncovals0", C.c_int32), ("nbufvals", C.c_int32), ("pad_diff", C.c_int32),
        ("nshells", _i32p), ("nshrows", _i32p), ("nshfix", _i32p), ("dconc", _i32p), ("fbufval", _i32p),
        ("fdiffs", _i32p), ("covals0", _f64p), ("bufvals", _f64p),
    ]


class HshOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("batch_k", C.c_int32), ("batch_waves", C.c_int32),
                ("flags", C.c_int32), ("tile_max_rows", C.c_int32), ("pack_cap", C.c_int32),
                ("pack_waves", C.c_int32), ("pack_stage", C.c_int32)]


class HshMember(C.Structure):
    _fields_ = [(f, C.c_int32) for f in ("row0", "nrows", "chip0", "nchips", "op0", "nops", "conc0", "nconcs",
                                         "giv0", "ngivals", "stab0", "pad")]


class HshRefusal(C.Structure):
    _fields_ = [("comp", C.c_int32), ("opcode", C.c_int32), ("op_index", C.c_int32),
                ("reason", C.c_int32)]


class HshStepOut(C.Structure):
    _fields_ = [("vm_ready", C.c_int32), ("nspikes", C.c_int32), ("launches", C.c_int32),
                ("pad", C.c_int32)]


class HshInfo(C.Structure):
    _fields_ = [("ntiles", C.c_int32), ("nblocks", C.c_int32), ("ntemplates", C.c_int32),
                ("ntrees", C.c_int32), ("max_rows", C.c_int32), ("batch_k", C.c_int32),
                ("batch_waves", C.c_int32), ("refused", C.c_int32),
                ("bytes_per_step", C.c_double), ("lane_fill", C.c_double),
                ("jit", C.c_int32), ("npacks", C.c_int32), ("jit_compile_s", C.c_double),
                ("tile_max_rows", C.c_int32), ("pack_cap", C.c_int32), ("pack_waves", C.c_int32),
                ("pack_stage", C.c_int32), ("max_tree_rows", C.c_int32), ("autotuned", C.c_int32),
                ("autotune_s", C.c_double)]


# every symbol include/hsolve_hip.h declares
EXPORTS = ("hsh_create", "hsh_refusals", "hsh_step", "hsh_sync", "hsh_get_state",
           "hsh_put_state", "hsh_get_conc", "hsh_put_conc", "hsh_create_group", "hsh_group_member", "hsh_get_givals", "hsh_set_chip", "hsh_drain_spikes", "hsh_info_get", "hsh_autotune",
           "hsh_last_time", "hsh_set_timing", "hsh_jit_source", "hsh_set_synapses", "hsh_post_events",
           "hsh_set_inputs", "hsh_set_inputs_static",
           "hsh_set_time", "hsh_set_ravals",
           "hsh_adj_forward",
           "hsh_adj_backward", "hsh_destroy", "hsh_last_error")

_lib = None


def lib():
    """Load libhsolve_hip.so (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        L = C.CDLL(lib_path("libhsolve_hip.so"))
        L.hsh_create.argtypes = [C.POINTER(HshModel), C.POINTER(HshOptions), C.POINTER(C.c_void_p)]
        L.hsh_create.restype = C.c_int32
        L.hsh_create_group.argtypes = [C.POINTER(C.POINTER(HshModel)), C.c_int32, C.POINTER(HshOptions),
                                       C.POINTER(C.c_void_p)]
        L.hsh_create_group.restype = C.c_int32
        L.hsh_group_member.argtypes = [C.c_void_p, C.c_int32, C.POINTER(HshMember)]
        L.hsh_group_member.restype = C.c_int32
        L.hsh_refusals.argtypes = [C.c_void_p, C.POINTER(HshRefusal), C.c_int32, _i32p]
        L.hsh_refusals.restype = C.c_int32
        L.hsh_step.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(HshStepOut)]
        L.hsh_step.restype = C.c_int32
        L.hsh_sync.argtypes = [C.c_void_p]
        L.hsh_sync.restype = C.c_int32
        L.hsh_get_state.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _i32p]
        L.hsh_get_state.restype = C.c_int32
        L.hsh_put_state.argtypes = [C.c_void_p, _f64p, _f64p, _i32p]
        L.hsh_put_state.restype = C.c_int32
        L.hsh_set_chip.argtypes = [C.c_void_p, C.c_int32, _i32p, _f64p]
        L.hsh_set_chip.restype = C.c_int32
        L.hsh_drain_spikes.argtypes = [C.c_void_p, _i32p, _i32p, C.c_int32, _i32p]
        L.hsh_drain_spikes.restype = C.c_int32
        L.hsh_info_get.argtypes = [C.c_void_p, C.POINTER(HshInfo)]
        L.hsh_info_get.restype = C.c_int32
        L.hsh_autotune.argtypes = [C.c_void_p]
        L.hsh_autotune.restype = C.c_int32
        L.hsh_set_synapses.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _f64p, _f64p]
        L.hsh_set_synapses.restype = C.c_int32
        L.hsh_post_events.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _f64p, C.c_int32]
        L.hsh_post_events.restype = C.c_int32
        L.hsh_set_inputs.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.hsh_set_inputs.restype = C.c_int32
        L.hsh_set_inputs_static.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.hsh_set_inputs_static.restype = C.c_int32
        L.hsh_set_time.argtypes = [C.c_void_p, C.c_double]
        L.hsh_set_time.restype = C.c_int32
        L.hsh_set_ravals.argtypes = [C.c_void_p, _f64p]
        L.hsh_set_ravals.restype = C.c_int32
        L.hsh_adj_forward.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _i32p, _f64p]
        L.hsh_adj_forward.restype = C.c_int32
        L.hsh_adj_backward.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        L.hsh_adj_backward.restype = C.c_int32
        L.hsh_last_time.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        L.hsh_last_time.restype = C.c_int32
        L.hsh_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.hsh_set_timing.restype = C.c_int32
        L.hsh_jit_source.argtypes = [C.POINTER(HshModel), C.c_char_p, C.c_int64, C.POINTER(C.c_int64)]
        L.hsh_jit_source.restype = C.c_int32
        L.hsh_destroy.argtypes = [C.c_void_p]
        L.hsh_destroy.restype = None
        L.hsh_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class HshError(RuntimeError):
    pass


class Refused(Exception):
    """The device path refuses this model; ``refusals`` lists why."""

    def __init__(self, msg, refusals):
        super().__init__(msg)
        self.refusals = refusals


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc, what):
    if rc < 0:
        raise HshError(f"{what}: {lib().hsh_last_error().decode()} (status {rc})")
    return rc


def _hsh_model(model, givals: bool = False):
    """Build the hsh_model struct (and the arrays it points to) for a Model."""
    k = dict(ops=np.ascontiguousarray(model.ops, np.int32),
             chip=np.ascontiguousarray(model.chip, np.float64),
             tablist=np.ascontiguousarray(model.tablist, np.float64),
             xvals=np.ascontiguousarray(model.xvals, np.float64),
             stablist=np.ascontiguousarray(model.stablist, np.float64),
             funcs=np.ascontiguousarray(model.funcs, np.int32),
             ravals=np.ascontiguousarray(model.ravals, np.float64),
             vm=np.ascontiguousarray(model.vm, np.float64),
             conc=np.ascontiguousarray(model.conc, np.float64),
             concops=np.ascontiguousarray(model.concops, np.int32),
             concchip=np.ascontiguousarray(model.concchip, np.float64))
    mm = HshModel()
    mm.ncompts, mm.nchips, mm.nops = model.ncompts, k["chip"].size, k["ops"].size
    mm.nfuncs, mm.nravals, mm.ncols = k["funcs"].size, k["ravals"].size, model.ncols
    mm.xdivs, mm.nstab = model.xdivs, k["stablist"].size // 4
    mm.method, mm.chanmode, mm.calcmode = model.method, model.chanmode, 1
    mm.dt, mm.xmin, mm.invdx = model.dt, model.xmin, model.invdx
    mm.ops, mm.chip = _p(k["ops"], C.c_int32), _p(k["chip"], C.c_double)
    mm.tablist = _p(k["tablist"], C.c_double) if k["tablist"].size else None
    mm.xvals = _p(k["xvals"], C.c_double) if k["xvals"].size else None
    mm.stablist = _p(k["stablist"], C.c_double) if k["stablist"].size else None
    mm.funcs, mm.ravals = _p(k["funcs"], C.c_int32), _p(k["ravals"], C.c_double)
    mm.vm = _p(k["vm"], C.c_double)
    mm.nconcs, mm.nconcops, mm.nconcchip = k["conc"].size, k["concops"].size, k["concchip"].size
    mm.ydivs, mm.ymin, mm.invdy = model.ydivs, model.ymin, model.invdy
    mm.conc = _p(k["conc"], C.c_double) if k["conc"].size else None
    mm.concops = _p(k["concops"], C.c_int32) if k["concops"].size else None
    mm.concchip = _p(k["concchip"], C.c_double) if k["concchip"].size else None
    if givals and np.size(model.givals):
        k["givals"] = np.ascontiguousarray(model.givals, np.float64)
        mm.ngivals, mm.givals = k["givals"].size, _p(k["givals"], C.c_double)
    if np.size(getattr(model, "nshells", ())):
        for f in ("nshells", "nshrows", "nshfix", "dconc", "fbufval", "fdiffs"):
            k[f] = np.ascontiguousarray(getattr(model, f), np.int32)
            setattr(mm, f, _p(k[f], C.c_int32))
        for f in ("covals0", "bufvals"):
            k[f] = np.ascontiguousarray(getattr(model, f), np.float64)
            setattr(mm, f, _p(k[f], C.c_double) if k[f].size else None)
        mm.ndiffs, mm.ncovals0, mm.nbufvals = k["nshells"].size, k["covals0"].size, k["bufvals"].size
    return mm, k


def jit_source(model, givals: bool = False) -> str:
    """Device source generated for this model's channel programs."""
    mm, keep = _hsh_model(model, givals)
    n = C.c_int64()
    _check(lib().hsh_jit_source(C.byref(mm), None, 0, C.byref(n)), "hsh_jit_source")
    buf = C.create_string_buffer(n.value + 1)
    _check(lib().hsh_jit_source(C.byref(mm), buf, n.value + 1, C.byref(n)), "hsh_jit_source")
    return buf.value.decode()


class _GroupShape:
    """Array shapes of a group handle (the members' arrays concatenated)."""

    def __init__(self, models, givals):
        self.ncompts = sum(int(m.ncompts) for m in models)
        self.chip = np.concatenate([np.asarray(m.chip) for m in models])
        self.ops = np.concatenate([np.asarray(m.ops)[:-1] for m in models] + [np.array([102], np.int32)])
        self.conc = np.concatenate([np.asarray(m.conc) for m in models])
        self.givals = np.concatenate([np.asarray(m.givals) for m in models]) if givals else np.zeros(0)


class HipSolver:
    """One hsolve on the GPU.  ``model`` is a :class:`hsolve_hip.model.Model`."""

    def __init__(self, model, *, device: int = 0, batch_k: int = 0, batch_waves: int = 0,
                 flags: int = 0, host_only: bool = False, tile_max_rows: int = 0, pack_cap: int = 0,
                 pack_waves: int = 0, pack_stage: int = 0, givals: bool = False):
        """Options as in hsh_options: 0 = autotuned (or the default under
        F_NO_AUTOTUNE); ``tile_max_rows < 0`` puts every tree in packs;
        ``pack_stage`` 1 = schedule read from memory, 2 = staged in LDS, 3 = its
        indices staged in LDS (coefficients read from memory);
        ``givals``: pass the model's givals so the device writes that stream."""
        L = lib()
        if host_only:
            flags |= F_HOST_ONLY
        opt = HshOptions(device, batch_k, batch_waves, flags, tile_max_rows, pack_cap, pack_waves,
                         pack_stage)
        h = C.c_void_p()
        if isinstance(model, (list, tuple)):
            # a group: several solvers in one handle (hsh_create_group); the
            # state arrays are the members' concatenated in member order
            built = [_hsh_model(mi, givals) for mi in model]
            self._keep = built
            arr = (C.POINTER(HshModel) * len(built))(*[C.pointer(b[0]) for b in built])
            self.m = _GroupShape(model, givals)
            rc = L.hsh_create_group(arr, len(built), C.byref(opt), C.byref(h))
        else:
            self.m = model
            mm, self._keep = _hsh_model(model, givals)
            rc = L.hsh_create(C.byref(mm), C.byref(opt), C.byref(h))
        self.h = h
        if rc == HSH_REFUSED:
            refs = self.refusals()
            L.hsh_destroy(self.h)
            self.h = None
            raise Refused(L.hsh_last_error().decode(), refs)
        _check(rc, "hsh_create")

    # -- life cycle
    def close(self):
        if getattr(self, "h", None):
            lib().hsh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- queries
    def refusals(self):
        L = lib()
        n = C.c_int32()
        _check(L.hsh_refusals(self.h, None, 0, C.byref(n)), "hsh_refusals")
        arr = (HshRefusal * max(1, n.value))()
        _check(L.hsh_refusals(self.h, arr, n.value, C.byref(n)), "hsh_refusals")
        return [(r.comp, r.opcode, r.op_index, REASONS.get(r.reason, str(r.reason)))
                for r in arr[:n.value]]

    def info(self) -> dict:
        i = HshInfo()
        _check(lib().hsh_info_get(self.h, C.byref(i)), "hsh_info_get")
        return {f: getattr(i, f) for f, _ in HshInfo._fields_}

    def set_synapses(self, src_op, dst_op, weight, delay):
        """Synaptic connections and the synchans' pending lists on the device
        (hsh_set_synapses): SPIKE_OP ops index -> SYN2_OP ops index, weight,
        delay (s)."""
        a = [np.ascontiguousarray(x, t) for x, t in ((src_op, np.int32), (dst_op, np.int32),
                                                       (weight, np.float64), (delay, np.float64))]
        _check(lib().hsh_set_synapses(self.h, a[0].size, _p(a[0], C.c_int32), _p(a[1], C.c_int32),
                                      _p(a[2], C.c_double), _p(a[3], C.c_double)), "hsh_set_synapses")

    def post_events(self, syn_op, steps, weight, after=False):
        """EVENTs from outside the handle (hsh_post_events): SYN2_OP ops
        index, steps to the hit, weight; after the last step's ops or before
        the next step's."""
        a = [np.ascontiguousarray(np.atleast_1d(x), t) for x, t in ((syn_op, np.int32), (steps, np.int32),
                                                                     (weight, np.float64))]
        _check(lib().hsh_post_events(self.h, a[0].size, _p(a[0], C.c_int32), _p(a[1], C.c_int32),
                                     _p(a[2], C.c_double), 1 if after else 0), "hsh_post_events")

    def set_inputs(self, inputs):
        """h_in_msgs inputs before the next step (hsh_set_inputs): a
        sequence of (op, index, stab, per_dt, value)."""
        a = input_array(inputs)
        _check(lib().hsh_set_inputs(self.h, a.size, a.ctypes.data), "hsh_set_inputs")

    def set_time(self, t):
        """SimulationTime() of the next step (hsh_set_time)."""
        _check(lib().hsh_set_time(self.h, float(t)), "hsh_set_time")

    def set_ravals(self, ravals):
        r = np.ascontiguousarray(ravals, np.float64)
        assert r.size == self.m.ravals.size
        _check(lib().hsh_set_ravals(self.h, _p(r, C.c_double)), "hsh_set_ravals")

    def adj_forward(self, nsteps: int, ckpt: int, obs):
        """Run nsteps plain steps keeping checkpoints; returns vm[obs] after
        every step as an (nsteps, nobs) array."""
        obs = np.ascontiguousarray(obs, np.int32)
        tr = np.empty((nsteps, obs.size))
        rc = lib().hsh_adj_forward(self.h, nsteps, ckpt, obs.size, _p(obs, C.c_int32), _p(tr, C.c_double))
        if rc == HSH_REFUSED:
            raise Refused(lib().hsh_last_error().decode(), [])
        _check(rc, "hsh_adj_forward")
        self._adj_shape = tr.shape
        return tr

    def adj_backward(self, dtrace):
        """Gradients (dL/dchip, dL/dravals, dL/dvm0) of the last adj_forward
        run for dL/dtrace."""
        g = np.ascontiguousarray(dtrace, np.float64).reshape(self._adj_shape)
        gc = np.empty(self.m.chip.size)
        gr = np.empty(self.m.ravals.size)
        gv = np.empty(self.m.ncompts)
        _check(lib().hsh_adj_backward(self.h, _p(g, C.c_double), _p(gc, C.c_double), _p(gr, C.c_double),
                                      _p(gv, C.c_double)), "hsh_adj_backward")
        return gc, gr, gv

    def autotune(self):
        _check(lib().hsh_autotune(self.h), "hsh_autotune")
        return self.info()

    # -- stepping
    def step(self, n: int = 1, mode: int = PLAIN) -> dict:
        o = HshStepOut()
        _check(lib().hsh_step(self.h, n, mode, C.byref(o)), "hsh_step")
        return dict(vm_ready=o.vm_ready, launches=o.launches)

    def sync(self):
        _check(lib().hsh_sync(self.h), "hsh_sync")

    def set_timing(self, on: bool = True):
        _check(lib().hsh_set_timing(self.h, 1 if on else 0), "hsh_set_timing")

    def last_time(self, which: int = T_TOTAL) -> float:
        ms = C.c_double()
        _check(lib().hsh_last_time(self.h, which, C.byref(ms)), "hsh_last_time")
        return ms.value

    # -- state
    @property
    def vm(self):
        v = np.empty(self.m.ncompts)
        _check(lib().hsh_get_state(self.h, _p(v, C.c_double), None, None, None), "hsh_get_state")
        return v

    def state(self) -> dict:
        m = self.m
        vm = np.empty(m.ncompts)
        chip = np.empty(m.chip.size)
        res = np.empty(2 * m.ncompts)
        ops = np.empty(m.ops.size, np.int32)
        _check(lib().hsh_get_state(self.h, _p(vm, C.c_double), _p(chip, C.c_double),
                                   _p(res, C.c_double), _p(ops, C.c_int32)), "hsh_get_state")
        return dict(vm=vm, chip=chip, results=res, ops=ops)

    def member(self, i: int) -> dict:
        """Where member i of a group sits in the state arrays (hsh_group_member)."""
        e = HshMember()
        _check(lib().hsh_group_member(self.h, i, C.byref(e)), "hsh_group_member")
        return {f: getattr(e, f) for f, _ in HshMember._fields_}

    def givals(self) -> np.ndarray:
        """The chanmode-4/5 output stream givals[] as the last step wrote it."""
        out = np.empty(np.size(self.m.givals))
        _check(lib().hsh_get_givals(self.h, _p(out, C.c_double)), "hsh_get_givals")
        return out

    def conc(self) -> np.ndarray:
        """Pool concentrations conc[nconcs] (reference indices)."""
        out = np.empty(self.m.conc.size)
        if out.size:
            _check(lib().hsh_get_conc(self.h, _p(out, C.c_double)), "hsh_get_conc")
        return out

    def put_conc(self, conc):
        a = np.ascontiguousarray(conc, np.float64)
        if a.size != self.m.conc.size:
            raise ValueError(f"conc has {a.size} values, the model {self.m.conc.size}")
        if a.size:
            _check(lib().hsh_put_conc(self.h, _p(a, C.c_double)), "hsh_put_conc")

    def put_state(self, vm=None, chip=None, ops=None):
        keep = []

        def pp(a, dt, ct):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _p(a, ct)
        _check(lib().hsh_put_state(self.h, pp(vm, np.float64, C.c_double),
                                   pp(chip, np.float64, C.c_double), pp(ops, np.int32, C.c_int32)),
               "hsh_put_state")

    def set_chip(self, idx, val):
        idx = np.ascontiguousarray(idx, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        _check(lib().hsh_set_chip(self.h, idx.size, _p(idx, C.c_int32), _p(val, C.c_double)),
               "hsh_set_chip")

    def drain_spikes(self, cap: int = 1 << 20):
        a = np.empty(cap, np.int32)
        b = np.empty(cap, np.int32)
        n = C.c_int32()
        _check(lib().hsh_drain_spikes(self.h, _p(a, C.c_int32), _p(b, C.c_int32), cap, C.byref(n)),
               "hsh_drain_spikes")
        return a[:n.value].copy(), b[:n.value].copy()
