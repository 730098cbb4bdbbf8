"""C ABI of libhsolve_hip on the CPU: the library loads, exports every symbol
include/hsolve_hip.h declares, and its host-side analysers (ops walk, tree
decode, tiling, refusals) behave as specified.  No kernel is launched here
(HSH_F_HOST_ONLY handles allocate nothing on a device)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import ap_verify_model, c3_model, mixed_model, random_tree_model
from hsolve_hip import device as D
from hsolve_hip import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "hsolve_hip.h")).read()
    return set(re.findall(r"^\s*(?:hsh_status|void|const char \*)\s*(hsh_\w+)\s*\(", txt, re.M))


def test_exports_match_header():
    L = D.lib()
    names = header_functions()
    assert names == set(D.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def plan(m, **kw):
    with D.HipSolver(m, host_only=True, **kw) as h:
        return h.info()


def test_plan_random_trees():
    m = random_tree_model(130, 16)
    info = plan(m)
    assert info["ntrees"] == 130
    assert info["max_rows"] == 16
    assert info["ntiles"] == 3                  # ceil(130/64)
    assert info["nblocks"] == 3 * 16
    # one dendrite program (Ek materialised) and one soma program
    assert info["ntemplates"] == 2
    assert abs(info["lane_fill"] - 130 / 192) < 1e-12


def test_plan_bytes_formula():
    """Algorithmic bytes per compartment-step: 16 (Vm r+w) + 16 per gate
    (state r+w) + 8 per read-only slot + 28 per non-root row (three axial
    coefficients + int32 parent)."""
    m = random_tree_model(64, 16, seed=3)
    info = plan(m)
    # soma: 11 slots (incl. materialised K Ek), 3 gates -> 16 + 48 + 8*8 = 128
    # dendrite: 7 slots, 1 gate -> 16 + 16 + 6*8 + 28 = 108
    want = 64 * (128 + 15 * 108)
    assert info["bytes_per_step"] == want


def test_plan_mixed_sizes_groups_tiles():
    m = mixed_model(500)
    info = plan(m)
    assert info["max_tree_rows"] == 64
    assert info["ntrees"] == 500
    # trees above 16 rows go to packs by default
    assert info["max_rows"] == 16 and info["npacks"] > 0
    wide = plan(m, tile_max_rows=64)
    assert wide["max_rows"] == 64 and wide["npacks"] == 0
    allp = plan(m, tile_max_rows=-1)
    assert allp["ntiles"] == 0 and allp["npacks"] > 0
    assert allp["tile_max_rows"] == 0


def test_plan_packs_fill_blocks():
    """Packs group compartments by program: with two programs (soma,
    dendrite) a pack of 1024-row trees wastes at most two partial blocks."""
    m = c3_model(400)
    for cap in (512, 1024, 2048):
        info = plan(m, pack_cap=cap, tile_max_rows=16)
        assert info["pack_cap"] == cap
        assert info["lane_fill"] > 0.8
        assert info["npacks"] >= 1
    assert plan(m)["max_tree_rows"] == 1024


def test_rejects_bad_options():
    m = random_tree_model(3, 4)
    for kw in (dict(pack_waves=9), dict(pack_stage=4), dict(pack_stage=-1), dict(pack_cap=-1)):
        with pytest.raises(D.HshError):
            D.HipSolver(m, host_only=True, **kw)


def synchan_model():
    comps = dict(parent=[-1, -1], Cm=[1e-11] * 2, Rm=[1e8] * 2, Ra=[1e7] * 2, Em=[-0.065] * 2,
                 initVm=[-0.065] * 2, inject=[0.0] * 2)
    return M.build(comps, [], [], dt=1e-4, spikes=[(0, 0.0, 5e-3)],
                   syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(1, 0, 1e-9)])


def test_accepts_synchan_and_checks_connections():
    m = synchan_model()
    with D.HipSolver(m, host_only=True) as h:
        assert h.refusals() == []
        src, dst = int(m.spike_op[0]), int(m.syn_op[0])
        h.set_synapses([src], [dst], [1.0], [1.5e-3])
        # with pending lists every step is a launch of its own: an EVENT can
        # hit in the next step (the head's re-timing drifts)
        assert h.info()["batch_k"] == 1
        with pytest.raises(D.HshError):          # not a SPIKE_OP
            h.set_synapses([dst], [dst], [1.0], [1.5e-3])
        with pytest.raises(D.HshError):          # not a SYN2_OP
            h.set_synapses([src], [src], [1.0], [1.5e-3])
        with pytest.raises(D.HshError):          # shorter than a step
            h.set_synapses([src], [dst], [1.0], [0.0] if dst > src else [-1.0])


def test_accepts_mg_block():
    from helpers import with_mgblock
    m, remap = with_mgblock(synchan_model(), [0])
    with D.HipSolver(m, host_only=True) as h:
        assert h.refusals() == []
        h.set_synapses([int(m.spike_op[0])], [int(m.syn_op[0])], [1.0], [1.5e-3])


def test_refuses_pending_synaptic_event_with_reason():
    m = synchan_model()
    ops = np.array(m.ops)
    op = int(m.syn_op[0])
    ops[op + 2] = 5            # the reference's HSEVENT counter of a pending event
    m.ops = ops
    with pytest.raises(D.Refused) as e:
        D.HipSolver(m, host_only=True)
    refs = e.value.refusals
    assert len(refs) == 1
    comp, opcode, opidx, why = refs[0]
    assert opcode == 4100 and opidx == op and why == "pending synaptic event"
    assert comp == m.hnum[1]


def test_rejects_malformed_ops():
    m = random_tree_model(3, 4)
    bad = np.array(m.ops)
    bad[-1] = 100          # no LCOMPT terminator
    m.ops = bad
    with pytest.raises(D.HshError):
        D.HipSolver(m, host_only=True)


def test_rejects_truncated_chip():
    m = random_tree_model(3, 4)
    m.chip = m.chip[:-3]
    with pytest.raises(D.HshError):
        D.HipSolver(m, host_only=True)


def test_host_only_handle_refuses_stepping():
    m = ap_verify_model(2)
    with D.HipSolver(m, host_only=True) as h:
        with pytest.raises(D.HshError):
            h.step(1)


def test_generated_source_compiles_for_gfx950():
    """The device source hsh_create generates for a model's channel programs
    compiles with hipRTC for gfx950 (no device needed to compile)."""
    L = C.CDLL("/opt/rocm/lib/libhiprtc.so")
    m = random_tree_model(70, 16, seed=5)
    src = D.jit_source(m)
    assert "tmpl_0" in src and "tmpl_1" in src and "hsh_jit_batched_r4" in src
    prog = C.c_void_p()
    assert L.hiprtcCreateProgram(C.byref(prog), src.encode(), b"hsh_jit.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-ffp-contract=off", b"-fno-fast-math", b"-std=c++17",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    rc = L.hiprtcCompileProgram(prog, len(opts), (C.c_char_p * len(opts))(*opts))
    if rc != 0:
        n = C.c_size_t()
        L.hiprtcGetProgramLogSize(prog, C.byref(n))
        log = C.create_string_buffer(n.value + 1)
        L.hiprtcGetProgramLog(prog, log)
        pytest.fail(log.value.decode()[:2000])
    n = C.c_size_t()
    L.hiprtcGetCodeSize(prog, C.byref(n))
    assert n.value > 0
    L.hiprtcDestroyProgram(C.byref(prog))


def test_header_is_clean_c(tmp_path):
    """include/hsolve_hip.h compiles as C99 and C++ with warnings as errors,
    and a C caller can pass the opaque handle between calls."""
    import subprocess
    src = tmp_path / "use.c"
    src.write_text('#include "hsolve_hip.h"\n'
                   'int f(const hsh_model *m) {\n'
                   '    struct hsh_handle *h = 0; hsh_step_out o; double c[1];\n'
                   '    if (hsh_create(m, 0, &h) != HSH_OK) return 1;\n'
                   '    hsh_step(h, 1, HSH_PLAIN, &o); hsh_get_conc(h, c);\n'
                   '    hsh_destroy(h); return 0;\n}\n')
    inc = os.path.join(ROOT, "include")
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        r = subprocess.run([cc, std, "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", inc,
                            "-x", "c" if cc == "gcc" else "c++", str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_givals_plan_and_source():
    """With givals the programs carry the leak-current op and its Rm slot; the
    generated source (with the stream writes) compiles for gfx950; a givals
    array shorter than the ops[] walk writes is rejected."""
    m = random_tree_model(70, 16, seed=5)
    with D.HipSolver(m, host_only=True, givals=True) as h:
        with_giv = h.info()["bytes_per_step"]
    with D.HipSolver(m, host_only=True) as h:
        assert h.info()["bytes_per_step"] < with_giv        # the Rm slots are read
    src = D.jit_source(m, givals=True)
    # the generated templates (not the shared interpreter text) carry the writes
    assert "Im = s0 - Vm / " in src and "if (giv) d.giv[gi] = Im;" in src
    assert "Im = s0 - Vm / " not in D.jit_source(m)
    L = C.CDLL("/opt/rocm/lib/libhiprtc.so")
    prog = C.c_void_p()
    assert L.hiprtcCreateProgram(C.byref(prog), src.encode(), b"hsh_jit.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-ffp-contract=off", b"-fno-fast-math", b"-std=c++17",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    assert L.hiprtcCompileProgram(prog, len(opts), (C.c_char_p * len(opts))(*opts)) == 0
    L.hiprtcDestroyProgram(C.byref(prog))
    short = M.Model(**{**m.__dict__, "givals": m.givals[:-5]})
    with pytest.raises(D.HshError):
        D.HipSolver(short, host_only=True, givals=True)


def test_group_offsets_refusals_and_compatibility():
    a, b = random_tree_model(3, 8, seed=1), M.isopotential_hh(4, seed=2)
    with D.HipSolver([a, b, a], host_only=True) as h:
        info = h.info()
        assert info["ntrees"] == 3 + 4 + 3
        e = [h.member(i) for i in range(3)]
        assert [x["row0"] for x in e] == [0, 24, 28]
        assert e[1]["op0"] == a.ops.size - 1 and e[2]["chip0"] == a.chip.size + b.chip.size
        with pytest.raises(D.HshError):
            h.member(3)
    # members must share dt
    c = random_tree_model(2, 8, seed=3, dt=10e-6)
    with pytest.raises(D.HshError):
        D.HipSolver([a, c], host_only=True)
    # a refused member refuses the group; refusal rows are group rows
    a4 = random_tree_model(3, 8, seed=1, dt=1e-4)
    syn = synchan_model()
    ops = np.array(syn.ops)
    ops[int(syn.syn_op[0]) + 2] = 5            # a pending event
    syn.ops = ops
    with pytest.raises(D.Refused) as e:
        D.HipSolver([a4, syn], host_only=True)
    (comp, opcode, opidx, why), = e.value.refusals
    assert why == "pending synaptic event" and comp == 24 + syn.hnum[1]
    assert opidx == a4.ops.size - 1 + int(syn.syn_op[0])


def test_symmetric_solver_is_a_unit():
    """A solver of symmetric compartments (its funcs[] holds COPY_ARRAY,
    SIBARRAY_ELIM, FASTSIBARRAY_ELIM) is accepted: its rows form one unit in
    a tile of their own, after the other trees' tiles."""
    from helpers import symmetric_model
    m = symmetric_model(3, 30, seed=4)
    with D.HipSolver(m, host_only=True) as h:
        info = h.info()
        assert info["refused"] == 0 and info["ntrees"] == 1 and info["npacks"] == 0
    with D.HipSolver([random_tree_model(70, 16, seed=1), m], host_only=True) as h:
        info = h.info()
        assert info["refused"] == 0 and info["ntiles"] == 3     # two of 64 trees, the unit's


def _sibling_ops(funcs):
    """(index, opcode) of every op of an elimination program."""
    f, out, i, fwd = list(funcs), [], 0, True
    while i < len(f):
        out.append((i, f[i]))
        if f[i] == 7:
            if not fwd:
                break
            fwd = False
            i += 1
        else:
            i += {0: 2, 3: 2, 4: 2, 5: 3}.get(f[i], 1)
    return out


@pytest.mark.parametrize("damage", ["ravals-index", "row-index", "past-last-row", "missing-calc"])
def test_refuses_malformed_sibling_programs(damage):
    """The device interprets a sibling program only after the host has
    walked it: every ravals index and row in range, the row cursor on the
    last row at the forward FINISH and one CALC_RESULTS per row below it.
    A program that fails is refused with its reason (the CPU path keeps the
    solver)."""
    from helpers import symmetric_model
    m = symmetric_model(1, 30, seed=4)
    f = np.array(m.funcs, dtype=np.int32)
    ops = _sibling_ops(f)
    if damage == "ravals-index":
        i = next(i for i, op in ops if op == 4)
        f[i + 1] = m.ravals.size + 3
    elif damage == "row-index":
        i = next(i for i, op in ops if op == 0)
        f[i + 1] = 2 * m.ncompts
    elif damage == "past-last-row":
        i = next(i for i, op in ops if op == 1)
        f[i] = 2                     # SET_DIAG -> SKIP_DIAG: the cursor overshoots
    else:
        i = next(i for i, op in ops if op == 6)
        f = np.delete(f, i)
    m.funcs = f
    with pytest.raises(D.Refused) as e:
        D.HipSolver(m, host_only=True)
    assert any(r[3] == "funcs (elimination program the device cannot run)" for r in e.value.refusals)


def test_group_member_stablist_offsets():
    """hsh_group_member reports where a member's synapse types start in the
    group's stablist: equal stablists are stored once, a different one
    follows (a member's hsh_input stab is its own index plus stab0)."""
    a = synchan_model()
    comps = dict(parent=[-1, -1], Cm=[1e-11] * 2, Rm=[1e8] * 2, Ra=[1e7] * 2, Em=[-0.065] * 2,
                 initVm=[-0.065] * 2, inject=[0.0] * 2)
    b = M.build(comps, [], [], dt=1e-4, spikes=[(0, 0.0, 5e-3)],
                syntypes=[dict(Ek=-0.08, tau1=10e-3, tau2=1e-3)], syns=[(1, 0, 1e-9)])
    with D.HipSolver([a, a, b, a], host_only=True) as h:
        assert [h.member(i)["stab0"] for i in range(4)] == [0, 0, 4, 0]
