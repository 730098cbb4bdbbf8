"""Host-side analysis of diffusion shells, buffers, pumps and concpools (no
kernel launched): the shell workloads are accepted, the device source
generated for them compiles for gfx950, the algorithmic bytes leave out the
solve's per-step values, and what the device path cannot run is refused with
its reason (the CPU path keeps the solver)."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M

ALL = ("fixbuf", "difbuf", "ctaupump", "taupump", "mmpump", "hillpump", "elmmpump", "pool", "kc", "leak")
TAUPUMP, PHILLPUMP, CONC_VAL = 11000, 11012, 1000


def refusals(m):
    try:
        with D.HipSolver(m, host_only=True):
            return []
    except D.Refused as e:
        return e.refusals


@pytest.mark.parametrize("feats,method", [(ALL, M.CRANK), (("ctaupump", "taupump", "mmpump", "hillpump", "kc",
                                                            "leak"), M.BEULER), (("fixbuf",), M.CRANK)])
def test_shell_population_accepted(feats, method):
    m = M.shell_population(30, (1, 4, 16), (0.4, 0.3, 0.3), shells=4, features=feats, method=method, seed=3)
    with D.HipSolver(m, host_only=True) as h:
        assert h.refusals() == []
        assert h.info()["ntrees"] == 30


def test_shell_generated_source_compiles_for_gfx950():
    L = C.CDLL("/opt/rocm/lib/libhiprtc.so")
    m = M.shell_population(12, (1, 4), (0.5, 0.5), shells=4, features=ALL, method=M.CRANK, seed=1)
    src = D.jit_source(m)
    for frag in ("diff_solve_t<1>(Sb, lane, 4, 3, 1,", "conc_fluxconc(", "conc_ctaupump(", "conc_taupump(",
                 "conc_mmpump(", "conc_hillpump(", "conc_phillpump(", "conc_buffer(", "+= csum;", "= csum;"):
        assert frag in src, frag
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
        pytest.fail(log.value.decode()[:3000])
    L.hiprtcDestroyProgram(C.byref(prog))


def test_bytes_count_state_and_constants_not_scratch():
    """Per compartment-step the algorithmic bytes grow by 16 per
    concentration of its systems (read and written), 8 per covals0 value,
    buffer rate and concentration-program value (read), and nothing for the
    solve's per-step diagonals, results and elimination scratch."""
    base = dict(shells=3, method=M.BEULER, seed=2)
    m0 = M.shell_population(64, (1,), (1.0,), features=("ctaupump",), **base)
    with D.HipSolver(m0, host_only=True) as h:
        b0 = h.info()["bytes_per_step"]
    # ctaupump on each shell system: 3 conc rows r+w (48), 4 covals0 (32),
    # FLUXCONC + CTAUPUMP values (32), and the channel part of the program
    m1 = M.shell_population(64, (1,), (1.0,), features=("ctaupump", "mmpump"), **base)
    with D.HipSolver(m1, host_only=True) as h:
        b1 = h.info()["bytes_per_step"]
    assert b1 - b0 == pytest.approx(64 * 16.0)        # MMPUMP's two values, per compartment
    pop = M.population(64, (1,), (1.0,), seed=2, method=M.BEULER)
    with D.HipSolver(pop, host_only=True) as h:
        bp = h.info()["bytes_per_step"]
    # channel additions: Ca channel (Ek, gmax, m gate r+w: 8+8+16), K_C off
    per = (b0 - bp) / 64
    assert per == pytest.approx(3 * 16 + 4 * 8 + 4 * 8 + (8 + 8 + 16))


def test_refuses_single_shell_system():
    m = M.shell_population(4, (1,), (1.0,), shells=2, features=("ctaupump",), method=M.BEULER)
    ns = m.nshells.copy()
    ns[1] = 1
    refs = refusals(dataclasses.replace(m, nshells=ns))
    assert [(r[0], r[3]) for r in refs] == [(-1, D.REASONS[10])]


def test_refuses_system_no_compartment_owns():
    m = M.shell_population(4, (1,), (1.0,), shells=3, features=("ctaupump",), method=M.BEULER)
    fd = m.fdiffs.copy()
    fd[4] = 3                # the last compartment no longer lists system 3
    refs = refusals(dataclasses.replace(m, fdiffs=fd))
    assert [(r[0], r[3]) for r in refs] == [(-1, D.REASONS[10])]
    fd = m.fdiffs.copy()
    fd[2] = fd[1]            # system 1 listed by compartment 2, read by compartment 1's channel
    refs = refusals(dataclasses.replace(m, fdiffs=fd))
    assert refs and all(r[0] == 1 and r[3] == "shared pool" for r in refs)


def test_refuses_taupump_on_another_compartments_vm():
    m = M.shell_population(4, (1,), (1.0,), shells=3, features=("taupump",), method=M.BEULER)
    co = m.concops.copy()
    q = int(np.flatnonzero(co == TAUPUMP)[2])
    assert co[q + 1] == 2
    co[q + 1] = 0
    refs = refusals(dataclasses.replace(m, concops=co))
    assert [(r[0], r[1], r[3]) for r in refs] == [(2, TAUPUMP, D.REASONS[10])]


def test_refuses_shell_read_by_another_compartment():
    """A channel of compartment 1 reading compartment 0's outermost shell:
    the shell is updated by compartment 0's program."""
    m = M.shell_population(4, (1,), (1.0,), shells=3, features=("ctaupump", "kc"), method=M.BEULER)
    ops = m.ops.copy()
    q = [int(i) for i in np.flatnonzero(ops == CONC_VAL)]
    ops[q[1] + 1] = ops[q[0] + 1]
    refs = refusals(dataclasses.replace(m, ops=ops))
    assert (1, CONC_VAL, q[1], "shared pool") in [tuple(r) for r in refs]


def test_refuses_concpool_whose_pump_is_not_on_a_shell():
    m = M.shell_population(4, (1,), (1.0,), shells=3, features=("fixbuf", "pool"), method=M.CRANK)
    co = m.concops.copy()
    q = int(np.flatnonzero(co == PHILLPUMP)[0])
    co[q + 2] = int(m.dconc[0])              # a buffer row, not the shell
    refs = refusals(dataclasses.replace(m, concops=co))
    assert any(r[1] == PHILLPUMP and r[3] == D.REASONS[10] for r in refs)


def test_group_of_shell_solvers_is_an_error():
    m = M.shell_population(2, (1,), (1.0,), shells=3, features=("ctaupump",), method=M.BEULER)
    with pytest.raises(D.HshError):
        D.HipSolver([m, m], host_only=True)
