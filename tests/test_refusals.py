"""Every refusal reason of the analyser, forced on a small model on the CPU
(host-only handles, nothing launched): the solver is refused -- its CPU path
keeps running it -- with the compartment, the offending opcode and its
index, never skipped.  Reasons covered elsewhere: pending synaptic event and
symcompartment funcs (test_abi.py), shared / unfed pools and register
live-ins of the concentration ops (test_abi_conc.py)."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_tree_model, spliced
from hsolve_hip import device as D
from hsolve_hip import model as M

DT = 2e-5
SYN2, SYN3, ELMMPUMP, STORE_ALLI, STORE_ALLC = 4100, 4101, 500, 3501, 3502
CONC_VAL, NERNST, CHAN, CHAN_EK, ADD_CURR, IPOL1V = 1000, 2000, 3000, 3001, 3200, 4001


def passive(n):
    comps = dict(parent=[-1] * n, Cm=[1e-11] * n, Rm=[1e8] * n, Ra=[1e6] * n, Em=[-0.065] * n,
                 initVm=[-0.065] * n, inject=[0.0] * n)
    return M.build(comps, [], [], dt=DT, float_fields=False)


def refusals(m):
    try:
        with D.HipSolver(m, host_only=True):
            return []
    except D.Refused as e:
        return e.refusals


def test_refuses_calcmode_0():
    """calcmode 0 (no interpolation, hines_chip.c SOLVEINTERPOL off) is a
    whole-solver refusal."""
    m = random_tree_model(2, 4)
    mm, keep = D._hsh_model(m)
    mm.calcmode = 0
    opt = D.HshOptions()
    opt.flags = D.F_HOST_ONLY
    h = C.c_void_p()
    L = D.lib()
    rc = L.hsh_create(C.byref(mm), C.byref(opt), C.byref(h))
    assert rc == D.HSH_REFUSED
    n = C.c_int32()
    arr = (D.HshRefusal * 4)()
    assert L.hsh_refusals(h, arr, 4, C.byref(n)) == 0
    assert n.value == 1 and (arr[0].comp, arr[0].reason) == (-1, 7)
    L.hsh_destroy(h)


def test_refuses_syn3():
    """SYN3_OP draws urandom() in the reference's serial op order
    (hines_chip.c:514-520): refused with its opcode."""
    comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[0.0])
    m = M.build(comps, [], [], dt=1e-4, syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(0, 0, 1e-9)],
                float_fields=False)
    import dataclasses
    op, xs = int(m.syn_op[0]), int(m.syn_chip[0])
    ops = np.array(m.ops)
    ops[op] = SYN3
    chip = np.insert(np.asarray(m.chip), xs, 1e-4)        # dt * frequency before X
    m3 = dataclasses.replace(m, ops=ops, chip=chip,
                             compchips=np.where(m.compchips > xs, m.compchips + 1, m.compchips))
    refs = refusals(m3)
    assert [(r[1], r[2], r[3]) for r in refs] == [(SYN3, op, "opcode")]


@pytest.mark.parametrize("prog,why", [(([ELMMPUMP], [1e-3, 2e-3, 3e-3]), "register live-in"),
                                      (([STORE_ALLI, 0], []), "opcode"), (([STORE_ALLC, 0], []), "opcode")],
                         ids=["ELMMPUMP", "STORE_ALLI", "STORE_ALLC"])
def test_refuses_pump_and_storeall_ops(prog, why):
    """The STORE_ALLI/STORE_ALLC storemode ops (hines_chip.c:314-322) are
    refused with their opcode in the compartment that holds them; an
    ELMMPUMP_OP (:278-302) with no CONC_VAL_OP before it in its compartment
    would read the concentration register of an earlier one."""
    m = spliced(passive(3), lambda r: prog if r == 1 else ([], []))
    refs = refusals(m)
    assert len(refs) == 1
    comp, opcode, opidx, got = refs[0]
    assert (opcode, got) == (prog[0][0], why) and comp == 1
    assert int(m.ops[opidx]) == prog[0][0]


def test_refuses_channel_that_inherits_computed_ek():
    """A CHAN_OP whose Ek register was computed by a Nernst op in an earlier
    compartment (the reference carries it across compartments,
    hines_chip.c:381-393): refused as an Ek live-in."""
    # NERNST_OP: Ek = log(Cout/Cin) * const, Gk = gmax (hines_chip.c:367-379)
    progs = {0: ([CONC_VAL, 0, NERNST, ADD_CURR], [43.0, 1e-10]),
             1: ([CHAN, ADD_CURR], [2e-10])}
    m = spliced(passive(2), lambda r: progs[r], conc=[1e-4, 2.0], concops=[], concchip=[])
    refs = refusals(m)
    assert [(r[0], r[1], r[3]) for r in refs] == [(1, CHAN, "Ek live-in")]


def test_refuses_gate_without_newvolt():
    """A voltage gate in a compartment with no NEWVOLT_OP reads the table row
    the previous compartment left (hines_chip.c:174-203): refused as a table
    live-in."""
    m = spliced(passive(2), lambda r: ([CHAN_EK, IPOL1V, 0, 3, ADD_CURR], [0.05, 1e-9, 0.2]) if r == 1
                else ([], []), extra_cols=np.tile([[1.01, 0.001]], (12, 1)))
    refs = refusals(m)
    assert [(r[0], r[1], r[3]) for r in refs] == [(1, IPOL1V, "table live-in")]


def test_refuses_kids_out_of_order():
    """funcs[] that pulls a row's children in descending order (the
    reference's kid sort makes them ascending, hines_init.c:278-295) does not
    match the device's level schedule: refused as a tree layout."""
    m = random_tree_model(1, 16, seed=5)
    f = np.array(m.funcs, dtype=np.int32)
    # forward pass: FORWARD_ELIM (0) kid*2 entries of one row, then SET_DIAG
    # (1) or SKIP_DIAG (2) to the next row, FINISH (7); two FORWARD_ELIMs in
    # a row pull two kids of one parent: swap them
    i, pairs = 0, []
    while f[i] != 7:
        if f[i] == 0:
            if f[i + 2] == 0:
                pairs.append(i)
            i += 2
        else:
            i += 1
    assert pairs, "no row with two kids in this tree"
    i = pairs[0]
    f[i + 1], f[i + 3] = f[i + 3], f[i + 1]
    import dataclasses
    m2 = dataclasses.replace(m, funcs=f)
    refs = refusals(m2)
    assert refs and all(r[3] == "tree layout" for r in refs)
