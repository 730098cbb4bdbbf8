"""Host-side analysis of concentration ops (no kernel launched): the Ca
workload is accepted, its generated device source compiles for gfx950, and
what the device path cannot run is refused with its reason."""
import ctypes as C

import numpy as np
import pytest

from helpers import spliced
from hsolve_hip import device as D
from hsolve_hip import model as M

CONC_VAL, GHK, CHAN_EK, ADD_CURR, FLUX = 1000, 2010, 3001, 3200, 3300
IPOL1C, NEWCONC1, CACONC, LCACONC, CONC_OP = 4002, 5110, 10200, 10202, 10100
DT = 2e-5


def passive(n):
    comps = dict(parent=[-1] * n, Cm=[1e-11] * n, Rm=[1e8] * n, Ra=[1e6] * n, Em=[-0.065] * n,
                 initVm=[-0.065] * n, inject=[0.0] * n)
    return M.build(comps, [], [], dt=DT, float_fields=False)


def cols():
    return np.tile([[1.01, 0.001]], (12, 1))


def refusals(m):
    try:
        with D.HipSolver(m, host_only=True):
            return []
    except D.Refused as e:
        return e.refusals


@pytest.mark.parametrize("feats", [M.CA_FEATURES[:6] + ("nernst",), M.CA_FEATURES[:6] + ("ghk",),
                                   ("flux", "kc")])
def test_ca_population_accepted(feats):
    m = M.ca_population(40, (1, 8, 16), (0.4, 0.3, 0.3), features=feats, seed=2)
    with D.HipSolver(m, host_only=True) as h:
        assert h.refusals() == []
        info = h.info()
        assert info["ntrees"] == 40


def test_ca_generated_source_compiles_for_gfx950():
    L = C.CDLL("/opt/rocm/lib/libhiprtc.so")
    m = M.ca_population(20, (1, 8), (0.5, 0.5), features=M.CA_FEATURES[:6] + ("ghk",), seed=1)
    src = D.jit_source(m)
    for frag in ("caconc(", "newconc(", "ghk_call(", "ipol_xy(", "flux[0] += "):
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
        pytest.fail(log.value.decode()[:2000])
    L.hiprtcDestroyProgram(C.byref(prog))


def test_refuses_pool_fed_by_two_compartments():
    m = spliced(passive(2), lambda r: ([CHAN_EK, ADD_CURR, CONC_VAL, 0, FLUX], [0.1, 1e-9]),
                conc=[1e-4], concops=[CACONC, 0, LCACONC], concchip=[5e-5, 1.01, 1e6])
    refs = refusals(m)
    assert [(r[1], r[3]) for r in refs] == [(CONC_VAL, "shared pool")]
    assert refs[0][0] == 0 or refs[0][0] == 1


def test_constant_pool_may_be_shared():
    m = spliced(passive(2), lambda r: ([CONC_VAL, 0, NEWCONC1, CHAN_EK, IPOL1C, 0, 1, ADD_CURR],
                                       [-0.09, 1e-10, 0.5]),
                conc=[1e-4], extra_cols=cols(), ydivs=10, invdy=1e4)
    assert refusals(m) == []


def test_refuses_updated_pool_nobody_feeds():
    m = spliced(passive(1), lambda r: ([], []), conc=[1e-4], concops=[CACONC, 0, LCACONC],
                concchip=[5e-5, 1.01, 1e6])
    assert [(r[0], r[1], r[3]) for r in refusals(m)] == [(-1, CACONC, "shared pool")]


def test_refuses_undefined_concentration_op():
    """An opcode do_h4_conc_chip_update does not define (it errors out,
    hines_conc_chip.c:215-219) refuses the solver with that opcode."""
    m = spliced(passive(1), lambda r: ([CHAN_EK, ADD_CURR, CONC_VAL, 0, FLUX], [0.1, 1e-9]),
                conc=[1e-4], concops=[CONC_OP, 0, 10150, LCACONC], concchip=[5e-5, 1.01, 1e6])
    assert [(r[0], r[1], r[3]) for r in refusals(m)] == [(-1, 10150, "opcode")]


def test_refuses_ghk_on_updated_outside_pool():
    m = spliced(passive(2), lambda r: ([CHAN_EK, ADD_CURR, CONC_VAL, 1, FLUX], [0.1, 1e-9]) if r == 1 else
                ([CONC_VAL, 0, GHK, ADD_CURR], [75.0, 1.9e5, 1e-15]),
                conc=[1e-4, 2.0], concops=[CACONC, 1, LCACONC], concchip=[5e-5, 1.01, 1e6])
    assert [(r[1], r[3]) for r in refusals(m)] == [(GHK, "shared pool")]


def test_refuses_conc_register_from_earlier_compartment():
    progs = {0: ([CONC_VAL, 0], []), 1: ([NEWCONC1, CHAN_EK, IPOL1C, 0, 1, ADD_CURR], [-0.09, 1e-10, 0.5])}
    m = spliced(passive(2), lambda r: progs[r], conc=[1e-4], extra_cols=cols(), ydivs=10, invdy=1e4)
    refs = refusals(m)
    # the gate that reads the row is refused with it
    assert [(r[1], r[3]) for r in refs] == [(NEWCONC1, "register live-in"), (IPOL1C, "table live-in")]


def test_refuses_conc_row_from_earlier_compartment():
    progs = {0: ([CONC_VAL, 0, NEWCONC1], []), 1: ([CHAN_EK, IPOL1C, 0, 1, ADD_CURR], [-0.09, 1e-10, 0.5])}
    m = spliced(passive(2), lambda r: progs[r], conc=[1e-4], extra_cols=cols(), ydivs=10, invdy=1e4)
    refs = refusals(m)
    assert [(r[1], r[3]) for r in refs] == [(IPOL1C, "table live-in")]


def test_refuses_more_than_four_pools_in_a_compartment():
    ops, chip = [], []
    for k in range(5):
        ops += [CHAN_EK, ADD_CURR, CONC_VAL, k, FLUX]
        chip += [0.1, 1e-9]
    concops = sum(([CACONC, k] for k in range(5)), []) + [LCACONC]
    m = spliced(passive(1), lambda r: (ops, chip), conc=[1e-4] * 5, concops=concops,
                concchip=[5e-5, 1.01, 1e6] * 5)
    assert [(r[1], r[3]) for r in refusals(m)] == [(CONC_VAL, "shared pool")]


IPOL1V, IPOL2C, NEWCONC2, SAME_GHK, DO_TABGHK, NEWVOLT = 4001, 4004, 5111, 3002, 3003, 5100


def test_2d_and_table_ghk_accepted():
    m = M.ca_population(10, (1, 8), (0.5, 0.5), features=("flux", "kc", "kc2d", "c1c2", "tabghk", "ghk",
                                                             "sameghk"), seed=3, ydivs=20)
    assert refusals(m) == []


def test_refuses_2d_lookup_without_its_registers():
    # IPOL2C needs NEWCONC2 (first-concentration fraction, table register) in its compartment
    cols2 = np.tile([[1.01, 0.001]], (12, 12))
    progs = {0: ([CONC_VAL, 0, NEWCONC1, NEWCONC2, 1], []),
             1: ([CONC_VAL, 0, NEWCONC1, CHAN_EK, IPOL2C, 0, 1, ADD_CURR], [-0.09, 1e-10, 0.5])}
    m = spliced(passive(2), lambda r: progs[r], conc=[1e-4, 2e-4], extra_cols=cols2, ydivs=10, invdy=1e4)
    assert [(r[1], r[3]) for r in refusals(m)] == [(IPOL2C, "table live-in")]


def test_refuses_ghk_slope_from_earlier_compartment_and_bad_table_ghk():
    progs = {0: ([CONC_VAL, 0, GHK, ADD_CURR], [75.0, 1.9e5, 1e-15]), 1: ([SAME_GHK, ADD_CURR], [1e-15])}
    m = spliced(passive(2), lambda r: progs[r], conc=[1e-4, 2.0])
    assert [(r[1], r[3]) for r in refusals(m)] == [(SAME_GHK, "register live-in")]
    # DO_TABGHK_OP must be followed by a power-0 lookup
    m = spliced(passive(1), lambda r: ([NEWVOLT, DO_TABGHK, IPOL1V, 0, 1, ADD_CURR], [0.5]), extra_cols=cols())
    assert [(r[1], r[3]) for r in refusals(m)] == [(DO_TABGHK, "opcode")]
