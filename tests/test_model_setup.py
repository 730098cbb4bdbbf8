"""Setup compiler against the reference's SETUP/RESET rules.

Each check restates a rule of the reference code by an independent route
(hand-worked small cases, structural invariants, counts the survey derives
from the reference's layout), so a wrong index or order fails.
"""
import numpy as np

from hsolve_hip import model as M

COMPT, FCOMPT, LCOMPT = 100, 101, 102
SPIKE, CHAN, CHAN_EK, ADD_CURR, IPOL1V, SYN2, NEWVOLT = 200, 3000, 3001, 3200, 4001, 4100, 5100


def _passive(parent):
    n = len(parent)
    comps = dict(parent=parent, Cm=[1e-11] * n, Rm=[1e8] * n, Ra=[1e7] * n, Em=[-0.065] * n,
                 initVm=[-0.065] * n, inject=[0.0] * n)
    return M.build(comps, [], [], dt=1e-5, float_fields=False)


def test_hines_numbering_hand_worked():
    # element tree:  0 -> {1, 2},  1 -> {3},  2 leaf,  3 leaf
    # do_hnum (hines_init.c:317-345): root 0 gets n-1=3, then its leaf kid 2
    # gets 2, then recursion into kid 1 (has kids): 1 gets 1, its leaf 3 gets 0.
    m = _passive([-1, 0, 0, 1])
    assert list(m.hnum) == [3, 1, 2, 0]
    assert list(m.elmnum) == [3, 1, 2, 0]
    assert list(m.parent_row) == [1, 3, 3, -1]


def test_multiple_roots_contiguous_blocks():
    # first root in element order takes the highest rows
    m = _passive([-1, 0, -1, 2, 2])
    rows_a = sorted(m.hnum[:2])
    rows_b = sorted(m.hnum[2:])
    assert rows_a == [3, 4] and rows_b == [0, 1, 2]


def test_funcs_program_structure():
    # unbranched cable of 4 + a single-compartment root
    m = _passive([-1, 0, 1, 2, -1])
    f = list(m.funcs)
    assert f.count(7) == 2                          # exactly two FINISH
    # every non-root row has exactly one BACKWARD_ELIM after the forward FINISH
    fwd_end = f.index(7)
    nonroot = int(np.sum(m.parent_row >= 0))
    assert len(m.ravals) == 3 * nonroot


def test_kid_sort_ascending_rows():
    rng = np.random.default_rng(0)
    m = M.population(50, (16,), (1.0,), seed=4)
    # FORWARD_ELIM operands for each parent appear in ascending row order
    f = m.funcs
    cur = 1
    last = {}
    i = 0
    while f[i] != 7:
        op = f[i]
        if op == 0:
            kid = f[i + 1] // 2
            assert kid > last.get(cur, -1)
            last[cur] = kid
            assert m.parent_row[kid] == cur
            i += 2
            continue
        cur += 1 if op == 1 else 2
        i += 1


def test_chip_layout_counts_isopotential():
    # 1000 HH cells: per cell [Em/Rm, inject, K(Ek,gmax,X), Na(Ek,gmax,X,Y), tbyc, diag]
    m = M.isopotential_hh(1000)
    assert m.chip.size == 11 * 1000
    assert m.ops.size == 15 * 1000 + 1
    assert m.ops[0] == FCOMPT and m.ops[-1] == LCOMPT


def test_ek_sharing_and_channel_order():
    """Channels are ordered by ascending Ek within a compartment
    (hines_child.c:2061-2068) and a channel whose Ek equals the previous one
    (across compartments too) is emitted as CHAN_OP without an Ek slot
    (hines_chip_init.c:1029-1034)."""
    m = M.population(4, (16,), (1.0,), seed=2)
    ops = m.ops
    # first compartment's first channel: CHAN_EK (nothing before it)
    assert ops[m.compops[0] + 2] == CHAN_EK
    # dendrites after the first inherit EK through CHAN_OP
    inherits = [ops[m.compops[r] + 2] for r in range(1, 15)]
    assert all(x == CHAN for x in inherits)
    # soma (root, last row of the neuron): K (lower Ek) before Na
    soma = m.compops[15]
    seq = list(ops[soma:m.compops[16]])
    assert seq[:2] == [COMPT, NEWVOLT]
    ek_positions = [i for i, x in enumerate(seq) if x in (CHAN, CHAN_EK)]
    assert len(ek_positions) == 2
    # the soma's K channel (EK, first by ascending Ek) follows the dendrites'
    # Kdr with the same EK, so it too inherits Ek: soma chip = [Em/Rm, inject,
    # K gmax, K X, Na Ek, Na gmax, Na X, Na Y, tbyc, diag] = 10 slots;
    # a dendrite with inherited Ek = 6
    assert ops[soma + 2] == CHAN
    sizes = np.diff(np.append(m.compchips, m.chip.size))
    assert sizes[15] == 10 and sizes[3] == 6


def test_gate_initial_values_are_steady_state():
    """RESET: X = A(V0)/B(V0) on the untransformed tables (olf/tabchannel.c:274-285),
    i.e. alpha/(alpha+beta) at the rest potential."""
    E = -0.065
    m = M.isopotential_hh(1)
    na, k = M.hh_chantypes(E)

    def rate(a, x):
        return (a[0] + a[1] * x) / (a[2] + np.exp((x + a[3]) / a[4]))
    def xinf(g):
        a = g[3]
        al, be = rate(a[:5], E), rate(a[5:], E)
        return al / (al + be)
    chip = m.chip
    # soma chip: [EmRm, inject, K Ek, K gmax, K X, Na Ek, Na gmax, Na X, Na Y, tbyc, diag]
    assert abs(chip[4] - xinf(k["gates"][0])) < 1e-6
    assert abs(chip[7] - xinf(na["gates"][0])) < 1e-6
    assert abs(chip[8] - xinf(na["gates"][1])) < 1e-6


def test_population_subrange_is_identical():
    full = M.population(10, (16,), (1.0,), seed=9)
    part = M.population(4, (16,), (1.0,), seed=9, first_neuron=6)
    # the last 4 neurons of the full population (elements 96..159) equal the
    # sub-population; compare element-ordered voltages and per-element chip
    # layout sizes
    assert part.ncompts == 64
    fs = np.diff(np.append(full.compchips, full.chip.size))[full.hnum[96:]]
    ps = np.diff(np.append(part.compchips, part.chip.size))[part.hnum]
    # the very first compartment differs only by CHAN_EK vs CHAN_OP (Ek slot)
    assert np.all((fs == ps) | (fs + 1 == ps))


def test_float_fields_round_element_values():
    n = 1
    comps = dict(parent=[-1], Cm=[1.2345678901234e-11], Rm=[1.2345678901234e8], Ra=[1e7],
                 Em=[-0.0654321987654], initVm=[-0.07], inject=[1e-10])
    mf = M.build(comps, [], [], dt=1e-5, float_fields=True)
    md = M.build(comps, [], [], dt=1e-5, float_fields=False)
    f32 = np.float32
    assert mf.chip[0] == float(f32(f32(-0.0654321987654) / f32(1.2345678901234e8)))
    assert md.chip[0] == -0.0654321987654 / 1.2345678901234e8
    assert mf.vm[0] == float(f32(-0.07))
