"""Message inputs on the device (hsh_set_inputs, hsh_set_chip) against the
oracle's h_in_msgs (hines_msgs.c:351-405), and the HSH_CHANNELS return-0
protocol (results[] only) against the oracle's chip update.  Bit-exact
state after every step of a run whose inputs change mid-run."""
import numpy as np
import pytest

from hsolve_hip import device as D
from hsolve_hip import model as M
from oracle import Oracle

pytestmark = pytest.mark.gpu

INJECT, THRESH, FLUX, ACT = 1, 3, 3300, 4100


def check_same(o, g, keys=("vm", "chip", "results", "ops")):
    os_, gs = o.state(), g.state()
    for k in keys:
        assert np.array_equal(os_[k], gs[k]), k


def input_plan(m, rng):
    """Per step, a list of inputs: inject pulses on a few cells, ACTIVATIONs (both scalings, two to one synchan in one step, one
    below VTINY)."""
    n = m.ncompts
    inj = [int(m.compchips[r]) + 1 for r in range(n)]
    plan = {}
    for s in range(0, 300, 7):
        cells = rng.choice(n, 5, replace=False)
        plan[s] = [(INJECT, inj[c], 0, 0, float(rng.uniform(0.0, 2e-9))) for c in cells]
    for s in range(3, 300, 11):
        e = []
        for j in rng.choice(len(m.syn_op), 4, replace=False):
            op, xs = int(m.syn_op[j]), int(m.syn_chip[j])
            k = int(m.ops[op + 1])
            e.append((ACT, xs, k, int(rng.integers(0, 2)), float(rng.uniform(0.1, 3.0))))
        e.append((ACT, e[0][1], e[0][2], 1, 0.7))        # a second one to the same synchan
        e.append((ACT, e[1][1], e[1][2], 1, 5e-11))      # below VTINY
        plan.setdefault(s, []).extend(e)
    return plan


@pytest.mark.parametrize("mode,kw", [(D.PLAIN, dict(flags=D.F_NO_AUTOTUNE)),
                                     (D.BATCHED, dict(flags=D.F_NO_AUTOTUNE, batch_k=8))], ids=["plain", "batched"])
def test_inputs_match_h_in_msgs(gpu, mode, kw):
    m, table = M.network(120, seed=4, inject=(0.2e-9, 0.6e-9))
    rng = np.random.default_rng(8)
    plan = input_plan(m, rng)
    print(f"[case] inputs on a 120-cell network, mode {mode}", flush=True)
    o = Oracle(m, givals=False)
    o.set_synapses(*table)
    with D.HipSolver(m, **kw) as g:
        g.set_synapses(*table)
        s = 0
        for t in sorted(plan) + [300]:
            if t > s:
                o.step(t - s)
                g.step(t - s, mode)
                s = t
            if t in plan:
                o.in_msgs(plan[t])
                g.set_inputs(plan[t])
        check_same(o, g)
        ow, og = o.spikes(), g.drain_spikes()
        assert np.array_equal(ow[0], og[0]) and np.array_equal(ow[1], og[1])


def test_set_chip(gpu):
    """hsh_set_chip overwrites chip values the device holds in its slots (the
    overwrite path INJECT, BASE and THRESH inputs share)."""
    m = M.isopotential_hh(8, seed=2, inject=(1.0e-9, 1.5e-9))
    o = Oracle(m, givals=False)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        o.step(100)
        g.step(100, D.BATCHED)
        inj = [int(m.compchips[r]) + 1 for r in range(m.ncompts)]
        vals = np.linspace(0.0, 2e-9, len(inj))
        g.set_chip(inj, vals)
        chip = o.state()["chip"].copy()
        chip[inj] = vals
        o.put(chip=chip)
        o.step(150)
        g.step(150, D.BATCHED)
        check_same(o, g)


def test_flux_input_has_no_effect_like_chanmode4(gpu):
    m = M.ca_population(20, (1, 8), (0.5, 0.5), features=M.CA_FEATURES[:6], seed=3)
    o = Oracle(m, givals=False)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        o.step(10)
        g.step(10, D.PLAIN)
        o.in_msgs([(FLUX, 0, 0, 0, 1e-6)])
        g.set_inputs([(FLUX, 0, 0, 0, 1e-6)])
        o.step(20)
        g.step(20, D.PLAIN)
        check_same(o, g)
        assert np.array_equal(o.conc(), g.conc())


@pytest.mark.parametrize("maker", [lambda: M.population(40, (1, 8, 16), (0.3, 0.3, 0.4), seed=6),
                                   lambda: M.network(60, seed=2, inject=(0.3e-9, 0.8e-9))[0]],
                         ids=["trees", "network"])
def test_channels_mode_results(gpu, maker):
    """HSH_CHANNELS (the reference's return-0 protocol, hines.c:225-252):
    the channel programs run and results[] are the oracle's chip update;
    vm[] is left for the CPU elimination."""
    m = maker()
    o = Oracle(m, givals=False)
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as g:
        o.step(30)
        g.step(30, D.PLAIN)
        vm0 = g.vm.copy()
        o.chip_update()
        r = g.step(1, D.CHANNELS)
        assert r["vm_ready"] == 0
        os_, gs = o.state(), g.state()
        assert np.array_equal(os_["results"], gs["results"])
        assert np.array_equal(os_["chip"], gs["chip"]) and np.array_equal(os_["ops"], gs["ops"])
        assert np.array_equal(gs["vm"], vm0)


def test_group_member_inputs_with_offsets(gpu):
    """Inputs to members of a group handle take the member's chip and
    stablist offsets (hsh_group_member): the group's state equals each
    member run on its own with its own inputs."""
    def cell(ek, tau1, inj):
        comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[inj])
        return M.build(comps, [], [], dt=1e-4, syntypes=[dict(Ek=ek, tau1=tau1, tau2=1e-3)], syns=[(0, 0, 1e-9)],
                       float_fields=False)
    cells = [cell(0.0, 5e-3, 0.1e-9), cell(-0.08, 10e-3, 0.2e-9), cell(0.0, 5e-3, 0.3e-9)]

    def own_inputs(m, v):
        xs = int(m.syn_chip[0])
        k = int(m.ops[int(m.syn_op[0]) + 1])
        return [(INJECT, int(m.compchips[0]) + 1, 0, 0, 0.4e-9 * v), (ACT, xs, k, 1, 2.0 * v)]
    with D.HipSolver(cells, flags=D.F_NO_AUTOTUNE) as g:
        mem = [g.member(i) for i in range(3)]
        singles = [D.HipSolver(c, flags=D.F_NO_AUTOTUNE) for c in cells]
        for step in range(3):
            grp_in = []
            for i, c in enumerate(cells):
                e = own_inputs(c, step + i + 1)
                singles[i].set_inputs(e)
                for (op, idx, stab, per_dt, val) in e:
                    grp_in.append((op, idx + mem[i]["chip0"], stab + (mem[i]["stab0"] if op == ACT else 0),
                                   per_dt, val))
            g.set_inputs(grp_in)
            g.step(7, D.BATCHED)
            for s in singles:
                s.step(7, D.BATCHED)
        gs = g.state()
        for i, s in enumerate(singles):
            ss = s.state()
            e = mem[i]
            assert np.array_equal(gs["vm"][e["row0"]:e["row0"] + e["nrows"]], ss["vm"])
            assert np.array_equal(gs["chip"][e["chip0"]:e["chip0"] + e["nchips"]], ss["chip"])
            s.close()
