"""Oracle pins for message inputs (h_in_msgs, hines_msgs.c:351-405):

* INJECT overwrites the compartment's inject before the chip update: a
  passive cell settles at the closed-form steady state Em + I*Rm of the new
  current, and a change at step 10 equals a run whose inject was the new
  value from that state on;
* an ACTIVATION to a synchan adds value*stablist[k+3] (times dt for a
  synchan) to X *before* SYN2_OP's decay: X after the step is
  (value*c3*dt)*xconst2, not value*c3*dt; values up to VTINY (1e-10) add
  nothing;
* FLUX sets the pool's flux register, which CONC_VAL_OP of the program that
  feeds the pool clears before use: no effect in chanmode 4.
"""
import numpy as np
import pytest

from hsolve_hip import model as M
from oracle import Oracle

INJECT, THRESH, FLUX, ACT = 1, 3, 3300, 4100


def passive_cell(I0=0.1e-9):
    comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[I0])
    return M.build(comps, [], [], dt=1e-4, float_fields=False)


def test_inject_overwrite_settles_on_closed_form():
    m = passive_cell()
    inj = int(m.compchips[0]) + 1
    assert m.chip[inj] == 0.1e-9
    o = Oracle(m, givals=False)
    o.step(10)
    o.in_msgs([(INJECT, inj, 0, 0, 0.3e-9)])
    assert o.state()["chip"][inj] == 0.3e-9
    o.step(20000)                                  # tau = Rm Cm = 1 ms = 10 steps
    assert abs(o.vm[0] - (-0.065 + 0.3e-9 * 1e8)) <= 1e-12
    # the same as starting from the step-10 state with the new inject
    a = Oracle(m, givals=False)
    a.step(10)
    st = a.state()
    chip = st["chip"].copy()
    chip[inj] = 0.3e-9
    b = Oracle(m, givals=False)
    b.put(vm=st["vm"], chip=chip)
    c = Oracle(m, givals=False)
    c.step(10)
    c.in_msgs([(INJECT, inj, 0, 0, 0.3e-9)])
    b.step(50)
    c.step(50)
    assert np.array_equal(b.vm, c.vm)


def one_synchan():
    comps = dict(parent=[-1], Cm=[1e-11], Rm=[1e8], Ra=[1e6], Em=[-0.065], initVm=[-0.065], inject=[0.0])
    m = M.build(comps, [], [], dt=1e-4, syntypes=[dict(Ek=0.0, tau1=5e-3, tau2=1e-3)], syns=[(0, 0, 1e-9)],
                float_fields=False)
    k = int(m.ops[int(m.syn_op[0]) + 1])
    return m, int(m.syn_chip[0]), k


@pytest.mark.parametrize("per_dt", [0, 1])
def test_activation_adds_before_the_decay(per_dt):
    m, xs, k = one_synchan()
    x2, c3 = m.stablist[k], m.stablist[k + 3]
    o = Oracle(m, givals=False)
    o.in_msgs([(ACT, xs, k, per_dt, 2.5)])
    o.step(1)
    a = 2.5 * c3 * m.dt if per_dt else 2.5 * c3
    X = o.state()["chip"][xs]
    assert X == a * x2
    assert X != a                                   # not added after the decay


def test_activation_up_to_vtiny_adds_nothing():
    m, xs, k = one_synchan()
    o = Oracle(m, givals=False)
    o.in_msgs([(ACT, xs, k, 1, 1e-10), (ACT, xs, k, 1, -3.0)])
    o.step(5)
    assert o.state()["chip"][xs] == 0.0
    o.in_msgs([(ACT, xs, k, 1, 1.0000001e-10)])
    o.step(1)
    assert o.state()["chip"][xs] > 0.0


def test_flux_message_cleared_by_conc_val():
    m = M.ca_population(3, (1,), (1.0,), features=M.CA_FEATURES[:6], seed=1)
    a, b = Oracle(m, givals=False), Oracle(m, givals=False)
    a.step(5)
    b.step(5)
    b.in_msgs([(FLUX, 0, 0, 0, 1e-6)])
    a.step(5)
    b.step(5)
    assert np.array_equal(a.conc(), b.conc()) and np.array_equal(a.vm, b.vm)


def test_bad_inputs_are_errors():
    m, xs, k = one_synchan()
    o = Oracle(m, givals=False)
    with pytest.raises(ValueError):
        o.in_msgs([(7, 0, 0, 0, 1.0)])
    with pytest.raises(ValueError):
        o.in_msgs([(ACT, xs, 400, 0, 1.0)])
