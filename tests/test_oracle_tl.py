"""Pins of the tangent-linear oracle (oracle/hsolve_tl.c), the CPU reference
for adjoint gradients.

* its values are the oracle's: the trace it carries equals the oracle's
  trace bit for bit;
* closed form: for a passive isopotential compartment V_{n+1} = (V_n +
  (Em/Rm + I) tbyc) / d, so dV_N/dI = tbyc (1 - d^-N) / (d - 1) and
  dV_N/dV_0 = d^-N (backward Euler);
* central finite differences of the oracle itself along random directions of
  vm0, chip[] (gmax, Ek, Em/Rm, inject, tbyc, diagterm, gate states) and
  ravals[], on spiking HH trees with both integration methods.  The step is
  h = 1e-9 (relative): through a spike upstroke the trace is strongly
  nonlinear in the parameters, so central differences at h = 1e-6 still
  carry ~1e-3 truncation error there; at 1e-9 they agree with the tangent to
  1e-7..3e-6 (rounding in the differences, and a voltage that sits on a
  table breakpoint, where the interpolated rates change slope).  A dropped
  term or a wrong sign shows as an O(1) difference.
"""
import numpy as np
import pytest

from hsolve_hip import fit as F
from hsolve_hip import model as M
from oracle import Oracle, TangentLinear


def oracle_trace(m, nsteps, obs):
    o = Oracle(m, givals=False)
    out = np.empty((nsteps, len(obs)))
    for s in range(nsteps):
        o.step(1)
        out[s] = o.vm[obs]
    return out


def test_tl_values_are_the_oracles():
    pm = F.ParamMap(3, (8,), (1.0,), seed=4)
    m = pm.base
    tr, _ = TangentLinear(m).run(120, pm.soma)
    assert np.array_equal(tr, oracle_trace(m, 120, pm.soma))


def test_tl_passive_closed_form():
    Cm, Rm, Em, I, V0, dt = 2.8e-11, 1.2e8, -0.065, 0.3e-9, -0.07, 20e-6
    comps = dict(parent=[-1], Cm=[Cm], Rm=[Rm], Ra=[1e6], Em=[Em], initVm=[V0], inject=[I])
    m = M.build(comps, [], [], dt=dt, float_fields=False)
    # chip of a passive compartment: [Em/Rm, inject, tbyc, diagterm]
    assert m.chip.size == 4
    tbyc, d = m.chip[2], m.chip[3]
    N = 500
    dchip = np.zeros(4)
    dchip[1] = 1.0
    _, dtr = TangentLinear(m, dchip=dchip).run(N, [0])
    want = tbyc * (1.0 - d ** -N) / (d - 1.0)
    assert abs(dtr[-1, 0] - want) <= 1e-12 * abs(want)
    _, dtr = TangentLinear(m, dvm=np.ones(1)).run(N, [0])
    assert abs(dtr[-1, 0] - d ** -N) <= 1e-12 * d ** -N


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_tl_matches_central_differences(method):
    pm = F.ParamMap(4, (8, 16), (0.5, 0.5), seed=7, method=method)
    m = pm.base
    obs = pm.soma
    nsteps = 150
    rng = np.random.default_rng(3)
    for trial in range(3):
        # random relative direction over every entry
        dchip = rng.uniform(-1, 1, m.chip.size) * m.chip
        drav = rng.uniform(-1, 1, m.ravals.size) * m.ravals
        dvm = rng.uniform(-1, 1, m.ncompts) * 1e-3
        _, dtr = TangentLinear(m, dvm=dvm, dchip=dchip, dravals=drav).run(nsteps, obs)
        h = 1e-9
        import dataclasses
        mp = dataclasses.replace(m, chip=m.chip + h * dchip, ravals=m.ravals + h * drav, vm=m.vm + h * dvm)
        mm = dataclasses.replace(m, chip=m.chip - h * dchip, ravals=m.ravals - h * drav, vm=m.vm - h * dvm)
        fd = (oracle_trace(mp, nsteps, obs) - oracle_trace(mm, nsteps, obs)) / (2 * h)
        scale = np.max(np.abs(fd))
        assert scale > 0
        err = np.max(np.abs(dtr - fd)) / scale
        assert err < 1e-5, (trial, err)
