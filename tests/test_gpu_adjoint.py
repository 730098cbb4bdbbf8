"""Device adjoint (hsh_adj_forward / hsh_adj_backward) against the CPU
tangent-linear oracle and against central finite differences.

* dot-product identity: for any direction (dvm0, dchip, dravals),
  <dL/dtrace, TL(direction)> == <adjoint gradients, direction>;
* parameter gradients through the torch.autograd.Function harness equal the
  tangent-linear gradients (1e-7); central differences of the device forward
  run agree within 1e-4 at h = 1e-8 for every parameter and within 1e-5 at
  h = 1e-6 where the perturbation keeps spike upstrokes inside one gate-table
  interval (gK, gKdr, Ra): the adjoint differentiates the tabulated model,
  whose rates are piecewise linear in Vm;
* checkpoint interval does not change the gradients (recomputation is
  exact); the forward trace is the oracle's bit for bit; the state after the
  backward pass is the state the forward run left.
"""
import numpy as np
import pytest
import torch

from hsolve_hip import device as D
from hsolve_hip import fit as F
from hsolve_hip import model as M
from oracle import Oracle, TangentLinear

pytestmark = pytest.mark.gpu


def solver_for(m, rows):
    return D.HipSolver(m, flags=D.F_NO_AUTOTUNE, tile_max_rows=rows)


def oracle_trace(m, nsteps, obs):
    o = Oracle(m, givals=False)
    out = np.empty((nsteps, len(obs)))
    for s in range(nsteps):
        o.step(1)
        out[s] = o.vm[obs]
    return out


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
@pytest.mark.parametrize("sizes,rows", [((8, 16), 16), ((40,), 40), ((100,), 128)],
                         ids=["rows16", "rows40-lds", "rows100-global"])
def test_adjoint_dot_product_with_tangent_linear(gpu, method, sizes, rows):
    """Tiles of up to 64 rows run the reverse sweep with the tile in LDS,
    larger tiles in global memory."""
    pm = F.ParamMap(6, sizes, (1.0 / len(sizes),) * len(sizes), seed=5, method=method)
    m, obs, T = pm.base, pm.soma, 200
    print(f"[case] dot-product method={method} sizes={sizes} ncompts={m.ncompts}", flush=True)
    rng = np.random.default_rng(11)
    with solver_for(m, rows) as s:
        tr = s.adj_forward(T, 64, obs)
        assert np.array_equal(tr, oracle_trace(m, T, obs))
        w = rng.standard_normal(tr.shape)
        gc, gr, gv = s.adj_backward(w)
    for trial in range(3):
        dchip = rng.uniform(-1, 1, m.chip.size) * m.chip
        drav = rng.uniform(-1, 1, m.ravals.size) * m.ravals
        dvm = rng.uniform(-1, 1, m.ncompts) * 1e-3
        _, dtr = TangentLinear(m, dvm=dvm, dchip=dchip, dravals=drav).run(T, obs)
        lhs = float(np.sum(w * dtr))
        terms = np.concatenate([gc * dchip, gr * drav, gv * dvm])
        rhs = float(np.sum(terms))
        scale = max(abs(lhs), float(np.sum(np.abs(terms))))
        assert abs(lhs - rhs) <= 1e-8 * scale, (trial, lhs, rhs)


def test_parameter_gradients_tl_and_finite_differences(gpu):
    """64-compartment trees: autograd gradient of the soma-Vm MSE against
    targets from the CPU solver at hidden parameters (+-30%)."""
    N, T = 6, 400
    pm = F.ParamMap(N, (64,), (1.0,), seed=0)
    rng = np.random.default_rng(1)
    th_star = 1 + 0.3 * rng.uniform(-1, 1, (N, 5))
    tgt = torch.from_numpy(oracle_trace(pm.model_at(th_star), T, pm.soma))
    run = F.Runner(pm, T, ckpt=100)
    print(f"[case] theta gradients N={N} T={T}", flush=True)
    theta = torch.ones(N, 5, dtype=torch.float64, requires_grad=True)
    loss = torch.mean((run(theta) - tgt) ** 2)
    loss.backward()
    g = theta.grad.numpy().copy()
    with torch.no_grad():
        tr0 = run(torch.ones(N, 5, dtype=torch.float64))
    w = (2 * (tr0 - tgt) / tr0.numel()).numpy()
    for (n, j) in [(0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (5, 0)]:
        e = torch.zeros(N, 5, dtype=torch.float64)
        e[n, j] = 1
        _, (dc, dr) = torch.func.jvp(pm.arrays, (torch.ones(N, 5, dtype=torch.float64),), (e,))
        _, dtr = TangentLinear(pm.base, dchip=dc.numpy(), dravals=dr.numpy()).run(T, pm.soma)
        g_tl = float(np.sum(w * dtr))
        assert abs(g[n, j] - g_tl) <= 1e-7 * abs(g_tl), (n, j, g[n, j], g_tl)
        def fd(h):
            with torch.no_grad():
                tp = torch.ones(N, 5, dtype=torch.float64)
                tp[n, j] += h
                tm = torch.ones(N, 5, dtype=torch.float64)
                tm[n, j] -= h
                return (torch.mean((run(tp) - tgt) ** 2) - torch.mean((run(tm) - tgt) ** 2)).item() / (2 * h)
        # rounding floor of a difference quotient of the loss: ~eps * L / h
        L0 = loss.item()
        floor = lambda h: 10 * np.finfo(float).eps * L0 / h
        # h = 1e-8: the perturbed voltages stay inside one gate-table interval
        # (5e-5 V), so the difference quotient measures the same piecewise-
        # linear model the adjoint differentiates; the bound is rounding.
        assert abs(g[n, j] - fd(1e-8)) <= 1e-4 * abs(g[n, j]) + floor(1e-8), (n, j)
        # h = 1e-6: within 1e-5 where the perturbation does not move a spike
        # upstroke across table intervals (dV/dgNa and dV/dCm reach ~60 V
        # there, i.e. 6e-5 V at h = 1e-6, more than one interval)
        if F.PARAMS[j] in ("gK", "gKdr", "Ra"):
            assert abs(g[n, j] - fd(1e-6)) <= 1e-5 * abs(g[n, j]) + floor(1e-6), (n, j)
    run.close()


def test_checkpoint_interval_and_state(gpu):
    pm = F.ParamMap(5, (16,), (1.0,), seed=9)
    m, obs, T = pm.base, pm.soma, 150
    print("[case] checkpoint intervals", flush=True)
    w = np.random.default_rng(2).standard_normal((T, obs.size))
    outs = []
    for ck in (7, 50, 150):
        with solver_for(m, 16) as s:
            s.adj_forward(T, ck, obs)
            after_fwd = s.state()
            outs.append(s.adj_backward(w))
            after_bwd = s.state()
            for k in ("vm", "chip", "ops"):
                assert np.array_equal(after_fwd[k], after_bwd[k]), k
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
