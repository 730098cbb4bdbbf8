"""Device adjoint (hsh_adj_forward / hsh_adj_backward) against the CPU
tangent-linear oracle and against central finite differences.

* dot-product identity: for any direction (dvm0, dchip, dravals),
  <dL/dtrace, TL(direction)> == <adjoint gradients, direction>;
* parameter gradients through the torch.autograd.Function harness equal the
  tangent-linear gradients (1e-7); central differences of the device forward
  run at h = 1e-6 agree within 1e-5 for every (neuron, parameter) whose
  +-h runs interpolate the same gate-table rows at every step (the oracle's
  row trace decides), for all five parameters: the adjoint differentiates
  the tabulated model, whose rates are piecewise linear in Vm;
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


def solver_for(m, rows, layout="tiles", cap=256):
    if layout == "packs":
        return D.HipSolver(m, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=cap)
    if layout == "packs-interp":
        return D.HipSolver(m, flags=D.F_NO_AUTOTUNE | D.F_NO_JIT, tile_max_rows=-1, pack_cap=cap)
    if layout == "packs-idx":
        return D.HipSolver(m, flags=D.F_NO_AUTOTUNE, tile_max_rows=-1, pack_cap=cap, pack_waves=8, pack_stage=3)
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


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
@pytest.mark.parametrize("sizes,cap", [((8, 16), 64), ((40,), 256), ((100, 7), 128), ((30, 1, 12), 1024)],
                         ids=["rows16-cap64", "rows40-cap256", "rows100-cap128", "mixed-cap1024"])
def test_pack_adjoint_equals_tile_adjoint(gpu, method, sizes, cap):
    """The pack layout's adjoint (one recording launch and one reverse launch
    per checkpoint interval, level-scheduled sweeps; generated reverse
    programs, and the interpreter; the recording launch with the schedule
    read from memory and with its indices in LDS) gives the tile layout's trace and
    gradients bit for bit, and leaves the forward run's state."""
    pm = F.ParamMap(9, sizes, (1.0 / len(sizes),) * len(sizes), seed=7, method=method)
    m, obs, T = pm.base, pm.soma, 160
    rows = int(np.max(np.bincount(F.tree_of_row(m))))
    print(f"[case] packs vs tiles method={method} sizes={sizes} cap={cap}", flush=True)
    w = np.random.default_rng(3).standard_normal((T, obs.size))
    out = {}
    for layout in ("tiles", "packs", "packs-interp", "packs-idx"):
        with solver_for(m, rows, layout, cap) as s:
            tr = s.adj_forward(T, 50, obs)
            after_fwd = s.state()
            g = s.adj_backward(w)
            after_bwd = s.state()
            for k in ("vm", "chip", "ops"):
                assert np.array_equal(after_fwd[k], after_bwd[k]), (layout, k)
            out[layout] = (tr, g)
    assert np.array_equal(out["tiles"][0], oracle_trace(m, T, obs))
    for layout in ("packs", "packs-interp", "packs-idx"):
        assert np.array_equal(out["tiles"][0], out[layout][0]), layout
        for a, b in zip(out["tiles"][1], out[layout][1]):
            assert np.array_equal(a, b), layout


def test_mixed_layout_moves_to_packs(gpu):
    """A handle whose layout mixes tiles and packs runs the adjoint with
    every tree in packs: same gradients as the tile layout."""
    pm = F.ParamMap(8, (4, 40), (0.5, 0.5), seed=2)
    m, obs, T = pm.base, pm.soma, 80
    w = np.random.default_rng(4).standard_normal((T, obs.size))
    with D.HipSolver(m, flags=D.F_NO_AUTOTUNE) as s:
        assert s.info()["ntiles"] > 0 and s.info()["npacks"] > 0
        s.adj_forward(T, 40, obs)
        assert s.info()["ntiles"] == 0
        gm = s.adj_backward(w)
    with solver_for(m, 40) as s:
        s.adj_forward(T, 40, obs)
        gt = s.adj_backward(w)
    for a, b in zip(gm, gt):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kw", [dict(tile_max_rows=64), dict(tile_max_rows=-1, pack_cap=128),
                                dict(tile_max_rows=16)], ids=["tiles", "packs", "mixed"])
def test_set_ravals_equals_fresh_model(gpu, kw):
    """hsh_set_ravals updates the coefficient arrays of every layout in place:
    a run after it equals the oracle's run of the model built with those
    coefficients, plain and batched."""
    pm = F.ParamMap(6, (8, 40), (0.5, 0.5), seed=3)
    th = np.ones((6, 5))
    th[:, 3] = np.linspace(0.7, 1.4, 6)
    th[:, 4] = np.linspace(1.2, 0.8, 6)
    m2 = pm.model_at(th)
    assert not np.array_equal(m2.ravals, pm.base.ravals)
    o = Oracle(m2, givals=False)
    o.step(60)
    with D.HipSolver(pm.base, flags=D.F_NO_AUTOTUNE, **kw) as s:
        s.set_ravals(m2.ravals)
        s.put_state(vm=m2.vm, chip=m2.chip, ops=m2.ops)
        s.step(30, D.PLAIN)
        s.step(30, D.BATCHED)
        assert np.array_equal(s.state()["vm"], o.vm)


def oracle_row_crossings(pm, n, j, h, T):
    """Steps at which the central-difference runs theta[n, j] +- h
    interpolate a different gate-table row in some compartment of neuron n
    (the oracle's NEWVOLT rows)."""
    rows = np.nonzero(F.tree_of_row(pm.base) == n)[0]
    tp = np.ones((pm.n_neurons, 5))
    tp[n, j] += h
    tm = np.ones((pm.n_neurons, 5))
    tm[n, j] -= h
    rp = Oracle(pm.model_at(tp), givals=False).row_trace(T, rows)
    rm = Oracle(pm.model_at(tm), givals=False).row_trace(T, rows)
    return int(np.sum(np.any(rp != rm, axis=1)))


def test_parameter_gradients_tl_and_finite_differences(gpu):
    """64-compartment trees: autograd gradient of the soma-Vm MSE against
    targets from the CPU solver at hidden parameters (+-30%).

    * every gradient equals the CPU tangent-linear gradient (1e-7);
    * central differences at the specified h = 1e-6: the model GENESIS
      integrates is piecewise linear in Vm (interpolated gate tables), so a
      difference quotient measures its derivative only when the +h and -h
      runs interpolate the same table rows at every step.  For every
      (neuron, parameter) whose runs keep identical rows -- checked with the
      oracle's row trace -- the adjoint is within 1e-5 of the quotient (plus
      its rounding floor); such pairs exist for all five parameters."""
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

    def fd(n, j, h):
        with torch.no_grad():
            tp = torch.ones(N, 5, dtype=torch.float64)
            tp[n, j] += h
            tm = torch.ones(N, 5, dtype=torch.float64)
            tm[n, j] -= h
            return (torch.mean((run(tp) - tgt) ** 2) - torch.mean((run(tm) - tgt) ** 2)).item() / (2 * h)
    # rounding floor of a difference quotient of the loss: ~eps * L / h
    L0 = loss.item()
    h = 1e-6
    floor = 10 * np.finfo(float).eps * L0 / h
    covered, crossing = set(), []
    for n in range(N):
        for j in range(5):
            cross = oracle_row_crossings(pm, n, j, h, T)
            if cross:
                crossing.append((n, j, cross))
                continue
            q = fd(n, j, h)
            assert abs(g[n, j] - q) <= 1e-5 * abs(g[n, j]) + floor, (n, F.PARAMS[j], g[n, j], q)
            covered.add(j)
    print(f"[case] crossing pairs (neuron, param, steps): {crossing}", flush=True)
    assert covered == {0, 1, 2, 3, 4}
    # the device's rows of a run (from its voltages) are the oracle's
    rows = np.nonzero(F.tree_of_row(pm.base) == 4)[0]
    assert np.array_equal(run.table_rows(np.ones((N, 5)), rows),
                          Oracle(pm.model_at(np.ones((N, 5))), givals=False).row_trace(T, rows))
    run.close()


def test_stale_backward_is_refused(gpu):
    """A second forward pass replaces the solver's checkpoints: backward for
    the first graph raises instead of returning another run's gradients."""
    pm = F.ParamMap(2, (16,), (1.0,), seed=4)
    run = F.Runner(pm, 50, ckpt=25)
    th1 = torch.ones(2, 5, dtype=torch.float64, requires_grad=True)
    l1 = torch.sum(run(th1))
    with torch.no_grad():
        run(torch.ones(2, 5, dtype=torch.float64) * 1.1)
    with pytest.raises(RuntimeError):
        l1.backward()
    th2 = torch.ones(2, 5, dtype=torch.float64, requires_grad=True)
    torch.sum(run(th2)).backward()
    assert np.all(np.isfinite(th2.grad.numpy()))
    run.close()


@pytest.mark.parametrize("layout", ["tiles", "packs"])
def test_checkpoint_interval_and_state(gpu, layout):
    pm = F.ParamMap(5, (16,), (1.0,), seed=9)
    m, obs, T = pm.base, pm.soma, 150
    print(f"[case] checkpoint intervals {layout}", flush=True)
    w = np.random.default_rng(2).standard_normal((T, obs.size))
    outs = []
    for ck in (7, 50, 150):
        with solver_for(m, 16, layout, 64) as s:
            s.adj_forward(T, ck, obs)
            after_fwd = s.state()
            outs.append(s.adj_backward(w))
            after_bwd = s.state()
            for k in ("vm", "chip", "ops"):
                assert np.array_equal(after_fwd[k], after_bwd[k]), k
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
