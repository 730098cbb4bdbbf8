"""Conductance fitting with the device adjoint.

Per-neuron parameters theta[n] = (gNa, gK, gKdr, Ra, Cm) scale factors of a
population's base values.  ``ParamMap.arrays`` maps them, in torch and
differentiably, onto the Hsolve arrays the reference's setup would produce:

* gmax of a channel          = base * s_g        (Gbar * area)
* tbyc                       = base / s_Cm       (dt / Cm, hines_data.c:91-94)
* diagterm                   = 1 + L / s_Cm + X / (s_Cm * s_Ra)
                               (1 + dt/(Cm Rm) + axial terms, hines_init.c)
* ravals (ta, tcf, tcb)      = base / (s_Cm * s_Ra)   (-dt/(Cm Ra) forms)

L and X (leak and axial parts of diagterm) and the positions of every entry
come from building the base population again with one parameter doubled.
The population is built with double-precision element fields, so every
entry is a smooth function of theta.

``Trace`` is the torch.autograd.Function: forward = libhsolve_hip run of
``nsteps`` plain steps keeping a checkpoint every ``ckpt`` steps
(hsh_adj_forward), returning vm of the observed rows after each step;
backward = the library's adjoint (hsh_adj_backward), which recomputes each
segment from its checkpoint and runs the transposed Hines sweep backward
through the channel updates.
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as D
from . import model as M

PARAMS = ("gNa", "gK", "gKdr", "Ra", "Cm")
F_SET_DIAG, F_SKIP_DIAG, F_CALC, F_FINISH = 1, 2, 6, 7


def tree_of_row(m: M.Model) -> np.ndarray:
    """Neuron (tree) index of every Hines row: trees are contiguous row
    blocks ending at their root."""
    t = np.empty(m.ncompts, np.int64)
    k = 0
    for r in range(m.ncompts):
        t[r] = k
        if m.parent_row[r] < 0:
            k += 1
    return t


def ravals_rows(m: M.Model) -> np.ndarray:
    """Row (Hines number) that owns each ravals entry: the kid row of a
    FORWARD_ELIM pair, the current row of a BACKWARD_ELIM value
    (hines_solve.c:246-295 consume them in this order)."""
    f = m.funcs
    rows = np.full(m.ravals.size, -1, np.int64)
    fi = ri = 0
    if m.ncompts > 1:
        cur = 1
        while True:
            op = int(f[fi]); fi += 1
            if op == 0:
                kid = int(f[fi]) // 2; fi += 1
                rows[ri] = rows[ri + 1] = kid
                ri += 2
            elif op == F_SET_DIAG:
                cur += 1
            elif op == F_SKIP_DIAG:
                cur += 2
            else:
                break
        row = m.ncompts - 2
        while True:
            op = int(f[fi]); fi += 1
            if op == 0:
                fi += 1
                rows[ri] = row
                ri += 1
            elif op == F_CALC:
                row -= 1
            else:
                break
    return rows


class ParamMap:
    """theta (n_neurons x 5, scale factors) -> (chip, ravals) of a population."""

    def __init__(self, n_neurons: int, sizes=(64,), probs=(1.0,), *, seed: int = 0, **popkw):
        popkw = dict(popkw, float_fields=False)
        self.base_kw = dict(gNa=1200.0, gK=360.0, gKdr=50.0, RA=1.0, CM=0.01)
        self.base_kw.update({k: v for k, v in popkw.items() if k in self.base_kw})

        def build(**over):
            kw = dict(popkw)
            kw.update(self.base_kw)
            kw.update(over)
            return M.population(n_neurons, sizes, probs, seed=seed, **kw)

        b = build()
        self.base = b
        self.n_neurons = n_neurons
        bc = b.chip
        row_of_chip = np.searchsorted(b.compchips, np.arange(bc.size), side="right") - 1
        nrn_row = tree_of_row(b)
        nrn_chip = nrn_row[row_of_chip]

        def doubled(v):
            return np.nonzero((bc != 0) & (v != bc) & (np.abs(v - 2 * bc) <= 1e-12 * np.abs(bc)))[0]

        k = self.base_kw
        self.g_idx = [doubled(build(gNa=2 * k["gNa"]).chip), doubled(build(gK=2 * k["gK"]).chip),
                      doubled(build(gKdr=2 * k["gKdr"]).chip)]
        cm = build(CM=2 * k["CM"])
        ra = build(RA=2 * k["RA"])
        self.tbyc_idx = np.nonzero((bc != 0) & (np.abs(cm.chip - 0.5 * bc) <= 1e-12 * np.abs(bc)))[0]
        changed = np.nonzero(cm.chip != bc)[0]
        self.diag_idx = np.setdiff1d(changed, self.tbyc_idx)
        X = 2.0 * (bc[self.diag_idx] - ra.chip[self.diag_idx])
        L = bc[self.diag_idx] - 1.0 - X
        self.L, self.X = L, X
        assert np.allclose(cm.ravals, 0.5 * b.ravals, rtol=1e-12, atol=0)
        assert np.allclose(ra.ravals, 0.5 * b.ravals, rtol=1e-12, atol=0)
        self.rav_nrn = torch.from_numpy(nrn_row[ravals_rows(b)])
        self.g_nrn = [torch.from_numpy(nrn_chip[i]) for i in self.g_idx]
        self.tbyc_nrn = torch.from_numpy(nrn_chip[self.tbyc_idx])
        self.diag_nrn = torch.from_numpy(nrn_chip[self.diag_idx])
        self.soma = np.nonzero(b.parent_row < 0)[0].astype(np.int32)   # roots: one per neuron

    def arrays(self, theta: torch.Tensor):
        """(chip, ravals) as float64 tensors, differentiable in theta."""
        b = self.base
        chip = torch.from_numpy(b.chip.copy())
        idx, vals = [], []
        for j, (ix, nr) in enumerate(zip(self.g_idx, self.g_nrn)):
            idx.append(torch.from_numpy(ix))
            vals.append(torch.from_numpy(b.chip[ix]) * theta[nr, j])
        scm = theta[:, 4]
        sra = theta[:, 3]
        idx.append(torch.from_numpy(self.tbyc_idx))
        vals.append(torch.from_numpy(b.chip[self.tbyc_idx]) / scm[self.tbyc_nrn])
        idx.append(torch.from_numpy(self.diag_idx))
        c = scm[self.diag_nrn]
        vals.append(1.0 + torch.from_numpy(self.L) / c + torch.from_numpy(self.X) / (c * sra[self.diag_nrn]))
        chip = chip.index_put((torch.cat(idx),), torch.cat(vals))
        rav = torch.from_numpy(b.ravals.copy()) / (scm[self.rav_nrn] * sra[self.rav_nrn])
        return chip, rav

    def model_at(self, theta) -> M.Model:
        """The base Model with the arrays for theta (for the CPU oracle)."""
        import dataclasses
        with torch.no_grad():
            chip, rav = self.arrays(torch.as_tensor(theta, dtype=torch.float64))
        return dataclasses.replace(self.base, chip=chip.numpy().copy(), ravals=rav.numpy().copy())


class Trace(torch.autograd.Function):
    """vm of the observed rows after each step of a device run, as a function
    of the chip[] and ravals[] arrays."""

    @staticmethod
    def forward(ctx, chip, ravals, runner):
        tr = runner.forward_arrays(chip.detach().cpu().numpy(), ravals.detach().cpu().numpy())
        ctx.runner = runner
        ctx.generation = runner.generation
        return torch.from_numpy(tr)

    @staticmethod
    def backward(ctx, g):
        # the solver keeps the checkpoints of its last forward run only
        if ctx.runner.generation != ctx.generation:
            raise RuntimeError("Trace.backward: the Runner ran another forward pass after this one; its "
                               "checkpoints are gone (call backward before the next forward, or use a "
                               "second Runner)")
        gc, gr, _ = ctx.runner.solver.adj_backward(g.detach().cpu().contiguous().numpy())
        return torch.from_numpy(gc), torch.from_numpy(gr), None


class Runner:
    """A device solver for a ParamMap's population that runs nsteps from the
    initial state for given arrays.

    layout "packs" (default): whole neurons in packs of about ``pack_cap``
    compartments, one workgroup of ``pack_waves`` waves each (``pack_stage``
    2: level schedule in LDS, 1: read from memory); the forward pass runs one
    batched launch per checkpoint interval and the backward pass one
    recording launch and one reverse launch per interval.  "tiles": 64 neurons per tile, one
    lane each, three launches per step.  Both give the same gradients bit for
    bit."""

    def __init__(self, pmap: ParamMap, nsteps: int, *, ckpt: int = 100, obs=None, device: int = 0,
                 layout: str = "packs", pack_cap: int = 128, pack_waves: int = 4, pack_stage: int = 2):
        self.pmap = pmap
        self.nsteps = nsteps
        self.ckpt = ckpt
        self.obs = pmap.soma if obs is None else np.ascontiguousarray(obs, np.int32)
        b = pmap.base
        rows = int(np.max(np.bincount(tree_of_row(b))))
        if layout == "packs":
            opts = dict(tile_max_rows=-1, pack_cap=max(pack_cap, rows), pack_waves=pack_waves,
                        pack_stage=pack_stage)
        elif layout == "tiles":
            opts = dict(tile_max_rows=rows)
        else:
            raise ValueError(f"Runner: layout {layout!r} is not 'packs' or 'tiles'")
        self.solver = D.HipSolver(b, device=device, flags=D.F_NO_AUTOTUNE, **opts)
        # forward runs so far: a backward pass must belong to the latest one
        self.generation = 0

    def forward_arrays(self, chip, ravals):
        b = self.pmap.base
        s = self.solver
        s.put_state(vm=b.vm, chip=chip, ops=b.ops)
        s.set_ravals(ravals)
        self.generation += 1
        return s.adj_forward(self.nsteps, self.ckpt, self.obs)

    def table_rows(self, theta, rows):
        """The gate-table row (filo, hines_chip.c:174-190) every compartment
        in ``rows`` interpolates at each step of a run at theta: the row of
        its voltage at the start of the step, clamp(int((Vm - xmin)*invdx),
        0, xdivs).  Shape (nsteps, len(rows)).  Invalidates a pending
        backward pass (the run replaces the checkpoints)."""
        b = self.pmap.base
        rows = np.ascontiguousarray(rows, np.int32)
        with torch.no_grad():
            chip, rav = self.pmap.arrays(torch.as_tensor(theta, dtype=torch.float64))
        s = self.solver
        s.put_state(vm=b.vm, chip=chip.numpy(), ops=b.ops)
        s.set_ravals(rav.numpy())
        self.generation += 1
        tr = s.adj_forward(self.nsteps, self.ckpt, rows)
        V = np.vstack([b.vm[rows][None, :], tr[:-1]])
        return np.clip(np.floor((V - b.xmin) * b.invdx), 0, b.xdivs).astype(np.int32)

    def __call__(self, theta: torch.Tensor) -> torch.Tensor:
        chip, rav = self.pmap.arrays(theta)
        return Trace.apply(chip, rav, self)

    def close(self):
        self.solver.close()
