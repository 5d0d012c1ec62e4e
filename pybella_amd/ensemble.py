"""Ensemble time stepping on the GPU: the members of one grid stepped
together, their pressure solves batched.

The reference runs its ensemble members one after the other, each through the
whole time-step driver (src/__main__.py:223-254: ``dis_time_update.do`` per
member).  Here every member keeps its own device state and operator, the
members advance in lockstep (each with its own dt and time, exactly as its own
``do`` would), and at each of a step's two projections the solves of all
members run as one batched solve (pb_batch_solve: one hipGraph per chunk of
iterations with a branch per member).  Every member's result equals its
single-member GPU step with the same solver mode, bit for bit.
"""
from __future__ import annotations

import copy

from . import second_projection as sp
from . import time_update as tu
from .batch import BatchSolver


class Ensemble:
    """``members``: a list of (Sol, flux, mpv) host objects on one grid
    (elem, node, ud, th).  ``mode``: the batched solver mode, "f64fast"
    (FP64) or "mixed"."""

    def __init__(self, members, elem, node, ud, th, mode="f64fast"):
        if mode not in ("f64fast", "mixed"):
            raise ValueError("batched solves support the modes 'f64fast' and 'mixed'")
        self.elem, self.node, self.th = elem, node, th
        self.ud = ud
        self.mode = mode
        grid = sp.grid_of(node, ud)
        self.models = [tu.DeviceModel(elem, node, ud, th, op=sp.Operator2D(grid)) for _ in members]
        self.batch = BatchSolver([M.op for M in self.models])
        for M, (Sol, flux, mpv) in zip(self.models, members):
            M.upload(Sol, flux, mpv)

    def download(self, members):
        for M, (Sol, flux, mpv) in zip(self.models, members):
            M.download(Sol, flux, mpv)

    def _solve(self, prs):
        ud = self.ud
        xs, infos = self.batch.solve([pr.b for pr in prs], tol=float(ud.tol), maxiter=int(ud.max_iterations),
                                     mode=self.mode, inner_tol=sp.solver_options.get("inner_tol", 1e-3),
                                     max_outer=sp.solver_options.get("max_outer", 12),
                                     check_every=sp.solver_options.get("check_every", 32))
        for info in infos:
            sp._count(info)
        return xs, infos

    def step(self, ts, tout, step_no, window_step):
        """One time step of every member (time_update.py:121-690, member m
        at time ts[m]).  Returns the per-member (dt, [info of both solves])."""
        begin = [tu.step_begin(M, t, tout, step_no, window_step) for M, t in zip(self.models, ts)]
        xs, infos1 = self._solve([b[3] for b in begin])
        prs2 = [tu.step_middle(M, b[0], step_no, b[3], x) for M, b, x in zip(self.models, begin, xs)]
        xs2, infos2 = self._solve(prs2)
        for pr, x in zip(prs2, xs2):
            sp._project_update(pr, x)
        return [(b[0], [i1, i2]) for b, i1, i2 in zip(begin, infos1, infos2)]

    def run(self, t, tout, steps, nsteps):
        """``nsteps`` steps of every member from time t (all members start at
        t); returns the per-member times and the per-step records."""
        window_step, step_no = steps
        ts = [t] * len(self.models)
        recs = []
        for _ in range(nsteps):
            r = self.step(ts, tout, step_no, window_step)
            ts = [tm + dt for tm, (dt, _) in zip(ts, r)]
            recs.append(r)
            step_no += 1
            window_step += 1
        return ts, recs


def rising_bubble_members(count, nx=161, ny=81, base_seed=20261204):
    """Rising-bubble ensemble members on one grid (inputs/rising_bubble.py
    sol_init), member m with delth ~ U(1, 3) K, xc ~ U(-3, 3) km and
    r0 ~ U(1.5, 2.5) km from numpy default_rng(base_seed + m).  Returns
    ([(Sol, flux, mpv)], elem, node, ud, th)."""
    import numpy as np
    from . import workloads as W
    ud = W.rising_bubble_ud(nx, ny)
    members = []
    for m in range(count):
        rng = np.random.default_rng(base_seed + m)
        delth, xc, r0 = rng.uniform(1.0, 3.0), rng.uniform(-0.3, 0.3), rng.uniform(0.15, 0.25)
        Sol, flux, mpv, elem, node, ud_m, th = W.rising_bubble_init(copy.deepcopy(ud), delth=delth, xc=xc, r0=r0)
        members.append((Sol, flux, mpv))
    return members, elem, node, ud_m, th
