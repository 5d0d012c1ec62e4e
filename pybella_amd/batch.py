"""Batched uneven 2D solves (pb_batch_*): many assembled operators, possibly
of different grid sizes, solved together with one hipGraph per chunk of
iterations (one branch per member).  No reference counterpart: the reference
solves ensemble members one after the other (src/__main__.py:223-254)."""
from __future__ import annotations

import ctypes as C

from . import _abi
from .second_projection import _MODES


class BatchSolver:
    def __init__(self, ops, ctx: _abi.Context | None = None):
        self.ctx = ctx or ops[0].ctx
        self.ops = list(ops)
        arr = (_abi.VP * len(self.ops))(*[op.h for op in self.ops])
        h = _abi.VP()
        self.ctx.check(self.ctx.lib.pb_batch_create(self.ctx.h, arr, len(self.ops), C.byref(h)), "pb_batch_create")
        self.h = h

    def solve(self, bs, xs=None, tol=1e-8, maxiter=6000, mode="mixed", inner_tol=1e-3, max_outer=12,
              check_every=32):
        """bs: per-member FP64 device Krylov vectors (ny, nx).  mode "mixed" or
        "f64fast".  Returns (xs, infos)."""
        import torch
        assert len(bs) == len(self.ops)
        if xs is None:
            xs = [torch.empty(op.ny, op.nx, dtype=torch.float64, device="cuda") for op in self.ops]
        bp = (_abi.VP * len(bs))(*[_abi.VP(b.data_ptr()) for b in bs])
        xp = (_abi.VP * len(xs))(*[_abi.VP(x.data_ptr()) for x in xs])
        opts = _abi.SolveOpts(_MODES[mode], float(tol), int(maxiter), float(inner_tol), int(max_outer),
                              int(check_every), 0, 0, 0, -1, 0.0, 0)
        infos = (_abi.SolveInfo * len(bs))()
        self.ctx.sync_stream()
        rc = self.ctx.lib.pb_batch_solve(self.h, bp, xp, C.byref(opts), infos)
        self.ctx.check(rc, "pb_batch_solve")
        return xs, [i.as_dict() for i in infos]

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.pb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
