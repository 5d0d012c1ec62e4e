"""Host side of the 3D nodal 27-point operator (stencil_27pt / lap3D,
laplacian.py:872-1026) and its solve.

Device layout: nodes (nz, ny, nx) and operator cells (nz+1, ny+1, nx+1), x
fastest -- the layout of ``workloads.State3D`` arrays, uploaded as they are.
"""
from __future__ import annotations

import ctypes as C

from . import _abi
from .second_projection import _MODES


def _torch():
    import torch
    return torch


def grid3_of(st) -> _abi.Grid3D:
    bc = lambda periodic: _abi.PB_BC_PERIODIC if periodic else _abi.PB_BC_WALL
    return _abi.Grid3D(st.nx, st.ny, st.nz, st.dx, st.dy, st.dz, bc(st.periodic[0]), bc(st.periodic[1]),
                       bc(st.periodic[2]))


def phys3_of(st) -> _abi.Phys:
    ud, th = st.ud, st.th
    return _abi.Phys(th.gamm, ud.Msq, ud.gravity_strength[1], (C.c_double * 3)(*ud.coriolis_strength),
                     ud.nonhydrostasy, int(ud.is_compressible), float(ud.compressibility),
                     int(ud.is_ArakawaKonor))


class Operator3D:
    """A device 27-point operator handle (pb_op3d) with its solver workspace."""

    def __init__(self, grid: _abi.Grid3D, ctx: _abi.Context | None = None):
        self.ctx = ctx or _abi.default_context()
        self.grid = grid
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        h = _abi.VP()
        self.ctx.check(self.ctx.lib.pb_op3d_create(self.ctx.h, C.byref(grid), C.byref(h)), "pb_op3d_create")
        self.h = h

    @classmethod
    def from_state(cls, st, dt=0.05):
        """Operator assembled on a workloads.State3D."""
        torch = _torch()
        op = cls(grid3_of(st))
        rho = torch.from_numpy(st.rho_c).to("cuda")
        rhoY = torch.from_numpy(st.rhoY_c).to("cuda")
        op.assemble(phys3_of(st), dt, rho, rhoY)
        return op

    def assemble(self, phys, dt, rho_c, rhoY_c, wcenter=None, diag_inv=None):
        P = _abi.ptr
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_op3d_assemble(self.h, C.byref(phys), float(dt), P(rho_c), P(rhoY_c),
                                                     P(wcenter), P(diag_inv)), "pb_op3d_assemble")

    def matvec(self, x, y=None):
        torch = _torch()
        if y is None:
            y = torch.empty_like(x)
        prec = _abi.PB_F64 if x.dtype == torch.float64 else _abi.PB_F32
        assert x.numel() == self.nx * self.ny * self.nz and y.dtype == x.dtype
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_op3d_matvec(self.h, _abi.ptr(x), _abi.ptr(y), prec), "pb_op3d_matvec")
        return y

    def solve(self, b, x=None, tol=1e-8, maxiter=6000, mode="mixed", inner_tol=1e-5, max_outer=20,
              check_every=32):
        """mode "f64" (reference legacy tolerance) or "mixed" (FP32 inner
        solves in FP64 refinement to true relative residual ``tol``)."""
        torch = _torch()
        if x is None:
            x = torch.empty(self.nz, self.ny, self.nx, dtype=torch.float64, device="cuda")
        opts = _abi.SolveOpts(_MODES[mode], float(tol), int(maxiter), float(inner_tol), int(max_outer),
                              int(check_every), 0, 0, 0, -1, 0.0, 0)
        info = _abi.SolveInfo()
        self.ctx.sync_stream()
        rc = self.ctx.lib.pb_op3d_solve(self.h, _abi.ptr(b), _abi.ptr(x), C.byref(opts), C.byref(info))
        if rc not in (-10, -11):
            self.ctx.check(rc, "pb_op3d_solve")
        return x, info.as_dict()

    def time_kernel(self, kernel, precision="f32", reps=100):
        us = C.c_double()
        prec = _abi.PB_F32 if precision == "f32" else _abi.PB_F64
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_op3d_time_kernel(self.h, int(kernel), prec, int(reps), C.byref(us)),
                       "pb_op3d_time_kernel")
        return us.value

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.pb_op3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
