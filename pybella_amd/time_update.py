"""GPU time step of pyBELLA's 2D vertical slice, one ensemble member.

``do`` keeps the signature and return value of the reference's time-step
driver (src/dycore/discretisation/time_update.py:59-708): it takes the host
``Sol, flux, mpv`` objects, runs every step on the GPU with the state resident
in device memory, and writes the state back into them at the end.  Between
the first upload and the last download the host only reads the CFL maxima (one
48-byte copy per step) and the solver's convergence state.

The functions below mirror the reference's explicit-part functions one to
one, each a thin wrapper over one C-ABI entry point (include/pybella_amd.h):

  recompute_advective_fluxes   numerical_flux.py:5-48      pb_advective_fluxes
  explicit_step_and_flux       explicit.py:75-165          pb_advect_split
  advect                       explicit.py:5-72            (composition)
  advect_rk                    explicit.py:168-257         pb_advect_split,
                                                           pb_advect_rk_update
  euler_backward_non_advective_expl_part
                               second_projection.py:72-90  pb_expl_part
  euler_forward_non_advective  second_projection.py:22-69  pb_euler_forward
  dynamic_timestep             cfl.py:5-74                 pb_cfl_maxima
  the implicit projection      second_projection.py:97-277 (second_projection)

Scope: the reference's path for one member as src/__main__.py runs it with
one member, i.e. blending off (the Blend object it passes is inert without
continuous or initial blending, and blending.py reduces to the eos regime
functions), the non-hydrostatic regime, no Rayleigh damping or forcing
(time_update.py:296-348, 478-524) and no writer; do() raises
NotImplementedError outside that.  There is no CPU
path: without the built library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from . import second_projection as sp

FIELDS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")
_EPS = float(np.finfo(float).eps)


def _torch():
    import torch
    return torch


def _ptrs(ts):
    return (_abi.VP * 6)(*[_abi.VP(t.data_ptr()) for t in ts])


class DeviceModel:
    """The device-resident state of one member (cells [y][x], x faces
    [y][face], y faces [face][x], nodes [y][x]) and the work arrays of a
    step."""

    def __init__(self, elem, node, ud, th, op=None):
        torch = _torch()
        self.elem, self.node, self.ud, self.th = elem, node, ud, th
        self.grid = sp.grid_of(node, ud)
        self.op = op if op is not None else sp.operator_for(node, ud)
        self.ctx = self.op.ctx
        nx, ny = self.grid.nx, self.grid.ny
        cx, cy = nx + 3, ny + 3
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")
        self.sol = sp._DevState(tensors=[z(cy, cx) for _ in FIELDS])
        self.sol0 = sp._DevState(tensors=[z(cy, cx) for _ in FIELDS])
        self.half = [z(cy, cx) for _ in FIELDS]
        self.fx = [z(cy, cx + 1) for _ in FIELDS]
        self.fy = [z(cy + 1, cx) for _ in FIELDS]
        self.p2n, self.p2n0, self.p2_half = z(ny + 4, nx + 4), z(ny + 4, nx + 4), z(ny + 4, nx + 4)
        self.rhs = z(ny + 2, nx + 2)
        self.wplus0, self.wplus1, self.nu_c = z(cy, cx), z(cy, cx), z(cy, cx)
        self.wcenter, self.diag_inv = z(ny + 2, nx + 2), z(ny + 2, nx + 2)
        self.u, self.v, self.w = z(cy, cx), z(cy, cx), z(cy, cx)
        self.b = z(ny, nx)
        self.dS = z(cy, cx)
        self.S0c = z(cy)
        self.rhoY0 = z(cy)
        self.cfl_work = z(int(self.ctx.lib.pb_cfl_work_words(C.byref(self.grid))))
        self.infos = []

    # ---- host <-> device -------------------------------------------------
    def upload(self, Sol, flux, mpv):
        """Host objects (reference layout) -> device state."""
        torch = _torch()
        host = [getattr(Sol, k) for k in FIELDS]
        host += [np.ascontiguousarray(getattr(flux[0], k).T) for k in FIELDS]   # [y][face] as it lies
        host += [getattr(flux[1], k) for k in FIELDS]
        host += [mpv.p2_nodes, mpv.p2_nodes0, sp.dSdy_cells(mpv, self.elem, self.node)]
        dev = sp._staging.upload(host)
        for t, d in zip(list(self.sol.f.values()) + self.fx + self.fy + [self.p2n, self.p2n0, self.dS], dev):
            t.copy_(d)
        self.S0c.copy_(torch.from_numpy(np.ascontiguousarray(mpv.HydroState.S0, dtype=np.float64)))
        self.rhoY0.copy_(torch.from_numpy(np.ascontiguousarray(mpv.HydroState.rhoY0, dtype=np.float64)))

    def download(self, Sol, flux, mpv, halves=True):
        """Device state -> the host objects, as the reference leaves them:
        Sol's fields (and its *_half copies), both flux containers, p2_nodes,
        p2_nodes0 (and p2_nodes_half), rhs and the last projection's wplus,
        wcenter, nu_c, u, v, w."""
        pairs = [(t, getattr(Sol, k)) for k, t in self.sol.f.items()]
        pairs += [(t.t(), getattr(flux[0], k)) for k, t in zip(FIELDS, self.fx)]
        pairs += [(t, getattr(flux[1], k)) for k, t in zip(FIELDS, self.fy)]
        pairs += [(self.p2n, mpv.p2_nodes), (self.p2n0, mpv.p2_nodes0), (self.rhs, mpv.rhs),
                  (self.wplus0, mpv.wplus[0]), (self.wplus1, mpv.wplus[1]), (self.wcenter, None), (self.nu_c, None),
                  (self.u, None), (self.v, None), (self.w, None)]
        if halves:
            pairs += [(t, None) for t in self.half] + [(self.p2_half, None)]
        res = sp._staging.download(pairs)
        n0 = 6 + 12
        mpv.wcenter, mpv.nu_c, mpv.u, mpv.v, mpv.w = res[n0 + 5:n0 + 10]
        if halves:
            for k, a in zip(FIELDS, res[n0 + 10:n0 + 16]):
                setattr(Sol, k + "_half", a)
            mpv.p2_nodes_half = res[n0 + 16]

    # ---- helpers -----------------------------------------------------------
    def phys(self):
        return sp.phys_of(self.ud, self.th)

    def cells(self):
        return self.sol.ptrs()

    def sync(self):
        self.ctx.sync_stream()


# ---------------------------------------------------------------------------
# the reference's explicit-part functions
# ---------------------------------------------------------------------------
def recompute_advective_fluxes(M: DeviceModel):
    """numerical_flux.py:5-48 into M.fx[4] / M.fy[4] (the rhoY entries)."""
    M.sync()
    F, P = M.sol.f, _abi.ptr
    M.ctx.check(M.ctx.lib.pb_advective_fluxes(M.ctx.h, C.byref(M.grid), P(F["rho"]), P(F["rhou"]), P(F["rhov"]),
                                              P(F["rhoY"]), P(M.fx[4]), P(M.fy[4])), "pb_advective_fluxes")


def explicit_step_and_flux(M: DeviceModel, split, lmbda, tag=None, phys=None):
    """explicit.py:75-165 along axis ``split``."""
    M.sync()
    flux = M.fx if split == 0 else M.fy
    M.ctx.check(M.ctx.lib.pb_advect_split(M.ctx.h, C.byref(M.grid), C.byref(phys or M.phys()), int(split),
                                          float(lmbda), 1 if tag == "rk" else 0, M.cells(), _ptrs(flux),
                                          _abi.ptr(M.rhoY0)), "pb_advect_split")


def advect(M: DeviceModel, dt, odd):
    """explicit.py:5-72: double Strang sweep, x y y x on odd steps and
    y x x y on even ones, each split a half step dt/2; then every ghost cell."""
    time_step = 0.5 * dt
    phys = M.phys()
    order = (0, 1) if odd else (1, 0)
    dxyz = M.elem.dxyz
    for split in order + order[::-1]:
        explicit_step_and_flux(M, split, time_step / dxyz[split], phys=phys)
    M.ctx.check(M.ctx.lib.pb_bc_cells(M.ctx.h, C.byref(M.grid), C.byref(phys), M.cells(), _abi.ptr(M.rhoY0)),
                "pb_bc_cells")


def advect_rk(M: DeviceModel, dt, odd):
    """explicit.py:168-257: both directions' fluxes from the same state, then
    the x and y updates and every ghost cell."""
    phys = M.phys()
    dxyz = M.elem.dxyz
    for split in (0, 1):
        explicit_step_and_flux(M, split, dt / dxyz[split], tag="rk", phys=phys)
    M.ctx.check(M.ctx.lib.pb_advect_rk_update(M.ctx.h, C.byref(M.grid), C.byref(phys), float(dt / dxyz[0]),
                                              float(dt / dxyz[1]), M.cells(), _ptrs(M.fx), _ptrs(M.fy),
                                              _abi.ptr(M.rhoY0)), "pb_advect_rk_update")


def euler_backward_non_advective_expl_part(M: DeviceModel, dt):
    """second_projection.py:72-90."""
    M.sync()
    M.ctx.check(M.ctx.lib.pb_expl_part(M.ctx.h, C.byref(M.grid), C.byref(M.phys()), float(dt), M.cells(),
                                       _abi.ptr(M.dS), _abi.ptr(M.rhoY0)), "pb_expl_part")


def euler_forward_non_advective(M: DeviceModel, dt):
    """second_projection.py:22-69 (mpv.rhs receives the wall-scaled nodal
    divergence, p2_nodes the compressible Exner increment)."""
    M.sync()
    M.ctx.check(M.ctx.lib.pb_euler_forward(M.ctx.h, C.byref(M.grid), C.byref(M.phys()), float(dt), M.cells(),
                                           _abi.ptr(M.p2n), _abi.ptr(M.rhs), _abi.ptr(M.S0c), _abi.ptr(M.dS),
                                           _abi.ptr(M.rhoY0)), "pb_euler_forward")


def cfl_maxima(M: DeviceModel):
    """max |u|, |v|, |w|, |u|+c, |v|+c, |w|+c over every cell (device)."""
    M.sync()
    out = (C.c_double * 6)()
    F, P = M.sol.f, _abi.ptr
    M.ctx.check(M.ctx.lib.pb_cfl_maxima(M.ctx.h, C.byref(M.grid), float(M.th.gamm), float(1.0 / np.sqrt(M.ud.Msq)),
                                        P(F["rho"]), P(F["rhou"]), P(F["rhov"]), P(F["rhow"]), P(F["rhoY"]),
                                        P(M.cfl_work), out), "pb_cfl_maxima")
    return list(out)


def dynamic_timestep(M: DeviceModel, time, time_output, step):
    """cfl.py:5-74 (the acoustic_timestep == 0 branch): (dt, cfl, cfl_ac)."""
    ud, elem = M.ud, M.elem
    if getattr(ud, "acoustic_timestep", 0) == 1:
        raise NotImplementedError("acoustic_timestep == 1: the reference's branch returns dt alone and its caller "
                                  "(time_update.py:126) cannot unpack it")
    u, v, w, upc, vpc, wpc = cfl_maxima(M)
    u_max, v_max, w_max = max(u, _EPS), max(v, _EPS), max(w, _EPS)
    upc, vpc, wpc = max(upc, _EPS), max(vpc, _EPS), max(wpc, _EPS)
    CFL = ud.CFL
    dt_cfl = min(min(CFL * elem.dx / u_max, CFL * elem.dy / v_max), CFL * elem.dz / w_max)
    dt = min(dt_cfl, ud.dtfixed0 + min(step, 1.0) * (ud.dtfixed - ud.dtfixed0))
    dt *= min(float(step + 1), 1.0)
    if dt > time_output - time:
        dt = time_output - time
    cfl = CFL * dt / dt_cfl
    cfl_ac = max(dt * upc / elem.dx, dt * vpc / elem.dy, dt * wpc / elem.dz)
    return dt, cfl, cfl_ac


def project_prepare(M: DeviceModel, dt, with_sol0):
    """The part of the projection before the solve (second_projection.py:
    113-214) on the device state: returns the prepared _Projection (its
    right-hand side in M.b)."""
    pr = sp._Projection()
    pr.grid, pr.phys, pr.op, pr.ctx, pr.dt = M.grid, M.phys(), M.op, M.ctx, dt
    pr.compressible = int(M.ud.is_compressible)
    pr.compressibility = float(M.ud.compressibility)
    pr.dS, pr.rhoY0, pr.p2n = M.dS, M.rhoY0, M.p2n
    pr.sol = M.sol
    pr.sol0 = M.sol0 if with_sol0 else None
    pr.u, pr.v, pr.w, pr.rhs = M.u, M.v, M.w, M.rhs
    pr.wplus0, pr.wplus1, pr.nu_c = M.wplus0, M.wplus1, M.nu_c
    pr.wcenter, pr.diag_inv, pr.b = M.wcenter, M.diag_inv, M.b
    sp._project_prepare(pr)
    return pr


def project(M: DeviceModel, dt, with_sol0):
    """euler_backward_non_advective_impl_part (second_projection.py:97-277)
    on the device state, Sol0 = the state at the start of the step when
    ``with_sol0``.  Returns the solver info dict."""
    pr = project_prepare(M, dt, with_sol0)
    x, info = sp._solve(pr, M.ud)
    sp._count(info)
    sp._project_update(pr, x)
    return info


# eos.py:5-66 with continuous_blending off
def _compressibility(ud, step):
    if step >= 0:
        return 0.0 if ud.is_compressible == 0 else 1.0
    return ud.compressibility


def _nonhydrostasy(ud, step):
    return float(ud.is_nonhydrostatic)


def step_begin(M: DeviceModel, t, tout, step_no, window_step):
    """time_update.py:121-250 up to the first projection's solve: ghost
    cells, dt, the regime functions, Sol0, the advective fluxes, the RK half
    step and the explicit part.  Returns (dt, cfl, cfl_ac, prepared
    projection)."""
    ud = M.ud
    F = M.sol.f
    M.sync()
    M.ctx.check(M.ctx.lib.pb_bc_cells(M.ctx.h, C.byref(M.grid), C.byref(M.phys()), M.cells(), _abi.ptr(M.rhoY0)),
                "pb_bc_cells")
    dt, cfl, cfl_ac = dynamic_timestep(M, t, tout, step_no)
    if "CFLfixed" in getattr(ud, "aux", ""):
        if step_no < 2:
            dt = 21.69 / ud.t_ref
    ud.compressibility = _compressibility(ud, window_step)
    ud.nonhydrostasy = _nonhydrostasy(ud, window_step)
    for d, s in zip(M.sol0.f.values(), F.values()):
        d.copy_(s)
    recompute_advective_fluxes(M)
    if getattr(ud, "do_advection", True):
        advect_rk(M, 0.5 * dt, step_no % 2)
    M.p2n0.copy_(M.p2n)
    euler_backward_non_advective_expl_part(M, 0.5 * dt)
    return dt, cfl, cfl_ac, project_prepare(M, 0.5 * dt, with_sol0=(ud.is_compressible == 0))


def step_middle(M: DeviceModel, dt, step_no, pr, x):
    """time_update.py:251-462 after the first solve: the projection's update,
    the half-step fluxes and copies, the restart from Sol0, the explicit
    forward step, the Strang advection and the explicit part.  Returns the
    second projection, prepared."""
    ud = M.ud
    F = M.sol.f
    sp._project_update(pr, x)
    recompute_advective_fluxes(M)
    for d, s in zip(M.half, F.values()):
        d.copy_(s)
    M.p2_half.copy_(M.p2n)
    if ud.is_compressible == 1 and ud.is_nonhydrostatic == 1:
        M.p2n.copy_(M.p2n0)
        for d, s in zip(F.values(), M.sol0.f.values()):
            d.copy_(s)
    elif ud.is_nonhydrostatic == 0:
        for d, s in zip(F.values(), M.sol0.f.values()):
            d.copy_(s)
        M.p2n.copy_(M.p2n0)
    elif ud.is_compressible == 0:
        for d, s in zip(F.values(), M.sol0.f.values()):
            d.copy_(s)
    euler_forward_non_advective(M, 0.5 * dt)
    if getattr(ud, "do_advection", True):
        advect(M, dt, step_no % 2)
    euler_backward_non_advective_expl_part(M, 0.5 * dt)
    return project_prepare(M, 0.5 * dt, with_sol0=False)


def step(M: DeviceModel, t, tout, step_no, window_step):
    """One pass of the loop body of time_update.py:121-690.  Returns
    (dt, cfl, cfl_ac, [info of both solves])."""
    dt, cfl, cfl_ac, pr = step_begin(M, t, tout, step_no, window_step)
    x, info1 = sp._solve(pr, M.ud)
    sp._count(info1)
    pr2 = step_middle(M, dt, step_no, pr, x)
    x2, info2 = sp._solve(pr2, M.ud)
    sp._count(info2)
    sp._project_update(pr2, x2)
    return dt, cfl, cfl_ac, [info1, info2]


_models: dict = {}


def model_for(elem, node, ud, th) -> DeviceModel:
    """The cached DeviceModel of a grid (the ud and th of the latest call)."""
    g = sp.grid_of(node, ud)
    key = (g.nx, g.ny, g.dx, g.dy, g.bc_x, g.bc_y)
    M = _models.get(key)
    if M is None:
        M = DeviceModel(elem, node, ud, th)
        _models[key] = M
    M.elem, M.node, M.ud, M.th = elem, node, ud, th
    return M


def do(Sol, flux, mpv, t, tout, ud, elem, node, steps, th, bld=None, writer=None, debug=False):
    """time_update.py:59-708 for one member on the GPU: steps from t until
    tout or ud.stepmax, updating Sol, flux and mpv as the reference does.
    Returns [Sol, flux, mpv, [window_step, step]]."""
    blending = getattr(ud, "continuous_blending", False) or getattr(ud, "initial_blending", False)
    if bld is not None:   # blending.Blend: active only with continuous blending (blending.py:16-32)
        blending = blending or getattr(bld, "cb", False) or getattr(bld, "bb", False)
    if blending:
        raise NotImplementedError("regime blending (data_assimilation/blending.py) is not part of this path")
    if getattr(ud, "is_nonhydrostatic", 1) != 1:
        raise NotImplementedError("the hydrostatic regime (time_update.py:137-152, 232-255) is not part of this path")
    if writer is not None or debug:
        raise NotImplementedError("the HDF5 writer is not part of this path")
    if ud.bdry_type[1].name == "RAYLEIGH" or getattr(ud, "rayleigh_forcing", False):
        raise NotImplementedError("Rayleigh damping / forcing is not part of this path")
    window_step, step_no = steps
    M = model_for(elem, node, ud, th)
    M.upload(Sol, flux, mpv)
    M.infos = []
    ran = False
    while t < tout and step_no < ud.stepmax:
        dt, cfl, cfl_ac, infos = step(M, t, tout, step_no, window_step)
        M.infos.append(dict(step=step_no, t=t, dt=dt, cfl=cfl, cfl_ac=cfl_ac, solves=infos))
        t += dt
        step_no += 1
        window_step += 1
        ran = True
    M.download(Sol, flux, mpv, halves=ran)
    return [Sol, flux, mpv, [window_step, step_no]]


def steps_per_second(M: DeviceModel, t, tout, n, step0=0):
    """Wall-clock time steps per second of ``n`` device-resident steps (the
    state stays on the GPU).  Returns (steps/s, last t, per-step records)."""
    import time
    torch = _torch()
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        dt, cfl, cfl_ac, infos = step(M, t, tout, step0 + k, step0 + k)
        recs.append(dict(dt=dt, iters=[i["iters"] for i in infos]))
        t += dt
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return (n / el if el > 0 else math.inf), t, recs
