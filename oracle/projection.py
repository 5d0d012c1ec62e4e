"""Second-projection pieces around the solve (oracle; test infrastructure only).

Restates src/dycore/physics/low_mach/second_projection.py of pyBELLA for the 2D
vertical slice, following the reference's operation order:
  operator_coefficients_nodes  :308-356     scale_wall_node_values  :433-464
  grad_nodes                   :496-521     divergence_nodes        :525-575
  rhs_from_p_old               :578-601     multiply_inverse_coriolis :604-652
  correction_nodes             :280-306
  euler_backward_non_advective_impl_part :97-277 (2D branch)
The solve itself uses ``oracle.bicgstab.bicgstab`` on ``oracle.laplacian.lap2d``.
"""
import numpy as np
from scipy import signal

from . import boundary as bdry
from . import laplacian as lp
from .bicgstab import bicgstab


def _wall(bt):
    return getattr(bt, "name", str(bt)) in ("WALL", "RAYLEIGH")


def _periodic(bt):
    return getattr(bt, "name", str(bt)) == "PERIODIC"


def dSdy_cells(mpv, elem, node):
    """variable.py:162-177: (S0[j+1]-S0[j])/dy repeated over x (cells)."""
    S0 = mpv.HydroState_n.S0
    d = signal.convolve(S0, [1.0, -1.0], mode="valid") / node.dy
    return np.repeat(d[None, :], elem.sc[0], axis=0)


def scale_wall_node_values(rhs, node, ud, factor=0.5):
    """second_projection.py:433-464: scale node rows next to walls; other axes
    are restricted to [igs:-igs], so periodic end nodes are not scaled."""
    ndim = node.ndim
    igs = node.igs
    wall_idx = [slice(igs[d], -igs[d]) for d in range(ndim)]
    for dim in range(ndim):
        if _wall(ud.bdry_type[dim]):
            for direction in (-1, 1):
                wall_idx[dim] = (igs[dim] - 1) * direction
                if direction == -1:
                    wall_idx[dim] -= 1
                rhs[tuple(wall_idx)] *= factor


def operator_coefficients_nodes(elem, node, Sol, mpv, ud, th, dt):
    """second_projection.py:308-356."""
    g = ud.gravity_strength[1]
    Msq = ud.Msq
    Gammainv = th.Gammainv
    ndim = node.ndim
    ccenter = -ud.Msq * th.gm1inv / (dt ** 2)
    cexp = 2.0 - th.gamm
    strat = dSdy_cells(mpv, elem, node)
    Y = Sol.rhoY / Sol.rho
    coeff = Gammainv * Sol.rhoY * Y
    nu = -dt ** 2 * (g / Msq) * strat * Y
    mpv.nu_c = nu
    for dim in range(ndim):
        mpv.wplus[dim][...] = coeff
    kernel = np.ones([2] * ndim)
    mpv.wcenter = ccenter * signal.fftconvolve(Sol.rhoY ** cexp, kernel, mode="valid") / kernel.sum()
    if not hasattr(ud, "LAMB_BDRY"):
        scale_wall_node_values(mpv.wcenter, node, ud)


def multiply_inverse_coriolis(Vec, Sol, mpv, ud, elem, node, dt, attrs=("rhou", "rhov", "rhow"), get_coeffs=False):
    """second_projection.py:604-652 (in-place on Vec's attributes)."""
    nonhydro = ud.nonhydrostasy
    g = ud.gravity_strength[1]
    Msq = ud.Msq
    wh1, wv, wh2 = dt * ud.coriolis_strength
    strat = dSdy_cells(mpv, elem, node)
    Y = Sol.rhoY / Sol.rho
    nu = -dt ** 2 * (g / Msq) * strat * Y
    denom = 1.0 / (wh1 ** 2 + wh2 ** 2 + (nu + nonhydro) * (wv ** 2 + 1.0))
    coeff_uu = wh1 ** 2 + nu + nonhydro
    coeff_uv = nonhydro * (wh1 * wv + wh2)
    coeff_uw = wh1 * wh2 - (nu + nonhydro) * wv
    coeff_vu = wh1 * wv - wh2
    coeff_vv = nonhydro * (1 + wv ** 2)
    coeff_vw = wh2 * wv + wh1
    coeff_wu = wh1 * wh2 + (nu + nonhydro) * wv
    coeff_wv = nonhydro * (wh2 * wv - wh1)
    coeff_ww = nu + nonhydro + wh2 ** 2
    U0 = getattr(Vec, attrs[0])
    V0 = getattr(Vec, attrs[1])
    W0 = getattr(Vec, attrs[2])
    U = denom * (coeff_uu * U0 + coeff_uv * V0 + coeff_uw * W0)
    V = denom * (coeff_vu * U0 + coeff_vv * V0 + coeff_vw * W0)
    W = denom * (coeff_wu * U0 + coeff_wv * V0 + coeff_ww * W0)
    U0[...] = U
    V0[...] = V
    W0[...] = W
    if get_coeffs:
        return ((coeff_uu * denom).T, (coeff_vv * denom).T, (coeff_uv * denom).T, (coeff_vu * denom).T)


def grad_nodes(p, ndim, dxy):
    """second_projection.py:496-521 (2D): cell gradients from corner nodes."""
    dx, dy, dz = dxy
    a00, a01, a10, a11 = p[:-1, :-1], p[:-1, 1:], p[1:, :-1], p[1:, 1:]
    Dpx = 0.0
    Dpy = 0.0
    Dpz = 0.0
    for sx, sy, a in ((-1.0, -1.0, a00), (-1.0, 1.0, a01), (1.0, -1.0, a10), (1.0, 1.0, a11)):
        Dpx = Dpx + sx * a
        Dpy = Dpy + sy * a
        Dpz = Dpz + 0.0 * a
    Dpx = Dpx * (0.5 ** (ndim - 1) / dx)
    Dpy = Dpy * (0.5 ** (ndim - 1) / dy)
    Dpz = Dpz * (0.5 ** (ndim - 1) / dz)
    return Dpx, Dpy, Dpz


def divergence_nodes(rhs, elem, node, Sol, ud):
    """second_projection.py:525-575 (2D). Zeroes the momenta in the two cell
    rows next to each y wall first (a side effect on Sol)."""
    if not hasattr(ud, "LAMB_BDRY") and _wall(ud.bdry_type[1]):
        for f in (Sol.rhou, Sol.rhov, Sol.rhow):
            f[:, :2, ...] = 0.0
            f[:, -2:, ...] = 0.0
    Y = Sol.rhoY / Sol.rho
    Ux = np.diff(Sol.rhou * Y, axis=0) / elem.dx
    Ux = 0.5 * (Ux[:, :-1, ...] + Ux[:, 1:, ...])
    Vy = np.diff(Sol.rhov * Y, axis=1) / elem.dy
    Vy = 0.5 * (Vy[:-1, ...] + Vy[1:, ...])
    return Ux + Vy


def rhs_from_p_old(rhs, node, mpv):
    """second_projection.py:578-601."""
    rhs_hh = mpv.wcenter * mpv.p2_nodes[node.i1]
    return rhs + 0.0 * rhs_hh


def correction_nodes(Sol, elem, node, mpv, p, dt, ud, th, updt_chi):
    """second_projection.py:280-306 (2D)."""
    Gammainv = th.Gammainv
    dSdy = dSdy_cells(mpv, elem, node)
    Dpx, Dpy, Dpz = grad_nodes(p, elem.ndim, (node.dx, node.dy, node.dz))
    thinv = Sol.rho / Sol.rhoY
    Y = Sol.rhoY / Sol.rho
    coeff = Gammainv * Sol.rhoY * Y
    mpv.u = -dt * coeff * Dpx
    mpv.v = -dt * coeff * Dpy
    mpv.w = -dt * coeff * Dpz
    multiply_inverse_coriolis(mpv, Sol, mpv, ud, elem, elem, dt, attrs=("u", "v", "w"))
    Sol.rhou += thinv * mpv.u
    Sol.rhov += thinv * mpv.v
    Sol.rhow += 0.0
    Sol.rhoX += -updt_chi * dt * dSdy * Sol.rhov


def assemble_system(Sol, mpv, elem, node, ud, th, dt, Sol0=None):
    """The part of impl_part that builds the linear system (:113-175), returning
    (op arrays, scaled rhs vector b~, diag_inv).  Mutates Sol/mpv like the
    reference."""
    S = Sol0 if Sol0 is not None else Sol
    bdry.set_explicit_boundary_data(S, elem, ud, th, mpv)
    operator_coefficients_nodes(elem, node, S, mpv, ud, th, dt)
    bdry.set_ghostnodes_p2(mpv.p2_nodes, node, ud)
    correction_nodes(Sol, elem, node, mpv, mpv.p2_nodes, dt, ud, th, 0)
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    rhs = mpv.rhs
    rhs[...] = divergence_nodes(rhs, elem, node, Sol, ud)
    rhs /= dt
    if ud.is_compressible == 1:
        rhs = rhs_from_p_old(rhs, node, mpv)
    elif ud.is_compressible == 0:
        if ud.is_ArakawaKonor:
            rhs -= mpv.wcenter * mpv.dp2_nodes
            mpv.wcenter[...] = 0.0
        else:
            rhs_new = rhs_from_p_old(rhs, node, mpv)
            rhs = ud.compressibility * rhs_new + (1.0 - ud.compressibility) * rhs
            mpv.wcenter[...] *= ud.compressibility
    else:
        mpv.wcenter *= ud.compressibility
    mpv.rhs[...] = rhs
    coriolis = multiply_inverse_coriolis(mpv, Sol, mpv, ud, elem, node, dt, attrs=("u", "v", "w"), get_coeffs=True)
    diag_inv = lp.precon_diag_prepare(mpv, elem, node, ud)
    rhs = rhs * diag_inv
    op = lp.op2d_arrays(mpv, node, coriolis, diag_inv, ud)
    b = rhs[1:-1, 1:-1].T.ravel().copy()
    return op, b, diag_inv


def finish_projection(Sol, mpv, elem, node, ud, th, dt, p2):
    """:241-277: scatter the solution, correct momenta, update p2_nodes."""
    nx, ny = node.iicx, node.iicy
    p2_full = np.zeros(node.sc[:2])
    p2_full[node.i2] = p2.reshape(ny, nx).T
    bdry.set_ghostnodes_p2(p2_full, node, ud)
    correction_nodes(Sol, elem, node, mpv, p2_full, dt, ud, th, 1)
    mpv.p2_nodes[...] += p2_full
    bdry.set_ghostnodes_p2(mpv.p2_nodes, node, ud)
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    return p2_full


def euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, t, dt, alpha_diff, Sol0=None, writer=None, label=None, debug=False):
    """second_projection.py:97-277 (2D).  Returns the BiCGStab result record."""
    op, b, diag_inv = assemble_system(Sol, mpv, elem, node, ud, th, dt, Sol0)
    res = bicgstab(lambda v: lp.lap2d(v, op), b, tol=ud.tol, maxiter=ud.max_iterations)
    finish_projection(Sol, mpv, elem, node, ud, th, dt, res.x)
    return res
