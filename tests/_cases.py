"""Shared builders for the CPU and GPU tests: seeded synthetic states and the
oracle's assembly of the 2D nodal operator on them."""
import copy

import numpy as np

from oracle import laplacian as olp
from oracle import projection as opj
from pybella_amd import workloads as W


def rel(a, b):
    """Max-norm relative difference of a against b."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def make_case(kind, nx, ny):
    if kind == "strat_psinc":
        return W.random_momentum_state(nx, ny, seed=1)
    if kind == "strat_comp":
        return W.random_momentum_state(nx, ny, seed=2, compressible=True)
    if kind == "coriolis":
        case = W.random_momentum_state(nx, ny, seed=3, compressible=True)
        case[4].omega = 7.292e-5 * 50
        case[4].derive()
        return case
    if kind == "xwall":
        case = W.random_momentum_state(nx, ny, seed=4, compressible=True)
        case[4].bdry_type[0] = W.BdryType.WALL
        return case
    if kind == "yper":   # the travelling vortex: 65 x 65 nodes, periodic on both axes (nx, ny ignored)
        Sol, flux, mpv, elem, node, ud, th = W.travelling_vortex_init(opj.euler_backward_non_advective_impl_part)
        return Sol, mpv, elem, node, ud, th
    raise ValueError(kind)


def oracle_op(Sol, mpv, elem, node, ud, th, dt):
    """Oracle assembly on deep copies: (op arrays, assembled mpv, diag_inv)."""
    Sol, mpv = copy.deepcopy(Sol), copy.deepcopy(mpv)
    opj.operator_coefficients_nodes(elem, node, Sol, mpv, ud, th, dt)
    if ud.is_compressible == 0:
        mpv.wcenter[...] *= ud.compressibility
    cor = opj.multiply_inverse_coriolis(mpv, Sol, mpv, ud, elem, node, dt, attrs=("u", "v", "w"), get_coeffs=True)
    dinv = olp.precon_diag_prepare(mpv, elem, node, ud)
    return olp.op2d_arrays(mpv, node, cor, dinv, ud), mpv, dinv


def random_op(nx, ny, seed, x_wall=False, y_wall=True, coriolis=True, center=True):
    """An operator dict with arbitrary positive weights (for invariants that hold
    for any coefficients)."""
    rng = np.random.default_rng(seed)
    cell = lambda: 0.5 + rng.random((ny + 1, nx + 1))
    op = dict(nx=nx, ny=ny, dx=0.7 / (nx - 1), dy=0.3 / (ny - 1), hx=cell(), hy=cell(),
              hc=(-rng.random((ny, nx)) if center else np.zeros((ny, nx))),
              cxx=cell(), cyy=cell(),
              cxy=(0.3 * rng.standard_normal((ny + 1, nx + 1)) if coriolis else np.zeros((ny + 1, nx + 1))),
              cyx=(0.3 * rng.standard_normal((ny + 1, nx + 1)) if coriolis else np.zeros((ny + 1, nx + 1))),
              dinv=-(0.5 + rng.random((ny, nx))), x_wall=x_wall, y_wall=y_wall)
    return op


def checkerboard(ny, nx):
    return np.where((np.arange(ny)[:, None] + np.arange(nx)[None, :]) % 2 == 0, 1.0, -1.0)


def timestep_state(kind, nx=33, ny=17, seed=3):
    """A generic state for the time-step kernels: (Sol, flux, mpv, elem, node,
    ud, th), ghost cells filled by the oracle.
      "bubble"   the rising bubble (x periodic, y walls with gravity, psinc)
                 with smooth random momenta and rhoX;
      "xwall"    the same with walls in x, compressible;
      "periodic" doubly periodic, no gravity, compressible, wind (0.3, -0.2),
                 smooth random fields."""
    from oracle import boundary as obd
    from pybella_amd.state import MPV, Thermo, Vars, flux_init, grid_init, hydrostatic_state
    rng = np.random.default_rng(seed)
    if kind in ("bubble", "xwall"):
        ud = W.rising_bubble_ud(nx, ny)
        if kind == "xwall":
            ud.bdry_type[0] = W.BdryType.WALL
            ud.is_compressible, ud.compressibility = 1, 1.0
        Sol, flux, mpv, elem, node, ud, th = W.rising_bubble_init(ud)
        X, Y = np.meshgrid(elem.x, elem.y, indexing="ij")
        Lx = 2.0
    else:
        ud = W.travelling_vortex_ud(nx)
        ud.u_wind_speed, ud.v_wind_speed = 0.3, -0.2
        th = Thermo(ud)
        elem, node = grid_init(ud)
        Sol = Vars(elem.sc, ud)
        mpv = MPV(elem, node, ud)
        hydrostatic_state(mpv, elem, node, th, ud)
        flux = flux_init(elem)
        X, Y = np.meshgrid(elem.x, elem.y, indexing="ij")
        Lx = 1.0
        Sol.rho[...] = 1.0
        Sol.rhoY[...] = 1.0

    def smooth(amp):
        f = np.zeros_like(X)
        for _ in range(4):
            kx, ky = rng.integers(1, 3, 2)
            f += amp * rng.standard_normal() * np.cos(2 * np.pi * kx * X / Lx + rng.uniform(0, 6.3)) * \
                np.cos(np.pi * ky * Y + rng.uniform(0, 6.3))
        return f
    Sol.rho[...] *= 1.0 + smooth(0.02)
    Sol.rhoY[...] *= 1.0 + smooth(0.01)
    Sol.rhou[...] = Sol.rho * (ud.u_wind_speed + smooth(0.1))
    Sol.rhov[...] = Sol.rho * (ud.v_wind_speed + smooth(0.1))
    Sol.rhoX[...] = Sol.rho * smooth(0.01)
    mpv.p2_nodes[...] = 0.01 * rng.standard_normal(mpv.p2_nodes.shape)
    obd.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    obd.set_ghostnodes_p2(mpv.p2_nodes, node, ud)
    return Sol, flux, mpv, elem, node, ud, th


def strip_node_null(p, q):
    """Max-norm relative difference of node arrays p and q (reference layout,
    two ghost layers) on the inner nodes after removing the difference's
    components along the constant and the checkerboard."""
    a, b = p[2:-2, 2:-2], q[2:-2, 2:-2]
    nx, ny = a.shape
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    B = np.stack([np.ones(a.size), np.where((i + j) % 2 == 0, 1.0, -1.0).ravel()], axis=1)
    d = (a - b).ravel()
    c, *_ = np.linalg.lstsq(B, d, rcond=None)
    return float(np.max(np.abs(d - B @ c)) / max(np.max(np.abs(b)), 1e-300))
