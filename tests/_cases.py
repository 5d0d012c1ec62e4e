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
