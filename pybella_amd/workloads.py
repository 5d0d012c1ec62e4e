"""Synthetic inputs for the nodal elliptic solve (host-side, float64 numpy).

Every builder returns pyBELLA-shaped objects ``(Sol, mpv, elem, node, ud, th)``
(see ``pybella_amd.state``) so the same inputs feed the drop-in solver and the
CPU checker.

* ``stratified_state``   hydrostatically stratified atmosphere, theta linear
  from 300 K at y=0 to 330 K at y=10 km, Exner from hydrostatic balance
  d(pi)/dy = -Gamma g / theta integrated in closed form:
  pi(y) = 1 - Gamma g (H/0.1) ln(1 + 0.1 y/H) (nondimensional, H = 1).
  P = pi^(1/(gamma-1)) = rhoY, rho = P/theta; S0 = 1/theta at nodes.
  Periodic in x, wall in y, pseudo-incompressible (wcenter = 0).
* ``fourier_rhs``        seeded smooth random node field: sum of modes
  a cos(2 pi kx (x-xmin)/Lx + phi) cos(pi ky (y-ymin)/H), integer (kx, ky) !=
  (0, 0) with |k| <= kmax, a ~ N(0,1), phi ~ U(0, 2 pi).  The cosine in y
  satisfies the wall Neumann condition; integer kx duplicates the periodic end
  column exactly.
* ``project_range``      makes a node RHS consistent with the singular
  pseudo-incompressible operator by removing its components along the left
  null vectors (trapezoid weights w: 1/2 on the duplicated periodic end
  columns, 1 elsewhere; and w times the checkerboard (-1)^(i+j)).
* ``rising_bubble_state`` the rising-bubble initial state of
  inputs/rising_bubble.py:42-99 (cos^2 theta perturbation of radius r0 centred
  at (xc, y0) on the isothermal background of hydrostatics.py:79-132), with
  the buoyancy impulse of one explicit half step applied to rhov so the
  second-projection RHS is non-zero (rhoX = rho (rho/rhoY - S0c) as in
  second_projection.py:66; rhov = -dt (g/Msq) rhoY rhoX/rho as in
  second_projection.py:82-83).
"""
from __future__ import annotations

import numpy as np

from .state import (MPV, BdryType, Thermo, UserData, Vars, grid_init,
                    hydrostatic_state)


def _objects(nx, ny, **kw):
    ud = UserData(inx=nx, iny=ny, inz=1, **kw)
    th = Thermo(ud)
    elem, node = grid_init(ud)
    Sol = Vars(elem.sc, ud)
    mpv = MPV(elem, node, ud)
    return Sol, mpv, elem, node, ud, th


def stratified_state(nx=1025, ny=513, theta0=300.0, theta1=330.0, compressible=False):
    Sol, mpv, elem, node, ud, th = _objects(nx, ny, is_compressible=1 if compressible else 0,
                                            compressibility=1.0 if compressible else 0.0)
    g = ud.gravity_strength[1]
    a = (theta1 - theta0) / theta0          # theta/theta0 = 1 + a y  (y in units of 10 km)

    def theta(y):
        return 1.0 + a * y

    def exner(y):
        return 1.0 - th.Gamma * g * np.log1p(a * y) / a

    yc = elem.y
    P = exner(yc) ** th.gm1inv
    rho = P / theta(yc)
    Sol.rhoY[...] = P[None, :]
    Sol.rho[...] = rho[None, :]
    mpv.HydroState_n.S0[...] = 1.0 / theta(node.y)
    mpv.HydroState.S0[...] = 1.0 / theta(yc)
    mpv.HydroState.rhoY0[...] = P
    mpv.HydroState.rho0[...] = rho
    mpv.HydroState_n.rhoY0[...] = exner(node.y) ** th.gm1inv
    return Sol, mpv, elem, node, ud, th


def fourier_rhs(node, seed, nmodes=64, kmax=16):
    """Node field of shape (NX, NY) in the reference's [x, y] layout."""
    rng = np.random.default_rng(seed)
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1)
          if (kx, ky) != (0, 0) and kx * kx + ky * ky <= kmax * kmax]
    pick = rng.choice(len(ks), size=nmodes, replace=False)
    amp = rng.standard_normal(nmodes)
    phi = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    x = node.x[2:-2]
    y = node.y[2:-2]
    Lx = x[-1] - x[0]
    H = y[-1] - y[0]
    f = np.zeros((x.size, y.size))
    for m, k in enumerate(pick):
        kx, ky = ks[k]
        f += amp[m] * np.cos(2 * np.pi * kx * (x - x[0]) / Lx + phi[m])[:, None] * \
            np.cos(np.pi * ky * (y - y[0]) / H)[None, :]
    return f


def left_null_vectors(nx, ny):
    """(2, NX, NY) left null vectors of the unscaled psinc operator."""
    w = np.ones((nx, ny))
    w[0, :] = 0.5
    w[-1, :] = 0.5
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    c = np.where((i + j) % 2 == 0, 1.0, -1.0)
    return np.stack([w, w * c])


def project_range(f):
    """Remove the components of a (NX, NY) node field along the left null
    vectors (orthogonal projection onto their complement)."""
    W = left_null_vectors(*f.shape).reshape(2, -1)
    G = W @ W.T
    coef = np.linalg.solve(G, W @ f.ravel())
    return (f.ravel() - coef @ W).reshape(f.shape)


def stratified_rhs(node, seed=20261202, nmodes=64, kmax=16):
    """Consistent node RHS (NX+2, NY+2) with a zero ghost ring, matching the
    shape of mpv.rhs."""
    f = project_range(fourier_rhs(node, seed, nmodes, kmax))
    out = np.zeros(node.isc[:2])
    out[1:-1, 1:-1] = f
    return out


def rising_bubble_state(nx=161, ny=81, delth=2.0, xc=0.0, y0=0.2, r0=0.2, dt=0.05):
    """Rising-bubble state (inputs/rising_bubble.py:42-99) after a buoyancy
    impulse; returns (Sol, Sol0, mpv, elem, node, ud, th)."""
    Sol, mpv, elem, node, ud, th = _objects(nx, ny, is_compressible=0, compressibility=0.0)
    hydrostatic_state(mpv, elem, node, th, ud)
    x, y = np.meshgrid(elem.x, elem.y)
    r = np.sqrt((x - xc) ** 2 + (y - y0) ** 2) / r0
    rhoY = np.repeat(mpv.HydroState.rhoY0.reshape(1, -1), elem.icx, axis=0)
    pert = (delth / 300.0) * (np.cos(0.5 * np.pi * r) ** 2)
    pert[np.where(r > 1.0)] = 0.0
    rho = rhoY / (ud.stratification(y) + pert.T)
    Sol.rho[...] = rho
    Sol.rhoY[...] = rhoY
    import copy
    Sol0 = copy.deepcopy(Sol)
    S0c = mpv.HydroState.get_S0c(elem)
    Sol.rhoX[...] = Sol.rho * (Sol.rho / Sol.rhoY - S0c)
    g = ud.gravity_strength[1]
    Sol.rhov[...] = -dt * (g / ud.Msq) * Sol.rhoY * (Sol.rhoX / Sol.rho)
    return Sol, Sol0, mpv, elem, node, ud, th


def random_momentum_state(nx, ny, seed, dt=0.05, compressible=False):
    """Stratified state with seeded smooth momenta and rhoX, used to exercise
    every term of the projection (divergence, both corrections, rhoX update)."""
    Sol, mpv, elem, node, ud, th = stratified_state(nx, ny, compressible=compressible)
    rng = np.random.default_rng(seed)
    X, Yg = np.meshgrid(elem.x, elem.y, indexing="ij")
    for f in (Sol.rhou, Sol.rhov, Sol.rhoX):
        acc = np.zeros_like(f)
        for _ in range(6):
            kx = rng.integers(1, 5)
            ky = rng.integers(1, 5)
            acc += rng.standard_normal() * np.cos(np.pi * kx * X + rng.uniform(0, 6.3)) * \
                np.sin(np.pi * ky * Yg + rng.uniform(0, 6.3))
        f[...] = 0.1 * acc * Sol.rho
    mpv.p2_nodes[...] = 0.01 * rng.standard_normal(mpv.p2_nodes.shape)
    return Sol, mpv, elem, node, ud, th


def batch_member_bubble(nx, ny, member, base_seed=20261204, dt=0.05):
    """Ensemble-member rising bubble: delth ~ U(1,3) K, xc ~ U(-3,3) km,
    r0 ~ U(1.5, 2.5) km from default_rng(base_seed + member)."""
    rng = np.random.default_rng(base_seed + member)
    delth = rng.uniform(1.0, 3.0)
    xc = rng.uniform(-0.3, 0.3)
    r0 = rng.uniform(0.15, 0.25)
    return rising_bubble_state(nx, ny, delth=delth, xc=xc, r0=r0, dt=dt)


class State3D:
    """Inputs of the isolated 3D nodal solve: operator-cell fields rho, rhoY
    of shape (nz+1, ny+1, nx+1) ([z][y][x]; cell (k, j, i) is centred between
    nodes k-1..k, j-1..j, i-1..i, the outer layers being the ghost cells),
    grid spacings, periodicity (x, y, z) and the reference's constants."""

    def __init__(self, nx, ny, nz, dx, dy, dz, periodic, rho_c, rhoY_c, ud, th):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        self.periodic = periodic
        self.rho_c, self.rhoY_c = rho_c, rhoY_c
        self.ud, self.th = ud, th


def stratified_state_3d(nx=257, ny=129, nz=129, theta0=300.0, theta1=330.0, compressible=True):
    """The stratified atmosphere of ``stratified_state`` on a 3D node grid:
    x in [-1, 1] and z in [-0.5, 0.5] periodic (duplicated end nodes), y in
    [0, 1] (gravity) with walls; compressible regime by default (the
    pseudo-incompressible 27-point operator is singular on every in-plane
    checkerboard mode)."""
    ud = UserData(inx=nx, iny=ny, inz=nz, zmin=-0.5, zmax=0.5, is_compressible=1 if compressible else 0,
                  compressibility=1.0 if compressible else 0.0)
    ud.bdry_type[2] = BdryType.PERIODIC
    th = Thermo(ud)
    g = ud.gravity_strength[1]
    a = (theta1 - theta0) / theta0
    dx = (ud.xmax - ud.xmin) / (nx - 1)
    dy = (ud.ymax - ud.ymin) / (ny - 1)
    dz = (ud.zmax - ud.zmin) / (nz - 1)
    yc = ud.ymin + (np.arange(ny + 1) - 0.5) * dy            # cell centres incl. ghosts
    exner = 1.0 - th.Gamma * g * np.log1p(a * yc) / a
    P = exner ** th.gm1inv
    rho = P / (1.0 + a * yc)
    shape = (nz + 1, ny + 1, nx + 1)
    rho_c = np.broadcast_to(rho[None, :, None], shape).copy()
    rhoY_c = np.broadcast_to(P[None, :, None], shape).copy()
    return State3D(nx, ny, nz, dx, dy, dz, (True, False, True), rho_c, rhoY_c, ud, th)


def fourier_rhs_3d(nx, ny, nz, seed, nmodes=64, kmax=12):
    """Seeded smooth zero-mean node field (nz, ny, nx): sum of modes
    a cos(2 pi kx s + phi) cos(pi ky t) cos(2 pi kz u + psi) with integer
    (kx, ky, kz) != 0, |k| <= kmax, s, t, u in [0, 1] over the nodes; the
    trapezoid-weighted mean (weights 1/2 on the duplicated periodic end
    nodes) is removed."""
    rng = np.random.default_rng(seed)
    ks = [(kx, ky, kz) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1) for kz in range(0, kmax + 1)
          if (kx, ky, kz) != (0, 0, 0) and kx * kx + ky * ky + kz * kz <= kmax * kmax]
    pick = rng.choice(len(ks), size=nmodes, replace=False)
    amp = rng.standard_normal(nmodes)
    phi = rng.uniform(0.0, 2.0 * np.pi, (nmodes, 2))
    s = np.linspace(0.0, 1.0, nx)
    t = np.linspace(0.0, 1.0, ny)
    u = np.linspace(0.0, 1.0, nz)
    f = np.zeros((nz, ny, nx))
    for m, k in enumerate(pick):
        kx, ky, kz = ks[k]
        f += amp[m] * np.cos(2 * np.pi * kz * u + phi[m, 1])[:, None, None] * \
            np.cos(np.pi * ky * t)[None, :, None] * np.cos(2 * np.pi * kx * s + phi[m, 0])[None, None, :]
    w = np.ones((nz, ny, nx))
    w[:, :, 0] = w[:, :, -1] = 0.5
    w[0] *= 0.5
    w[-1] *= 0.5
    return f - (w * f).sum() / w.sum()


__all__ = ["stratified_state", "fourier_rhs", "left_null_vectors", "project_range",
           "stratified_rhs", "rising_bubble_state", "random_momentum_state",
           "batch_member_bubble", "BdryType", "State3D", "stratified_state_3d", "fourier_rhs_3d"]
