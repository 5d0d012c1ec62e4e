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

from .state import (MPV, BdryType, Thermo, UserData, Vars, flux_init, grid_init,
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
    """Node field of shape (NX, NY) in the reference's [x, y] layout.  The
    last column duplicates the first (the periodic end node), bit for bit."""
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
    f[-1, :] = f[0, :]
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


def rising_bubble_ud(nx=161, ny=81, **kw):
    """UserData of the rising-bubble time step: inputs/rising_bubble.py:5-39
    (CFL 0.5, dtfixed0 = dtfixed = 100, 'CFLfixed' in aux, so dt = 21.69/t_ref
    on the first two steps) composed with the UserDataInit defaults
    (user_data.py:28-129) that the shipped file omits, pseudo-incompressible,
    g = 9.81 m/s^2 and t_ref = 100 s."""
    ud = UserData(inx=nx, iny=ny, inz=1, is_compressible=0, compressibility=0.0, CFL=0.5, dtfixed0=100.0,
                  dtfixed=100.0, stepmax=10000, aux="debug_imbal_CFLfixed", **kw)
    return ud


def rising_bubble_init(ud=None, delth=2.0, xc=0.0, y0=0.2, r0=0.2):
    """sol_init of inputs/rising_bubble.py:42-99: the isothermal background
    (hydrostatics.py:79-132) with rho = rhoY / (1 + delth/300 cos^2(pi r/2))
    inside r < 1 (r the distance to (xc, y0) in units of r0), zero wind, zero
    rhoX and p2_nodes.  The reference's closing ghost fill is left to the
    time step, whose first action is the same fill (time_update.py:122).
    Returns (Sol, flux, mpv, elem, node, ud, th)."""
    if ud is None:
        ud = rising_bubble_ud()
    th = Thermo(ud)
    elem, node = grid_init(ud)
    Sol = Vars(elem.sc, ud)
    mpv = MPV(elem, node, ud)
    hydrostatic_state(mpv, elem, node, th, ud)
    x, y = np.meshgrid(elem.x, elem.y)
    r = np.sqrt((x - xc) ** 2 + (y - y0) ** 2) / r0
    rhoY = np.repeat(mpv.HydroState.rhoY0.reshape(1, -1), elem.icx, axis=0)
    pert = (delth / 300.0) * (np.cos(0.5 * np.pi * r) ** 2)
    pert[np.where(r > 1.0)] = 0.0
    rho = rhoY / (ud.stratification(y) + pert.T)
    Sol.rho[...] = rho
    Sol.rhou[...] = rho * ud.u_wind_speed
    Sol.rhov[...] = rho * ud.v_wind_speed
    Sol.rhow[...] = rho * ud.w_wind_speed
    Sol.rhoY[...] = rhoY
    mpv.p2_nodes[...] = 0.0
    return Sol, flux_init(elem), mpv, elem, node, ud, th


def travelling_vortex_ud(n=65):
    """UserData of src/tests/test_travelling_vortex.py:10-158: 64 x 64 cells
    on [-1/2, 1/2]^2, doubly periodic, no gravity, background wind (1, 1),
    compressible, CFL 0.45, dt at most 0.01, 101 steps."""
    ud = UserData(inx=n, iny=n, inz=1, grav=0.0, omega=0.0, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5,
                  zmin=-0.5, zmax=0.5, u_wind_speed=1.0, v_wind_speed=1.0, w_wind_speed=0.0, CFL=0.45,
                  dtfixed=0.01, dtfixed0=0.01, is_compressible=1, compressibility=1.0, is_nonhydrostatic=1,
                  tout=[1.0], stepmax=101, aux="")
    ud.bdry_type[1] = BdryType.PERIODIC
    ud.initial_projection = True
    return ud


def internal_long_wave_ud():
    """UserData of src/tests/test_internal_long_wave.py:8-180: 301 x 10 cells
    on x in [-300, 300] (periodic) and y in [0, 1] (walls), g = 9.81 m/s^2,
    Coriolis omega t_ref on the first axis only, stratification
    exp(N^2 y / g) with N^2 = 1e-4 s^-2, compressible, dt = 1 (CFL 0.9), solves
    to 1e-12, 31 steps.  ``molly`` is the test's window function."""
    ud = UserData(inx=302, iny=11, inz=1, grav=9.81, omega=7.292e-5, xmin=-300.0, xmax=300.0, ymin=0.0,
                  ymax=1.0, zmin=-1.0, zmax=1.0, u_wind_speed=0.0, v_wind_speed=0.0, w_wind_speed=0.0, CFL=0.9,
                  dtfixed0=1.0, dtfixed=1.0, is_compressible=1, compressibility=0.0, is_nonhydrostatic=1,
                  tol=1.0e-12, max_iterations=6000, tout=[20.0 * 3000.0 / 100.0], stepmax=31,
                  no_of_pi_initial=0, no_of_hy_initial=1, aux="")
    ud.bdry_type[2] = BdryType.PERIODIC
    ud.coriolis_strength[2] = 0.0     # the test sets only the first component
    ud.scale_factor = 20.0
    ud.Nsq_ref = 1.0e-4
    Nsq = ud.Nsq_ref * ud.t_ref * ud.t_ref
    g = ud.gravity_strength[1] / ud.Msq
    ud.stratification = lambda y: np.exp(Nsq * y / g)

    def molly(x):
        del0 = 0.25
        L = ud.xmax - ud.xmin
        xi_l = np.minimum(1.0, (x - ud.xmin) / (del0 * L))
        xi_r = np.minimum(1.0, (ud.xmax - x) / (del0 * L))
        return 0.5 * np.minimum(1.0 - np.cos(np.pi * xi_l), 1.0 - np.cos(np.pi * xi_r))
    ud.molly = molly
    return ud


_TV_COE = (1.0 / 24.0, -6.0 / 13.0, 15.0 / 7.0, -74.0 / 15.0, 57.0 / 16.0, 174.0 / 17.0, -269.0 / 9.0,
           450.0 / 19.0, 1071.0 / 40.0, -1564.0 / 21.0, 510.0 / 11.0, 1020.0 / 23.0, -1105.0 / 12.0,
           204.0 / 5.0, 510.0 / 13.0, -1564.0 / 27.0, 153.0 / 8.0, 450.0 / 29.0, -269.0 / 15.0, 174.0 / 31.0,
           57.0 / 32.0, -74.0 / 33.0, 15.0 / 17.0, -6.0 / 35.0, 1.0 / 72.0)
_TV_CONST = (1.0 / 24, -6.0 / 13, 33.0 / 14, -22.0 / 3, 495.0 / 32, -396.0 / 17, +77.0 / 3, -396.0 / 19,
             99.0 / 8, -110.0 / 21, +3.0 / 2, -6.0 / 23, +1.0 / 48)


def travelling_vortex_init(project, ud=None):
    """sol_init of src/tests/test_travelling_vortex.py:163-355 (no seed): a
    vortex of radius 0.4 with a density dip and its balancing pressure in the
    background wind, then the initial pseudo-incompressible projection with
    dt = dtfixed (the wind removed around it, p2_nodes restored after it).
    ``project`` is the projection to use (the signature of
    euler_backward_non_advective_impl_part).  Returns (Sol, flux, mpv, elem,
    node, ud, th)."""
    if ud is None:
        ud = travelling_vortex_ud()
    th = Thermo(ud)
    elem, node = grid_init(ud)
    Sol = Vars(elem.sc, ud)
    mpv = MPV(elem, node, ud)
    u0, v0, w0 = ud.u_wind_speed, ud.v_wind_speed, 0.0
    rotdir, p0, alpha, alpha_const = 1.0, 1.0, -1.0, 3.0
    rho0, del_rho, R0, fac = 1.0, -0.5, 0.4, 1.0 * 1024.0
    xc = yc = 0.0
    xcm = xc - np.sign(xc) * (ud.xmax - ud.xmin)
    ycm = yc - np.sign(yc) * (ud.ymax - ud.ymin)
    igy = elem.igs[1]
    igxn, igyn = node.igx, node.igy
    hydrostatic_state(mpv, elem, node, th, ud)
    xs = elem.x.reshape(-1, 1)
    ys = elem.y[igy:-igy].reshape(1, -1)
    xccs = np.zeros_like(xs)
    yccs = np.zeros_like(ys)
    xccs[...] = xc * (np.abs(xs - xc) < np.abs(xs - xcm))
    xccs[...] += xcm * (np.abs(xs - xc) > np.abs(xs - xcm))
    yccs[...] = yc * (np.abs(ys - yc) < np.abs(ys - ycm))
    yccs[...] += ycm * (np.abs(ys - yc) > np.abs(ys - ycm))
    r = np.sqrt((xs - xccs) ** 2 + (ys - yccs) ** 2)
    uth = (rotdir * fac * (1.0 - r / R0) ** 6 * (r / R0) ** 6) * (r < R0)
    u = u0 + uth * (-(ys - yccs) / r)
    v = v0 + uth * (+(xs - xccs) / r)
    w = w0
    rho = np.zeros_like(r)
    rho[...] += (rho0 + del_rho * (1.0 - (r / R0) ** 2) ** 6) * (r < R0)
    rho[...] += rho0 * (r >= R0)
    dp2c = np.zeros_like(r)
    dp2c_const = np.zeros_like(r)
    for ip in range(25):
        dp2c += (_TV_COE[ip] * ((r / R0) ** (12 + ip) - 1.0) * rotdir ** 2) * (r / R0 < 1.0)
    for ip in range(13):
        dp2c_const += (_TV_CONST[ip] * ((r / R0) ** (12 + ip) - 1.0) * rotdir ** 2) * (r / R0 < 1.0)
    dp2c = alpha * dp2c + alpha_const * dp2c_const
    Sol.rho[:, igy:-igy] = rho
    Sol.rhou[:, igy:-igy] = rho * u
    Sol.rhov[:, igy:-igy] = rho * v
    Sol.rhow[:, igy:-igy] = rho * w
    if ud.is_compressible:
        p = p0 + ud.Msq * fac ** 2 * dp2c
        Sol.rhoY[:, igy:-igy] = p ** th.gamminv
    else:
        Sol.rhoY[:, igy:-igy] = mpv.HydroState.rhoY0[igy:-igy]
    xs = node.x[igxn:-igxn].reshape(-1, 1)
    ys = node.y[igyn:-igyn].reshape(1, -1)
    xccs = np.zeros_like(xs)
    yccs = np.zeros_like(ys)
    yccs[np.where(np.abs(ys - yc) < np.abs(ys - ycm))] = yc
    yccs[np.where(np.abs(ys - yc) > np.abs(ys - ycm))] = ycm
    xccs[np.where(np.abs(xs - xc) < np.abs(xs - xcm))] = xc
    xccs[np.where(np.abs(xs - xc) > np.abs(xs - xcm))] = xcm
    r = np.sqrt((xs - xccs) ** 2 + (ys - yccs) ** 2)
    inner = mpv.p2_nodes[igxn:-igxn, igyn:-igyn]
    for ip in range(25):
        inner += alpha * _TV_COE[ip] * ((r / R0) ** (12 + ip) - 1.0) * rotdir ** 2
    for ip in range(13):
        inner += alpha_const * _TV_CONST[ip] * ((r / R0) ** (12 + ip) - 1.0) * rotdir ** 2
    inner *= r / R0 < 1.0
    inner[...] = th.Gamma * fac ** 2 * np.divide(inner, mpv.HydroState.rhoY0[igyn:-igyn + 1])
    ud.nonhydrostasy = float(ud.is_nonhydrostatic)
    ud.compressibility = float(ud.is_compressible)
    if getattr(ud, "initial_projection", False):
        is_comp, comp = ud.is_compressible, ud.compressibility
        ud.is_compressible = 0
        ud.compressibility = 0.0
        p2aux = np.copy(mpv.p2_nodes)
        Sol.rhou -= u0 * Sol.rho
        Sol.rhov -= v0 * Sol.rho
        project(Sol, mpv, elem, node, ud, th, 0.0, ud.dtfixed, 0.5)
        mpv.p2_nodes[...] = p2aux
        mpv.dp2_nodes[...] = 0.0
        Sol.rhou += u0 * Sol.rho
        Sol.rhov += v0 * Sol.rho
        ud.is_compressible = is_comp
        ud.compressibility = comp
    return Sol, flux_init(elem), mpv, elem, node, ud, th


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
           "batch_member_bubble", "rising_bubble_ud", "rising_bubble_init", "travelling_vortex_ud",
           "travelling_vortex_init", "internal_long_wave_ud", "BdryType", "State3D", "stratified_state_3d", "fourier_rhs_3d"]
