"""Host-side containers with the attribute names of pyBELLA's solver state.

These hold plain float64 numpy arrays in the reference's ``arr[x, y]`` layout
(ghost cells/nodes included) so that ``Sol, mpv, elem, node, ud, th`` objects
built here can be passed to the drop-in
``pybella_amd.second_projection.euler_backward_non_advective_impl_part`` exactly
as pyBELLA passes its own objects.  They are data holders: no solver arithmetic
lives here.

Reference counterparts (file:line, relative to the pyBELLA tree):
  * ``BdryType``            src/dycore/utils/options.py:42-49
  * ``global constants``    src/utils/sim_params.py:11-28
  * ``UserData``            src/utils/user_data.py:17-321 (derived Msq, gravity,
                            Coriolis strengths)
  * ``Thermo``              src/dycore/physics/gas_dynamics/thermodynamic.py:1-21
  * ``Grid`` / ``ElemSpaceDiscr`` / ``NodeSpaceDiscr``
                            src/dycore/discretisation/grid.py:39-297
  * ``Vars`` / ``States``   src/dycore/utils/variable.py:5-191
  * ``MPV``                 src/dycore/physics/low_mach/mpv.py:4-35
  * ``hydrostatic_state``   src/dycore/physics/hydrostatics.py:79-132
"""
from __future__ import annotations

from enum import Enum

import numpy as np


class BdryType(Enum):
    """Boundary types; values equal the reference's (options.py:42-49)."""

    WALL = "symmetric"
    PERIODIC = "wrap"
    RAYLEIGH = "radiation"


def is_wall(bt) -> bool:
    """True for WALL or RAYLEIGH of either this enum or the reference's."""
    return getattr(bt, "name", str(bt)) in ("WALL", "RAYLEIGH")


def is_periodic(bt) -> bool:
    return getattr(bt, "name", str(bt)) == "PERIODIC"


class UserData:
    """Configuration with the reference's default constants.

    Global constants follow sim_params.py:11-28; grid, boundary, regime and
    solver defaults follow user_data.py:28-129; derived quantities follow the
    property setters at user_data.py:132-321:
      u_ref = h_ref/t_ref, Msq = u_ref^2/(R_gas T_ref),
      gravity_strength[1] = grav h_ref/(R_gas T_ref),
      coriolis_strength[0] = coriolis_strength[2] = omega t_ref.
    Keyword arguments override any attribute; derived values are recomputed.
    """

    def __init__(self, **kw):
        # sim_params.global_constants
        self.nspec = 1
        self.buoy = 0
        self.grav = 9.81
        self.omega = 0.0
        self.R_gas = 287.4
        self.R_vap = 461.0
        self.Q_vap = 2.53e06
        self.gamm = 1.4
        self.p_ref = 8.61e4
        self.T_ref = 300.0
        self.h_ref = 10000.0
        self.t_ref = 100.0
        # user_data defaults
        self.inx, self.iny, self.inz = 64 + 1, 64 + 1, 1
        self.xmin, self.xmax = -1.0, 1.0
        self.ymin, self.ymax = 0.0, 1.0
        self.zmin, self.zmax = -1.0, 1.0
        self.bdry_type = np.empty(3, dtype=object)
        self.bdry_type[0] = BdryType.PERIODIC
        self.bdry_type[1] = BdryType.WALL
        self.bdry_type[2] = BdryType.WALL
        self.CFL = 0.5
        self.dtfixed0 = 100.0
        self.dtfixed = 100.0
        self.is_ArakawaKonor = 0
        self.is_nonhydrostatic = 1
        self.is_compressible = 1
        self.compressibility = 0.0
        self.nonhydrostasy = 1.0
        self.u_wind_speed = 0.0
        self.v_wind_speed = 0.0
        self.w_wind_speed = 0.0
        self.tol = 1.0e-8
        self.max_iterations = 6000
        # time stepping (user_data.py:57-64, 87-101, 105-118)
        self.acoustic_timestep = 0
        self.tout = np.arange(0.0, 1.01, 0.01)[10:]
        self.stepmax = 10000
        self.do_advection = True
        self.continuous_blending = False
        self.initial_blending = False
        self.no_of_pi_initial = 1
        self.no_of_pi_transition = 0
        self.no_of_hy_initial = 0
        self.no_of_hy_transition = 0
        self.aux = ""
        for k, v in kw.items():
            setattr(self, k, v)
        self.derive()

    @staticmethod
    def stratification(y):
        """Default stratification function (user_data.py:180-182)."""
        return 1.0

    def derive(self):
        self.u_ref = self.h_ref / self.t_ref
        self.Msq = self.u_ref * self.u_ref / (self.R_gas * self.T_ref)
        self.gravity_strength = np.zeros(3)
        self.gravity_strength[1] = self.grav * self.h_ref / (self.R_gas * self.T_ref)
        self.coriolis_strength = np.zeros(3)
        self.coriolis_strength[0] = self.omega * self.t_ref
        self.coriolis_strength[2] = self.omega * self.t_ref
        self.cp_gas = self.gamm * self.R_gas / (self.gamm - 1.0)


class Thermo:
    """gamma-derived constants (thermodynamic.py:14-20)."""

    def __init__(self, ud):
        g = ud.gamm
        self.gamm = g
        self.gamminv = 1.0 / g
        self.gm1 = g - 1.0
        self.gm1inv = 1.0 / (g - 1.0)
        self.Gamma = (g - 1.0) / g
        self.Gammainv = g / (g - 1.0)


class Grid:
    """Node counts and spacings (grid.py:45-101); ``inx`` counts nodes."""

    def __init__(self, inx, iny, inz, x0, x1, y0, y1, z0, z1):
        assert inx > 1 and iny >= 1 and inz >= 1
        self.ndim = 3 if inz > 1 else (2 if iny > 1 else 1)
        self.inx, self.iny, self.inz = inx, iny, inz
        self.dx = (x1 - x0) / (inx - 1.0)
        self.dy = (y1 - y0) / (iny - 1.0) if iny > 1 else 0.0
        self.dz = (z1 - z0) / (inz - 1.0) if inz > 1 else 0.0
        self.x0, self.x1, self.y0, self.y1, self.z0, self.z1 = x0, x1, y0, y1, z0, z1


class _SpaceDiscr:
    """Cell-grid sizes with two ghost layers (grid.py:122-189)."""

    def __init__(self, g: Grid):
        self.ndim = g.ndim
        self.igx = 2
        self.igy = 2 if g.iny > 1 else 0
        self.igz = 2 if g.inz > 1 else 0
        self.icx = g.inx - 1 + 2 * self.igx
        self.icy = g.iny - 1 + 2 * self.igy if g.iny > 1 else 1
        self.icz = g.inz - 1 + 2 * self.igz if g.inz > 1 else 1
        self.iicx = g.inx - 1
        self.iicy = g.iny - 1 if g.iny > 1 else 1
        self.iicz = g.inz - 1 if g.inz > 1 else 1
        self.igs = [self.igx, self.igy, self.igz]
        self.dx = g.dx
        self.dy = g.dy if self.icy > 1 else 1.0
        self.dz = g.dz if self.icz > 1 else 1.0
        self.dxyz = np.array([self.dx, self.dy, self.dz])
        self._finish()

    def _finish(self):
        self.nc = self.icx * self.icy * self.icz
        self.iisc = (self.iicx, self.iicy, self.iicz)
        self.isc = (self.iicx + self.igx, self.iicy + self.igy, self.iicz + self.igz)
        self.sc = (self.icx, self.icy, self.icz)
        self.i1 = tuple(slice(1, -1) for _ in range(self.ndim))
        self.i2 = tuple(slice(self.igs[d], -self.igs[d]) for d in range(self.ndim))


class ElemSpaceDiscr(_SpaceDiscr):
    """Cell grid; cell centres at half offsets (grid.py:192-215)."""

    def __init__(self, g: Grid, ud=None):
        super().__init__(g)
        x0 = g.x0 - self.igx * self.dx + 0.5 * self.dx
        y0 = g.y0 - self.igy * self.dy + 0.5 * self.dy if self.icy > 1 else g.y0
        z0 = g.z0 - self.igz * self.dz + 0.5 * self.dz if self.icz > 1 else g.z0
        self.x = x0 + self.dx * np.arange(self.icx)
        self.y = y0 + self.dy * np.arange(self.icy)
        self.z = z0 + self.dz * np.arange(self.icz)


class NodeSpaceDiscr(_SpaceDiscr):
    """Node grid: one more point per active axis (grid.py:241-275)."""

    def __init__(self, g: Grid, ud=None):
        super().__init__(g)
        x0 = g.x0 - self.igx * self.dx
        y0 = g.y0 - self.igy * self.dy if self.icy > 1 else g.y0
        z0 = g.z0 - self.igz * self.dz if self.icz > 1 else g.z0
        self.icx += 1
        self.icy += 1 if g.iny > 1 else 0
        self.icz += 1 if g.inz > 1 else 0
        self.iicx += 1
        self.iicy += 1 if g.iny > 1 else 0
        self.iicz += 1 if g.inz > 1 else 0
        self._finish()
        self.x = x0 + self.dx * np.arange(self.icx)
        self.y = y0 + self.dy * np.arange(self.icy)
        self.z = z0 + self.dz * np.arange(self.icz)


def grid_init(ud):
    """(elem, node) for a UserData (grid.py:5-37)."""
    g = Grid(ud.inx, ud.iny, ud.inz, ud.xmin, ud.xmax, ud.ymin, ud.ymax, ud.zmin, ud.zmax)
    return ElemSpaceDiscr(g, ud), NodeSpaceDiscr(g, ud)


def _squeeze_shape(size):
    return tuple(int(s) for s in size if s != 1)


class Vars:
    """Cell state rho, rhou, rhov, rhow, rhoY, rhoX (variable.py:5-39)."""

    def __init__(self, size, ud=None):
        shp = _squeeze_shape(size)
        for k in ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX"):
            setattr(self, k, np.zeros(shp))


class States:
    """Hydrostatic background column plus cached derivatives (variable.py:118-191)."""

    def __init__(self, n):
        for k in ("p0", "p20", "rho0", "S0", "S10", "pi0", "rhoY0", "Y0"):
            setattr(self, k, np.zeros(n))

    def get_dSdy(self, elem, node):
        """dS/dy between consecutive nodes, repeated over x (variable.py:162-177)."""
        if hasattr(self, "dSdy"):
            return self.dSdy
        d = (self.S0[1:] - self.S0[:-1]) / node.dy
        d = np.repeat(d[None, :], elem.sc[0], axis=0)
        if node.ndim == 3:
            d = np.repeat(d[:, :, None], elem.sc[2], axis=2)
        self.dSdy = d
        return d

    def get_S0c(self, elem):
        if hasattr(self, "S0c"):
            return self.S0c
        s = np.repeat(self.S0[None, :], elem.sc[0], axis=0)
        self.S0c = s
        return s


class Flux:
    """Face fluxes of one direction (variable.States as src/__main__.py:70-75
    allocates them): rho, rhou, rhov, rhow, rhoY, rhoX of shape elem.sfx =
    (icy, icx+1) for the x faces ([y, face]) or elem.sfy = (icx, icy+1) for
    the y faces ([x, face]) (grid.py:164-166)."""

    def __init__(self, shape):
        for k in ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX"):
            setattr(self, k, np.zeros(shape))


def flux_init(elem):
    """[x-face, y-face] flux containers of a 2D grid."""
    return [Flux((elem.icy, elem.icx + 1)), Flux((elem.icx, elem.icy + 1))]


class MPV:
    """Solver-side fields (mpv.py:5-29)."""

    def __init__(self, elem, node, ud=None):
        sc = _squeeze_shape(elem.sc)
        sn = _squeeze_shape(node.sc)
        self.p0 = 1.0
        self.p00 = 1.0
        self.p2_cells = np.zeros(sc)
        self.dp2_cells = np.zeros(sc)
        self.p2_nodes = np.zeros(sn)
        self.p2_nodes0 = np.zeros(sn)
        self.dp2_nodes = np.zeros(sn)
        self.u = np.zeros(sc)
        self.v = np.zeros(sc)
        self.w = np.zeros(sc)
        isc = _squeeze_shape(node.isc)
        self.rhs = np.zeros(isc)
        self.wcenter = np.zeros(isc)
        self.wplus = np.zeros([elem.ndim] + list(sc))
        self.HydroState = States(elem.sc[1])
        self.HydroState_n = States(node.sc[1])


def hydrostatic_state(mpv, elem, node, th, ud):
    """Isothermal (exponential-Exner) background column (hydrostatics.py:79-132)."""
    g = ud.gravity_strength[1]
    if g == 0.0:
        for st in (mpv.HydroState, mpv.HydroState_n):
            for k in ("p20", "p0", "rho0", "rhoY0", "Y0", "S0"):
                getattr(st, k)[...] = 1.0
        return
    Hex = 1.0 / (th.Gamma * g)
    dy = elem.dy
    for st, y in ((mpv.HydroState_n, node.y), (mpv.HydroState, elem.y)):
        pi_p = np.exp(-(y + 0.5 * dy) / Hex)
        pi_m = np.exp(-(y - 0.5 * dy) / Hex)
        pi = np.exp(-y / Hex)
        Yc = -th.Gamma * g * dy / (pi_p - pi_m)
        P = pi ** th.gm1inv
        st.p20[...] = pi / ud.Msq
        st.p0[...] = pi ** th.Gammainv
        st.rho0[...] = P / Yc
        st.rhoY0[...] = P
        st.Y0[...] = Yc
        st.S0[...] = 1.0 / Yc
