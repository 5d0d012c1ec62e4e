"""The explicit parts of the time step and the single-member time-step driver
(oracle; TEST INFRASTRUCTURE ONLY: imported by tests/, smoke() and the CPU
baseline of bench.py, never by pybella_amd).

Restates, for the 2D vertical slice (x without gravity, periodic or wall;
y the gravity axis, or an axis without gravity), following the reference's
operation order (pyBELLA tree, file:line):

  recompute_advective_fluxes   src/dycore/physics/gas_dynamics/numerical_flux.py:5-48
  hll_solver                   numerical_flux.py:50-99
  recovery (do, slopes, limiters, get_conservatives)
                               src/dycore/physics/gas_dynamics/recovery.py:6-210
  explicit_step_and_flux, advect, advect_rk
                               src/dycore/physics/gas_dynamics/explicit.py:5-257
  euler_forward_non_advective, euler_backward_non_advective_expl_part
                               src/dycore/physics/low_mach/second_projection.py:22-93
  dynamic_timestep             src/dycore/physics/gas_dynamics/cfl.py:5-74
  nonhydrostasy, compressibility, is_compressible, is_nonhydrostatic
                               src/dycore/physics/gas_dynamics/eos.py:5-66
  do (one member)              src/dycore/discretisation/time_update.py:59-708

Layouts are the reference's: cell arrays [x, y] with two ghost layers, the x
face fluxes flux[0] of shape (icy, icx+1) ([y, face]) and the y face fluxes
flux[1] of shape (icx, icy+1) ([x, face]).  The reference flips its arrays so
that the split axis is last (variable.py:99-105); here the same arrays are
viewed with the split axis last (``_line``), which leaves the memory and the
arithmetic unchanged.

Deliberate differences from the reference, none of which changes a value by
more than rounding:
  * the 2x3 and 2x2 box filters that the reference evaluates with
    scipy.signal.fftconvolve (numerical_flux.py:34-44, second_projection.py:49)
    are direct weighted sums here (FFT round-off, ~1e-16 relative);
  * time_update.do is restated for one member as the reference runs it with
    one ensemble member: the blending calls (time_update.py:154-171, 531-548)
    reduce to the eos regime functions with continuous_blending and
    initial_blending off; Rayleigh damping and forcing (time_update.py:
    296-348, 478-524) and the writer are not restated.

Pinned by: the reference's own whole-run target sums for the travelling vortex
(src/tests/test_targets.yml:17-24, checked as src/tests/diagnostics.py:42-58
checks them), and by exact invariants of the advection: conservation of
rho, rhoY and momentum sums on a doubly periodic grid, preservation of a
uniform state, and x-y symmetry of a transposed state.
"""
import copy

import numpy as np

from . import boundary as bdry
from . import projection as opj

FIELDS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")


def _line(a, axis):
    """View of a 2D array with ``axis`` last (the reference's flip)."""
    return np.moveaxis(a, 0, -1) if axis == 0 else a


# ---------------------------------------------------------------------------
# numerical_flux.py
# ---------------------------------------------------------------------------
def recompute_advective_fluxes(flux, Sol):
    """numerical_flux.py:5-48 (2D): rhoY-weighted face velocities.  The x
    face between cells f-1 and f of row j averages rhoY u / rho over the 2x3
    cells (f-1..f, j-1..j+1) with weights (1/2, 1, 1/2) along y, divided by
    the weight sum 4; the y faces likewise with x and y exchanged.  Only the
    inner faces and rows are written."""
    a = Sol.rhoY * Sol.rhou / Sol.rho
    fx = (0.5 * a[:-1, :-2] + a[:-1, 1:-1] + 0.5 * a[:-1, 2:]
          + 0.5 * a[1:, :-2] + a[1:, 1:-1] + 0.5 * a[1:, 2:]) / 4.0
    flux[0].rhoY[1:-1, 1:-1] = fx.T
    b = Sol.rhoY * Sol.rhov / Sol.rho
    fy = (0.5 * b[:-2, :-1] + b[1:-1, :-1] + 0.5 * b[2:, :-1]
          + 0.5 * b[:-2, 1:] + b[1:-1, 1:] + 0.5 * b[2:, 1:]) / 4.0
    flux[1].rhoY[1:-1, 1:-1] = fy


class _Face:
    """Lefts / Rights of recovery.py:50-57 (only the members used)."""

    def __init__(self, shape):
        for k in FIELDS + ("u", "v", "w", "Y", "X"):
            setattr(self, k, np.zeros(shape))

    def primitives(self):
        """variable.py:44-89: primitives where rho != 0, zero elsewhere."""
        nz = np.nonzero(self.rho)
        for k in ("u", "v", "w", "Y", "X"):
            setattr(self, k, np.zeros_like(self.rho))
        self.u[nz] = self.rhou[nz] / self.rho[nz]
        self.v[nz] = self.rhov[nz] / self.rho[nz]
        self.w[nz] = self.rhow[nz] / self.rho[nz]
        self.Y[nz] = self.rhoY[nz] / self.rho[nz]
        self.X[nz] = self.rhoX[nz] / self.rho[nz]


def _primitives(L):
    """variable.py:44-89 on a dict of line views."""
    rho = L["rho"]
    nz = np.nonzero(rho)
    P = {}
    for k, src in (("u", "rhou"), ("v", "rhov"), ("w", "rhow"), ("Y", "rhoY"), ("X", "rhoX")):
        P[k] = np.zeros_like(rho)
        P[k][nz] = L[src][nz] / rho[nz]
    return P


def _slopes(d):
    """recovery.py:106-156 with limiter NONE (recovery.py:186-190): the slope
    of cell c in 1..n-2 is (d[c-1] + d[c]) / 2 of the forward differences; the
    end cells have zero slope."""
    s = np.zeros_like(d)
    s[..., 1:-1] = 0.5 * (d[..., :-2] + d[..., 1:-1])
    return s


def recovery(L, F, lmbda, split_step, tag):
    """recovery.py:6-104 on line views L (split axis last) and the split's
    flux container F.  Returns (Lefts, Rights)."""
    P = _primitives(L)
    if tag == "rk":
        lmbda = 0.0
    rhoY = L["rhoY"]
    u = np.zeros_like(rhoY)
    Fi = F.rhoY[1:-1, 1:-1]
    u[1:-1, 1:-1] = 0.5 * (Fi[..., :-1] + Fi[..., 1:]) / rhoY[1:-1, 1:-1]
    shape = rhoY.shape
    Lefts, Rights = _Face(shape), _Face(shape)
    diffs = {}
    for k in ("u", "v", "w", "X"):
        d = np.zeros(shape)
        d[..., :-1] = P[k][..., 1:] - P[k][..., :-1]
        diffs[k] = d
    d = np.zeros(shape)
    d[..., :-1] = 1.0 / P["Y"][..., 1:] - 1.0 / P["Y"][..., :-1]
    diffs["Y"] = d
    slopes = {k: _slopes(v) for k, v in diffs.items()}
    for k in ("u", "v", "w", "X"):
        setattr(Lefts, k, P[k] + 0.5 * slopes[k] * (1.0 - lmbda * u))
    Lefts.Y = 1.0 / (1.0 / P["Y"] + 0.5 * slopes["Y"] * (1.0 - lmbda * u))
    for k in ("u", "v", "w", "X"):
        setattr(Rights, k, P[k] + -0.5 * slopes[k] * (1.0 + lmbda * u))
    Rights.Y = 1.0 / (1.0 / P["Y"] + -0.5 * slopes["Y"] * (1.0 + lmbda * u))
    vel = (P["u"], P["v"], P["w"])[split_step]
    face = 0.5 * (rhoY[..., :-1] + rhoY[..., 1:]) \
        - 0.5 * lmbda * (vel[..., 1:] * rhoY[..., 1:] - vel[..., :-1] * rhoY[..., :-1])
    Lefts.rhoY[..., :-1] = face
    Rights.rhoY[..., 1:] = face
    for U in (Rights, Lefts):
        _get_conservatives(U)
    return Lefts, Rights


def _get_conservatives(U):
    """recovery.py:192-210."""
    U.rho = U.rhoY / U.Y
    U.rhou = U.u * U.rho
    U.rhov = U.v * U.rho
    U.rhow = U.w * U.rho
    U.rhoY = U.Y * U.rho
    U.rhoX = U.X * U.rho


def hll_solver(F, Lefts, Rights):
    """numerical_flux.py:50-99: upwinded face fluxes of rho, rhou, rhov, rhow
    and rhoX from the advective rhoY flux (faces 1..n-1)."""
    Lefts.primitives()
    Rights.primitives()
    upwind = 0.5 * (1.0 + np.sign(F.rhoY))
    upl = upwind[..., 1:]
    upr = 1.0 - upwind[..., :-1]
    Fy = F.rhoY[..., 1:-1]
    a = upl[..., :-1] / Lefts.Y[..., :-1]
    b = upr[..., 1:] / Rights.Y[..., 1:]
    F.rhou[..., 1:-1] = Fy * (a * Lefts.u[..., :-1] + b * Rights.u[..., 1:])
    F.rho[..., 1:-1] = Fy * (a * 1.0 + b * 1.0)
    F.rhov[..., 1:-1] = Fy * (a * Lefts.v[..., :-1] + b * Rights.v[..., 1:])
    F.rhow[..., 1:-1] = Fy * (a * Lefts.w[..., :-1] + b * Rights.w[..., 1:])
    F.rhoX[..., 1:-1] = Fy * (a * Lefts.X[..., :-1] + b * Rights.X[..., 1:])
    return F


# ---------------------------------------------------------------------------
# explicit.py
# ---------------------------------------------------------------------------
def explicit_step_and_flux(Sol, F, lmbda, elem, split_step, ud, th, mpv, tag=None):
    """explicit.py:75-165: ghost cells along the split axis, recovery, HLL,
    the conservative update (not for tag 'rk'), ghost cells again."""
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv, step=split_step)
    L = {k: _line(getattr(Sol, k), split_step) for k in FIELDS}
    Lefts, Rights = recovery(L, F, lmbda, split_step, tag)
    hll_solver(F, Lefts, Rights)
    if tag != "rk":
        for k in FIELDS:
            f = getattr(F, k)
            L[k] += lmbda * (f[..., :-1] - f[..., 1:])
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv, step=split_step)
    return F


def advect(Sol, flux, dt, elem, odd, ud, th, mpv):
    """explicit.py:5-72: double Strang sweep with half steps dt/2; x then y
    then y then x on odd steps, y x x y on even steps."""
    time_step = 0.5 * dt
    order = (0, 1) if odd else (1, 0)
    for split in order + order[::-1]:
        if elem.iisc[split] > 1:
            explicit_step_and_flux(Sol, flux[split], time_step / elem.dxyz[split], elem, split, ud, th, mpv)
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)


def advect_rk(Sol, flux, dt, elem, odd, ud, th, mpv):
    """explicit.py:168-257: one-stage Runge-Kutta: both directions' fluxes
    from the same state (recovery without the time-centring term), then the
    x update and the y update."""
    for split in range(elem.ndim):
        if elem.iisc[split] > 1:
            explicit_step_and_flux(Sol, flux[split], dt / elem.dxyz[split], elem, split, ud, th, mpv, tag="rk")
    for dim in range(elem.ndim):
        lmbda = dt / elem.dxyz[dim]
        for k in FIELDS:
            f = getattr(flux[dim], k)
            _line(getattr(Sol, k), dim)[...] += lmbda * (f[..., :-1] - f[..., 1:])
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)


# ---------------------------------------------------------------------------
# second_projection.py:22-93
# ---------------------------------------------------------------------------
def _mod_bg_wind(Sol, ud, fac):
    """variable.py:107-114."""
    Sol.rhou = Sol.rhou + fac * ud.u_wind_speed * Sol.rho
    Sol.rhov = Sol.rhov + fac * ud.v_wind_speed * Sol.rho
    Sol.rhow = Sol.rhow + fac * ud.w_wind_speed * Sol.rho


def euler_backward_non_advective_expl_part(Sol, mpv, elem, dt, ud, th):
    """second_projection.py:72-90: buoyancy, then the inverse of the implicit
    Coriolis/buoyancy matrix on the momenta without the background wind."""
    nonhydro = ud.nonhydrostasy
    g = ud.gravity_strength[1]
    Msq = ud.Msq
    dbuoy = Sol.rhoY * (Sol.rhoX / Sol.rho)
    Sol.rhov = (nonhydro * Sol.rhov) - dt * (g / Msq) * dbuoy
    _mod_bg_wind(Sol, ud, -1.0)
    opj.multiply_inverse_coriolis(Sol, Sol, mpv, ud, elem, elem, dt)
    _mod_bg_wind(Sol, ud, +1.0)
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)


def box_mean2x2(a):
    """2x2 box mean of a cell array -> node inner array (the 'valid'
    convolution with ones((2, 2)) / 4 of second_projection.py:47-49)."""
    return (a[:-1, :-1] + a[:-1, 1:] + a[1:, :-1] + a[1:, 1:]) / 4.0


def euler_forward_non_advective(Sol, mpv, elem, node, dt, ud, th):
    """second_projection.py:22-69: explicit pressure-gradient, buoyancy and
    Coriolis forcing, rhoX reset to the perturbation of 1/theta advected by
    v, and the compressible Exner increment (weight ud.compressibility)."""
    nonhydro = ud.nonhydrostasy
    g = ud.gravity_strength[1]
    Msq = ud.Msq
    Ginv = th.Gammainv
    corr_h1, corr_v, corr_h2 = ud.coriolis_strength
    u0, v0, w0 = ud.u_wind_speed, ud.v_wind_speed, ud.w_wind_speed
    p2n = np.copy(mpv.p2_nodes)
    dp2n = np.zeros_like(p2n)
    ndim = elem.ndim
    S0c = np.repeat(mpv.HydroState.S0[None, :], elem.sc[0], axis=0)
    dSdy = opj.dSdy_cells(mpv, elem, node)
    mpv.rhs[...] = opj.divergence_nodes(mpv.rhs, elem, node, Sol, ud)
    opj.scale_wall_node_values(mpv.rhs, node, ud, 2.0)
    div = mpv.rhs
    rhoY = Sol.rhoY ** (th.gamm - 2.0)
    dpidP = (th.gm1 / ud.Msq) * box_mean2x2(rhoY)
    rhoYovG = Ginv * Sol.rhoY
    dbuoy = Sol.rhoY * (Sol.rhoX / Sol.rho)
    dpdx, dpdy, dpdz = opj.grad_nodes(p2n, elem.ndim, (node.dx, node.dy, node.dz))
    drhou = Sol.rhou - u0 * Sol.rho
    drhov = Sol.rhov - v0 * Sol.rho
    drhow = Sol.rhow - w0 * Sol.rho
    v = Sol.rhov / Sol.rho
    Sol.rhou = Sol.rhou - dt * (rhoYovG * dpdx - corr_h2 * drhov + corr_v * drhow)
    Sol.rhov = Sol.rhov - dt * (rhoYovG * dpdy + (g / Msq) * dbuoy * nonhydro - corr_h1 * drhow
                                + corr_h2 * drhou) * (1 - ud.is_ArakawaKonor)
    Sol.rhow = Sol.rhow - dt * (rhoYovG * dpdz - corr_v * drhou + corr_h1 * drhov) * (ndim == 3)
    Sol.rhoX = (Sol.rho * (Sol.rho / Sol.rhoY - S0c)) - dt * (v * dSdy) * Sol.rho
    dp2n[1:-1, 1:-1] -= dt * dpidP * div
    mpv.p2_nodes += ud.compressibility * dp2n
    bdry.set_ghostnodes_p2(mpv.p2_nodes, node, ud)
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)


# ---------------------------------------------------------------------------
# cfl.py, eos.py
# ---------------------------------------------------------------------------
def dynamic_timestep(Sol, time, time_output, elem, ud, th, step):
    """cfl.py:5-74 (acoustic_timestep == 0 branch unless set): returns
    (dt, cfl, cfl_ac); every maximum runs over the whole cell array."""
    eps = np.finfo(float).eps
    gamm = th.gamm
    Minv = 1.0 / np.sqrt(ud.Msq)
    CFL = ud.CFL
    p = Sol.rhoY ** gamm
    c = np.sqrt(gamm * p / Sol.rho) * Minv
    u = np.abs(Sol.rhou / Sol.rho)
    v = np.abs(Sol.rhov / Sol.rho)
    w = np.abs(Sol.rhow / Sol.rho)
    u_max, v_max, w_max = max(u.max(), eps), max(v.max(), eps), max(w.max(), eps)
    upc, vpc, wpc = max((u + c).max(), eps), max((v + c).max(), eps), max((w + c).max(), eps)
    if getattr(ud, "acoustic_timestep", 0) == 1:
        dt_cfl = min(min(CFL * elem.dx / upc, CFL * elem.dy / vpc), CFL * elem.dz / wpc)
        dt = min(dt_cfl, ud.dtfixed0 + min(step, 1.0) * (ud.dtfixed - ud.dtfixed0))
        if dt > time_output - time:
            dt = (time_output - time) + eps
        return dt, None, None
    dt_cfl = min(min(CFL * elem.dx / u_max, CFL * elem.dy / v_max), CFL * elem.dz / w_max)
    dt = min(dt_cfl, ud.dtfixed0 + min(step, 1.0) * (ud.dtfixed - ud.dtfixed0))
    dt *= min(float(step + 1), 1.0)
    if dt > time_output - time:
        dt = time_output - time
    cfl = CFL * dt / dt_cfl
    cfl_ac = max(dt * upc / elem.dx, dt * vpc / elem.dy, dt * wpc / elem.dz)
    return dt, cfl, cfl_ac


def is_compressible(ud, step):
    """eos.py:38-50 with continuous_blending off."""
    return ud.is_compressible


def compressibility(ud, t, step):
    """eos.py:20-36 with continuous_blending off."""
    if step >= 0:
        return 0.0 if ud.is_compressible == 0 else 1.0
    return ud.compressibility


def is_nonhydrostatic(ud, step):
    """eos.py:53-66 with continuous_blending off."""
    return ud.is_nonhydrostatic


def nonhydrostasy(ud, t, step):
    """eos.py:5-18 for is_nonhydrostatic in {0, 1}."""
    if step >= 0:
        return float(ud.is_nonhydrostatic)
    return float(ud.is_nonhydrostatic)


# ---------------------------------------------------------------------------
# time_update.py
# ---------------------------------------------------------------------------
class StepRecord:
    """What one time step did: dt, cfl numbers and both solver results."""

    def __init__(self, step, t, dt, cfl, cfl_ac):
        self.step, self.t, self.dt, self.cfl, self.cfl_ac = step, t, dt, cfl, cfl_ac
        self.solves = []


def do(Sol, flux, mpv, t, tout, ud, elem, node, steps, th, project=None, record=None):
    """time_update.py:59-708 for one member: time steps from t until tout or
    ud.stepmax.  ``project`` is the implicit projection (default: the
    oracle's euler_backward_non_advective_impl_part).  Returns
    [Sol, flux, mpv, [window_step, step]] as the reference; ``record`` (a
    list) receives a StepRecord per step."""
    if project is None:
        project = opj.euler_backward_non_advective_impl_part
    window_step, step = steps
    while t < tout and step < ud.stepmax:
        bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
        dt, cfl, cfl_ac = dynamic_timestep(Sol, t, tout, elem, ud, th, step)
        if "CFLfixed" in getattr(ud, "aux", ""):
            if step < 2:
                dt = 21.69 / ud.t_ref
        # blending_before_timestep with blending off (blending.py:386-545)
        ud.is_compressible = is_compressible(ud, window_step)
        ud.compressibility = compressibility(ud, t, window_step)
        ud.is_nonhydrostatic = is_nonhydrostatic(ud, window_step)
        ud.nonhydrostasy = nonhydrostasy(ud, t, window_step)
        # time_update.py:173-174
        ud.is_nonhydrostatic = is_nonhydrostatic(ud, window_step)
        ud.nonhydrostasy = nonhydrostasy(ud, t, window_step)
        rec = StepRecord(step, t, dt, cfl, cfl_ac)
        Sol0 = copy.deepcopy(Sol)
        recompute_advective_fluxes(flux, Sol)
        if getattr(ud, "do_advection", True):
            advect_rk(Sol, flux, 0.5 * dt, elem, step % 2, ud, th, mpv)
        mpv.p2_nodes0[...] = mpv.p2_nodes
        euler_backward_non_advective_expl_part(Sol, mpv, elem, 0.5 * dt, ud, th)
        if ud.is_compressible == 0:
            rec.solves.append(project(Sol, mpv, elem, node, ud, th, t, 0.5 * dt, 1.0, Sol0=Sol0))
        else:
            rec.solves.append(project(Sol, mpv, elem, node, ud, th, t, 0.5 * dt, 1.0))
        recompute_advective_fluxes(flux, Sol)
        half = {k: np.copy(getattr(Sol, k)) for k in FIELDS}
        p2_half = np.copy(mpv.p2_nodes)
        if ud.is_compressible == 1 and ud.is_nonhydrostatic == 1:
            mpv.p2_nodes[...] = mpv.p2_nodes0
            Sol = copy.deepcopy(Sol0)
        elif ud.is_nonhydrostatic == 0:
            Sol = copy.deepcopy(Sol0)
            mpv.p2_nodes[...] = mpv.p2_nodes0
        elif ud.is_compressible == 0:
            Sol = copy.deepcopy(Sol0)
        for k in FIELDS:
            setattr(Sol, k + "_half", half[k])
        mpv.p2_nodes_half = p2_half
        euler_forward_non_advective(Sol, mpv, elem, node, 0.5 * dt, ud, th)
        if getattr(ud, "do_advection", True):
            advect(Sol, flux, dt, elem, step % 2, ud, th, mpv)
        euler_backward_non_advective_expl_part(Sol, mpv, elem, 0.5 * dt, ud, th)
        rec.solves.append(project(Sol, mpv, elem, node, ud, th, t, 0.5 * dt, 2.0))
        t += dt
        step += 1
        window_step += 1
        if record is not None:
            record.append(rec)
    return [Sol, flux, mpv, [window_step, step]]
