"""Initial states the reference's regression tests start from (oracle; TEST
INFRASTRUCTURE ONLY).

Restates, following the reference's operation order (pyBELLA tree):
  hydrostatic_column     src/dycore/physics/hydrostatics.py:5-76 (column)
  initial_pressure       src/dycore/physics/hydrostatics.py:134-240 with the
                         recursion of __loop_over_array (:242-246)
  internal_long_wave     sol_init of src/tests/test_internal_long_wave.py:182-251

Pinned by: the internal long wave's 31 steps reproducing the reference's
regression sums (src/tests/test_targets.yml:1-8; tests/test_oracle_timestep.py).
"""
import numpy as np

from . import boundary as bdry


class _Column:
    """var.States(node.sc) as the reference allocates HySt / HyStn: every
    member a zero array of the node-grid shape."""

    def __init__(self, shape):
        for k in ("p0", "p20", "rho0", "S0", "S10", "rhoY0", "Y0"):
            setattr(self, k, np.zeros(shape))


def hydrostatic_column(HydroState, HydroState_n, Y, Y_n, elem, node, th, ud):
    """hydrostatics.py:5-76: the hydrostatic state of every column of a
    2D potential-temperature field Y (cells) / Y_n (nodes), pi = 1 at the
    bottom node row."""
    Gamma = th.gm1 / th.gamm
    gamm = th.gamm
    Gamma_inv = 1.0 / Gamma
    gm1_inv = 1.0 / th.gm1
    icy, igy = elem.icy, elem.igy
    xc_idx = slice(0, -1)
    c_idx = (slice(0, -1), slice(0, -1))
    rhoY0 = 1.0
    g = ud.gravity_strength[1]
    p0 = rhoY0 ** gamm
    pi0 = rhoY0 ** th.gm1
    HydroState_n.rho0[xc_idx, igy] = rhoY0 / Y_n[:, igy]
    HydroState_n.rhoY0[xc_idx, igy] = rhoY0
    HydroState_n.Y0[xc_idx, igy] = Y_n[:, igy]
    HydroState_n.S0[xc_idx, igy] = 1.0 / Y_n[:, igy]
    HydroState_n.p0[xc_idx, igy] = p0
    HydroState_n.p20[xc_idx, igy] = pi0 / ud.Msq

    dys = np.array([-elem.dy] + [-elem.dy / 2] + [elem.dy / 2] + list(np.ones((icy - 3)) * elem.dy))
    S_p = 1.0 / Y[:, :]
    S_m = np.zeros_like(S_p)
    S_m[:, igy - 1:igy + 1] = 1.0 / Y_n[:, igy].reshape(-1, 1)
    S_m[:, 0] = 1.0 / Y[:, igy - 1]
    S_m[:, igy + 1:] = 1.0 / Y[:, igy:-1]
    S_integral_p = dys * 0.5 * (S_p + S_m)
    S_integral_p[:, :igy] = np.cumsum(S_integral_p[:, :igy][:, ::-1], axis=1)[:, ::-1]
    S_integral_p[:, igy:] = np.cumsum(S_integral_p[:, igy:], axis=1)
    pi_hydro = pi0 - Gamma * g * S_integral_p
    p_hydro = pi_hydro ** Gamma_inv
    rhoY_hydro = pi_hydro ** gm1_inv
    HydroState.rho0[c_idx] = rhoY_hydro * S_p
    HydroState.p0[c_idx] = p_hydro
    HydroState.p20[c_idx] = pi_hydro / ud.Msq
    HydroState.S0[c_idx] = S_p
    HydroState.S10[c_idx] = 0.0
    HydroState.Y0[c_idx] = 1.0 / S_p
    HydroState.rhoY0[c_idx] = rhoY_hydro

    Sn_p = 1.0 / Y[:, :]
    dys = np.ones((icy)) * elem.dy
    dys[:igy] *= -1
    Sn_integral_p = dys * Sn_p
    Sn_integral_p[:, :igy] = np.cumsum(Sn_integral_p[:, :igy][:, ::-1], axis=1)[:, ::-1]
    Sn_integral_p[:, igy:] = np.cumsum(Sn_integral_p[:, igy:], axis=1)
    pi_hydro_n = pi0 - Gamma * g * Sn_integral_p
    rhoY_hydro_n = pi_hydro_n ** gm1_inv
    HydroState_n.rhoY0[xc_idx, :igy] = rhoY_hydro_n[:, :igy]
    HydroState_n.Y0[xc_idx, :igy] = Y_n[0, :igy]
    HydroState_n.S0[xc_idx, :igy] = 1.0 / Y_n[:, :igy]
    HydroState_n.p0[xc_idx, :igy] = rhoY_hydro_n[:, :igy] ** th.gamm
    HydroState_n.p20[xc_idx, :igy] = pi_hydro_n[:, :igy] / ud.Msq
    HydroState_n.rhoY0[xc_idx, igy + 1:] = rhoY_hydro_n[:, igy:]
    HydroState_n.Y0[xc_idx, igy + 1:] = Y_n[0, igy:]
    HydroState_n.S0[xc_idx, igy + 1:] = 1.0 / Y_n[:, igy:]
    HydroState_n.p0[xc_idx, igy + 1:] = rhoY_hydro_n[:, igy:] ** th.gamm
    HydroState_n.p20[xc_idx, igy + 1:] = pi_hydro_n[:, igy:] / ud.Msq


def initial_pressure(Sol, mpv, elem, node, ud, th):
    """hydrostatics.py:134-240: balance the horizontal pressure gradient of
    the columns (cells, then nodes by the alternating recursion), then reset
    rhoY and rho from the cell Exner pressure."""
    Gammainv = th.Gammainv
    igy, igx = node.igy, node.igx
    icx = elem.icx
    dx, dy = node.dx, node.dy

    beta = np.zeros((node.icx))
    bdpdx = np.zeros((node.icx))
    x_idx_m = slice(0, -1)
    x_idx_c = slice(1, None)
    y_idx = slice(igy, -igy)
    xn_idx = slice(1, -1)
    height = node.y[-igy - 1]
    Pc = Sol.rhoY[x_idx_c, y_idx]
    Pm = Sol.rhoY[x_idx_m, y_idx]
    thc = Pc / Sol.rho[x_idx_c, y_idx]
    thm = Pm / Sol.rho[x_idx_m, y_idx]
    beta[xn_idx] = np.sum(0.5 * (Pm * thm + Pc * thc) * dy, axis=1)
    bdpdx[xn_idx] = np.sum(0.5 * (Pm * thm + Pc * thc) *
                           (mpv.p2_cells[x_idx_c, y_idx] - mpv.p2_cells[x_idx_m, y_idx]) * dy, axis=1)
    beta *= Gammainv / height
    bdpdx *= Gammainv / height / dx
    coeff = np.zeros((elem.icx))
    pibot = np.zeros((elem.icx))
    coeff[igx + 1:-igx + 1] = np.cumsum(coeff[igx:-igx] + dx / beta[igx + 1:-igx])
    pibot[igx + 1:-igx + 1] = np.cumsum(pibot[igx:-igx] - dx * bdpdx[igx + 1:-igx] / beta[igx + 1:-igx])
    dotPU = pibot[icx - igx] / coeff[icx - igx]
    pibot[igx:-igx] -= dotPU * coeff[igx:-igx]
    x_idx = slice(igx, -igx + 1)
    y_idx = slice(igy, -igy + 1)
    mpv.p2_cells[x_idx, y_idx] += pibot[x_idx].reshape(-1, 1) - 1.0 * mpv.HydroState.p20[y_idx].reshape(1, -1)
    bdry.set_ghostcells_p2(mpv.p2_cells, elem, ud)

    icxn, icyn = node.icx, node.icy
    y_idx = slice(igy, -igy + 1)
    height = node.y[-igy]
    Pc = Sol.rhoY[1:, y_idx]
    thc = Pc / Sol.rho[1:, y_idx]
    beta = np.zeros((elem.icx,))
    bdpdx = np.zeros((elem.icx))
    beta[1:] = np.sum(Pc * thc * dy, axis=1)
    beta *= Gammainv / height
    bdpdx[1:] = np.sum(Pc * thc * (mpv.p2_nodes[1:-1, igy:-igy] - mpv.p2_nodes[:-2, igy:-igy]) * dy, axis=1)
    bdpdx *= Gammainv / height / dx
    coeff = np.zeros((node.icx))
    pibot = np.zeros((node.icx))
    coeff[igx + 1:-igx + 1] = np.cumsum(coeff[igx:-igx] + dx / beta[igx + 1:])
    pibot[igx + 1:-igx + 1] = np.cumsum(pibot[igx:-igx] - dx * bdpdx[igx + 1:] / beta[igx + 1:])
    dotPU = pibot[icx - igx] / coeff[icx - igx]
    pibot[igx:-igx] -= dotPU * coeff[igx:-igx]
    x_idx = slice(igx, -igx + 1)
    y_idx = slice(igy, -igy + 1)
    mpv.p2_nodes[x_idx, y_idx] += pibot[x_idx].reshape(-1, 1) - 1.0 * mpv.HydroState_n.p20[y_idx].reshape(1, -1)
    mpv.dp2_nodes[:, :] = mpv.p2_nodes
    mpv.p2_nodes[igx, igy:-igy] = mpv.dp2_nodes[igx, igy:-igy]
    p, dp = mpv.p2_nodes, mpv.dp2_nodes
    for j in range(igy, icyn - igy):          # __loop_over_array (hydrostatics.py:242-246)
        for i in range(igx + 1, icxn - igx):
            p[i, j] = 2.0 * dp[i - 1, j] - p[i - 1, j]
    assert ((node.icx + 1) % 2) == 1
    delp2 = 0.5 * (mpv.p2_nodes[-igx - 1, igy:-igy] - mpv.p2_nodes[igx, igy:-igy])
    delp2 = delp2.reshape(1, -1)
    sgn = np.ones_like(mpv.p2_nodes[:, 0][igy:-igy]).reshape(-1, 1)
    sgn[1::2] *= -1
    mpv.p2_nodes[igx:-igx, igy:-igy] += sgn * delp2
    bdry.set_ghostnodes_p2(mpv.p2_nodes, node, ud)
    mpv.dp2_nodes[:, :] = 0.0

    inner = (slice(igx, -igx), slice(igy, -igy))
    pi = ud.Msq * (mpv.p2_cells[inner] + 1.0 * mpv.HydroState.p20[igy:-igy])
    Y = Sol.rhoY[inner] / Sol.rho[inner]
    rhoold = np.copy(Sol.rho[inner])
    Sol.rhoY[inner] = pi ** th.gm1inv
    Sol.rho[inner] = Sol.rhoY[inner] / Y
    Sol.rhou[inner] *= Sol.rho[inner] / rhoold
    Sol.rhov[inner] *= Sol.rho[inner] / rhoold
    Sol.rhow[inner] *= Sol.rho[inner] / rhoold
    Sol.rhoX[inner] *= Sol.rho[inner] / rhoold


def internal_long_wave(Sol, mpv, elem, node, th, ud, hydrostatic_state):
    """sol_init of src/tests/test_internal_long_wave.py:182-251 (no seeds):
    a long internal gravity wave (theta perturbation 0.01 K under a
    molly-windowed sin(pi y) / (1 + x^2/a^2) envelope) on a stably
    stratified column in hydrostatic and horizontal pressure balance.
    ``hydrostatic_state`` is the isothermal background routine
    (hydrostatics.py:79-132).  Mutates and returns Sol."""
    u0, v0, w0 = ud.u_wind_speed, ud.v_wind_speed, ud.w_wind_speed
    delth = 0.01 / ud.T_ref
    xc = 0.0
    a = ud.scale_factor * 5.0e3 / ud.h_ref
    hydrostatic_state(mpv, elem, node, th, ud)
    HySt = _Column(node.sc[:2])
    HyStn = _Column(node.sc[:2])
    x = elem.x.reshape(-1, 1)
    y = elem.y.reshape(1, -1)
    Y = ud.stratification(y) + delth * ud.molly(x) * np.sin(np.pi * y) / (1.0 + (x - xc) ** 2 / (a ** 2))
    xn = node.x[:-1].reshape(-1, 1)
    yn = node.y[:-1].reshape(1, -1)
    Yn = ud.stratification(yn) + delth * ud.molly(xn) * np.sin(np.pi * yn) / (1.0 + (xn - xc) ** 2 / (a ** 2))
    hydrostatic_column(HySt, HyStn, Y, Yn, elem, node, th, ud)
    x_idx = slice(None)
    y_idx = slice(elem.igy, -elem.igy + 1)
    c_idx = (slice(0, -1), slice(0, -1))
    if ud.is_compressible:
        rhoY = HySt.rhoY0[:, y_idx][c_idx]
    else:
        rhoY = mpv.HydroState.rhoY0[y_idx]
    rho = rhoY / Y[:, y_idx]
    Sol.rho[x_idx, y_idx] = rho
    Sol.rhou[x_idx, y_idx] = rho * u0
    Sol.rhov[x_idx, y_idx] = rho * v0
    Sol.rhow[x_idx, y_idx] = rho * w0
    Sol.rhoY[x_idx, y_idx] = rhoY
    mpv.p2_cells[x_idx, y_idx] = HySt.p20[x_idx, y_idx][c_idx]
    Sol.rhoX[x_idx, y_idx] = Sol.rho[x_idx, y_idx] * (1.0 / Y[:, y_idx] - mpv.HydroState.S0[y_idx])
    mpv.p2_nodes[:, elem.igy:-elem.igy] = HyStn.p20[:, elem.igy:-elem.igy]
    initial_pressure(Sol, mpv, elem, node, ud, th)
    ud.nonhydrostasy = 1.0 if ud.is_nonhydrostatic == 1 else 0.0
    ud.compressibility = 1.0 if ud.is_compressible == 1 else 0.0
    if "imbal" in ud.aux:
        mpv.p2_nodes[...] = 0.0
    bdry.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    return Sol
