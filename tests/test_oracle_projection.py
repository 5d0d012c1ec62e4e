"""CPU pins of the oracle's projection pieces and ghost fills
(second_projection.py:280-306, 496-575, 604-652; boundary.py:7-132, 296-369).

* grad_nodes: zero on constants and on the checkerboard, exact on linear fields;
* divergence_nodes: exact on linear momenta, and the discrete divergence
  theorem (trapezoid- and checkerboard-weighted sums of the node divergence
  vanish after the ghost fill) -- the consistency the singular solve relies on;
* multiply_inverse_coriolis: its coefficients invert the implicit
  Coriolis/buoyancy matrix [[1, -wh2, wv], [wh2, nu + nh, -wh1], [-wv, wh1, 1]];
* ghost fills: periodic axes reproduce a periodic function at the ghost
  coordinates, the gravity-axis image mirrors the source rows about the wall
  face (rhov flipped) and rebuilds the hydrostatic background.
"""
import copy

import numpy as np
import pytest

from oracle import boundary as obd
from oracle import projection as opj
from pybella_amd import workloads as W

from _cases import make_case


def test_grad_nodes_constant_checkerboard_linear():
    Sol, mpv, elem, node, ud, th = make_case("strat_psinc", 17, 9)
    d = (node.dx, node.dy, node.dz)
    shp = node.sc[:2]
    c = np.where((np.arange(shp[0])[:, None] + np.arange(shp[1])[None, :]) % 2 == 0, 1.0, -1.0)
    for p in (np.full(shp, 3.7), c):
        gx, gy, _ = opj.grad_nodes(p, 2, d)
        assert np.max(np.abs(gx)) < 1e-12 and np.max(np.abs(gy)) < 1e-12
    X, Y = np.meshgrid(node.x, node.y, indexing="ij")
    gx, gy, _ = opj.grad_nodes(2.5 * X - 1.25 * Y, 2, d)
    assert np.max(np.abs(gx - 2.5)) < 1e-11 and np.max(np.abs(gy + 1.25)) < 1e-11


def test_divergence_linear_momenta():
    Sol, mpv, elem, node, ud, th = make_case("strat_comp", 17, 9)
    ud.bdry_type[1] = W.BdryType.PERIODIC     # no wall-row zeroing for this closed form
    Sol.rho[...] = 1.0
    Sol.rhoY[...] = 1.0
    X, Y = np.meshgrid(elem.x, elem.y, indexing="ij")
    Sol.rhou[...] = 0.8 * X + 0.1 * Y
    Sol.rhov[...] = -0.3 * Y + 0.2 * X
    div = opj.divergence_nodes(np.zeros(node.isc[:2]), elem, node, Sol, ud)
    assert np.max(np.abs(div - 0.5)) < 1e-11


@pytest.mark.parametrize("seed", [5, 6])
def test_divergence_theorem(seed):
    Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=seed)
    obd.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    div = opj.divergence_nodes(np.zeros(node.isc[:2]), elem, node, Sol, ud)
    r = div[1:-1, 1:-1]                                # (NX, NY) solve nodes
    nx, ny = r.shape
    w = np.ones((nx, ny))
    w[0, :] = w[-1, :] = 0.5
    c = np.where((np.arange(nx)[:, None] + np.arange(ny)[None, :]) % 2 == 0, 1.0, -1.0)
    scale = np.abs(r).sum()
    assert scale > 1.0
    assert abs((w * r).sum()) < 1e-13 * scale
    assert abs((w * c * r).sum()) < 1e-13 * scale


def test_inverse_coriolis_closed_form():
    rng = np.random.default_rng(0)
    Sol, mpv, elem, node, ud, th = make_case("coriolis", 17, 9)
    ud.coriolis_strength = np.array([0.4, 0.9, 0.2])
    ud.nonhydrostasy = 1.0
    dt = 0.3
    wh1, wv, wh2 = dt * ud.coriolis_strength
    U = rng.standard_normal(Sol.rho.shape)
    V = rng.standard_normal(Sol.rho.shape)
    Wm = rng.standard_normal(Sol.rho.shape)

    class Vec:
        pass
    v = Vec()
    v.rhou, v.rhov, v.rhow = U.copy(), V.copy(), Wm.copy()
    opj.multiply_inverse_coriolis(v, Sol, mpv, ud, elem, node, dt)
    Y = Sol.rhoY / Sol.rho
    nu = -dt ** 2 * (ud.gravity_strength[1] / ud.Msq) * opj.dSdy_cells(mpv, elem, node) * Y
    nh = ud.nonhydrostasy
    # K (u, v, w)_new == (u, v, w)_old
    ru = v.rhou - wh2 * v.rhov + wv * v.rhow
    rv = wh2 * v.rhou + (nu + nh) * v.rhov - wh1 * v.rhow
    rw = -wv * v.rhou + wh1 * v.rhov + v.rhow
    for a, b in ((ru, U), (rv, V), (rw, Wm)):
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))


def test_correction_nodes_linear_pressure():
    """No Coriolis, nh = 1: rhou += -dt Gammainv rhoY Y a / (1) * rho/rhoY for
    p2 = a x + b y (grad exact), rhov gets the buoyancy factor 1/(1 + nu)."""
    Sol, mpv, elem, node, ud, th = make_case("strat_comp", 17, 9)
    ud.coriolis_strength = np.zeros(3)
    ud.nonhydrostasy = 1.0
    dt = 0.05
    X, Y = np.meshgrid(node.x, node.y, indexing="ij")
    p = 0.7 * X - 0.4 * Y
    S = copy.deepcopy(Sol)
    opj.correction_nodes(S, elem, node, mpv, p, dt, ud, th, 1.0)
    Yc = Sol.rhoY / Sol.rho
    coeff = th.Gammainv * Sol.rhoY * Yc
    nu = -dt ** 2 * (ud.gravity_strength[1] / ud.Msq) * opj.dSdy_cells(mpv, elem, node) * Yc
    du = (1.0 / Yc) * (-dt * coeff * 0.7)
    dv = (1.0 / Yc) * (-dt * coeff * -0.4) / (1.0 + nu)
    assert np.max(np.abs((S.rhou - Sol.rhou) - du)) < 1e-12 * np.max(np.abs(du))
    assert np.max(np.abs((S.rhov - Sol.rhov) - dv)) < 1e-12 * np.max(np.abs(dv))
    dX = -dt * opj.dSdy_cells(mpv, elem, node) * S.rhov
    assert np.max(np.abs((S.rhoX - Sol.rhoX) - dX)) < 1e-12 * np.max(np.abs(dX))


def test_ghostnodes_periodic_and_reflect():
    """x periodic with a duplicated end node: ghosts carry the periodic
    function at their coordinates; y reflect: ghosts mirror about the wall node."""
    Sol, mpv, elem, node, ud, th = make_case("strat_psinc", 17, 9)
    X, Y = np.meshgrid(node.x, node.y, indexing="ij")
    x0, x1 = node.x[2], node.x[-3]
    y0 = node.y[2]
    L = x1 - x0
    H = node.y[-3] - y0
    f = lambda x, y: np.sin(2 * np.pi * (x - x0) / L + 0.3) * np.cos(np.pi * (y - y0) / H) + \
        0.5 * np.cos(4 * np.pi * (x - x0) / L) * np.cos(2 * np.pi * (y - y0) / H)
    p = np.zeros(node.sc[:2])
    p[2:-2, 2:-2] = f(X, Y)[2:-2, 2:-2]
    obd.set_ghostnodes_p2(p, node, ud)
    assert np.max(np.abs(p - f(X, Y))) < 1e-12


def test_cells_periodic_wrap_and_gravity_image():
    Sol, mpv, elem, node, ud, th = make_case("strat_psinc", 33, 17)
    ref = copy.deepcopy(Sol)
    X, Yc = np.meshgrid(elem.x, elem.y, indexing="ij")
    Lx = elem.x[-3] - elem.x[1]             # NX cells span the period
    g = lambda x: np.cos(2 * np.pi * (x - elem.x[2]) / Lx + 0.2)
    Sol.rhou[...] = g(X) * (1 + Yc)
    Sol.rhou[:2] = 0.0
    Sol.rhou[-2:] = 0.0
    obd.set_explicit_boundary_data(Sol, elem, ud, th, mpv)
    inner = (slice(None), slice(2, -2))
    assert np.max(np.abs(Sol.rhou[inner] - (g(X) * (1 + Yc))[inner])) < 1e-12
    # gravity axis, bottom: image row 1 mirrors row 2, row 0 mirrors row 3
    for img, src in ((1, 2), (0, 3), (-2, -3), (-1, -4)):
        Yimg = Sol.rhoY[:, img] / Sol.rho[:, img]
        Ysrc = Sol.rhoY[:, src] / Sol.rho[:, src]
        assert np.allclose(Sol.rhov[:, img] * Yimg, -Sol.rhov[:, src] * Ysrc, rtol=1e-13, atol=1e-15)
        assert np.allclose(Sol.rhou[:, img] / Sol.rho[:, img], Sol.rhou[:, src] / Sol.rho[:, src], rtol=1e-13)
        assert np.allclose(Sol.rhoX[:, img] / Sol.rho[:, img], Sol.rhoX[:, src] / Sol.rho[:, src], rtol=1e-13)
        # psinc: the ghost background is the hydrostatic rhoY0 of the image row
        assert np.allclose(Sol.rhoY[:, img], mpv.HydroState.rhoY0[img % elem.icy], rtol=1e-14)
