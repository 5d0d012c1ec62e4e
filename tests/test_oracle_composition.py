"""CPU pins of the oracle's composition of the whole projection
(oracle/projection.py: assemble_system, finish_projection,
euler_backward_non_advective_impl_part; second_projection.py:97-277), by
what the mathematics fixes rather than by the oracle's own output.

* The pseudo-incompressible projection makes the corrected momenta
  divergence free up to the solver's residual: the correction adds
  -dt Gamma^-1 rhoY Y grad(p) to Y rhou, and the nodal operator is the
  discrete div(hx grad) with that same coefficient, so divergence_nodes of
  the corrected state equals dt (rhs - A p), i.e. it falls by about the
  solve's relative residual.  A wrong sign, a wrong dt power, a transposed
  scatter or a coefficient from another state would leave it O(1).
* With Sol0 given, the operator's coefficients (wplus, wcenter) are Sol0's
  while the correction uses Sol's (second_projection.py:113-118, 280-306):
  checked against the closed form Gamma^-1 rhoY0 Y0 on a Sol0 that differs
  from Sol.
"""
import copy

import numpy as np

from oracle import boundary as obd
from oracle import projection as opj
from pybella_amd import workloads as W


def _divergence(Sol, elem, node, ud, th, mpv):
    S = copy.deepcopy(Sol)
    obd.set_explicit_boundary_data(S, elem, ud, th, mpv)
    d = opj.divergence_nodes(np.zeros(node.isc[:2]), elem, node, S, ud)
    return d[1:-1, 1:-1]


def test_projection_removes_the_divergence():
    Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=9)
    ud.tol = 1e-12
    ud.max_iterations = 20000
    dt = 0.05
    before = _divergence(Sol, elem, node, ud, th, mpv)
    res = opj.euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, 0.0, dt, 1.0)
    assert res.info == 0
    after = _divergence(Sol, elem, node, ud, th, mpv)
    ratio = np.linalg.norm(after) / np.linalg.norm(before)
    # the unscaled residual is the scaled one times the diagonal's spread
    assert ratio <= 1e-8, ratio
    # and the divergence was not small to begin with
    assert np.linalg.norm(before) > 1.0


def test_projection_with_a_wrong_sign_fails_the_pin():
    """The same pin on a deliberately broken composition (the correction's
    sign flipped, in a copy of the oracle's step): the divergence does not
    fall, so the pin above can fail."""
    Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=9)
    ud.tol = 1e-12
    ud.max_iterations = 20000
    dt = 0.05
    before = _divergence(Sol, elem, node, ud, th, mpv)
    op, b, _ = opj.assemble_system(Sol, mpv, elem, node, ud, th, dt)
    from oracle import bicgstab as obs
    from oracle import laplacian as olp
    r = obs.bicgstab(lambda v: olp.lap2d(v, op), b, tol=ud.tol, maxiter=ud.max_iterations)
    opj.finish_projection(Sol, mpv, elem, node, ud, th, dt, -r.x)   # wrong sign
    after = _divergence(Sol, elem, node, ud, th, mpv)
    assert np.linalg.norm(after) / np.linalg.norm(before) > 0.5


def test_sol0_drives_the_coefficients():
    Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=4, compressible=True)
    Sol0 = copy.deepcopy(Sol)
    rng = np.random.default_rng(1)
    Sol0.rhoY[...] *= 1.0 + 0.1 * rng.random(Sol0.rhoY.shape)
    Sol0.rho[...] *= 1.0 + 0.05 * rng.random(Sol0.rho.shape)
    S0 = copy.deepcopy(Sol0)
    opj.assemble_system(Sol, mpv, elem, node, ud, th, 0.05, Sol0=Sol0)
    # the ghost-filled Sol0 the reference assembles from
    obd.set_explicit_boundary_data(S0, elem, ud, th, mpv)
    Y0 = S0.rhoY / S0.rho
    want = th.Gammainv * S0.rhoY * Y0
    assert np.max(np.abs(mpv.wplus[0] - want)) <= 1e-13 * np.max(np.abs(want))
    S1 = copy.deepcopy(Sol)
    obd.set_explicit_boundary_data(S1, elem, ud, th, mpv)
    other = th.Gammainv * S1.rhoY * (S1.rhoY / S1.rho)
    assert np.max(np.abs(mpv.wplus[0] - other)) > 1e-3 * np.max(np.abs(other))
