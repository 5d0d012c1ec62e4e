"""CPU pins of the oracle's 27-point operator (laplacian.lap3d) and its 3D
assembly (laplacian.assemble3d).  The reference's lap3D cannot run
(second_projection.py:183-203 uses names before assignment), so these are the
pins of its restatement:

* unit weights: each axis contributes the 1D second difference along it,
  smoothed by (1/16)[1 2 1] x [1 2 1] across the other two axes (the 4-edge
  flux average of laplacian.py:983-1013 counted per edge);
* exact on x^2, y^2, z^2 with distinct per-axis weights; no cross terms (xy);
* A*1 and A*(in-plane checkerboard x any profile along the normal) reduce to
  the centre term for any weights -- which is why the pseudo-incompressible
  27-point operator is singular on O(nx + ny + nz) modes;
* assembly closed forms on a uniform state.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import laplacian as olp
from pybella_amd import workloads as W


def rand_op(nx, ny, nz, seed, periodic=(True, False, True), center=True):
    rng = np.random.default_rng(seed)
    cell = lambda: 0.5 + rng.random((nz + 1, ny + 1, nx + 1))
    return dict(nx=nx, ny=ny, nz=nz, dx=0.3, dy=0.2, dz=0.25, periodic=periodic, hx=cell(), hy=cell(), hz=cell(),
                hc=(-rng.random((nz, ny, nx)) if center else np.zeros((nz, ny, nx))),
                dinv=-(0.5 + rng.random((nz, ny, nx))))


def test_unit_weights_smoothed_second_differences():
    nx, ny, nz, h = 9, 8, 7, 0.5
    one = np.ones((nz + 1, ny + 1, nx + 1))
    op = dict(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h, periodic=(False, False, False), hx=one, hy=one, hz=one,
              hc=np.zeros((nz, ny, nx)), dinv=np.ones((nz, ny, nx)))
    P = np.random.default_rng(0).standard_normal((nz, ny, nx))
    Y = olp.lap3d(P.ravel(), op).reshape(nz, ny, nx)
    w = np.array([1.0, 2.0, 1.0]) / 4.0
    ref = np.zeros_like(P)
    for axis in range(3):
        d2 = np.zeros_like(P)
        sl = lambda a, b: tuple(slice(a, b) if ax == axis else slice(None) for ax in range(3))
        d2[sl(1, -1)] = P[sl(2, None)] - 2 * P[sl(1, -1)] + P[sl(None, -2)]
        sm = d2
        for other in range(3):
            if other != axis:
                sm = ndimage.correlate1d(sm, w, axis=other, mode="constant")
        ref += sm / h ** 2
    inner = (slice(2, -2),) * 3
    assert np.max(np.abs(Y[inner] - ref[inner])) < 1e-12 * np.max(np.abs(ref[inner]))


def test_quadratics_exact_per_axis_weights():
    nx, ny, nz = 9, 8, 7
    c = lambda v: np.full((nz + 1, ny + 1, nx + 1), v)
    hx, hy, hz = 1.5, 0.7, 2.5
    op = dict(nx=nx, ny=ny, nz=nz, dx=0.3, dy=0.2, dz=0.25, periodic=(False, False, False), hx=c(hx), hy=c(hy),
              hz=c(hz), hc=np.zeros((nz, ny, nx)), dinv=np.ones((nz, ny, nx)))
    K, J, I = np.meshgrid(np.arange(nz) * 0.25, np.arange(ny) * 0.2, np.arange(nx) * 0.3, indexing="ij")
    inner = (slice(1, -1),) * 3
    for p, ref in ((I * I, 2 * hx), (J * J, 2 * hy), (K * K, 2 * hz), (I * J, 0.0), (J * K, 0.0)):
        out = olp.lap3d(p.ravel(), op).reshape(nz, ny, nx)[inner]
        assert np.max(np.abs(out - ref)) < 1e-10, (ref, out.min(), out.max())


@pytest.mark.parametrize("periodic", [(True, False, True), (False, False, False), (True, True, True)])
def test_null_modes_reduce_to_centre(periodic):
    nx, ny, nz = 9, 7, 5          # odd counts keep the checkerboards periodic-consistent
    op = rand_op(nx, ny, nz, seed=1, periodic=periodic)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(2)
    modes = [np.ones((nz, ny, nx)), (-1.0) ** (I + J) * rng.standard_normal(nz)[:, None, None],
             (-1.0) ** (J + K) * rng.standard_normal(nx)[None, None, :],
             (-1.0) ** (I + K) * rng.standard_normal(ny)[None, :, None], (-1.0) ** (I + J + K)]
    centre = op["hc"] * op["dinv"]
    scale = np.max(op["hx"]) / op["dz"] ** 2
    for v in modes:
        out = olp.lap3d(v.ravel(), op).reshape(nz, ny, nx)
        assert np.max(np.abs(out - centre * v)) < 1e-12 * scale
    ramp = I.astype(float)
    assert np.max(np.abs(olp.lap3d(ramp.ravel(), op).reshape(nz, ny, nx) - centre * ramp)) > 1e-3


def test_assemble3d_uniform_state():
    st = W.stratified_state_3d(9, 7, 5)
    st.rho_c[...] = 1.3
    st.rhoY_c[...] = 1.1
    dt = 0.05
    op = olp.assemble3d(st, dt)
    th, ud = st.th, st.ud
    h = th.Gammainv * 1.1 * (1.1 / 1.3)
    assert np.allclose(op["hx"], h, rtol=1e-15, atol=0)
    wc = -ud.Msq * th.gm1inv / dt ** 2 * 1.1 ** (2.0 - th.gamm)
    hc = op["hc"]
    assert np.allclose(hc[1:-1, 1:-1, 1:-1], wc, rtol=1e-14)
    assert np.allclose(hc[1:-1, 0, 1:-1], 0.5 * wc, rtol=1e-14)       # y wall planes, inner x and z
    assert np.allclose(hc[1:-1, -1, 1:-1], 0.5 * wc, rtol=1e-14)
    assert np.allclose(hc[0, 0, 1:-1], wc, rtol=1e-14)                 # z end: periodic axis, not restricted
    d = wc - 0.0625 * 8 * h * (1 / st.dx ** 2 + 1 / st.dy ** 2 + 1 / st.dz ** 2 + 2 / (st.dx * st.dy))
    inner = op["dinv"][1:-1, 1:-1, 1:-1]
    assert np.allclose(inner, 1.0 / d, rtol=1e-14)


def test_compressible_3d_oracle_solve():
    from oracle import bicgstab as obs
    st = W.stratified_state_3d(17, 9, 9)
    op = olp.assemble3d(st, 0.05)
    b = W.fourier_rhs_3d(17, 9, 9, seed=1, nmodes=16, kmax=4).ravel() * op["dinv"].ravel()
    r = obs.bicgstab(lambda v: olp.lap3d(v, op), b, tol=1e-10, maxiter=3000)
    assert r.info == 0 and obs.true_rel_resid(lambda v: olp.lap3d(v, op), b, r.x) < 1e-9
