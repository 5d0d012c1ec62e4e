"""CPU pins of the oracle's 2D nodal operator and diagonal preconditioner.

Each check is chosen so that a plausible restatement error (a dropped flux
term, a swapped corner, a wrong wrap index, a missing wall zeroing) fails it:

* the literal per-node loop (laplacian.py:546-746 restated line by line) and
  the vectorised form agree to rounding on every boundary combination;
* for ANY coefficients, A*1 and A*c (c the checkerboard) reduce to the centre
  term, because every cell-averaged difference of 1 and of c vanishes;
* unit coefficients give the textbook rotated 9-point Laplacian;
* the probed CSR matrix equals dense columns, and a manufactured solution is
  recovered by a sparse direct solve (SciPy splu) modulo the null space;
* the trapezoid weights are a left null vector of the unscaled
  pseudo-incompressible operator;
* the C restatement (oracle/c/pb_oracle.c, built without FP contraction) is
  bitwise equal to the vectorised NumPy form.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import laplacian as olp
from oracle import projection as opj

from _cases import checkerboard, make_case, oracle_op, random_op, rel

BCS = [(False, True), (True, True), (True, False), (False, False)]


@pytest.mark.parametrize("x_wall,y_wall", BCS)
def test_loop_equals_vectorised(x_wall, y_wall):
    op = random_op(13, 9, seed=11, x_wall=x_wall, y_wall=y_wall)
    p = np.random.default_rng(5).standard_normal(13 * 9)
    a = olp.lap2d_loop(p, op)
    b = olp.lap2d(p, op)
    assert rel(b, a) < 1e-15


@pytest.mark.parametrize("x_wall,y_wall", BCS)
def test_constant_and_checkerboard_reduce_to_centre(x_wall, y_wall):
    nx, ny = 17, 9      # odd node counts: the periodic wrap keeps c consistent
    op = random_op(nx, ny, seed=3, x_wall=x_wall, y_wall=y_wall)
    ones = np.ones(nx * ny)
    c = checkerboard(ny, nx).ravel()
    scale = np.max(np.abs(op["hx"])) / op["dx"] ** 2
    y1 = olp.lap2d(ones, op)
    yc = olp.lap2d(c, op)
    ref1 = (op["hc"] * op["dinv"]).ravel()
    refc = ref1 * c
    assert np.max(np.abs(y1 - ref1)) < 1e-12 * scale
    assert np.max(np.abs(yc - refc)) < 1e-12 * scale
    # a non-trivial vector does feel the flux terms (guards a vacuous pass)
    ramp = np.tile(np.arange(nx, dtype=float), ny)
    assert np.max(np.abs(olp.lap2d(ramp, op) - ref1 * ramp)) > 1e-3 * scale


def test_unit_coefficients_give_rotated_9pt():
    nx, ny, h = 9, 7, 0.25
    one = lambda: np.ones((ny + 1, nx + 1))
    zero = lambda: np.zeros((ny + 1, nx + 1))
    op = dict(nx=nx, ny=ny, dx=h, dy=h, hx=one(), hy=one(), hc=np.zeros((ny, nx)),
              cxx=one(), cyy=one(), cxy=zero(), cyx=zero(), dinv=np.ones((ny, nx)),
              x_wall=False, y_wall=True)
    rng = np.random.default_rng(0)
    P = rng.standard_normal((ny, nx))
    Y = olp.lap2d(P.ravel(), op).reshape(ny, nx)
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            ref = (0.5 * (P[j - 1, i - 1] + P[j - 1, i + 1] + P[j + 1, i - 1] + P[j + 1, i + 1])
                   - 2.0 * P[j, i]) / h ** 2
            assert abs(Y[j, i] - ref) < 1e-12 * np.max(np.abs(P)) / h ** 2


@pytest.mark.parametrize("x_wall,y_wall", [(False, True), (True, True)])
def test_probe_csr_equals_dense(x_wall, y_wall):
    nx, ny = 11, 9
    op = random_op(nx, ny, seed=8, x_wall=x_wall, y_wall=y_wall)
    n = nx * ny
    apply = lambda v: olp.lap2d(v, op)
    A = olp.probe_csr(apply, n, (ny, nx)).toarray()
    D = np.stack([apply(e) for e in np.eye(n)], axis=1)
    # the colour probe misses the periodic wrap couplings (distance nx-2);
    # compare only columns/rows whose stencil does not wrap
    jj, ii = np.divmod(np.arange(n), nx)
    inner = (ii >= 2) & (ii <= nx - 3) & (jj >= 2) & (jj <= ny - 3)
    assert rel(A[np.ix_(inner, inner)], D[np.ix_(inner, inner)]) < 1e-14
    assert np.count_nonzero(np.abs(D[inner]) > 0, axis=1).max() == 9


def test_manufactured_solution_splu():
    """b = A x*; a sparse direct solve recovers x* modulo {1, c} (psinc
    operator: singular with exactly that null space)."""
    Sol, mpv, elem, node, ud, th = make_case("strat_psinc", 17, 9)
    op, _, _ = oracle_op(Sol, mpv, elem, node, ud, th, 0.05)
    nx, ny = op["nx"], op["ny"]
    n = nx * ny
    A = np.stack([olp.lap2d(e, op) for e in np.eye(n)], axis=1)
    s = np.linalg.svd(A, compute_uv=False)
    assert s[-2] < 1e-12 * s[0] and s[-3] > 1e-6 * s[0]        # rank n-2
    c = checkerboard(ny, nx).ravel()
    xs = np.random.default_rng(2).standard_normal(n)
    b = A @ xs
    # pin the null components by replacing two rows with x[0] = xs[0], x[1] = xs[1]
    import scipy.sparse as sp
    B = A.copy()
    B[0, :] = 0.0
    B[0, 0] = 1.0
    B[1, :] = 0.0
    B[1, 1] = 1.0
    bb = b.copy()
    bb[0], bb[1] = xs[0], xs[1]
    x = spla.splu(sp.csc_matrix(B)).solve(bb)
    assert rel(A @ x, b) < 1e-10
    d = x - xs
    N = np.stack([np.ones(n), c], axis=1)
    coef = np.linalg.lstsq(N, d, rcond=None)[0]
    assert np.max(np.abs(d - N @ coef)) < 1e-8 * np.max(np.abs(xs))


def test_left_null_vector_trapezoid():
    """For the psinc operator without Coriolis, w (1/2 on the duplicated
    periodic end columns, 1 elsewhere) is a left null vector of the unscaled
    operator A_hat = A / dinv, and so is w*c."""
    Sol, mpv, elem, node, ud, th = make_case("strat_psinc", 17, 9)
    op, _, _ = oracle_op(Sol, mpv, elem, node, ud, th, 0.05)
    nx, ny = op["nx"], op["ny"]
    n = nx * ny
    A = np.stack([olp.lap2d(e, op) for e in np.eye(n)], axis=1)
    Ahat = A / op["dinv"].ravel()[:, None]
    w = np.ones((ny, nx))
    w[:, 0] = w[:, -1] = 0.5
    c = checkerboard(ny, nx)
    scale = np.max(np.abs(Ahat))
    for v in (w, w * c):
        assert np.max(np.abs(Ahat.T @ v.ravel())) < 1e-12 * scale * n
    # and it is not a left null vector of the compressible operator
    Sol, mpv, elem, node, ud, th = make_case("strat_comp", 17, 9)
    op2, _, _ = oracle_op(Sol, mpv, elem, node, ud, th, 0.05)
    A2 = np.stack([olp.lap2d(e, op2) for e in np.eye(n)], axis=1) / op2["dinv"].ravel()[:, None]
    assert np.max(np.abs(A2.T @ w.ravel())) > 1e-6 * np.max(np.abs(A2))


def test_box_sum_fft_vs_direct():
    a = np.random.default_rng(1).standard_normal((14, 9))
    assert rel(olp.box_sum(a, 2), olp.box_sum_direct(a)) < 1e-13


def test_diag_closed_form_uniform():
    """Uniform wplus = h, wcenter = w0: every 2x2 box sum is 4h, so
    diag_inv = 1 / (w0 - 0.75 * 4h * (1/dx^2 + 1/dy^2 + 2/(dx dy)))
    (laplacian.py:1284-1356)."""
    Sol, mpv, elem, node, ud, th = make_case("strat_comp", 17, 9)
    h, w0 = 1.7, -3.0
    mpv.wplus[0][...] = h
    mpv.wplus[1][...] = h
    mpv.wcenter = np.full_like(mpv.wcenter, w0)
    d = olp.precon_diag_prepare(mpv, elem, node, ud)
    dx, dy = node.dx, node.dy
    ref = 1.0 / (w0 - 0.75 * 4 * h * (1 / dx ** 2 + 1 / dy ** 2 + 2 / (dx * dy)))
    assert rel(d, np.full_like(d, ref)) < 1e-14


def test_uniform_state_uniform_coefficients():
    """Constant rho, rhoY and no stratification: wplus = Gammainv*rhoY*Y
    everywhere, nu = 0, wcenter = ccenter*rhoY^(2-gamma) with wall rows halved."""
    Sol, mpv, elem, node, ud, th = make_case("strat_comp", 17, 9)
    Sol.rho[...] = 1.3
    Sol.rhoY[...] = 1.1
    mpv.HydroState_n.S0[...] = 0.8
    dt = 0.05
    opj.operator_coefficients_nodes(elem, node, Sol, mpv, ud, th, dt)
    Y = 1.1 / 1.3
    assert rel(mpv.wplus[0], np.full_like(mpv.wplus[0], th.Gammainv * 1.1 * Y)) < 1e-15
    assert np.max(np.abs(mpv.nu_c)) == 0.0
    wc = -ud.Msq * th.gm1inv / dt ** 2 * 1.1 ** (2.0 - th.gamm)
    inner = mpv.wcenter[2:-2, 2:-2]
    assert rel(inner, np.full_like(inner, wc)) < 1e-14
    igs = node.igs[1]
    assert rel(mpv.wcenter[2:-2, igs - 1], np.full(mpv.wcenter.shape[0] - 4, 0.5 * wc)) < 1e-14


def test_c_port_bitwise():
    from oracle import cport
    for kind in ("strat_psinc", "coriolis", "xwall"):
        Sol, mpv, elem, node, ud, th = make_case(kind, 33, 17)
        op, _, _ = oracle_op(Sol, mpv, elem, node, ud, th, 0.05)
        p = np.random.default_rng(4).standard_normal(op["nx"] * op["ny"])
        a = olp.lap2d(p, op)
        b = cport.COp(op).matvec(p)
        assert np.array_equal(a, b), (kind, rel(b, a))


def test_polynomial_closed_forms_distinct_coefficients():
    """Uniform but distinct coefficients: the interior rows are exact on
    quadratics, A(x^2) = 2 cxx hx, A(y^2) = 2 cyy hy and A(xy) = cxy hy + cyx hx
    (y grows with the row index j, second_projection.py:245)."""
    nx, ny = 11, 9
    dx, dy = 0.3, 0.2
    full = lambda v: np.full((ny + 1, nx + 1), v)
    hx, hy, cxx, cyy, cxy, cyx = 1.5, 0.7, 2.0, 3.0, 0.5, 0.25
    op = dict(nx=nx, ny=ny, dx=dx, dy=dy, hx=full(hx), hy=full(hy), hc=np.zeros((ny, nx)),
              cxx=full(cxx), cyy=full(cyy), cxy=full(cxy), cyx=full(cyx), dinv=np.ones((ny, nx)),
              x_wall=False, y_wall=True)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy)
    inner = (slice(1, -1), slice(1, -1))
    for p, ref in ((X * X, 2 * cxx * hx), (Y * Y, 2 * cyy * hy), (X * Y, cxy * hy + cyx * hx)):
        out = olp.lap2d(p.ravel(), op).reshape(ny, nx)[inner]
        assert np.max(np.abs(out - ref)) < 1e-10, (ref, out.min(), out.max())
