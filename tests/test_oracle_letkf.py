"""CPU pins of the LETKF oracle (oracle/letkf.py).

* One grid point's analysis (analysis.analyse, letkf.py:120-166) equals the
  Kalman-filter update computed in observation space, with the forecast
  covariance P = rho X^T X / (k-1), the linear observation operator that
  maps the members' anomalies X to Y, and the localised covariance R / L:
  mean  x_a = x_f + K (y - y_f),  K = P H^T (H P H^T + R/L)^-1,
  spread X_a^T X_a / (k-1) = (I - K H) P.
  The two formulations share no code: one works in the k-dimensional
  ensemble space with an eigendecomposition, the other in the
  l-dimensional observation space with a linear solve.
* The windows of analyse_grid (sliding-window views over 'wrap'-padded
  fields) equal a direct modular-index gather, and an all-masked window
  leaves the members unchanged (rho = 1).
"""
import numpy as np
import pytest

from oracle import letkf as ol


def _kalman(Xf, Hx, obs, R, L, rho):
    """Observation-space Kalman update of an ensemble (k, m) whose members'
    observed values are Hx (k, l) (a linear function of the members)."""
    k = Xf.shape[0]
    xm, ym = Xf.mean(axis=0), Hx.mean(axis=0)
    A, B = Xf - xm, Hx - ym
    P_xy = rho * A.T @ B / (k - 1)
    P_yy = rho * B.T @ B / (k - 1)
    P_xx = rho * A.T @ A / (k - 1)
    S = P_yy + np.diag(R / L)
    K = np.linalg.solve(S, P_xy.T).T
    mean = xm + K @ (obs - ym)
    cov = P_xx - K @ P_xy.T
    return mean, cov


@pytest.mark.parametrize("k,m,l,rho", [(8, 3, 20, 1.0), (16, 4, 50, 1.1), (5, 1, 3, 1.0)])
def test_point_analysis_equals_kalman_update(k, m, l, rho):
    rng = np.random.default_rng(k * 100 + l)
    Xf = rng.standard_normal((k, m)) + rng.standard_normal(m)
    H = rng.standard_normal((l, m))
    Hx = Xf @ H.T + 0.1 * rng.standard_normal((k, l))   # members' values in observation space
    obs = Hx.mean(axis=0) + rng.standard_normal(l)
    R = 0.5 + rng.random(l)
    L = 0.2 + rng.random(l)
    xm, ym = Xf.mean(axis=0), Hx.mean(axis=0)
    Xa = ol.analyse_point(Xf - xm, xm, Hx - ym, ym, obs, R, L, rho)
    mean, cov = _kalman(Xf, Hx, obs, R, L, rho)
    assert np.allclose(Xa.mean(axis=0), mean, rtol=1e-11, atol=1e-11)
    Aa = Xa - Xa.mean(axis=0)
    assert np.allclose(Aa.T @ Aa / (k - 1), cov, rtol=1e-10, atol=1e-11)


def test_point_analysis_catches_a_planted_mistake():
    """The same comparison with the innovation's sign flipped in a copy of
    the analysis fails."""
    rng = np.random.default_rng(3)
    k, m, l = 8, 2, 12
    Xf = rng.standard_normal((k, m))
    Hx = Xf @ rng.standard_normal((m, l))
    obs = Hx.mean(axis=0) + rng.standard_normal(l)
    R, L = np.ones(l), np.ones(l)
    xm, ym = Xf.mean(axis=0), Hx.mean(axis=0)
    Xa = ol.analyse_point(Xf - xm, xm, Hx - ym, 2 * obs - ym, obs, R, L, 1.0)   # innovation ym - obs
    mean, _ = _kalman(Xf, Hx, obs, R, L, 1.0)
    assert not np.allclose(Xa.mean(axis=0), mean, rtol=1e-6)


def _gather(field_padded, ix, iy, oX, oY):
    return field_padded[..., ix:ix + oX, iy:iy + oY]


def test_windows_equal_a_direct_gather():
    rng = np.random.default_rng(5)
    k, attr, Nx, Ny, oX, oY = 6, 2, 9, 7, 5, 3
    state = rng.standard_normal((k, attr, Nx, Ny)) + 3.0
    obs = rng.standard_normal((attr, Nx, Ny))
    mask = (rng.random((attr, Nx, Ny)) < 0.5).astype(float)
    covar = np.array([0.3, 0.7])
    locm = ol.loc_matrix(oX, oY)[:oX, :oY] if oX == oY else np.ones((oX, oY))
    bc = ol.boundary_mask((Nx, Ny), (True, False), (oX - 1) // 2, (oY - 1) // 2)
    got = ol.analyse_grid(state, obs, mask, covar, locm, bc, oX, oY, 1.0)
    px, py = (oX - 1) // 2, (oY - 1) // 2
    xm = state.mean(axis=0)
    for ix in range(Nx):
        for iy in range(Ny):
            Yc, yo, R, L, ybar = [], [], [], [], []
            for a in range(attr):
                for wx in range(oX):
                    for wy in range(oY):
                        gx, gy = (ix + wx - px) % Nx, (iy + wy - py) % Ny   # 'wrap' padding
                        if mask[a, gx, gy] != 0:
                            continue
                        vals = state[:, a, gx, gy]
                        Yc.append(vals - vals.mean())
                        ybar.append(vals.mean())
                        yo.append(obs[a, gx, gy])
                        R.append(covar[a])
                        inside_y = 0 <= iy + wy - py < Ny                    # walls in y
                        L.append(locm[wx, wy] * (1.0 if inside_y else 0.0))
            Yn = np.array(Yc).T.reshape(k, -1)
            ref = ol.analyse_point(state[:, :, ix, iy] - xm[:, ix, iy], xm[:, ix, iy], Yn,
                                   np.array(ybar), np.array(yo), np.array(R), np.array(L), 1.0)
            assert np.allclose(got[:, :, ix, iy], ref, rtol=1e-12, atol=1e-12), (ix, iy)


def test_no_observation_leaves_members_unchanged():
    rng = np.random.default_rng(7)
    state = rng.standard_normal((5, 1, 6, 6))
    mask = np.ones((1, 6, 6))
    bc = ol.boundary_mask((6, 6), (True, True), 1, 1)
    got = ol.analyse_grid(state, np.zeros((1, 6, 6)), mask, None, np.ones((3, 3)), bc, 3, 3, 1.0)
    assert np.allclose(got, state, rtol=1e-13, atol=1e-13)
