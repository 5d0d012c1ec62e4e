"""LETKF local analysis with grid-point localisation (oracle; TEST
INFRASTRUCTURE ONLY).

Restates, for the 2D vertical slice and one grid type (cells or nodes), the
reference's r-localised LETKF (Hunt et al. 2007) with NumPy in place of dask
(same arithmetic, no chunking):
  analysis.analyse          src/data_assimilation/letkf.py:120-166
  prepare_rloc.analyse_by_grid_type / do_analysis
                            letkf.py:288-426
  get_state, get_obs, get_state_in_obs_space, get_bc_mask, get_loc_mat
                            letkf.py:452-622, 688-700
  boundary_mask             src/data_assimilation/utils.py:96-118
  the gaussian localisation weights  src/data_assimilation/params.py:192-216

Pinned by (tests/test_oracle_letkf.py): the analysis of one grid point equals
the Kalman-filter update in observation space (gain
P H^T (H P H^T + R/L)^-1 with P = rho X^T X / (k-1)) in mean and covariance,
an independent formulation of the same estimate; and the localised windows
equal a direct modular-index gather.
"""
import numpy as np
from numpy.lib.stride_tricks import sliding_window_view
from scipy import linalg, stats


def loc_matrix(obs_X, obs_Y, kind="gaussian", c=1.0):
    """params.py:192-216."""
    x = np.linspace(0, 1, obs_X) - 0.5
    y = np.linspace(0, 1, obs_Y) - 0.5
    X, Y = np.meshgrid(x, y)
    if kind == "constants":
        return np.ones((obs_X, obs_Y)).flatten() * c
    if kind == "linear":
        Z = np.sqrt(X ** 2 + Y ** 2)
        Z = Z.max() - Z
        return Z.flatten() * c
    pos = np.array([X.flatten(), Y.flatten()]).T
    norm = stats.multivariate_normal([0, 0], [[0.05, 0.0], [0.0, 0.05]])
    Z = norm.pdf(pos).reshape(X.shape[0], X.shape[1])
    Z = Z - Z.min()
    Z /= Z.max()
    return Z * c


def boundary_mask(shape, periodic, pad_X, pad_Y):
    """utils.py:96-118 for one grid of inner size ``shape``: ones padded
    with 1 on periodic axes and 0 on wall axes."""
    m = np.ones(shape)
    for dim, pad in enumerate((pad_X, pad_Y)):
        gp = [[0, 0], [0, 0]]
        gp[dim] = [pad, pad]
        m = np.pad(m, gp, mode="constant", constant_values=(1.0 if periodic[dim] else 0.0))
    return m


def analyse_point(X, X_mean, Y, Y_mean, obs, R, L, rho):
    """analysis.analyse (letkf.py:120-166) of one grid point: X (k, m)
    anomalies and X_mean (m); Y (k, l) observation-space anomalies and Y_mean
    (l); obs, R (variances), L (localisation weights), each (l,).  Returns
    the analysis ensemble (k, m)."""
    k = X.shape[0]
    C = (Y.T / R[:, None]).T                 # spsolve(diag(R), Y.T).T
    C = ((np.diag(L)) @ C.T).T
    Pa = (k - 1.0) * np.eye(k) / rho + (C @ Y.T)
    Lam, P = linalg.eigh(Pa)
    Pa = P @ (np.diag(1.0 / Lam) @ P.T)
    Wa = np.sqrt(k - 1.0) * P @ (np.diag(np.sqrt(1.0 / Lam)) @ P.T)
    wa = Pa @ (C @ (obs - Y_mean))
    Wa += wa
    return (np.dot(X.T, Wa.T) + X_mean.reshape(-1, 1)).T


def analyse_grid(state, obs, mask, covar, loc_mat, bc_base, obs_X, obs_Y, rho):
    """analyse_by_grid_type (letkf.py:288-426) for one grid type.

    state (k, attr, Nx, Ny): the members' inner fields of the observed
    attributes; obs, mask (attr, Nx, Ny) (mask 1: no observation); covar:
    (attr,) observation variances or None (identity); loc_mat (obs_X,
    obs_Y); bc_base: boundary_mask of the grid.  Returns the analysis
    (k, attr, Nx, Ny)."""
    k, attr_len, Nx, Ny = state.shape
    px, py = (obs_X - 1) // 2, (obs_Y - 1) // 2
    pads = [[px, px], [py, py]]
    # get_state (:452-478)
    orig = np.array(state, dtype=np.float64)
    state = orig.copy()
    X_bar = state.mean(axis=0, keepdims=True)
    state -= X_bar
    Xl = np.swapaxes(state.reshape(k, attr_len, Nx * Ny), 1, 2)          # (k, n, attr)
    Xl = np.swapaxes(Xl, 0, 1)                                            # (n, k, attr)
    Xb = np.swapaxes(X_bar.reshape(1, attr_len, Nx * Ny), 1, 2)[0]        # (n, attr)
    # get_obs (:480-504)
    obs_p = np.array([np.pad(o, pads, mode="wrap") for o in obs])
    mask_p = np.array([np.pad(m, pads, mode="wrap") for m in mask])
    obs_p = np.ma.array(obs_p, mask=mask_p).filled(fill_value=np.nan)
    obs_w = sliding_window_view(obs_p, (obs_X, obs_Y), axis=(1, 2))
    obs_w = np.swapaxes(obs_w, 0, 2)
    obs_w = np.swapaxes(obs_w, 0, 1).reshape(Nx * Ny, attr_len, obs_X, obs_Y)
    # get_state_in_obs_space (:507-616)
    sios = np.array([[np.pad(f, pads, mode="wrap") for f in mem] for mem in orig])
    sios = np.ma.array([np.ma.array(mem, mask=mask_p) for mem in sios])
    Y_bar = sios.mean(axis=0, keepdims=True)
    sios = sios - Y_bar
    sios = sios.filled(np.nan)
    Y_bar = Y_bar.filled(np.nan)
    sios = sliding_window_view(sios, (obs_X, obs_Y), axis=(2, 3)).reshape(k, attr_len, Nx * Ny, obs_X, obs_Y)
    Y_bar = sliding_window_view(Y_bar, (obs_X, obs_Y), axis=(2, 3)).reshape(1, attr_len, Nx * Ny, obs_X, obs_Y)
    sios = np.swapaxes(np.swapaxes(sios, 1, 2), 0, 1)                     # (n, k, attr, oX, oY)
    Y_bar = np.swapaxes(Y_bar, 1, 2)[0]                                   # (n, attr, oX, oY)
    # get_bc_mask (:619-656)
    bc = np.pad(np.ones((Nx, Ny)), pads, mode="constant", constant_values=(1.0)) * bc_base
    bc_w = sliding_window_view(bc, (obs_X, obs_Y)).reshape(Nx * Ny, obs_X, obs_Y)
    mask_w = sliding_window_view(mask_p, (obs_X, obs_Y), axis=(1, 2)).reshape(attr_len, Nx * Ny, obs_X, obs_Y)
    mask_w = np.swapaxes(mask_w, 0, 1)
    # covariance of the local window (:331-344)
    if covar is not None:
        cov = np.repeat(np.repeat(np.asarray(covar, dtype=np.float64)[:, None, None], obs_X, axis=1), obs_Y, axis=2)
    out = np.zeros((Nx * Ny, k, attr_len))
    for n in range(Nx * Ny):                                              # do_analysis (:381-426)
        Yn = np.array([mem[~np.isnan(mem)] for mem in sios[n]])
        Y_mean = Y_bar[n][~np.isnan(Y_bar[n])]
        lm = np.repeat((loc_mat * bc_w[n])[None], attr_len, axis=0)       # get_loc_mat (:688-700)
        lm = np.ma.array(lm, mask=mask_w[n]).filled(fill_value=np.nan)
        L = lm[~np.isnan(lm)].flatten()
        if covar is None:
            R = np.ones(L.size)
        else:
            cc = np.ma.array(cov, mask=mask_w[n]).filled(fill_value=np.nan)
            R = cc[~np.isnan(cc)].flatten()
        on = obs_w[n][~np.isnan(obs_w[n])]
        out[n] = analyse_point(Xl[n], Xb[n], Yn, Y_mean, on, R, L, rho)
    res = np.swapaxes(out, 0, 2)                                          # (attr, k, n)
    return np.swapaxes(res.reshape(attr_len, k, Nx, Ny), 0, 1)
