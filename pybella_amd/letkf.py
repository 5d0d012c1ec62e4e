"""LETKF data assimilation of an ensemble on the GPU (r-localisation).

``LETKF`` mirrors the reference's prepare_rloc (src/data_assimilation/
letkf.py:190-700): built from the grid, the DA parameters (observed
attributes, window size obs_X x obs_Y, localisation weights, inflation) and
the ensemble size, it analyses the members' cell attributes and node
attributes (grid types analysed separately, as analyse() does, letkf.py:
288-297).  The members' device states are analysed in place: every grid
point of a grid type is one 64-lane wave of pb_letkf_analyse (the window
gather, Pa = (k-1)/rho I + Y^T W Y, its Jacobi eigendecomposition in LDS and
the analysis weights), then the analysed inner fields are written back with
zero ghost layers, as the reference writes np.pad(data, 2, mode='constant')
(letkf.py:367-376).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .state import is_periodic

CELL_ATTRS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")


def _torch():
    import torch
    return torch


def gaussian_weights(obs_X, obs_Y):
    """The localisation weights of dap.get_loc_mat('gaussian')
    (params.py:192-216): the isotropic normal density of variance 0.05 on
    the window's [-1/2, 1/2]^2 coordinates, shifted and scaled to [0, 1]."""
    x = np.linspace(0, 1, obs_X) - 0.5
    y = np.linspace(0, 1, obs_Y) - 0.5
    X, Y = np.meshgrid(x, y)
    Z = np.exp(-(X ** 2 + Y ** 2) / (2 * 0.05)) / (2 * np.pi * 0.05)
    Z = Z - Z.min()
    return (Z / Z.max()).T        # indexed [window x][window y]


class LETKF:
    """attributes: the observed attributes (cell fields of Sol, or
    "p2_nodes"); obs_X, obs_Y: odd window sizes; loc: (obs_X, obs_Y)
    weights (default: the reference's gaussian); inflation: rho."""

    def __init__(self, ud, attributes=("rho", "rhou", "rhov", "rhoY", "p2_nodes"), obs_X=11, obs_Y=11, loc=None,
                 inflation=1.0, ctx=None):
        self.ctx = ctx or _abi.default_context()
        self.cell_attrs = [a for a in attributes if a in CELL_ATTRS]
        self.node_attrs = [a for a in attributes if a == "p2_nodes"]
        self.obs_X, self.obs_Y = int(obs_X), int(obs_Y)
        self.loc = np.asarray(gaussian_weights(obs_X, obs_Y) if loc is None else loc, dtype=np.float64)
        self.inflation = float(inflation)
        self.periodic = (is_periodic(ud.bdry_type[0]), is_periodic(ud.bdry_type[1]))

    def _analyse_type(self, fields, obs, mask, covar):
        """fields: per member, per attribute, the device [y][x] field with
        two ghost layers; obs, mask, covar: host arrays of the inner grid in
        the reference's [x, y] layout (obs, mask per attribute; covar per
        attribute or None).  Writes the analysis into the fields."""
        torch = _torch()
        k, na = len(fields), len(fields[0])
        inner = [[f[2:-2, 2:-2] for f in mem] for mem in fields]
        ny, nx = inner[0][0].shape
        state = torch.stack([torch.stack(mem) for mem in inner]).contiguous()          # (k, attrs, ny, nx)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).cuda()
        obs_d = dev(np.stack([np.asarray(o).T for o in obs]))
        mask_d = dev(np.stack([np.asarray(m).T != 0 for m in mask]).astype(np.float64))
        cov_d = dev(np.asarray(covar, dtype=np.float64)) if covar is not None else None
        loc_d = dev(self.loc)
        out = torch.empty_like(state)
        cfg = _abi.LetkfCfg(k, na, nx, ny, self.obs_X, self.obs_Y, int(self.periodic[0]), int(self.periodic[1]),
                            self.inflation)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_letkf_analyse(self.ctx.h, C.byref(cfg), _abi.ptr(state), _abi.ptr(obs_d),
                                                     _abi.ptr(mask_d), _abi.ptr(cov_d), _abi.ptr(loc_d),
                                                     _abi.ptr(out)), "pb_letkf_analyse")
        for m, mem in enumerate(fields):
            for a, f in enumerate(mem):
                f.zero_()
                f[2:-2, 2:-2].copy_(out[m, a])

    def analyse(self, models, obs, mask, covar=None):
        """The analysis of the members' device states (time_update.DeviceModel
        objects, e.g. Ensemble.models).  obs, mask: dicts attribute -> host
        array of the inner grid ([x, y]; mask != 0 where not observed);
        covar: dict attribute -> observation variance (None: 1)."""
        for attrs, get in ((self.cell_attrs, lambda M, a: M.sol.f[a]), (self.node_attrs, lambda M, a: M.p2n)):
            if not attrs:
                continue
            fields = [[get(M, a) for a in attrs] for M in models]
            cv = None if covar is None else [covar[a] for a in attrs]
            self._analyse_type(fields, [obs[a] for a in attrs], [mask[a] for a in attrs], cv)


class RLoc:
    """Drop-in for the reference's prepare_rloc (src/data_assimilation/
    letkf.py:190-260) on host containers: ``RLoc(ud, elem, node, dap, N)``
    and ``analyse(results, obs, covar, mask, N, tout)`` with the reference's
    arguments and return value (letkf.py:288-297): ``results`` the members'
    [Sol, flux, mpv, steps] lists, ``obs`` / ``mask`` per DA time a dict of
    attribute -> full array (ghosts included), ``covar`` the pair (cell,
    node) of per-time variance arrays or None.  The analysed attributes are
    set on the members as new arrays with zero ghost layers, as the
    reference sets them."""

    def __init__(self, ud, elem, node, dap, N):
        self.N = N
        self.da_times = list(dap.da_times)
        self.attrs = list(dap.obs_attributes)
        self.loc_of = dict(dap.loc)
        self.obs_X, self.obs_Y = int(dap.obs_X), int(dap.obs_Y)
        assert self.obs_X < elem.iicx and self.obs_Y < elem.iicy, "window larger than the inner grid"
        assert self.obs_X % 2 == 1 and self.obs_Y % 2 == 1, "window sizes must be odd"
        loc = np.asarray(dap.localisation_matrix, dtype=np.float64).reshape(self.obs_Y, self.obs_X).T
        self.engine = LETKF(ud, attributes=self.attrs, obs_X=self.obs_X, obs_Y=self.obs_Y, loc=loc,
                            inflation=float(dap.inflation_factor))

    def analyse(self, results, obs, covar, mask, N, tout):
        from . import second_projection as sp
        ti = self.da_times.index(tout)
        i2 = (slice(2, -2), slice(2, -2))
        eng = self.engine
        for grid, attrs, cov in (("cell", eng.cell_attrs, None if covar is None else covar[0]),
                                 ("node", eng.node_attrs, None if covar is None else covar[1])):
            if not attrs:
                continue
            holder = (lambda mem: mem[0]) if grid == "cell" else (lambda mem: mem[2])
            fields = [[sp.to_dev(getattr(holder(mem), a)) for a in attrs] for mem in results]
            cv = None if cov is None else [float(v) for v in np.asarray(cov)[ti]]
            eng._analyse_type(fields, [np.asarray(obs[ti][a])[i2] for a in attrs],
                              [np.asarray(mask[ti][a])[i2] for a in attrs], cv)
            for mem, fs in zip(results, fields):
                for a, f in zip(attrs, fs):
                    setattr(holder(mem), a, sp.from_dev(f))
        return results
