"""Ghost-cell and ghost-node fills (oracle; test infrastructure only).

Restates src/dycore/utils/boundary.py of pyBELLA for the 2D vertical slice
(x periodic or wall without gravity, y the gravity axis).
"""
import numpy as np


def _name(bt):
    return getattr(bt, "name", str(bt))


def _wall(bt):
    return _name(bt) in ("WALL", "RAYLEIGH")


def _periodic(bt):
    return _name(bt) == "PERIODIC"


def _negative_symmetric_pad(a, pads):
    """np.pad 'symmetric' with the sign of the ghost values flipped (boundary.py:219-245)."""
    out = np.pad(a, pads, "symmetric")
    for ax, (lo, hi) in enumerate(pads):
        if lo > 0 or hi > 0:
            sl = [slice(None)] * a.ndim
            sl[ax] = slice(0, lo)
            out[tuple(sl)] *= -1.0
            sl[ax] = slice(out.shape[ax] - hi, out.shape[ax])
            out[tuple(sl)] *= -1.0
    return out


def _set_boundary(Sol, pads, btype, idx):
    """boundary.py:135-216."""
    Sol.rho[...] = np.pad(Sol.rho[idx], pads, btype)
    if btype == "symmetric":
        Sol.rhov[...] = _negative_symmetric_pad(Sol.rhov[idx], pads)
        Sol.rho[...] = np.pad(Sol.rho[idx], pads, "symmetric")
        Sol.rhou[...] = np.pad(Sol.rhou[idx], pads, "symmetric")
    else:
        Sol.rhou[...] = np.pad(Sol.rhou[idx], pads, btype)
        Sol.rhov[...] = np.pad(Sol.rhov[idx], pads, btype)
        Sol.rhow[...] = np.pad(Sol.rhow[idx], pads, btype)
    Sol.rhoY[...] = np.pad(Sol.rhoY[idx], pads, btype)
    Sol.rhoX[...] = np.pad(Sol.rhoX[idx], pads, btype)


def set_explicit_boundary_data(Sol, elem, ud, th, mpv, step=None):
    """Cell ghost fill (boundary.py:7-132): wrap/mirror on an axis without
    gravity, hydrostatic image rows on the gravity axis (y).  ``step`` (an
    axis) restricts the fill to that axis, as the advection's split steps
    call it (boundary.py:30-39 with the arrays flipped so that the split axis
    is last; unflipped here, the same axis is filled)."""
    ndim = elem.ndim
    igs = elem.igs
    for dim in (range(ndim) if step is None else (step,)):
        pads = [(0, 0)] * ndim
        pads[dim] = (igs[dim], igs[dim])
        pads = tuple(pads)
        idx = [slice(None)] * ndim
        idx[dim] = slice(igs[dim], -igs[dim])
        idx = tuple(idx)
        if ud.gravity_strength[dim] == 0.0:
            if _periodic(ud.bdry_type[dim]):
                _set_boundary(Sol, pads, "wrap", idx)
            elif _name(ud.bdry_type[dim]) == "WALL":
                _set_boundary(Sol, pads, "symmetric", idx)
            else:
                raise AssertionError("Rayleigh boundary not defined on x-direction.")
            continue
        # gravity axis (boundary.py:56-132)
        g = ud.gravity_strength[dim]
        direction = -1.0
        offset = 0
        for side in pads[dim]:
            direction *= -1
            for cur in np.arange(side)[::-1]:
                cur = int(cur)
                cur_i = cur
                c = cur + offset * ((elem.icy - 1) - 2 * cur)
                nlast = c + int(direction)
                nsource = int(offset * elem.icy + direction * (2 * elem.igy - (1 - offset) - cur_i))
                nimage = c
                L = (slice(None), nlast)
                S = (slice(None), nsource)
                I = (slice(None), nimage)
                Y_last = Sol.rhoY[L] / Sol.rho[L]
                rhoYv_image = -Sol.rhov[S] * Sol.rhoY[S] / Sol.rho[S]
                Sstrat = 1.0 / ud.stratification(elem.y[nimage])
                dpi = direction * (th.Gamma * g) * 0.5 * elem.dy * (1.0 / Y_last + Sstrat)
                if ud.is_compressible == 1:
                    rhoY = ((Sol.rhoY[L] ** th.gm1) + dpi) ** th.gm1inv
                else:
                    rhoY = mpv.HydroState.rhoY0[nimage]
                rho = rhoY * Sstrat
                v = rhoYv_image / rhoY
                u = Sol.rhou[S] / Sol.rho[S]
                w = Sol.rhow[S] / Sol.rho[S]
                X = Sol.rhoX[S] / Sol.rho[S]
                Sol.rho[I] = rho
                Sol.rhou[I] = rho * u
                Sol.rhov[I] = rho * v
                Sol.rhow[I] = rho * w
                Sol.rhoY[I] = rhoY
                Sol.rhoX[I] = rho * X
            offset += 1


def set_ghostcells_p2(p, elem, ud):
    """Cell ghost fill of a scalar cell array (boundary.py:284-293): 'wrap'
    on periodic axes, 'symmetric' otherwise, axis by axis."""
    igs = elem.igs
    for dim in range(elem.ndim):
        pads = [(0, 0)] * elem.ndim
        pads[dim] = (igs[dim], igs[dim])
        idx = [slice(None)] * elem.ndim
        idx[dim] = slice(igs[dim], -igs[dim])
        p[...] = np.pad(p[tuple(idx)], pads, "wrap" if _periodic(ud.bdry_type[dim]) else "symmetric")


def _periodic_plus_one_1d(vec, lo, hi):
    """boundary.py:348-369 applied to one padded vector (in place)."""
    if lo > 0 and hi > 0:
        first = vec[-hi - hi - 1:-hi]
        last = vec[lo:lo + lo + 1].copy()
        vec[:lo + 1] = first
        vec[-hi - 1:] = last


def set_ghostnodes_p2(p, node, ud):
    """Node ghost fill (boundary.py:296-311): periodic_plus_one (which also swaps
    the two duplicated end nodes) or 'reflect' per axis, axis by axis."""
    igs = node.igs
    ndim = node.ndim
    for dim in range(ndim):
        pads = [(0, 0)] * ndim
        pads[dim] = (igs[dim], igs[dim])
        idx = [slice(None)] * ndim
        idx[dim] = slice(igs[dim], -igs[dim])
        inner = p[tuple(idx)]
        if _periodic(ud.bdry_type[dim]):
            padded = np.pad(inner, pads, "constant")
            view = np.moveaxis(padded, dim, -1)
            for ind in np.ndindex(view.shape[:-1]):
                _periodic_plus_one_1d(view[ind], igs[dim], igs[dim])
            p[...] = padded
        else:
            p[...] = np.pad(inner, pads, "reflect")
    if node.iicy == 2:
        pn = p[:, 2, ...][:, None, ...]
        p[...] = np.repeat(pn, node.icy, axis=1)
