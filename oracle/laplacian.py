"""Nodal operators and the diagonal preconditioner (oracle; test infrastructure only).

Restates src/dycore/physics/low_mach/laplacian.py of pyBELLA:
  * precon_diag_prepare      laplacian.py:1284-1356
  * stencil_9pt closure      laplacian.py:496-528 (-> ``op2d_arrays``)
  * lap2D_gather_new         laplacian.py:546-746 (``lap2d_loop``: literal
                             per-node loop; ``lap2d``: vectorised, same
                             operation order per node)
  * lap3D / stencil_27pt     laplacian.py:872-1026 (``lap3d``: this build's
                             27-point definition following lap3D's structure;
                             see the notes there)
"""
import numpy as np
from scipy import signal


def _name(bt):
    return getattr(bt, "name", str(bt))


def _wall(bt):
    return _name(bt) in ("WALL", "RAYLEIGH")


def box_sum(a, ndim):
    """2x2 (2x2x2) 'valid' box sum, the reference's fftconvolve with ones."""
    return signal.fftconvolve(a, np.ones([2] * ndim), mode="valid")


def box_sum_direct(a):
    """Direct 4-point sum (closed form used to pin ``box_sum``)."""
    return a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:]


def precon_diag_prepare(mpv, elem, node, ud):
    """laplacian.py:1284-1356.  Heuristic diagonal: Coriolis factors unused
    (commented out in the reference), box *sum*, no wall zeroing,
    coeff = 1 - 0.25 in 2D and 1/16 in 3D."""
    dx, dy, dz = node.dx, node.dy, node.dz
    ndim = elem.ndim
    nine_pt = 0.5 * 0.5
    coeff = 1.0 - nine_pt if ndim == 2 else 0.0625
    wxx = coeff / (dx ** 2)
    wyy = coeff / (dy ** 2)
    wzz = coeff / (dz ** 2)
    wxy = coeff / (dx * dy)
    wyx = coeff / (dy * dx)
    hxx = mpv.wplus[0]
    hyy = mpv.wplus[1]
    hxy = mpv.wplus[0]
    hyx = mpv.wplus[1]
    diag = np.zeros_like(mpv.wcenter)
    diag[...] = -wxx * box_sum(hxx, ndim)
    diag[...] -= wyy * box_sum(hyy, ndim)
    diag[...] -= wxy * box_sum(hxy, ndim)
    diag[...] -= wyx * box_sum(hyx, ndim)
    if ndim == 3:
        diag[...] -= wzz * box_sum(mpv.wplus[2], ndim)
    diag[...] += mpv.wcenter
    diag[...] = 1.0 / diag
    return diag


def op2d_arrays(mpv, node, coriolis, diag_inv, ud):
    """The arrays the stencil_9pt closure captures (laplacian.py:496-528),
    reshaped to 2D: cell weights (NY+1, NX+1), node arrays (NY, NX)."""
    nx, ny = node.iicx, node.iicy
    hx = mpv.wplus[0][1:-1, 1:-1].T.copy()
    hy = mpv.wplus[1][1:-1, 1:-1].T.copy()
    hc = mpv.wcenter[1:-1, 1:-1].T.copy()
    cor = tuple(np.ascontiguousarray(c[1:-1, 1:-1]) for c in coriolis)
    dinv = diag_inv[1:-1, 1:-1].T.copy()
    assert hx.shape == (ny + 1, nx + 1) and hc.shape == (ny, nx)
    return dict(nx=nx, ny=ny, dx=node.dx, dy=node.dy, hx=hx, hy=hy, hc=hc,
                cxx=cor[0], cyy=cor[1], cxy=cor[2], cyx=cor[3], dinv=dinv,
                x_wall=_wall(ud.bdry_type[0]), y_wall=_wall(ud.bdry_type[1]))


def lap2d_loop(p, op):
    """Literal per-node restatement of lap2D_gather_new (laplacian.py:546-746).
    Slow; for small grids only."""
    nx, ny = op["nx"], op["ny"]
    hx, hy, hc = op["hx"].ravel(), op["hy"].ravel(), op["hc"].ravel()
    cxx, cyy, cxy, cyx = (op[k].ravel() for k in ("cxx", "cyy", "cxy", "cyx"))
    dinv = op["dinv"].ravel()
    x_wall, y_wall = op["x_wall"], op["y_wall"]
    oodx, oody = 1.0 / op["dx"], 1.0 / op["dy"]
    lap = np.zeros(nx * ny)
    cnt_x = cnt_y = 0
    for idx in range(nx * ny):
        row = idx // nx
        col = idx - row * nx
        ne = row * (nx + 1) + col
        ne_tl, ne_tr, ne_bl, ne_br = ne, ne + 1, ne + nx + 1, ne + nx + 2
        tl, ml, bl = idx - nx - 1, idx - 1, idx + nx - 1
        tm, mm, bm = idx - nx, idx, idx + nx
        tr, mr, br = idx - nx + 1, idx + 1, idx + nx + 1
        if cnt_x == 0:
            tl += nx - 1; ml += nx - 1; bl += nx - 1
        if cnt_x == nx - 1:
            tr -= nx - 1; mr -= nx - 1; br -= nx - 1
        if cnt_y == 0:
            tl += nx * (ny - 1); tm += nx * (ny - 1); tr += nx * (ny - 1)
        if cnt_y == ny - 1:
            bl -= nx * (ny - 1); bm -= nx * (ny - 1); br -= nx * (ny - 1)
        topleft, midleft, botleft = p[tl], p[ml], p[bl]
        topmid, midmid, botmid = p[tm], p[mm], p[bm]
        topright, midright, botright = p[tr], p[mr], p[br]
        hx_tl, hx_bl, hy_tl, hy_bl = hx[ne_tl], hx[ne_bl], hy[ne_tl], hy[ne_bl]
        hx_tr, hx_br, hy_tr, hy_br = hx[ne_tr], hx[ne_br], hy[ne_tr], hy[ne_br]
        if x_wall and cnt_x == 0:
            hx_tl = hy_tl = hx_bl = hy_bl = 0.0
        if x_wall and cnt_x == nx - 1:
            hx_tr = hy_tr = hx_br = hy_br = 0.0
        if y_wall and cnt_y == 0:
            hx_tl = hy_tl = hx_tr = hy_tr = 0.0
        if y_wall and cnt_y == ny - 1:
            hx_bl = hy_bl = hx_br = hy_br = 0.0
        Dx_tl = 0.5 * (topmid - topleft + midmid - midleft) * hx_tl
        Dx_tr = 0.5 * (topright - topmid + midright - midmid) * hx_tr
        Dx_bl = 0.5 * (botmid - botleft + midmid - midleft) * hx_bl
        Dx_br = 0.5 * (botright - botmid + midright - midmid) * hx_br
        Dy_tl = 0.5 * (midmid - topmid + midleft - topleft) * hy_tl
        Dy_tr = 0.5 * (midright - topright + midmid - topmid) * hy_tr
        Dy_bl = 0.5 * (botmid - midmid + botleft - midleft) * hy_bl
        Dy_br = 0.5 * (botright - midright + botmid - midmid) * hy_br
        Dxx = 0.5 * (cxx[ne_tr] * Dx_tr - cxx[ne_tl] * Dx_tl + cxx[ne_br] * Dx_br - cxx[ne_bl] * Dx_bl) * oodx * oodx
        Dyy = 0.5 * (cyy[ne_br] * Dy_br - cyy[ne_tr] * Dy_tr + cyy[ne_bl] * Dy_bl - cyy[ne_tl] * Dy_tl) * oody * oody
        Dyx = 0.5 * (cxy[ne_br] * Dy_br - cxy[ne_bl] * Dy_bl + cxy[ne_tr] * Dy_tr - cxy[ne_tl] * Dy_tl) * oody * oodx
        Dxy = 0.5 * (cyx[ne_br] * Dx_br - cyx[ne_tr] * Dx_tr + cyx[ne_bl] * Dx_bl - cyx[ne_tl] * Dx_tl) * oodx * oody
        lap[idx] = Dxx + Dyy + Dyx + Dxy + hc[idx] * p[idx]
        lap[idx] *= dinv[idx]
        cnt_x += 1
        if cnt_x % nx == 0:
            cnt_y += 1
            cnt_x = 0
    return lap


def lap2d(p, op):
    """Vectorised lap2D_gather_new: same per-node operation order as the loop."""
    nx, ny = op["nx"], op["ny"]
    P = np.asarray(p, dtype=np.float64).reshape(ny, nx)
    xm = np.r_[nx - 2, np.arange(nx - 1)]          # i-1 with wrap onto NX-2
    xp = np.r_[np.arange(1, nx), 1]                 # i+1 with wrap onto 1
    ym = np.r_[ny - 2, np.arange(ny - 1)]
    yp = np.r_[np.arange(1, ny), 1]
    topleft, topmid, topright = P[ym][:, xm], P[ym], P[ym][:, xp]
    midleft, midmid, midright = P[:, xm], P, P[:, xp]
    botleft, botmid, botright = P[yp][:, xm], P[yp], P[yp][:, xp]

    def quad(a):
        return a[:-1, :-1], a[:-1, 1:], a[1:, :-1], a[1:, 1:]

    hx_tl, hx_tr, hx_bl, hx_br = (q.copy() for q in quad(op["hx"]))
    hy_tl, hy_tr, hy_bl, hy_br = (q.copy() for q in quad(op["hy"]))
    if op["x_wall"]:
        for a in (hx_tl, hy_tl, hx_bl, hy_bl):
            a[:, 0] = 0.0
        for a in (hx_tr, hy_tr, hx_br, hy_br):
            a[:, -1] = 0.0
    if op["y_wall"]:
        for a in (hx_tl, hy_tl, hx_tr, hy_tr):
            a[0, :] = 0.0
        for a in (hx_bl, hy_bl, hx_br, hy_br):
            a[-1, :] = 0.0
    cxx_tl, cxx_tr, cxx_bl, cxx_br = quad(op["cxx"])
    cyy_tl, cyy_tr, cyy_bl, cyy_br = quad(op["cyy"])
    cxy_tl, cxy_tr, cxy_bl, cxy_br = quad(op["cxy"])
    cyx_tl, cyx_tr, cyx_bl, cyx_br = quad(op["cyx"])
    oodx, oody = 1.0 / op["dx"], 1.0 / op["dy"]
    Dx_tl = 0.5 * (topmid - topleft + midmid - midleft) * hx_tl
    Dx_tr = 0.5 * (topright - topmid + midright - midmid) * hx_tr
    Dx_bl = 0.5 * (botmid - botleft + midmid - midleft) * hx_bl
    Dx_br = 0.5 * (botright - botmid + midright - midmid) * hx_br
    Dy_tl = 0.5 * (midmid - topmid + midleft - topleft) * hy_tl
    Dy_tr = 0.5 * (midright - topright + midmid - topmid) * hy_tr
    Dy_bl = 0.5 * (botmid - midmid + botleft - midleft) * hy_bl
    Dy_br = 0.5 * (botright - midright + botmid - midmid) * hy_br
    Dxx = 0.5 * (cxx_tr * Dx_tr - cxx_tl * Dx_tl + cxx_br * Dx_br - cxx_bl * Dx_bl) * oodx * oodx
    Dyy = 0.5 * (cyy_br * Dy_br - cyy_tr * Dy_tr + cyy_bl * Dy_bl - cyy_tl * Dy_tl) * oody * oody
    Dyx = 0.5 * (cxy_br * Dy_br - cxy_bl * Dy_bl + cxy_tr * Dy_tr - cxy_tl * Dy_tl) * oody * oodx
    Dxy = 0.5 * (cyx_br * Dx_br - cyx_tr * Dx_tr + cyx_bl * Dx_bl - cyx_tl * Dx_tl) * oodx * oody
    lap = Dxx + Dyy + Dyx + Dxy + op["hc"] * P
    lap *= op["dinv"]
    return lap.ravel()


def probe_csr(apply, n, shape2d, colors=3):
    """Assemble the matrix of a stencil operator by probing with coloured unit
    vectors (colors x colors colouring on a (ny, nx) grid); exact for stencils
    of radius < colors/2 apart from the periodic wrap, which is why the default
    colouring is checked against dense columns in the tests."""
    import scipy.sparse as sp
    ny, nx = shape2d
    rows, cols, vals = [], [], []
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    for cy in range(colors):
        for cx in range(colors):
            mask = ((jj % colors) == cy) & ((ii % colors) == cx)
            e = mask.astype(np.float64).ravel()
            y = apply(e).reshape(ny, nx)
            src_j = jj[mask]
            src_i = ii[mask]
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    tj, ti = src_j + dj, src_i + di
                    ok = (tj >= 0) & (tj < ny) & (ti >= 0) & (ti < nx)
                    rows.append((tj[ok] * nx + ti[ok]))
                    cols.append((src_j[ok] * nx + src_i[ok]))
                    vals.append(y[tj[ok], ti[ok]])
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()
    A.eliminate_zeros()
    return A


# ---------------------------------------------------------------------------
# 3D: the 27-point operator.  lap3D (laplacian.py:898-1026) cannot run in the
# reference (its branch uses diag_inv before assignment, second_projection.py:
# 183-203, and its reshape order disagrees with hcenter's shape), so this build
# defines the operator from lap3D's structure on a clean layout:
#   * nodes (nz, ny, nx), x fastest; periodic axes wrap onto the duplicated end
#     node exactly as the 2D operator does (i-1 -> n-2 at i=0, i+1 -> 1 at
#     i=n-1); wall axes zero the weights of the outside cells;
#   * cell weights hx, hy, hz on (nz+1, ny+1, nx+1) cells (cell (k,j,i) has
#     corner nodes k-1..k, j-1..j, i-1..i);
#   * per cell, the face-averaged difference along each axis is the sum of the
#     4 edge differences of that cube, times 1/4 (lap3D's x_flx etc.);
#   * node value = coeff * sum over the 8 surrounding cells of
#     sign * weight * difference / d^2 (coeff = 1/4 per the 4-edge average and
#     the 1/2 of the node's half-cell, so 1/16 overall as in lap3D's coeff),
#     plus hcenter * p, times diag_inv.  No Coriolis cross terms
#     (coriolis_strength = 0 in every configuration here).
# ---------------------------------------------------------------------------

def _wrap_index(n, periodic):
    lo = np.r_[n - 2 if periodic else 0, np.arange(n - 1)]
    hi = np.r_[np.arange(1, n), 1 if periodic else n - 1]
    return lo, hi


def lap3d(p, op):
    """Build-defined 27-point nodal operator (see module notes)."""
    nx, ny, nz = op["nx"], op["ny"], op["nz"]
    P = np.asarray(p, dtype=np.float64).reshape(nz, ny, nx)
    per = op["periodic"]  # (x, y, z)
    xm, xp = _wrap_index(nx, per[0])
    ym, yp = _wrap_index(ny, per[1])
    zm, zp = _wrap_index(nz, per[2])
    # extended array with one halo layer on every side, halo = wrapped node
    ix = np.r_[xm[0], np.arange(nx), xp[-1]]
    iy = np.r_[ym[0], np.arange(ny), yp[-1]]
    iz = np.r_[zm[0], np.arange(nz), zp[-1]]
    E = P[np.ix_(iz, iy, ix)]
    # per cell (nz+1, ny+1, nx+1): differences averaged over the 4 edges
    dX = E[:, :, 1:] - E[:, :, :-1]
    gx = dX[:-1, :-1] + dX[:-1, 1:] + dX[1:, :-1] + dX[1:, 1:]
    dY = E[:, 1:, :] - E[:, :-1, :]
    gy = dY[:-1, :, :-1] + dY[:-1, :, 1:] + dY[1:, :, :-1] + dY[1:, :, 1:]
    dZ = E[1:, :, :] - E[:-1, :, :]
    gz = dZ[:, :-1, :-1] + dZ[:, :-1, 1:] + dZ[:, 1:, :-1] + dZ[:, 1:, 1:]
    hx, hy, hz = op["hx"].copy(), op["hy"].copy(), op["hz"].copy()
    for axis, periodic in ((2, per[0]), (1, per[1]), (0, per[2])):
        if not periodic:  # wall: zero the outside cell layers
            for h in (hx, hy, hz):
                sl = [slice(None)] * 3
                sl[axis] = 0
                h[tuple(sl)] = 0.0
                sl[axis] = -1
                h[tuple(sl)] = 0.0
    Fx = hx * gx / op["dx"] ** 2
    Fy = hy * gy / op["dy"] ** 2
    Fz = hz * gz / op["dz"] ** 2

    def oct_sum(F, axis):
        # sum over the 8 cells around each node with sign + for cells on the
        # high side along ``axis`` and - on the low side
        s = 0.0
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    sub = F[a:a + nz, b:b + ny, c:c + nx]
                    sign = (a, b, c)[axis] * 2 - 1
                    s = s + sign * sub
        return s

    lap = (oct_sum(Fz, 0) + oct_sum(Fy, 1) + oct_sum(Fx, 2)) / 16.0
    lap = lap + op["hc"] * P
    lap = lap * op["dinv"]
    return lap.ravel()


def assemble3d(st, dt):
    """3D operator arrays for ``lap3d`` from a workloads.State3D, restating
    operator_coefficients_nodes (second_projection.py:308-356: coeff =
    Gammainv rhoY Y on cells for every wplus[d], wcenter = ccenter * 2x2x2
    mean of rhoY^(2-gamma) at nodes, wall node planes halved with the other
    axes restricted to their inner nodes, :433-464), the regime scaling
    (:139-158) and precon_diag_prepare in 3D (laplacian.py:1284-1356: coeff
    1/16, diag = wcenter - sum of the box sums of wplus times 1/dx^2, 1/dy^2,
    1/dz^2 and twice 1/(dx dy), on the wall-unzeroed weights)."""
    th, ud = st.th, st.ud
    Y = st.rhoY_c / st.rho_c
    coeff = th.Gammainv * st.rhoY_c * Y
    ccenter = -ud.Msq * th.gm1inv / dt ** 2
    r = st.rhoY_c ** (2.0 - th.gamm)
    box = (r[:-1, :-1, :-1] + r[:-1, :-1, 1:] + r[:-1, 1:, :-1] + r[:-1, 1:, 1:]
           + r[1:, :-1, :-1] + r[1:, :-1, 1:] + r[1:, 1:, :-1] + r[1:, 1:, 1:])
    wc = ccenter * box / 8.0
    per = st.periodic
    for axis, periodic in ((2, per[0]), (1, per[1]), (0, per[2])):
        if not periodic:
            for end in (0, -1):
                sl = [slice(1, -1)] * 3
                sl[axis] = end
                wc[tuple(sl)] *= 0.5
    if ud.is_compressible == 0:
        wc = wc * ud.compressibility
    hb = (coeff[:-1, :-1, :-1] + coeff[:-1, :-1, 1:] + coeff[:-1, 1:, :-1] + coeff[:-1, 1:, 1:]
          + coeff[1:, :-1, :-1] + coeff[1:, :-1, 1:] + coeff[1:, 1:, :-1] + coeff[1:, 1:, 1:])
    c = 0.0625
    diag = wc - c * hb * (1.0 / st.dx ** 2 + 1.0 / st.dy ** 2 + 1.0 / st.dz ** 2 + 2.0 / (st.dx * st.dy))
    return dict(nx=st.nx, ny=st.ny, nz=st.nz, dx=st.dx, dy=st.dy, dz=st.dz, periodic=per,
                hx=coeff, hy=coeff, hz=coeff, hc=wc, dinv=1.0 / diag)
