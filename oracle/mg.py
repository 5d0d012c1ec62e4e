"""Two-branch multigrid cycle of the 9-point nodal operator (oracle; test
infrastructure only).

No reference counterpart: pyBELLA preconditions its nodal BiCGStab with the
diagonal only (laplacian.py:1284-1356, second_projection.py:205-230).  This is
the checker of the build's multigrid preconditioner, the cycle around the
learned coarse solver.

The operator is the scaled 9-point operator D^-1 A of lap2D_gather_new
(laplacian.py:546-746).  With zero Coriolis and wx = wy it couples only nodes
of equal parity i + j, so it is two nearly independent operators on the two
checkerboard sublattices: a smooth f and the checkerboard c * f are equally
low-energy modes.  A coarse grid of the even-even nodes sees only the sum of
the two sublattices' smooth envelopes through full weighting (c * f restricts
to ~0), so the fine level restricts twice, r and c * r, and adds back
P e_s + c * P e_d: the "two branches".  Below the fine level the Galerkin
operators are ordinary 9-point stencils and the cycle is a standard V-cycle.

Levels live on unique node grids: a periodic axis of the Krylov layout
(which duplicates the periodic end node) keeps its first n - 1 nodes and
wraps; a wall axis keeps all nodes.  Stencils are (9, ny, nx) arrays, entry
k = 3 (dj + 1) + (di + 1) multiplying z[j + dj, i + di].

Pinned (tests/test_oracle_mg.py): expand9 == lap2d; the Galerkin energy
identity <e, Ac f> = <P e, A P f> / 4; R = P^T / 4 (adjoint identity);
bilinear prolongation exact on linear / constant fields; the constant null
vector preserved; the two-grid error contraction with and without the second
branch.
"""
import numpy as np
import scipy.sparse as sp

OFFS = [(dj, di) for dj in (-1, 0, 1) for di in (-1, 0, 1)]
CENTER = 4


class Level:
    """One grid level: node counts, periodic flags, stencil (9, ny, nx)."""

    def __init__(self, ny, nx, py, px, st):
        self.ny, self.nx, self.py, self.px = ny, nx, py, px
        self.st = st

    @property
    def diag(self):
        return self.st[CENTER]


def _nbr(n, periodic, d):
    """Neighbour index (wrapped) and validity mask along an axis of n nodes."""
    idx = np.arange(n) + d
    if periodic:
        return idx % n, np.ones(n, bool)
    ok = (idx >= 0) & (idx < n)
    return np.clip(idx, 0, n - 1), ok


def expand9(op):
    """Stencil of D^-1 A (lap2D_gather_new, laplacian.py:546-746) on the unique
    node grid, from the flux form: per op-cell gx = (tr - tl) + (br - bl),
    gy = (bl - tl) + (br - tr), Fx = Wxx gx + Wyx gy, Fy = Wyy gy + Wxy gx,
    node = sum over its 4 cells of sx Fx + sy Fy, plus hc p, times dinv."""
    nx, ny = op["nx"], op["ny"]
    dx, dy = op["dx"], op["dy"]
    hx, hy = op["hx"].copy(), op["hy"].copy()
    if op["x_wall"]:
        for a in (hx, hy):
            a[:, 0] = 0.0
            a[:, -1] = 0.0
    if op["y_wall"]:
        for a in (hx, hy):
            a[0, :] = 0.0
            a[-1, :] = 0.0
    Wxx = hx * op["cxx"] / (4 * dx * dx)
    Wyy = hy * op["cyy"] / (4 * dy * dy)
    Wyx = hy * op["cxy"] / (4 * dx * dy)
    Wxy = hx * op["cyx"] / (4 * dx * dy)
    st = np.zeros((9, ny, nx))
    # a node's four cells: (row offset, col offset, sx, sy) of cell (j+dcj, i+dci)
    for dcj, dci, sx, sy in ((0, 0, -1, -1), (0, 1, 1, -1), (1, 0, -1, 1), (1, 1, 1, 1)):
        a = Wxx[dcj:dcj + ny, dci:dci + nx]
        b = Wyy[dcj:dcj + ny, dci:dci + nx]
        c = Wyx[dcj:dcj + ny, dci:dci + nx]
        d = Wxy[dcj:dcj + ny, dci:dci + nx]
        for ca in (0, 1):          # corner row within the cell (0 top, 1 bottom)
            for cb in (0, 1):      # corner col (0 left, 1 right)
                gxc = 1.0 if cb else -1.0
                gyc = 1.0 if ca else -1.0
                coef = sx * (a * gxc + c * gyc) + sy * (b * gyc + d * gxc)
                oj, oi = dcj - 1 + ca, dci - 1 + cb
                st[3 * (oj + 1) + (oi + 1)] += coef
    st[CENTER] += op["hc"]
    st *= op["dinv"][None]
    px, py = not op["x_wall"], not op["y_wall"]
    if px:
        st = st[:, :, :nx - 1]
    if py:
        st = st[:, :ny - 1, :]
    return Level(st.shape[1], st.shape[2], py, px, st)


def apply(lv, z):
    """(A z) of a level; z (ny, nx)."""
    out = np.zeros((lv.ny, lv.nx))
    for k, (dj, di) in enumerate(OFFS):
        jj, jok = _nbr(lv.ny, lv.py, dj)
        ii, iok = _nbr(lv.nx, lv.px, di)
        zz = z[jj][:, ii] * jok[:, None] * iok[None, :]
        out += lv.st[k] * zz
    return out


def to_unique(v, py, px):
    v = np.asarray(v)
    if px:
        v = v[:, :-1]
    if py:
        v = v[:-1, :]
    return v


def to_dup(v, py, px):
    if px:
        v = np.concatenate([v, v[:, :1]], axis=1)
    if py:
        v = np.concatenate([v, v[:1, :]], axis=0)
    return v


def coarse_dims(lv):
    if lv.px:
        assert lv.nx % 2 == 0
        nxc = lv.nx // 2
    else:
        assert lv.nx % 2 == 1
        nxc = (lv.nx + 1) // 2
    if lv.py:
        assert lv.ny % 2 == 0
        nyc = lv.ny // 2
    else:
        assert lv.ny % 2 == 1
        nyc = (lv.ny + 1) // 2
    return nyc, nxc


def _p1d(n_f, n_c, periodic):
    """1D linear interpolation coarse (n_c) -> fine (n_f), fine node 2I = coarse I."""
    rows, cols, vals = [], [], []
    for i in range(n_f):
        if i % 2 == 0:
            rows.append(i); cols.append(i // 2); vals.append(1.0)
        else:
            for c in ((i - 1) // 2, (i + 1) // 2):
                if periodic:
                    c %= n_c
                rows.append(i); cols.append(c); vals.append(0.5)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n_f, n_c))


def prolong_matrix(lv):
    nyc, nxc = coarse_dims(lv)
    return sp.kron(_p1d(lv.ny, nyc, lv.py), _p1d(lv.nx, nxc, lv.px)).tocsr()


def prolong(lv, e):
    """Bilinear interpolation of a coarse field e (nyc, nxc) to level lv."""
    return (prolong_matrix(lv) @ e.ravel()).reshape(lv.ny, lv.nx)


def restrict(lv, r):
    """Full weighting R = P^T / 4 of a field r on level lv."""
    nyc, nxc = coarse_dims(lv)
    return (0.25 * (prolong_matrix(lv).T @ r.ravel())).reshape(nyc, nxc)


def matrix(lv):
    """Sparse matrix of a level's stencil."""
    n = lv.ny * lv.nx
    J, I = np.meshgrid(np.arange(lv.ny), np.arange(lv.nx), indexing="ij")
    rows, cols, vals = [], [], []
    for k, (dj, di) in enumerate(OFFS):
        jj, jok = _nbr(lv.ny, lv.py, dj)
        ii, iok = _nbr(lv.nx, lv.px, di)
        ok = (jok[:, None] & iok[None, :]).ravel()
        rows.append((J * lv.nx + I).ravel()[ok])
        cols.append((jj[J] * lv.nx + ii[I]).ravel()[ok])
        vals.append(lv.st[k].ravel()[ok])
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()


def galerkin(lv):
    """Coarse level with stencil R A P (R = P^T / 4), read back into (9, nyc, nxc)."""
    nyc, nxc = coarse_dims(lv)
    P = prolong_matrix(lv)
    Ac = (0.25 * (P.T @ matrix(lv) @ P)).tocsr()
    st = np.zeros((9, nyc, nxc))
    J, I = np.meshgrid(np.arange(nyc), np.arange(nxc), indexing="ij")
    for k, (dj, di) in enumerate(OFFS):
        jj, jok = _nbr(nyc, lv.py, dj)
        ii, iok = _nbr(nxc, lv.px, di)
        ok = jok[:, None] & iok[None, :]
        rows = (J * nxc + I)
        cols = jj[J] * nxc + ii[I]
        vals = np.asarray(Ac[rows.ravel(), cols.ravel()]).reshape(nyc, nxc)
        st[k] = np.where(ok, vals, 0.0)
    return Level(nyc, nxc, lv.py, lv.px, st)


def hierarchy(op, nlev):
    """Level 0 (fine, unique grid) and nlev - 1 Galerkin levels."""
    levels = [expand9(op)]
    for _ in range(nlev - 1):
        levels.append(galerkin(levels[-1]))
    return levels


def checkerboard(ny, nx):
    return np.where((np.arange(ny)[:, None] + np.arange(nx)[None, :]) % 2 == 0, 1.0, -1.0)


def jacobi(lv, v, z, sweeps, omega):
    """z <- z + omega (v - A z) / diag, ``sweeps`` times; z None = start from 0
    (the first sweep is then omega v / diag)."""
    D = lv.diag
    if z is None:
        z = omega * v / D
        sweeps -= 1
    for _ in range(sweeps):
        z = z + omega * (v - apply(lv, z)) / D
    return z


def vcycle(levels, l, v, coarse_solve, nu=(2, 2), omega=0.8):
    """Standard V-cycle from level l (l >= 1) for A_l z = v; the last level is
    solved by ``coarse_solve(v)``."""
    lv = levels[l]
    if l == len(levels) - 1:
        return coarse_solve(v)
    z = jacobi(lv, v, None, nu[0], omega)
    r = v - apply(lv, z)
    z = z + prolong(lv, vcycle(levels, l + 1, restrict(lv, r), coarse_solve, nu, omega))
    return jacobi(lv, v, z, nu[1], omega)


def precondition(levels, v, coarse_solve, nu=(2, 2), omega=0.8, branches=2):
    """The two-branch cycle z = M v on the fine level (unique grid)."""
    lv = levels[0]
    z = jacobi(lv, v, None, nu[0], omega)
    r = v - apply(lv, z)
    c = checkerboard(lv.ny, lv.nx)
    e = prolong(lv, vcycle(levels, 1, restrict(lv, r), coarse_solve, nu, omega))
    if branches == 2:
        e = e + c * prolong(lv, vcycle(levels, 1, restrict(lv, c * r), coarse_solve, nu, omega))
    z = z + e
    return jacobi(lv, v, z, nu[1], omega)


def exact_coarse(lv):
    """Least-squares solve of the last level (singular for the pseudo-
    incompressible operator: its null space is the constant)."""
    A = matrix(lv).toarray()
    pinv = np.linalg.pinv(A)
    return lambda v: (pinv @ v.ravel()).reshape(lv.ny, lv.nx)
