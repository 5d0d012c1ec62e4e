"""GPU multigrid cycle of the learned mode vs the multigrid oracle (C ABI).

  * the hierarchy: level-0 stencil (expand9) and every Galerkin level R A P
    against oracle/mg.py in FP64, relative <= 1e-12, on the stratified
    pseudo-incompressible operator, a compressible one with Coriolis cross
    terms, one with walls on both axes and the doubly periodic travelling
    vortex;
  * one application z = M v of the FP32 cycle against the FP64 oracle cycle
    with the same coarse solve (Jacobi sweeps on the network's level, or the
    cycle continued to the coarsest level), with one and two branches and 1
    or 2 sweeps: relative <= 2e-5 (FP32 arithmetic through ~20 launches);
  * FGMRES with the cycle (no network: Jacobi on the coarse level, or
    multigrid to the bottom) converges to the FP64 true residual 1e-8 and
    matches the oracle's BiCGStab solution modulo the null space.
"""
import copy

import numpy as np
import pytest

from oracle import bicgstab as obs
from oracle import laplacian as olp
from oracle import mg
from pybella_amd import workloads as W

from _cases import make_case, oracle_op

pytestmark = pytest.mark.gpu


def _sp():
    from pybella_amd import second_projection as sp
    return sp


def gpu_op(Sol, mpv, elem, node, ud, th, dt=0.05):
    sp = _sp()
    op = sp.Operator2D(sp.grid_of(node, ud))
    rho, rhoY = sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY)
    op.assemble(sp.phys_of(ud, th), dt, rho, rhoY, rho, rhoY, sp.to_dev(sp.dSdy_cells(mpv, elem, node)))
    return op


def case(kind, nx, ny):
    if kind == "strat":
        Sol, mpv, elem, node, ud, th = W.stratified_state(nx, ny)
        return Sol, mpv, elem, node, ud, th
    Sol, mpv, elem, node, ud, th = make_case(kind, nx, ny)[:6]
    return Sol, mpv, elem, node, ud, th


@pytest.mark.parametrize("kind,nx,ny", [("strat", 257, 129), ("coriolis", 129, 65), ("xwall", 65, 33),
                                        ("yper", 65, 65), ("strat", 1025, 513)])
def test_hierarchy_stencils(kind, nx, ny):
    print(f"case hierarchy {kind} {nx}x{ny}", flush=True)
    st = case(kind, nx, ny)
    op_o, _, _ = oracle_op(*st, 0.05)
    op = gpu_op(*st)
    levels, nnl = op.mg_info()
    L = [mg.expand9(op_o)]
    while len(L) < len(levels):
        L.append(mg.galerkin(L[-1]))
    for l, lv in enumerate(L):
        assert levels[l] == (lv.ny, lv.nx, int(lv.py), int(lv.px)), (l, levels[l])
        got = op.mg_stencil(l).cpu().numpy()
        scale = np.max(np.abs(lv.st))
        assert np.max(np.abs(got - lv.st)) <= 1e-12 * scale, (l, np.max(np.abs(got - lv.st)) / scale)
    if nx == 1025:   # the network's level: 33 x 64 unique nodes, 33 x 65 with the periodic column
        assert nnl == 4 and levels[4][:2] == (33, 64)
    op.close()


def _oracle_cycle(L, nnl, v_dup, coarse, nu, omega, branches):
    lv0 = L[0]
    v = mg.to_unique(v_dup, lv0.py, lv0.px)
    if coarse == "smooth":
        levels = L[:nnl + 1]
        cs = lambda r: mg.jacobi(levels[-1], r, None, 2, omega)
    else:
        levels = L
        cs = lambda r: mg.jacobi(levels[-1], r, None, 40, omega)
    z = mg.precondition(levels, v, cs, nu=(nu, nu), omega=omega, branches=branches)
    return mg.to_dup(z, lv0.py, lv0.px)


@pytest.mark.parametrize("kind,nx,ny,coarse,nu,branches", [
    ("strat", 257, 129, "smooth", 2, 2), ("strat", 257, 129, "mg", 2, 2), ("strat", 257, 129, "smooth", 1, 1),
    ("coriolis", 129, 65, "smooth", 2, 2), ("xwall", 65, 33, "mg", 2, 2), ("yper", 65, 65, "mg", 2, 2),
    ("yper", 65, 65, "smooth", 2, 2), ("strat", 1025, 513, "smooth", 2, 2)])
def test_cycle_matches_oracle(kind, nx, ny, coarse, nu, branches):
    import torch
    print(f"case cycle {kind} {nx}x{ny} {coarse} nu={nu} branches={branches}", flush=True)
    st = case(kind, nx, ny)
    op_o, _, _ = oracle_op(*st, 0.05)
    op = gpu_op(*st)
    levels, nnl = op.mg_info()
    L = [mg.expand9(op_o)]
    while len(L) < len(levels):
        L.append(mg.galerkin(L[-1]))
    rng = np.random.default_rng(7)
    v = rng.standard_normal((ny, nx))
    if not op_o["x_wall"]:
        v[:, -1] = v[:, 0]
    if not op_o["y_wall"]:
        v[-1, :] = v[0, :]
    vd = torch.from_numpy(v).cuda()
    z = op.precond(vd, coarse=coarse, nu=nu, omega=0.8, branches=branches)
    torch.cuda.synchronize()
    ref = _oracle_cycle(L, nnl, v, coarse, nu, 0.8, branches)
    got = z.double().cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= 2e-5, err
    op.close()


def strip_null(v2d):
    ny, nx = v2d.shape
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    B = np.stack([np.ones(ny * nx), np.where((ii + jj) % 2 == 0, 1.0, -1.0).ravel()], axis=1)
    c, *_ = np.linalg.lstsq(B, v2d.ravel(), rcond=None)
    return (v2d.ravel() - B @ c).reshape(ny, nx)


@pytest.mark.parametrize("nx,ny,use", [(257, 129, "mg"), (257, 129, False), (1025, 513, "mg")])
def test_fgmres_multigrid_converges(nx, ny, use):
    import torch
    sp = _sp()
    print(f"case fgmres {nx}x{ny} coarse={use}", flush=True)
    st = case("strat", nx, ny)
    op_o, _, dinv = oracle_op(*st, 0.05)
    op = gpu_op(*st)
    node = st[3]
    rhs = W.stratified_rhs(node, seed=20261202)
    b = op.scale_rhs(sp.to_dev(rhs))
    x, info = op.solve(b, tol=1e-8, maxiter=400, mode="nn", use_network=use, fallback_maxiter=1)
    torch.cuda.synchronize()
    assert info["info"] == 0 and info["fallback"] == 0, info
    assert info["true_rel_resid"] <= 1e-8
    if use == "mg":
        assert info["iters"] <= 20, info
    # oracle solution of the same system (FP64 BiCGStab to 1e-12), modulo null space
    if nx <= 257:
        bh = b.cpu().numpy().ravel()
        ref = obs.bicgstab(lambda q: olp.lap2d(q, op_o), bh, tol=1e-12, maxiter=20000)
        xs = strip_null(x.cpu().numpy())
        rs = strip_null(ref.x.reshape(ny, nx))
        assert np.linalg.norm(xs - rs) / np.linalg.norm(rs) <= 1e-6
    op.close()
