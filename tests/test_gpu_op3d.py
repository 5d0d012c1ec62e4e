"""GPU parity of the 3D 27-point operator and its solve against the oracle
(laplacian.assemble3d / lap3d, bicgstab), through the C ABI.

Tolerances: assembly and FP64 matvec <= 1e-13 relative (max-norm; FMA and
summation order differ), FP32 matvec <= 2e-5; BiCGStab iterates after k
steps <= 1e-9; converged solves reach the true relative residual asked for.
"""
import numpy as np
import pytest

from oracle import bicgstab as obs
from oracle import laplacian as olp
from pybella_amd import workloads as W

from _cases import rel

pytestmark = pytest.mark.gpu

SIZES = [(17, 9, 9), (33, 17, 19), (70, 11, 13)]


def _state(nx, ny, nz, periodic=(True, False, True), compressible=True):
    st = W.stratified_state_3d(nx, ny, nz, compressible=compressible)
    st.periodic = periodic
    # a horizontally varying perturbation so the weights differ cell by cell
    rng = np.random.default_rng(nx + ny + nz)
    bump = 1.0 + 0.05 * rng.standard_normal(st.rho_c.shape)
    st.rhoY_c = st.rhoY_c * bump
    return st


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("periodic", [(True, False, True), (False, False, False)])
@pytest.mark.parametrize("nx,ny,nz", SIZES)
def test_assemble_and_matvec(nx, ny, nz, periodic):
    import torch
    from pybella_amd.op3d import Operator3D, grid3_of, phys3_of
    print(f"case op3d {nx}x{ny}x{nz} periodic={periodic}", flush=True)
    st = _state(nx, ny, nz, periodic)
    ref = olp.assemble3d(st, 0.05)
    op = Operator3D(grid3_of(st))
    wc = torch.empty(nz, ny, nx, dtype=torch.float64, device="cuda")
    di = torch.empty_like(wc)
    op.assemble(phys3_of(st), 0.05, _dev(st.rho_c), _dev(st.rhoY_c), wc, di)
    assert rel(wc.cpu().numpy(), ref["hc"]) < 1e-13
    assert rel(di.cpu().numpy(), ref["dinv"]) < 1e-13
    p = np.random.default_rng(3).standard_normal(nx * ny * nz)
    y_ref = olp.lap3d(p, ref)
    y64 = op.matvec(_dev(p)).cpu().numpy()
    assert rel(y64, y_ref) < 1e-13
    y32 = op.matvec(_dev(p.astype(np.float32))).cpu().numpy()
    assert rel(y32, y_ref) < 2e-5


@pytest.mark.parametrize("k", [1, 2, 5])
def test_first_iterates(k):
    from pybella_amd.op3d import Operator3D
    print(f"case op3d iterates k={k}", flush=True)
    st = _state(33, 17, 19)
    ref = olp.assemble3d(st, 0.05)
    op = Operator3D.from_state(st)
    b = W.fourier_rhs_3d(33, 17, 19, seed=5, nmodes=32, kmax=6).ravel() * ref["dinv"].ravel()
    x, info = op.solve(_dev(b), tol=1e-30, maxiter=k, mode="f64")
    r = obs.bicgstab(lambda v: olp.lap3d(v, ref), b, tol=1e-30, maxiter=k)
    assert info["iters"] == k == r.niter
    assert rel(x.cpu().numpy().ravel(), r.x) < 1e-9


@pytest.mark.parametrize("check_every,k", [(2, 3), (2, 8), (4, 9), (32, 33)])
def test_iterates_across_chunks(check_every, k):
    """Each launch applies the scalar step of the launch before it, and a
    one-block fold applies the last pending step at the end of every chunk of
    iterations.  Iterates that cross chunk boundaries, stopping inside a chunk
    or on its last iteration, match the oracle's (a skipped or repeated step
    would change omega or beta at O(1))."""
    from pybella_amd.op3d import Operator3D
    print(f"case op3d iterates across chunks check_every={check_every} k={k}", flush=True)
    st = _state(33, 17, 19)
    ref = olp.assemble3d(st, 0.05)
    op = Operator3D.from_state(st)
    b = W.fourier_rhs_3d(33, 17, 19, seed=5, nmodes=32, kmax=6).ravel() * ref["dinv"].ravel()
    x, info = op.solve(_dev(b), tol=1e-30, maxiter=k, mode="f64", check_every=check_every)
    r = obs.bicgstab(lambda v: olp.lap3d(v, ref), b, tol=1e-30, maxiter=k)
    assert info["iters"] == k == r.niter
    assert rel(x.cpu().numpy().ravel(), r.x) < 1e-7


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_converged_solve_short_chunks(mode):
    """Two-iteration chunks: a fold after every second iteration; the solve
    converges inside a chunk and stops there."""
    from pybella_amd.op3d import Operator3D
    print(f"case op3d solve {mode} check_every=2", flush=True)
    st = _state(33, 17, 19)
    ref = olp.assemble3d(st, 0.05)
    op = Operator3D.from_state(st)
    b = W.fourier_rhs_3d(33, 17, 19, seed=6, nmodes=32, kmax=6).ravel() * ref["dinv"].ravel()
    x, info = op.solve(_dev(b), tol=1e-10, mode=mode, check_every=2)
    assert info["info"] == 0 and info["true_rel_resid"] <= 1e-10, info
    mv = lambda v: olp.lap3d(v, ref)
    assert obs.true_rel_resid(mv, b, x.cpu().numpy().ravel()) < 2e-10


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_converged_solve(mode):
    from pybella_amd.op3d import Operator3D
    print(f"case op3d solve {mode}", flush=True)
    st = _state(33, 17, 19)
    ref = olp.assemble3d(st, 0.05)
    op = Operator3D.from_state(st)
    b = W.fourier_rhs_3d(33, 17, 19, seed=6, nmodes=32, kmax=6).ravel() * ref["dinv"].ravel()
    x, info = op.solve(_dev(b), tol=1e-10, mode=mode)
    assert info["info"] == 0 and info["true_rel_resid"] <= 1e-10, info
    mv = lambda v: olp.lap3d(v, ref)
    assert obs.true_rel_resid(mv, b, x.cpu().numpy().ravel()) < 2e-10
    r = obs.bicgstab(mv, b, tol=1e-12, maxiter=5000)
    assert rel(x.cpu().numpy().ravel(), r.x) < 1e-7


def test_large_mixed_solve():
    """129 x 65 x 65 (a quarter of the 257 x 129 x 129 workload per axis
    pair): mixed FP32/FP64 solve to true relative residual 1e-8."""
    import time

    import torch
    from pybella_amd.op3d import Operator3D
    print("case op3d 129x65x65 mixed", flush=True)
    st = W.stratified_state_3d(129, 65, 65)
    op = Operator3D.from_state(st)
    ref_dinv = olp.assemble3d(st, 0.05)["dinv"]
    b = _dev(W.fourier_rhs_3d(129, 65, 65, seed=7).ravel() * ref_dinv.ravel())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = op.solve(b, tol=1e-8, mode="mixed")
    torch.cuda.synchronize()
    print("mixed 129x65x65:", info, f"{time.perf_counter() - t0:.3f} s", flush=True)
    assert info["info"] == 0 and info["true_rel_resid"] <= 1e-8


def test_reassembly_other_regime_replays_fresh_graph():
    """One handle: a compressible solve captures its iteration chunks as
    graphs; re-assembled in the pseudo-incompressible regime (no centre term),
    the solve's first iterates must be the oracle's, not a replay with the
    compressible launch arguments."""
    import torch
    from pybella_amd.op3d import Operator3D, phys3_of
    print("case op3d reassembly compressible -> psinc", flush=True)
    st_c = _state(33, 17, 19, compressible=True)
    op = Operator3D.from_state(st_c)
    ref_c = olp.assemble3d(st_c, 0.05)
    b_c = W.fourier_rhs_3d(33, 17, 19, seed=8, nmodes=32, kmax=6).ravel() * ref_c["dinv"].ravel()
    _, info = op.solve(_dev(b_c), tol=1e-10, mode="f64")
    assert info["info"] == 0
    st_p = _state(33, 17, 19, compressible=False)
    ref_p = olp.assemble3d(st_p, 0.05)
    op.assemble(phys3_of(st_p), 0.05, _dev(st_p.rho_c), _dev(st_p.rhoY_c))
    b_p = W.fourier_rhs_3d(33, 17, 19, seed=9, nmodes=32, kmax=6).ravel() * ref_p["dinv"].ravel()
    for k in (1, 5):
        x, info = op.solve(_dev(b_p), tol=1e-30, maxiter=k, mode="f64")
        torch.cuda.synchronize()
        r = obs.bicgstab(lambda v: olp.lap3d(v, ref_p), b_p, tol=1e-30, maxiter=k)
        assert info["iters"] == k == r.niter
        assert rel(x.cpu().numpy().ravel(), r.x) < 1e-9, k
