"""The allocation-free FP64 solve (pb_bicgstab_solve_noalloc) against the
allocating one (pb_bicgstab_solve), through the C ABI.

Both run the same kernels at the same tile height, so iterates and results are
bitwise equal; the test checks x, the iteration count, the info code and the
true residual on two calls in a row, the second on a smaller grid (the static
scratch keeps the first call's contents), for both FP64 modes, and that a grid
beyond the scratch's capacity is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from pybella_amd import workloads as W

pytestmark = pytest.mark.gpu


def _system(nx, ny, seed):
    import torch
    from pybella_amd import second_projection as sp
    Sol, mpv, elem, node, ud, th = W.stratified_state(nx, ny)
    op = sp.Operator2D(sp.grid_of(node, ud))
    rho, rhoY = sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY)
    op.assemble(sp.phys_of(ud, th), 0.05, rho, rhoY, rho, rhoY, sp.to_dev(sp.dSdy_cells(mpv, elem, node)))
    b = op.scale_rhs(sp.to_dev(W.stratified_rhs(node, seed=seed)))
    torch.cuda.synchronize()
    return op, b


@pytest.mark.parametrize("mode", ["f64fast", "f64"])
def test_noalloc_equals_allocating_solve(mode):
    import torch
    for nx, ny, seed in ((257, 129, 11), (129, 65, 12)):   # second call on a smaller grid
        print(f"case noalloc {mode} {nx}x{ny}", flush=True)
        op, b = _system(nx, ny, seed)
        x_ref, info_ref = op.solve(b, tol=1e-8, maxiter=6000, mode=mode)
        x = torch.full_like(b, np.nan)   # poisoned: the solve must overwrite it
        x, info = op.solve_noalloc(b, x, tol=1e-8, maxiter=6000, mode=mode)
        torch.cuda.synchronize()
        assert info["info"] == info_ref["info"] == 0, (info, info_ref)
        assert info["iters"] == info_ref["iters"]
        assert info["true_rel_resid"] == info_ref["true_rel_resid"] <= 1e-8
        assert torch.equal(x, x_ref), float((x - x_ref).abs().max())
        op.close()


def test_noalloc_first_iterates_and_guess():
    """Short runs (maxiter 3, not converged) and a given initial guess agree
    bitwise with the allocating solve as well."""
    import torch
    print("case noalloc short runs 129x65", flush=True)
    op, b = _system(129, 65, 13)
    x_ref, info_ref = op.solve(b, tol=1e-30, maxiter=3, mode="f64fast", check_every=2)
    x = torch.zeros_like(b)
    x, info = op.solve_noalloc(b, x, tol=1e-30, maxiter=3, mode="f64fast", check_every=2)
    torch.cuda.synchronize()
    assert info["info"] == info_ref["info"] == 3 and info["iters"] == 3
    assert torch.equal(x, x_ref)
    x0 = x_ref.clone()
    x_ref2, info_ref2 = op.solve(b, tol=1e-8, maxiter=6000, mode="f64fast", x0=x0)
    x2 = x0.clone()
    x2, info2 = op.solve_noalloc(b, x2, tol=1e-8, maxiter=6000, mode="f64fast", x0_given=True)
    torch.cuda.synchronize()
    assert info2["info"] == info_ref2["info"] == 0 and info2["iters"] == info_ref2["iters"]
    assert torch.equal(x2, x_ref2)
    op.close()


def test_noalloc_refuses_oversize_and_mixed():
    """A grid beyond the static scratch and a non-FP64 mode are refused with
    PB_ERR_ARG before any launch."""
    import torch
    from pybella_amd import _abi
    from pybella_amd import second_projection as sp
    print("case noalloc refusals", flush=True)
    op, b = _system(129, 65, 14)
    x = torch.zeros_like(b)
    opts = _abi.SolveOpts(_abi.PB_SOLVE_MIXED, 1e-8, 100, 1e-3, 4, 32, 0, 0, 0, 0, 0.0, 0)
    info = _abi.SolveInfo()
    rc = op.ctx.lib.pb_bicgstab_solve_noalloc(op.h, _abi.ptr(b), _abi.ptr(x), C.byref(opts), C.byref(info))
    assert rc == -1 and "only PB_SOLVE_F64" in op.ctx.lib.pb_last_error(op.ctx.h).decode()
    op.close()
    Sol, mpv, elem, node, ud, th = W.stratified_state(1025, 1025)
    big = sp.Operator2D(sp.grid_of(node, ud))
    rho, rhoY = sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY)
    big.assemble(sp.phys_of(ud, th), 0.05, rho, rhoY, rho, rhoY, sp.to_dev(sp.dSdy_cells(mpv, elem, node)))
    bb = torch.zeros(1025, 1025, dtype=torch.float64, device="cuda")
    xb = torch.zeros_like(bb)
    opts = _abi.SolveOpts(_abi.PB_SOLVE_F64_FAST, 1e-8, 100, 0.0, 0, 32, 0, 0, 0, 0, 0.0, 0)
    rc = big.ctx.lib.pb_bicgstab_solve_noalloc(big.h, _abi.ptr(bb), _abi.ptr(xb), C.byref(opts), C.byref(info))
    msg = big.ctx.lib.pb_last_error(big.ctx.h).decode()
    assert rc == -1 and "static scratch holds" in msg, msg
    big.close()
