"""GPU parity of the learned-preconditioner network against the NumPy oracle.

* forward: GPU (bf16 storage, FP32 accumulation) vs the oracle run in float64
  with bf16 rounding at the same storage points: relative L2 <= 2e-2 (the
  remaining difference is accumulation order flipping bf16 roundings);
* gradients: GPU parameter gradients (bf16 gradient tensors) vs the oracle's
  float64 backward on the same rounded forward: per-layer cosine >= 0.99 and
  relative L2 <= 0.1; the oracle's backward is itself pinned by finite
  differences in the CPU tests.
"""
import numpy as np
import pytest

from oracle import nn as onn

pytestmark = pytest.mark.gpu


def _net(seed=3):
    from pybella_amd.nn import Network
    return Network(seed)


def _inputs(nimg, h, w, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((nimg, h, w, 3)).astype(np.float32)
    X[..., 1:] = 0.2 + 0.05 * rng.standard_normal((nimg, h, w, 2))
    T = rng.standard_normal((nimg, h, w))
    T = onn.project_null(T).astype(np.float32)
    return X, T


@pytest.mark.parametrize("shape", [(1, 17, 17), (2, 33, 17), (1, 65, 33)])
def test_forward_matches_oracle(shape):
    import torch
    print(f"case nn forward {shape}", flush=True)
    net = _net()
    P = net.get_params().astype(np.float64)
    X, _ = _inputs(*shape, seed=sum(shape))
    z = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    zr, _ = onn.forward(P, X.astype(np.float64), bf16=True)
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    print("rel L2", err)
    assert err < 2e-2


def test_gradients_match_oracle():
    import torch
    print("case nn gradients 2x33x17", flush=True)
    net = _net(5)
    P = net.get_params().astype(np.float64)
    X, T = _inputs(2, 33, 17, seed=11)
    loss = net.step(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), lr=0.0)
    g = net.get_grads().astype(np.float64)
    z, cache = onn.forward(P, X.astype(np.float64), bf16=True)
    lr_, dz = onn.loss(z, T.astype(np.float64))
    gr = onn.backward(P, cache, dz)
    assert abs(loss - lr_) / lr_ < 2e-2
    off = 0
    stats = []
    for l, (ci, co, _) in enumerate(onn.LAYERS):
        nw = 9 * ci * co
        for part, (a, b) in (("W", (g[off:off + nw], gr[off:off + nw])), ("b", (g[off + nw:off + nw + co], gr[off + nw:off + nw + co]))):
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
            rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300))
            print(l, part, round(cos, 5), round(rel, 5), flush=True)
            stats.append((l, part, cos, rel))
        off += nw + co
    for l, part, cos, rel in stats:
        if l == 14 and part == "b":
            continue   # sum of a projected gradient: exactly zero in exact arithmetic
        assert cos > 0.99 and rel < 0.1, (l, part, cos, rel)


def test_training_reduces_loss():
    """Adam steps on a fixed batch lower the loss (the update path runs)."""
    import torch
    print("case nn adam", flush=True)
    net = _net(7)
    X, T = _inputs(2, 33, 17, seed=13)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    l0 = net.step(Xd, Td, lr=0.0)
    for _ in range(30):
        net.step(Xd, Td, lr=2e-3)
    l1 = net.step(Xd, Td, lr=0.0)
    assert l1 < 0.7 * l0, (l0, l1)


def test_image_weights_scale_gradients():
    """Per-image loss weights: loss and gradients of a weighted 2-image batch
    equal the weighted combination of the single-image ones (each single-image
    loss is normalised by its own pixel count, the batch by twice that)."""
    import torch
    print("case nn image weights", flush=True)
    net = _net(11)
    X, T = _inputs(2, 33, 17, seed=21)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    la = net.step(Xd[:1].contiguous(), Td[:1].contiguous(), lr=0.0)
    ga = net.get_grads().astype(np.float64)
    lb = net.step(Xd[1:].contiguous(), Td[1:].contiguous(), lr=0.0)
    gb = net.get_grads().astype(np.float64)
    w = torch.tensor([0.25, 3.0], dtype=torch.float32, device="cuda")
    lw = net.step(Xd, Td, lr=0.0, img_weight=w)
    gw = net.get_grads().astype(np.float64)
    assert abs(lw - 0.5 * (0.25 * la + 3.0 * lb)) < 1e-4 * lw
    ref = 0.5 * (0.25 * ga + 3.0 * gb)
    assert np.linalg.norm(gw - ref) < 0.03 * np.linalg.norm(ref)


def _stratified_system(nx, ny, seed=20261202):
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    Sol, mpv, elem, node, ud, th = W.stratified_state(nx, ny)
    op = sp.Operator2D(sp.grid_of(node, ud))
    rho, rhoY = sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY)
    op.assemble(sp.phys_of(ud, th), 0.05, rho, rhoY, rho, rhoY, sp.to_dev(sp.dSdy_cells(mpv, elem, node)))
    b = op.scale_rhs(sp.to_dev(W.stratified_rhs(node, seed=seed)))
    torch.cuda.synchronize()
    return op, b


def _trained():
    from pybella_amd.nn import Network, default_weights_path
    net = Network(0)
    net.load(default_weights_path())
    return net


@pytest.mark.parametrize("budget,sweeps", [(0, 2), (20, 1), (40, 0)])
def test_nn_solve_reaches_tolerance(budget, sweeps):
    """The learned-preconditioned solve (network guess, flexible GMRES budget,
    plain BiCGStab after it) reaches true relative residual 1e-8."""
    print(f"case nn solve 257x129 budget {budget} sweeps {sweeps}", flush=True)
    op, b = _stratified_system(257, 129)
    net = _trained()
    op.attach_nn(net)
    x, info = op.solve(b, tol=1e-8, mode="nn", maxiter=budget, sweeps=sweeps, fallback_maxiter=6000)
    assert info["info"] == 0 and info["true_rel_resid"] <= 1e-8, info
    # the same solution as plain BiCGStab, modulo the null space {1, checkerboard}
    xp, infop = op.solve(b, tol=1e-12, mode="f64", maxiter=6000)
    d = (x - xp).cpu().numpy()
    d = onn.project_null(d[None])[0]
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(onn.project_null(xp.cpu().numpy()[None])[0])


def test_network_gain_over_jacobi_only():
    """A/B: the same flexible GMRES (60 iterations, no fallback), with the
    trained network + one Jacobi sweep vs one Jacobi sweep alone; the network
    must lower the true residual reached by at least 5x."""
    print("case nn A/B 513x257", flush=True)
    op, b = _stratified_system(513, 257)
    net = _trained()
    op.attach_nn(net)
    _, with_nn = op.solve(b, tol=1e-8, mode="nn", maxiter=60, sweeps=1, fallback_maxiter=1)
    _, without = op.solve(b, tol=1e-8, mode="nn", maxiter=60, sweeps=1, fallback_maxiter=1, use_network=False)
    print("with network", with_nn, "without", without, flush=True)
    assert with_nn["true_rel_resid"] * 5 < without["true_rel_resid"]


def test_features_follow_reassembly():
    """The network's cached operator features are refreshed when the same
    operator handle is re-assembled on a different state."""
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    print("case nn features cache", flush=True)
    net = _net(1)
    Sol, mpv, elem, node, ud, th = W.stratified_state(65, 33)
    op = sp.Operator2D(sp.grid_of(node, ud))
    dS = sp.to_dev(sp.dSdy_cells(mpv, elem, node))
    op.assemble(sp.phys_of(ud, th), 0.05, sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY), sp.to_dev(Sol.rho),
                sp.to_dev(Sol.rhoY), dS)
    f1a = net.features(op)[0].clone()
    rhoY2 = Sol.rhoY * (1.0 + 0.2 * np.random.default_rng(0).random(Sol.rhoY.shape))
    op.assemble(sp.phys_of(ud, th), 0.05, sp.to_dev(Sol.rho), sp.to_dev(rhoY2), sp.to_dev(Sol.rho),
                sp.to_dev(rhoY2), dS)
    f1b = net.features(op)[0].clone()
    torch.cuda.synchronize()
    assert float((f1a - f1b).abs().max()) > 1e-4


def test_dropin_learned_mode():
    """The drop-in projection step with the learned preconditioner attached to
    the grid's operator (mode "nn": network guess + 2 damped Jacobi sweeps,
    then plain BiCGStab capped at ud.max_iterations) reaches the tolerance and
    updates the state like the FP64 mode: momenta and rhoX within 1e-5 of
    their magnitude (both solves stop at relative residual 1e-8)."""
    import copy
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    print("case nn drop-in 257x129 rising bubble", flush=True)
    Sol, Sol0, mpv, elem, node, ud, th = W.rising_bubble_state(257, 129)
    A = [copy.deepcopy(o) for o in (Sol, Sol0, mpv)]
    B = [copy.deepcopy(o) for o in (Sol, Sol0, mpv)]
    saved = dict(sp.solver_options)
    op = sp.operator_for(node, ud)
    try:
        sp.solver_options["mode"] = "f64"
        ia = sp.euler_backward_non_advective_impl_part(A[0], A[2], elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=A[1])
        op.attach_nn(_trained())
        sp.solver_options.update(mode="nn", nn_budget=0, nn_sweeps=2)
        ib = sp.euler_backward_non_advective_impl_part(B[0], B[2], elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=B[1])
    finally:
        sp.solver_options.clear()
        sp.solver_options.update(saved)
        op.attach_nn(None)
    print("f64", ia, "nn", ib, flush=True)
    assert ia["info"] == 0 and ib["info"] == 0 and ib["fallback"] == 1, (ia, ib)
    assert ib["true_rel_resid"] <= 1e-8
    for k in ("rhou", "rhov", "rhoX"):
        a, r = getattr(B[0], k), getattr(A[0], k)
        assert np.max(np.abs(a - r)) / np.max(np.abs(r)) < 1e-5, k
