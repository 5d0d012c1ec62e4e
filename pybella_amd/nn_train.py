"""Training of the learned preconditioner (pybella_amd.nn.Network) on the GPU.

Training pairs (b, x) with A x = b come from nodal operators assembled on
513 x 257 atmospheres (the hydrostatic stratified state, pseudo-incompressible
and compressible, and rising-bubble ensemble members), in two kinds:
  * solver pairs: b is a seeded smooth random right-hand side (64 Fourier
    modes, |k| <= 16, made consistent with the singular operator) scaled by
    the diagonal preconditioner, and x its solution by the plain GPU
    BiCGStab (relative residual 1e-7);
  * operator pairs: x is a random field with a random power-law spectrum,
    from rough (p ~ 0.5) to smooth (p ~ 4), and b = A x is one FP32 matvec,
    so these cost one stencil application instead of a solve and cover the
    rougher residuals seen inside the Krylov loop.
Each batch holds half of each kind.

Network input: [b / rms(b), f1, f2] (f1, f2 the operator's coefficient
features).  Target: x with its constant and checkerboard components removed,
scaled to unit rms.  The preconditioner only needs the *shape* of A^-1 b:
the flexible Krylov method is invariant to the scale of each preconditioned
direction, and the initial guess is rescaled by the alpha that minimises
||b - alpha A z|| on the device.  Loss: mean squared error (projected).

    python -m pybella_amd.nn_train            (writes pybella_amd/weights/precond_513x257.npz)
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import second_projection as sp
from . import workloads as W
from .nn import Network, default_weights_path



def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assembled_operator(case, dt=0.05):
    """Operator2D assembled on a workload state (Sol0 drives the coefficients
    when the state carries one, as in second_projection.py:113-140)."""
    if len(case) == 7:
        Sol, Sol0, mpv, elem, node, ud, th = case
        S = Sol0
    else:
        Sol, mpv, elem, node, ud, th = case
        S = Sol
    op = sp.Operator2D(sp.grid_of(node, ud))
    rho, rhoY = sp.to_dev(S.rho), sp.to_dev(S.rhoY)
    op.assemble(sp.phys_of(ud, th), dt, rho, rhoY, rho, rhoY, sp.to_dev(sp.dSdy_cells(mpv, elem, node)))
    return op


def training_cases(nx=513, ny=257, members=6):
    """[(operator, node grid)] for the training atmospheres."""
    cases = [W.stratified_state(nx, ny), W.stratified_state(nx, ny, compressible=True)]
    cases += [W.batch_member_bubble(nx, ny, m) for m in range(members)]
    return [(assembled_operator(c), c[-3]) for c in cases]


def fourier_rhs_dev(gen, nx, ny, nmodes=64, kmax=16, device="cuda"):
    """Device version of workloads.stratified_rhs: a sum of ``nmodes`` modes
    a cos(2 pi kx s + phi) cos(pi ky t) (s, t in [0, 1] over the nodes,
    integer (kx, ky) != 0 with |k| <= kmax, a ~ N(0, 1), phi ~ U(0, 2 pi)),
    made orthogonal to the operator's left null vectors, in the device node
    layout [y][x] (NY+2, NX+2) with a zero ghost ring (what second_projection.
    to_dev makes of the host (NX+2, NY+2) array)."""
    import torch
    ks = torch.tensor([(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1)
                       if (kx, ky) != (0, 0) and kx * kx + ky * ky <= kmax * kmax],
                      dtype=torch.float64, device=device)
    pick = torch.randperm(ks.shape[0], generator=gen, device=device)[:nmodes]
    k = ks[pick]
    amp = torch.randn(nmodes, generator=gen, device=device, dtype=torch.float64)
    phi = 2 * math.pi * torch.rand(nmodes, generator=gen, device=device, dtype=torch.float64)
    sx = torch.linspace(0.0, 1.0, nx, device=device, dtype=torch.float64)
    ty = torch.linspace(0.0, 1.0, ny, device=device, dtype=torch.float64)
    cx = torch.cos(2 * math.pi * k[:, 0, None] * sx[None] + phi[:, None])       # (m, nx)
    cy = torch.cos(math.pi * k[:, 1, None] * ty[None])                           # (m, ny)
    f = (amp[:, None] * cx).T @ cy                                                # (nx, ny)
    w = torch.ones(nx, ny, dtype=torch.float64, device=device)
    w[0] = w[-1] = 0.5
    c = torch.ones_like(w)
    c[(torch.arange(nx, device=device)[:, None] + torch.arange(ny, device=device)[None]) % 2 == 1] = -1.0
    Wn = torch.stack([w.reshape(-1), (w * c).reshape(-1)])
    coef = torch.linalg.solve(Wn @ Wn.T, Wn @ f.reshape(-1))
    f = f - (coef @ Wn).reshape(nx, ny)
    out = torch.zeros(ny + 2, nx + 2, dtype=torch.float64, device=device)
    out[1:-1, 1:-1] = f.T
    return out


def solver_pairs(cases, count, seed=1000, tol=1e-7):
    """``count`` (b, x) pairs, FP32 on the device, spread over the cases."""
    import torch
    B = torch.empty(count, cases[0][0].ny, cases[0][0].nx, dtype=torch.float32, device="cuda")
    X = torch.empty_like(B)
    which = np.zeros(count, dtype=np.int64)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    t0 = time.time()
    iters = []
    for k in range(count):
        c = k % len(cases)
        op, node = cases[c]
        b = op.scale_rhs(fourier_rhs_dev(gen, op.nx, op.ny))
        x, info = op.solve(b, tol=tol, maxiter=6000, mode="f64fast")
        if info["info"] != 0:
            raise RuntimeError(f"training solve {k} did not converge: {info}")
        B[k] = b.float()
        X[k] = x.float()
        which[k] = c
        iters.append(info["iters"])
        if (k + 1) % 500 == 0 or k + 1 == count:
            log(f"solver pairs {k + 1}/{count}  mean iterations {np.mean(iters[-500:]):.0f}  "
                f"{time.time() - t0:6.1f} s")
    return B, X, which


def random_fields(gen, nimg, ny, nx):
    """Random fields with spectrum (1 + k^2/k0^2)^(-p/2), p ~ U(0.5, 4),
    k0 ~ U(1, 24) (k in grid wavenumbers), normalised to unit rms."""
    import torch
    white = torch.randn(nimg, ny, nx, generator=gen, device="cuda", dtype=torch.float32)
    F = torch.fft.rfft2(white)
    ky = torch.fft.fftfreq(ny, device="cuda") * ny
    kx = torch.fft.rfftfreq(nx, device="cuda") * nx
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    p = 0.5 + 3.5 * torch.rand(nimg, 1, 1, generator=gen, device="cuda")
    k0 = 1.0 + 23.0 * torch.rand(nimg, 1, 1, generator=gen, device="cuda")
    x = torch.fft.irfft2(F * (1.0 + k2[None] / k0 ** 2) ** (-p / 2), s=(ny, nx))
    return x / x.pow(2).mean(dim=(1, 2), keepdim=True).sqrt()


def project_null(v):
    """v minus its least-squares constant and checkerboard components, per image."""
    import torch
    n, h, w = v.shape
    c = torch.ones(h, w, device=v.device, dtype=v.dtype)
    c[(torch.arange(h, device=v.device)[:, None] + torch.arange(w, device=v.device)[None, :]) % 2 == 1] = -1.0
    B = torch.stack([torch.ones(h * w, device=v.device, dtype=v.dtype), c.reshape(-1)], 1)
    coef = torch.linalg.solve(B.T @ B, B.T @ v.reshape(n, -1).T).T
    return v - (coef @ B.T).reshape(n, h, w)


class PairSource:
    """Batches of (inputs, target) on the device: half solver pairs drawn
    from a fixed set, half fresh operator pairs from one operator."""

    def __init__(self, cases, net, solver_set=None, seed=0):
        import torch
        self.ops = [op for op, _ in cases]
        self.gen = torch.Generator(device="cuda")
        self.gen.manual_seed(seed)
        self.feats = [net.features(op) for op in self.ops]
        self.solver_set = solver_set
        self.k = 0

    def batch(self, nimg):
        import torch
        c = self.k % len(self.ops)
        op = self.ops[c]
        self.k += 1
        n_op = nimg if self.solver_set is None else nimg // 2
        x = random_fields(self.gen, n_op, op.ny, op.nx)
        b = torch.empty_like(x)
        for i in range(n_op):
            op.matvec(x[i].reshape(-1), b[i].reshape(-1))
        f1 = self.feats[c][0].expand_as(b)
        f2 = self.feats[c][1].expand_as(b)
        if n_op < nimg:
            B, X, which = self.solver_set
            idx = torch.randint(0, B.shape[0], (nimg - n_op,), generator=self.gen, device="cuda")
            b = torch.cat([b, B[idx]])
            x = torch.cat([x, X[idx]])
            w = which[idx.cpu().numpy()]
            f1 = torch.cat([f1, torch.stack([self.feats[k][0] for k in w])])
            f2 = torch.cat([f2, torch.stack([self.feats[k][1] for k in w])])
        b = b / b.pow(2).mean(dim=(1, 2), keepdim=True).sqrt()
        t = project_null(x.double()).float()
        t = t / t.pow(2).mean(dim=(1, 2), keepdim=True).sqrt()
        inputs = torch.stack([b, f1, f2], dim=-1).contiguous()
        return inputs, t.contiguous()


def train(pairs=40000, batch=8, lr=1.5e-3, seed=0, nx=513, ny=257, log_every=25, net=None, solver_count=20000):
    import torch
    net = net or Network(seed=seed)
    cases = training_cases(nx, ny)
    solver_set = solver_pairs(cases, solver_count) if solver_count > 0 else None
    src = PairSource(cases, net, solver_set, seed=seed + 1)
    steps = max(1, pairs // batch)
    warm = min(100, steps // 10 + 1)
    t0 = time.time()
    hist = []
    for it in range(steps):
        sched = min(1.0, (it + 1) / warm) * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        inputs, target = src.batch(batch)
        loss = net.step(inputs, target, lr=lr * max(sched, 0.01))
        hist.append(loss)
        if it % log_every == 0 or it == steps - 1:
            log(f"step {it:5d}/{steps}  loss {np.mean(hist[-log_every:]):.4e}  lr {lr * sched:.2e}  "
                f"{time.time() - t0:6.1f} s")
    torch.cuda.synchronize()
    for op, _ in cases:
        op.close()
    return net, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=default_weights_path())
    ap.add_argument("--pairs", type=int, default=40000, help="training samples drawn (batches x batch size)")
    ap.add_argument("--solver-pairs", type=int, default=20000, help="(b, x) pairs made by the plain solver")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1.5e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights")
    args = ap.parse_args(argv)
    net = Network(seed=args.seed)
    if args.init:
        net.load(args.init)
    net, hist = train(args.pairs, args.batch, args.lr, args.seed, net=net, solver_count=args.solver_pairs)
    net.save(args.out)
    log(f"saved {args.out}; final loss {np.mean(hist[-50:]):.4e}")


if __name__ == "__main__":
    main()
