"""Training of the learned coarse solver (pybella_amd.nn.Network) on the GPU.

The network solves the coarse level of the multigrid cycle that
preconditions the learned mode (pb_mg2d.hpp): on a 513 x 257 operator that
level is 65 x 33 nodes (the periodic column duplicated), the same size as on
a 1025 x 513 operator one level further down.  Training pairs (b, x) with
A_c x = b on that level come from 513 x 257 operators assembled on the
hydrostatic stratified atmosphere (pseudo-incompressible and compressible)
and six rising-bubble ensemble members, in two kinds:
  * solver pairs: b is a seeded smooth random right-hand side (64 Fourier
    modes, |k| <= 16, made consistent with the singular operator) scaled by
    the diagonal preconditioner, x its 513 x 257 solution by the plain GPU
    BiCGStab (relative residual 1e-7); the pair is x injected onto the coarse
    level and b = A_c x there;
  * operator pairs: x is a random field with a random power-law spectrum on
    the coarse level and b = A_c x (cheap, and covering the rougher
    right-hand sides the cycle hands down).
Each batch holds half of each kind.

Network input: [b / rms(b), f1, f2] (f1, f2 the coarse stencil's x and y
coupling relative to its centre).  Target: x minus its mean (the constant is
the null space of the pseudo-incompressible coarse operator), scaled to unit
rms: the cycle fits the scale of the network's output in the energy norm
(alpha = z.b / z.A_c z), so the network learns the shape of A_c^-1 b.  Loss:
mean squared error of the output minus its mean.

    python -m pybella_amd.nn_train            (writes pybella_amd/weights/coarse_65x33.npz)
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


class CoarseLevel:
    """The network level of an operator's hierarchy: its FP64 stencil on the
    unique grid, the injection stride from the operator's grid and the
    network features on the duplicated grid."""

    def __init__(self, op):
        levels, nnl = op.mg_info()
        if nnl < 0:
            raise ValueError("no coarse level of this grid fits the network")
        self.op = op
        self.ny, self.nx, self.py, self.px = levels[nnl]
        self.stride = 2 ** nnl
        self.st = op.mg_stencil(nnl)
        self.f1, self.f2 = op.mg_features()

    def apply(self, x):
        """A_c x for a batch x (n, ny, nx) on the unique grid (periodic axes
        wrap; a wall's outside entries of the stencil are zero)."""
        import torch
        out = torch.zeros_like(x)
        for k in range(9):
            dj, di = k // 3 - 1, k % 3 - 1
            out += self.st[k] * torch.roll(x, shifts=(-dj, -di), dims=(1, 2))
        return out

    def inject(self, x_fine):
        """Krylov-layout fine vectors (n, NY, NX) -> coarse unique grid."""
        if self.px:
            x_fine = x_fine[..., :-1]
        if self.py:
            x_fine = x_fine[..., :-1, :]
        return x_fine[..., ::self.stride, ::self.stride]

    def dup(self, x):
        import torch
        if self.px:
            x = torch.cat([x, x[..., :1]], dim=-1)
        if self.py:
            x = torch.cat([x, x[..., :1, :]], dim=-2)
        return x


def training_cases(nx=513, ny=257, members=6):
    cases = [W.stratified_state(nx, ny), W.stratified_state(nx, ny, compressible=True)]
    cases += [W.batch_member_bubble(nx, ny, m) for m in range(members)]
    return [assembled_operator(c) for c in cases]


def fourier_rhs_batch(gen, count, nx, ny, nmodes=64, kmax=16, device="cuda"):
    """``count`` right-hand sides of workloads.stratified_rhs's kind on the
    device: each a sum of ``nmodes`` modes a cos(2 pi kx s + phi) cos(pi ky t)
    (s, t in [0, 1] over the nodes, integer (kx, ky) != 0 with |k| <= kmax,
    a ~ N(0, 1), phi ~ U(0, 2 pi)), made orthogonal to the operator's left
    null vectors, in the device node layout [y][x] (NY+2, NX+2) with a zero
    ghost ring: (count, NY+2, NX+2)."""
    import torch
    ks = torch.tensor([(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(0, kmax + 1)
                       if (kx, ky) != (0, 0) and kx * kx + ky * ky <= kmax * kmax],
                      dtype=torch.float64, device=device)
    pick = torch.argsort(torch.rand(count, ks.shape[0], generator=gen, device=device), dim=1)[:, :nmodes]
    k = ks[pick]                                                                   # (c, m, 2)
    amp = torch.randn(count, nmodes, generator=gen, device=device, dtype=torch.float64)
    phi = 2 * math.pi * torch.rand(count, nmodes, generator=gen, device=device, dtype=torch.float64)
    sx = torch.linspace(0.0, 1.0, nx, device=device, dtype=torch.float64)
    ty = torch.linspace(0.0, 1.0, ny, device=device, dtype=torch.float64)
    cx = torch.cos(2 * math.pi * k[..., 0, None] * sx + phi[..., None])           # (c, m, nx)
    cy = torch.cos(math.pi * k[..., 1, None] * ty)                                 # (c, m, ny)
    f = torch.einsum("cm,cmy,cmx->cyx", amp, cy, cx)                               # (c, ny, nx)
    w = torch.ones(ny, nx, dtype=torch.float64, device=device)
    w[:, 0] = w[:, -1] = 0.5
    chk = torch.ones_like(w)
    chk[(torch.arange(ny, device=device)[:, None] + torch.arange(nx, device=device)[None]) % 2 == 1] = -1.0
    Wn = torch.stack([w.reshape(-1), (w * chk).reshape(-1)])                      # (2, n)
    coef = torch.linalg.solve(Wn @ Wn.T, Wn @ f.reshape(count, -1).T).T            # (c, 2)
    f = f - (coef @ Wn).reshape(count, ny, nx)
    out = torch.zeros(count, ny + 2, nx + 2, dtype=torch.float64, device=device)
    out[:, 1:-1, 1:-1] = f
    return out


def solver_pairs(ops, levels, count, seed=1000, tol=1e-7):
    """``count`` coarse pairs (b_c, x_c) from fine solves, FP32 on the device."""
    import torch
    lv0 = levels[0]
    X = torch.empty(count, lv0.ny, lv0.nx, dtype=torch.float64, device="cuda")
    which = np.zeros(count, dtype=np.int64)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    t0 = time.time()
    iters = []
    rhs = None
    k = drawn = skipped = 0
    while k < count:
        c = k % len(ops)
        op = ops[c]
        if drawn % 256 == 0:   # right-hand sides made 256 at a time
            rhs = fourier_rhs_batch(gen, 256, op.nx, op.ny)
        b = op.scale_rhs(rhs[drawn % 256])
        drawn += 1
        x, info = op.solve(b, tol=tol, maxiter=6000, mode="f64fast")
        for _ in range(3):   # a breakdown (rho or omega ~ 0) restarts from the iterate reached
            if info["info"] == 0:
                break
            x, info2 = op.solve(b, tol=tol, maxiter=6000, mode="f64fast", x0=x.clone())
            info = dict(info2, iters=info["iters"] + info2["iters"])
        if info["info"] != 0 or info["true_rel_resid"] > 10 * tol:
            skipped += 1
            if skipped > count // 100:
                raise RuntimeError(f"too many training solves did not converge; last: {info}")
            continue
        X[k] = levels[c].inject(x)
        which[k] = c
        iters.append(info["iters"])
        k += 1
        if k % 1000 == 0 or k == count:
            log(f"solver pairs {k}/{count}  mean iterations {np.mean(iters[-1000:]):.0f}  skipped {skipped}  "
                f"{time.time() - t0:6.1f} s")
    B = torch.empty_like(X)
    for c in range(len(ops)):
        idx = torch.from_numpy(np.nonzero(which == c)[0]).cuda()
        if idx.numel():
            B[idx] = levels[c].apply(X[idx])
    return B.float(), X.float(), which


def random_fields(gen, nimg, ny, nx):
    """Random fields with spectrum (1 + k^2/k0^2)^(-p/2), p ~ U(0.5, 4),
    k0 ~ U(1, 16) (k in grid wavenumbers), normalised to unit rms."""
    import torch
    white = torch.randn(nimg, ny, nx, generator=gen, device="cuda", dtype=torch.float64)
    F = torch.fft.rfft2(white)
    ky = torch.fft.fftfreq(ny, device="cuda", dtype=torch.float64) * ny
    kx = torch.fft.rfftfreq(nx, device="cuda", dtype=torch.float64) * nx
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    p = 0.5 + 3.5 * torch.rand(nimg, 1, 1, generator=gen, device="cuda", dtype=torch.float64)
    k0 = 1.0 + 15.0 * torch.rand(nimg, 1, 1, generator=gen, device="cuda", dtype=torch.float64)
    x = torch.fft.irfft2(F * (1.0 + k2[None] / k0 ** 2) ** (-p / 2), s=(ny, nx))
    return x / x.pow(2).mean(dim=(1, 2), keepdim=True).sqrt()


def unit(v):
    """v minus its mean, scaled to unit rms, per image."""
    v = v - v.mean(dim=(1, 2), keepdim=True)
    return v / v.pow(2).mean(dim=(1, 2), keepdim=True).sqrt().clamp_min(1e-30)


class PairSource:
    """Batches of (inputs, target) on the device, network layout: half solver
    pairs drawn from a fixed set, half fresh operator pairs of one level."""

    def __init__(self, levels, solver_set=None, seed=0):
        import torch
        self.levels = levels
        self.gen = torch.Generator(device="cuda")
        self.gen.manual_seed(seed)
        self.solver_set = solver_set
        self.k = 0

    def batch(self, nimg):
        import torch
        c = self.k % len(self.levels)
        lv = self.levels[c]
        self.k += 1
        n_op = nimg if self.solver_set is None else nimg // 2
        x = random_fields(self.gen, n_op, lv.ny, lv.nx)
        b = lv.apply(x)
        cidx = np.full(n_op, c)
        if n_op < nimg:
            B, X, which = self.solver_set
            idx = torch.randint(0, B.shape[0], (nimg - n_op,), generator=self.gen, device="cuda")
            b = torch.cat([b, B[idx].double()])
            x = torch.cat([x, X[idx].double()])
            cidx = np.concatenate([cidx, which[idx.cpu().numpy()]])
        f1 = torch.stack([self.levels[k].f1 for k in cidx])
        f2 = torch.stack([self.levels[k].f2 for k in cidx])
        b = b / b.pow(2).mean(dim=(1, 2), keepdim=True).sqrt()
        t = unit(x)
        inputs = torch.stack([lv.dup(b).float(), f1, f2], dim=-1).contiguous()
        return inputs, lv.dup(t).float().contiguous()


def train(steps=6000, batch=32, lr=1.5e-3, seed=0, nx=513, ny=257, log_every=100, net=None, solver_count=20000):
    import torch
    net = net or Network(seed=seed)
    net.ctx.check(net.ctx.lib.pb_nn_set_loss_nullspace(net.h, 1), "pb_nn_set_loss_nullspace")
    ops = training_cases(nx, ny)
    levels = [CoarseLevel(op) for op in ops]
    solver_set = solver_pairs(ops, levels, solver_count) if solver_count > 0 else None
    src = PairSource(levels, solver_set, seed=seed + 1)
    warm = min(200, steps // 10 + 1)
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
    for op in ops:
        op.close()
    return net, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=default_weights_path())
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--solver-pairs", type=int, default=20000, help="(b, x) pairs made by the plain solver")
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1.5e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights")
    args = ap.parse_args(argv)
    net = Network(seed=args.seed)
    if args.init:
        net.load(args.init)
    net, hist = train(args.steps, args.batch, args.lr, args.seed, net=net, solver_count=args.solver_pairs)
    net.save(args.out)
    log(f"saved {args.out}; final loss {np.mean(hist[-100:]):.4e}")


if __name__ == "__main__":
    main()
