"""CPU pins of the network oracle (oracle/nn.py), which checks the BF16 MFMA
kernels.  There is no reference counterpart; the pins are independent
implementations:

* torch (CPU, float64): ``conv2d`` with padding 1, bilinear
  ``interpolate(align_corners=True)`` from 2^k + 1 to 2^(k+1) + 1 nodes,
  strided slicing, erf-GELU and autograd -- forward and every parameter
  gradient of the whole network;
* central finite differences of the projected loss along random directions;
* adjoint identities <up(c), f> = <c, up_adj(f)>, <down(f), c> = <f, down_adj(c)>;
* bf16 rounding against torch's float32 -> bfloat16 conversion.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import nn as onn


def torch_forward(params, X):
    """The same network written with torch ops (float64, NCHW)."""
    P = [(torch.tensor(W).permute(2, 1, 0).reshape(W.shape[2], W.shape[1], 3, 3).requires_grad_(True),
          torch.tensor(b).requires_grad_(True)) for W, b in onn.unpack(params)]
    x = torch.tensor(X).permute(0, 3, 1, 2)

    def cv(l, t):
        W, b = P[l]
        z = F.conv2d(t, W, b, padding=1)
        return F.gelu(z) if onn.LAYERS[l][2] else z

    def up(t):
        h, w = t.shape[2:]
        return F.interpolate(t, size=(2 * h - 1, 2 * w - 1), mode="bilinear", align_corners=True)

    e0 = cv(1, cv(0, x))
    e1 = cv(3, cv(2, e0[:, :, ::2, ::2]))
    e2 = cv(5, cv(4, e1[:, :, ::2, ::2]))
    b3 = cv(7, cv(6, e2[:, :, ::2, ::2]))
    d2 = cv(9, cv(8, torch.cat([up(b3), e2], 1)))
    d1 = cv(11, cv(10, torch.cat([up(d2), e1], 1)))
    d0 = cv(13, cv(12, torch.cat([up(d1), e0], 1)))
    z = cv(14, d0)[:, 0]
    return z, P


def torch_flat_grad(P):
    out = []
    for W, b in P:
        out.append(W.grad.reshape(W.shape[0], W.shape[1], 9).permute(2, 1, 0).reshape(-1).numpy())
        out.append(b.grad.numpy())
    return np.concatenate(out)


def rand_params(seed):
    """He-scaled random weights (activations stay O(1)), non-zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for ci, co, _ in onn.LAYERS:
        out.append(rng.standard_normal(9 * ci * co) * np.sqrt(2.0 / (9 * ci)))
        out.append(rng.standard_normal(co) * 0.1)
    return np.concatenate(out)


def test_num_params():
    assert onn.num_params() == 302977


@pytest.mark.parametrize("shape", [(1, 17, 17), (2, 17, 33)])
def test_forward_and_gradients_match_torch(shape):
    n, h, w = shape
    rng = np.random.default_rng(1)
    params = rand_params(2)
    X = rng.standard_normal((n, h, w, 3))
    target = onn.project_null(rng.standard_normal((n, h, w)))
    z, cache = onn.forward(params, X)
    zt, P = torch_forward(params, X)
    assert np.max(np.abs(z - zt.detach().numpy())) < 1e-10 * np.max(np.abs(z))
    L, dz = onn.loss(z, target)
    g = onn.backward(params, cache, dz)
    # torch loss: same projected MSE
    zz = zt
    c = torch.tensor(np.where((np.arange(h)[:, None] + np.arange(w)[None, :]) % 2 == 0, 1.0, -1.0))
    B = torch.stack([torch.ones(h * w, dtype=torch.float64), c.reshape(-1)], 1)
    flat = zz.reshape(n, -1)
    coef = torch.linalg.solve(B.T @ B, B.T @ flat.T).T
    d = flat - coef @ B.T - torch.tensor(target).reshape(n, -1)
    Lt = (d * d).sum() / zz.numel()
    Lt.backward()
    assert abs(L - Lt.item()) < 1e-12 * abs(L)
    gt = torch_flat_grad(P)
    assert np.max(np.abs(g - gt)) < 1e-9 * np.max(np.abs(gt))


def test_finite_difference_directions():
    rng = np.random.default_rng(3)
    params = rand_params(4)
    X = rng.standard_normal((1, 17, 17, 3))
    target = onn.project_null(rng.standard_normal((1, 17, 17)))

    def L(p):
        return onn.loss(onn.forward(p, X)[0], target)[0]

    z, cache = onn.forward(params, X)
    g = onn.backward(params, cache, onn.loss(z, target)[1])
    for k in range(3):
        d = rng.standard_normal(params.size)
        eps = 1e-5
        fd = (L(params + eps * d) - L(params - eps * d)) / (2 * eps)
        assert abs(fd - g @ d) < 1e-6 * max(abs(fd), 1e-8), (fd, g @ d)


def test_up_down_adjoints():
    rng = np.random.default_rng(5)
    c = rng.standard_normal((2, 9, 5, 3))
    f = rng.standard_normal((2, 17, 9, 3))
    assert abs((onn.up(c) * f).sum() - (c * onn.up_adj(f)).sum()) < 1e-12
    assert abs((onn.down(f) * c).sum() - (f * onn.down_adj(c, f.shape)).sum()) < 1e-12
    # linear interpolation reproduces linear fields exactly
    jj, ii = np.meshgrid(np.arange(9), np.arange(5), indexing="ij")
    lin = (0.3 * jj - 0.7 * ii)[None, :, :, None]
    J, I = np.meshgrid(np.arange(17) / 2, np.arange(9) / 2, indexing="ij")
    assert np.max(np.abs(onn.up(lin)[0, :, :, 0] - (0.3 * J - 0.7 * I))) < 1e-14


def test_conv_dgrad_is_adjoint():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1, 9, 7, 4))
    dz = rng.standard_normal((1, 9, 7, 5))
    W = rng.standard_normal((9, 4, 5))
    lhs = (onn.conv3x3(x, W, np.zeros(5)) * dz).sum()
    rhs = (x * onn.conv3x3_dgrad(dz, W)).sum()
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-8, 8, 4000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(onn.to_bf16(x), ref)


def test_projected_loss_ignores_null_space():
    rng = np.random.default_rng(8)
    z = rng.standard_normal((2, 9, 9))
    t = onn.project_null(rng.standard_normal((2, 9, 9)))
    c = np.where((np.arange(9)[:, None] + np.arange(9)[None, :]) % 2 == 0, 1.0, -1.0)
    L0 = onn.loss(z, t)[0]
    L1 = onn.loss(z + 3.0 - 2.0 * c, t)[0]
    assert abs(L0 - L1) < 1e-12 * L0
