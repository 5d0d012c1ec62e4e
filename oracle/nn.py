"""NumPy forward/backward of the learned-preconditioner network (oracle; test
infrastructure only).  No reference counterpart: pyBELLA has no learned part.

Same architecture and parameter layout as pybella_amd/csrc/pb_nn.hip:
15 3x3 convolutions (zero padding), GELU (erf form), injection down-sampling
(coarse(i, j) = fine(2i, 2j)), linear up-sampling on 2^k + 1 node grids,
channel concatenation [up-sampled, skip]; flat parameters per layer
W[9][cin][cout] (tap = 3 (dy + 1) + (dx + 1)) then bias[cout].  The loss is the
mean squared error of the output after removing its least-squares constant
and checkerboard components.

``bf16=True`` rounds (round-to-nearest-even) at the points where the GPU
stores bf16: the packed weights, the network input, every stored activation
and pre-activation; arithmetic stays float64 otherwise.

Pinned by: scipy.signal.correlate (convolution), math.erf (GELU), central
finite differences of the loss (backward), explicit small cases for up/down.
"""
import numpy as np
from scipy.special import erf

LAYERS = [(3, 16, 1), (16, 16, 1), (16, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 64, 1), (64, 64, 1),
          (128, 64, 1), (64, 64, 1), (96, 32, 1), (32, 32, 1), (48, 16, 1), (16, 16, 1), (16, 1, 0)]


def num_params():
    return sum(9 * ci * co + co for ci, co, _ in LAYERS)


def unpack(params):
    out, off = [], 0
    for ci, co, _ in LAYERS:
        W = params[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        b = params[off:off + co]
        off += co
        out.append((W, b))
    return out


def to_bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(finite, r, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def conv3x3(x, W, b):
    n, h, w, ci = x.shape
    co = W.shape[2]
    xp = np.zeros((n, h + 2, w + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((n, h, w, co)) + b
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] @ W[tap]
    return out


def conv3x3_dgrad(dz, W):
    """Adjoint of conv3x3 w.r.t. its input."""
    n, h, w, co = dz.shape
    ci = W.shape[1]
    dxp = np.zeros((n, h + 2, w + 2, ci))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dxp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += dz @ W[tap].T
    return dxp[:, 1:-1, 1:-1]


def conv3x3_wgrad(x, dz):
    n, h, w, ci = x.shape
    xp = np.zeros((n, h + 2, w + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros((9, ci, dz.shape[3]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        xs = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w].reshape(-1, ci)
        dW[tap] = xs.T @ dz.reshape(-1, dz.shape[3])
    return dW, dz.reshape(-1, dz.shape[3]).sum(axis=0)


def down(x):
    return x[:, ::2, ::2, :]


def down_adj(dc, shape):
    df = np.zeros(shape)
    df[:, ::2, ::2, :] = dc
    return df


def _up1(c, axis):
    n = c.shape[axis]
    f_shape = list(c.shape)
    f_shape[axis] = 2 * n - 1
    f = np.zeros(f_shape)
    sl = [slice(None)] * c.ndim
    sl[axis] = slice(0, None, 2)
    f[tuple(sl)] = c
    a = [slice(None)] * c.ndim
    b = [slice(None)] * c.ndim
    a[axis] = slice(0, -1)
    b[axis] = slice(1, None)
    sl[axis] = slice(1, None, 2)
    f[tuple(sl)] = 0.5 * (c[tuple(a)] + c[tuple(b)])
    return f


def up(c):
    return _up1(_up1(c, 1), 2)


def _up1_adj(df, axis):
    n = (df.shape[axis] + 1) // 2
    sh = list(df.shape)
    sh[axis] = n
    dc = np.zeros(sh)
    sl = [slice(None)] * df.ndim
    sl[axis] = slice(0, None, 2)
    dc += df[tuple(sl)]
    sl[axis] = slice(1, None, 2)
    odd = df[tuple(sl)]
    a = [slice(None)] * df.ndim
    b = [slice(None)] * df.ndim
    a[axis] = slice(0, -1)
    b[axis] = slice(1, None)
    dc[tuple(a)] += 0.5 * odd
    dc[tuple(b)] += 0.5 * odd
    return dc


def up_adj(df):
    return _up1_adj(_up1_adj(df, 2), 1)


def forward(params, X, bf16=False):
    """X: (n, h, w, 3) -> z (n, h, w); returns (z, cache)."""
    R = to_bf16 if bf16 else (lambda a: a)
    P = unpack(params)
    if bf16:
        P = [(R(W), b) for W, b in P]
    cache = {"P": P, "X": R(X)}

    def cv(l, x):
        W, b = P[l]
        pre = conv3x3(x, W, b)
        act = gelu(pre) if LAYERS[l][2] else pre
        cache[f"pre{l}"] = R(pre)
        cache[f"in{l}"] = x
        return R(act)

    e0a = cv(0, cache["X"])
    e0b = cv(1, e0a)
    e1a = cv(2, down(e0b))
    e1b = cv(3, e1a)
    e2a = cv(4, down(e1b))
    e2b = cv(5, e2a)
    b3a = cv(6, down(e2b))
    b3b = cv(7, b3a)
    cat2 = np.concatenate([R(up(b3b)), e2b], axis=3)
    d2a = cv(8, cat2)
    d2b = cv(9, d2a)
    cat1 = np.concatenate([R(up(d2b)), e1b], axis=3)
    d1a = cv(10, cat1)
    d1b = cv(11, d1a)
    cat0 = np.concatenate([R(up(d1b)), e0b], axis=3)
    d0a = cv(12, cat0)
    d0b = cv(13, d0a)
    W, b = P[14]
    z = conv3x3(d0b, W, b)[..., 0]
    cache["in14"] = d0b
    cache["shapes"] = (e0b.shape, e1b.shape, e2b.shape)
    return z, cache


def null_fit(z):
    """Least-squares (mu, kappa) of z on the constant and the checkerboard, per image."""
    n, h, w = z.shape
    c = np.where((np.arange(h)[:, None] + np.arange(w)[None, :]) % 2 == 0, 1.0, -1.0)
    B = np.stack([np.ones(h * w), c.ravel()], axis=1)
    coef = np.linalg.solve(B.T @ B, B.T @ z.reshape(n, -1).T).T
    return coef, c


def project_null(v):
    """v minus its least-squares constant and checkerboard components."""
    coef, c = null_fit(v)
    return v - coef[:, 0, None, None] - coef[:, 1, None, None] * c


def loss(z, target):
    """Mean squared error of the projected output; the gradient is projected
    as well (the projection is symmetric), so it is exact for any target."""
    d = project_null(z) - target
    N = z.size
    return float((d * d).sum() / N), project_null(2.0 * d / N)


def backward(params, cache, dz):
    """dz: dL/dz (n, h, w) -> flat gradient in the GPU layout."""
    P = cache["P"]
    grads = [None] * 15

    def bw(l, d_out):
        """d_out: gradient w.r.t. layer l's activation output."""
        d = d_out * gelu_grad(cache[f"pre{l}"]) if LAYERS[l][2] else d_out
        grads[l] = conv3x3_wgrad(cache[f"in{l}"], d)
        return conv3x3_dgrad(d, P[l][0])

    d = bw(14, dz[..., None])              # -> d d0b
    d = bw(13, d)                           # -> d d0a
    dcat0 = bw(12, d)
    d_d1b = up_adj(dcat0[..., :32])
    d_e0b = dcat0[..., 32:]
    d = bw(11, d_d1b)
    dcat1 = bw(10, d)
    d_d2b = up_adj(dcat1[..., :64])
    d_e1b = dcat1[..., 64:]
    d = bw(9, d_d2b)
    dcat2 = bw(8, d)
    d_b3b = up_adj(dcat2[..., :64])
    d_e2b = dcat2[..., 64:]
    d = bw(7, d_b3b)
    d_p3 = bw(6, d)
    e0s, e1s, e2s = cache["shapes"]
    d_e2b = d_e2b + down_adj(d_p3, e2s)
    d = bw(5, d_e2b)
    d_p2 = bw(4, d)
    d_e1b = d_e1b + down_adj(d_p2, e1s)
    d = bw(3, d_e1b)
    d_p1 = bw(2, d)
    d_e0b = d_e0b + down_adj(d_p1, e0s)
    d = bw(1, d_e0b)
    bw(0, d)
    flat = []
    for (dW, db) in grads:
        flat.append(dW.ravel())
        flat.append(db.ravel())
    return np.concatenate(flat)
