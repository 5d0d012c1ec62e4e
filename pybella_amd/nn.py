"""Learned preconditioner: host side of the network in pb_nn.hip.

``Network`` owns a device network (4-level 3x3 conv encoder-decoder,
16-32-64-64 channels, GELU, BF16 MFMA kernels, FP32 master weights).  Its
input is the residual r normalised by its rms plus two coefficient features
of the operator; its output, times rms(r), approximates the solution of
A z = r up to the operator's null space.  Every call runs on torch's current
stream (registered with the context first), so device tensors produced by
torch are ordered before the network reads them.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi


def default_weights_path() -> str:
    """The committed weights of the coarse solver trained by pybella_amd.nn_train."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "coarse_65x33.npz")


class Network:
    def __init__(self, seed: int = 0, ctx: _abi.Context | None = None):
        self.ctx = ctx or _abi.default_context()
        h = _abi.VP()
        self.ctx.check(self.ctx.lib.pb_nn_create(self.ctx.h, C.c_ulonglong(seed), C.byref(h)), "pb_nn_create")
        self.h = h
        self.nparam = int(self.ctx.lib.pb_nn_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_get_params(self.h, out.ctypes.data, None), "pb_nn_get_params")
        return out

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_get_params(self.h, None, out.ctypes.data), "pb_nn_get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_set_params(self.h, p.ctypes.data), "pb_nn_set_params")

    def step(self, inputs, target=None, lr=0.0, out_z=None, img_weight=None):
        """inputs: device float32 (nimg, h, w, 3); target: device float32 (nimg,
        h, w) free of constant/checkerboard components, or None; img_weight:
        device float32 (nimg,) loss weights or None.  Returns the loss."""
        nimg, h, w, c = inputs.shape
        assert c == 3 and inputs.is_contiguous()
        if img_weight is not None:
            assert img_weight.shape == (nimg,) and img_weight.is_contiguous()
        loss = C.c_double(0.0)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_train_step(
            self.h, nimg, h, w, _abi.ptr(inputs), _abi.ptr(target), _abi.ptr(img_weight), float(lr),
            _abi.ptr(out_z), C.byref(loss)), "pb_nn_train_step")
        return loss.value

    def forward(self, inputs):
        import torch
        nimg, h, w, _ = inputs.shape
        z = torch.empty(nimg, h, w, dtype=torch.float32, device="cuda")
        self.step(inputs, None, 0.0, z)
        return z

    def apply(self, op, r):
        """z = M(r) on operator ``op``'s grid (FP64 device vectors)."""
        import torch
        z = torch.empty_like(r)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_apply(self.h, op.h, _abi.ptr(r), _abi.ptr(z)), "pb_nn_apply")
        return z

    def features(self, op):
        import torch
        f1 = torch.empty(op.ny, op.nx, dtype=torch.float32, device="cuda")
        f2 = torch.empty_like(f1)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.lib.pb_nn_features(self.h, op.h, _abi.ptr(f1), _abi.ptr(f2)), "pb_nn_features")
        return f1, f2

    def save(self, path):
        np.savez(path, params=self.get_params())

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            self.set_params(f["params"])

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.pb_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
