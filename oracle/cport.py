"""ctypes access to the C restatement (oracle/c/pb_oracle.c) — oracle / CPU
baseline only.  Built by ``make -C oracle``."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class _Op(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("x_wall", C.c_int), ("y_wall", C.c_int)] + \
               [(k, C.c_void_p) for k in ("hx", "hy", "hc", "cxx", "cyy", "cxy", "cyx", "dinv")]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_lap2d.argtypes = [C.POINTER(_Op), C.c_void_p, C.c_void_p]
        _lib.orc_bicgstab.argtypes = [C.POINTER(_Op), C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        _lib.orc_bicgstab.restype = C.c_int
    return _lib


class COp:
    """Holds contiguous copies of an oracle op dict (laplacian.op2d_arrays)."""

    def __init__(self, op):
        self.arrs = {k: np.ascontiguousarray(op[k], dtype=np.float64)
                     for k in ("hx", "hy", "hc", "cxx", "cyy", "cxy", "cyx", "dinv")}
        self.s = _Op(op["nx"], op["ny"], op["dx"], op["dy"], int(op["x_wall"]), int(op["y_wall"]),
                     *[self.arrs[k].ctypes.data for k in ("hx", "hy", "hc", "cxx", "cyy", "cxy", "cyx", "dinv")])
        self.n = op["nx"] * op["ny"]

    def matvec(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty(self.n)
        lib().orc_lap2d(C.byref(self.s), p.ctypes.data, out.ctypes.data)
        return out

    def bicgstab(self, b, tol=1e-8, maxiter=6000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.n)
        it, half, res = C.c_int(), C.c_int(), C.c_double()
        info = lib().orc_bicgstab(C.byref(self.s), b.ctypes.data, x.ctypes.data, tol, maxiter,
                                  C.byref(it), C.byref(half), C.byref(res))
        return x, info, it.value, bool(half.value), res.value
