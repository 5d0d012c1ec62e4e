"""ctypes binding of the C ABI in include/pybella_amd.h.

The shared library is built in-tree by ``pybella_amd.build``.  There is no
fallback: if the library or a GPU is missing, every entry raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libpybella_amd.so"

PB_OK = 0
PB_BC_PERIODIC = 0
PB_BC_WALL = 1
PB_F64 = 0
PB_F32 = 1
PB_SOLVE_F64 = 0
PB_SOLVE_MIXED = 1
PB_SOLVE_NN = 2
PB_SOLVE_F64_FAST = 3
PB_SOLVE_F64_FUSED = 4


class Grid2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("bc_x", C.c_int), ("bc_y", C.c_int)]


class Grid3D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("bc_x", C.c_int), ("bc_y", C.c_int), ("bc_z", C.c_int)]


class Phys(C.Structure):
    _fields_ = [("gamm", C.c_double), ("Msq", C.c_double), ("g", C.c_double),
                ("coriolis", C.c_double * 3), ("nonhydro", C.c_double),
                ("is_compressible", C.c_int), ("compressibility", C.c_double),
                ("is_arakawa_konor", C.c_int), ("wind", C.c_double * 3), ("stratified", C.c_int),
                ("strat_inv", C.c_double * 4)]


class LetkfCfg(C.Structure):
    _fields_ = [("members", C.c_int), ("attrs", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("obs_x", C.c_int),
                ("obs_y", C.c_int), ("periodic_x", C.c_int), ("periodic_y", C.c_int), ("inflation", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("tol", C.c_double), ("maxiter", C.c_int),
                ("inner_tol", C.c_double), ("max_outer", C.c_int), ("check_every", C.c_int),
                ("x0_given", C.c_int), ("fallback_maxiter", C.c_int), ("krylov_m", C.c_int),
                ("sweeps", C.c_int), ("omega", C.c_double), ("use_network", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("info", C.c_int), ("iters", C.c_int), ("outer", C.c_int), ("half_step", C.c_int),
                ("resid", C.c_double), ("bnrm2", C.c_double), ("true_rel_resid", C.c_double),
                ("fallback", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


VP = C.c_void_p
I = C.c_int
D = C.c_double

# name -> (restype, argtypes); the list every test checks against the header
SIGNATURES = {
    "pb_version": (I, []),
    "pb_batch_create": (I, [VP, C.POINTER(VP), I, C.POINTER(VP)]),
    "pb_batch_destroy": (I, [VP]),
    "pb_batch_solve": (I, [VP, C.POINTER(VP), C.POINTER(VP), C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "pb_op3d_create": (I, [VP, C.POINTER(Grid3D), C.POINTER(VP)]),
    "pb_op3d_destroy": (I, [VP]),
    "pb_op3d_assemble": (I, [VP, C.POINTER(Phys), D, VP, VP, VP, VP]),
    "pb_op3d_matvec": (I, [VP, VP, VP, I]),
    "pb_op3d_solve": (I, [VP, VP, VP, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "pb_op3d_time_kernel": (I, [VP, I, I, I, C.POINTER(D)]),
    "pb_op3d_march_map": (I, [I, I, I, C.POINTER(I)]),
    "pb_ctx_create": (I, [I, C.POINTER(VP)]),
    "pb_ctx_destroy": (I, [VP]),
    "pb_ctx_set_stream": (I, [VP, VP]),
    "pb_ctx_synchronize": (I, [VP]),
    "pb_last_error": (C.c_char_p, [VP]),
    "pb_bc_cells": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), VP * 6, VP]),
    "pb_ghostnodes": (I, [VP, C.POINTER(Grid2D), VP]),
    "pb_ghostcells_p2": (I, [VP, C.POINTER(Grid2D), VP]),
    "pb_bc_cells_axis": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), VP * 6, VP, I]),
    "pb_advective_fluxes": (I, [VP, C.POINTER(Grid2D), VP, VP, VP, VP, VP, VP]),
    "pb_advect_split": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), I, D, I, VP * 6, VP * 6, VP]),
    "pb_advect_rk_update": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), D, D, VP * 6, VP * 6, VP * 6, VP]),
    "pb_expl_part": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), D, VP * 6, VP, VP]),
    "pb_euler_forward": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), D, VP * 6, VP, VP, VP, VP, VP]),
    "pb_cfl_work_words": (C.c_long, [C.POINTER(Grid2D)]),
    "pb_letkf_analyse": (I, [VP, C.POINTER(LetkfCfg), VP, VP, VP, VP, VP, VP]),
    "pb_cfl_maxima": (I, [VP, C.POINTER(Grid2D), D, D, VP, VP, VP, VP, VP, VP, C.POINTER(D)]),
    "pb_op2d_create": (I, [VP, C.POINTER(Grid2D), C.POINTER(VP)]),
    "pb_op2d_destroy": (I, [VP]),
    "pb_op2d_assemble": (I, [VP, C.POINTER(Phys), D, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "pb_op2d_matvec": (I, [VP, VP, VP, I]),
    "pb_op2d_scale_rhs": (I, [VP, VP, VP]),
    "pb_bicgstab_solve": (I, [VP, VP, VP, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "pb_bicgstab_solve_noalloc": (I, [VP, VP, VP, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "pb_op2d_time_kernel": (I, [VP, I, I, I, C.POINTER(D)]),
    "pb_op2d_time_in_solve": (I, [VP, VP, VP, D, I, C.POINTER(D), C.POINTER(I)]),
    "pb_op2d_set_tile": (I, [VP, I]),
    "pb_divergence_nodes": (I, [VP, C.POINTER(Grid2D), D, VP, VP, VP, VP, VP, VP]),
    "pb_correction_nodes": (I, [VP, C.POINTER(Grid2D), C.POINTER(Phys), D, D, VP, VP, VP, VP,
                                VP, VP, VP, VP, VP, VP]),
    "pb_scatter_update_p2": (I, [VP, C.POINTER(Grid2D), VP, VP, VP]),
    "pb_nn_create": (I, [VP, C.c_ulonglong, C.POINTER(VP)]),
    "pb_nn_destroy": (I, [VP]),
    "pb_nn_num_params": (C.c_long, [VP]),
    "pb_nn_get_params": (I, [VP, VP, VP]),
    "pb_nn_set_params": (I, [VP, VP]),
    "pb_nn_train_step": (I, [VP, I, I, I, VP, VP, VP, D, VP, C.POINTER(D)]),
    "pb_nn_apply": (I, [VP, VP, VP, VP]),
    "pb_nn_features": (I, [VP, VP, VP, VP]),
    "pb_op2d_attach_nn": (I, [VP, VP]),
    "pb_nn_set_loss_nullspace": (I, [VP, I]),
    "pb_op2d_precond": (I, [VP, VP, VP, I, I, D, I]),
    "pb_op2d_mg_info": (I, [VP, C.POINTER(I), I, C.POINTER(I)]),
    "pb_op2d_mg_stencil": (I, [VP, I, VP]),
    "pb_op2d_mg_features": (I, [VP, VP, VP]),
}

_lib = None


def load():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m pybella_amd.build` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HipError(RuntimeError):
    pass


class Context:
    """One pb_ctx on a device; work is issued on torch's current stream."""

    def __init__(self, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise HipError("pybella_amd needs a ROCm GPU; torch.cuda.is_available() is False")
        self.lib = load()
        self.device = device
        h = VP()
        rc = self.lib.pb_ctx_create(device, C.byref(h))
        if rc != 0:
            raise HipError(f"pb_ctx_create failed ({rc})")
        self.h = h
        self.sync_stream()

    def sync_stream(self):
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.lib.pb_ctx_set_stream(self.h, VP(s))

    def check(self, rc, what=""):
        if rc < 0:
            msg = self.lib.pb_last_error(self.h).decode(errors="replace")
            raise HipError(f"{what}: {msg} (code {rc})")
        return rc

    def synchronize(self):
        self.check(self.lib.pb_ctx_synchronize(self.h), "pb_ctx_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    else:
        _default_ctx.sync_stream()
    return _default_ctx


def ptr(t) -> VP:
    """Device pointer of a torch tensor (None -> NULL)."""
    return VP(0) if t is None else VP(t.data_ptr())
