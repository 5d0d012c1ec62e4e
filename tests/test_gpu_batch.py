"""GPU tests of the batched uneven solves (pb_batch_*).

* FP64 (two-launch) batch: every member's solution and iteration count are
  bitwise those of the same member solved alone with the same tile height;
* mixed batch: every member reaches true relative residual 1e-8, and agrees
  with its single mixed solve to the tolerance;
* the batched projection step equals the single drop-in per member (<= 1e-7
  on the corrected momenta, as the drop-in's own parity bound);
* per member, the oracle's FP64 operator confirms the residual.
"""
import copy

import numpy as np
import pytest

from oracle import laplacian as olp
from pybella_amd import workloads as W

from _cases import oracle_op, rel

pytestmark = pytest.mark.gpu

SIZES = [(33, 17), (65, 33), (129, 65), (257, 129)]


def _members(per_size=2):
    out = []
    for nx, ny in SIZES:
        for m in range(per_size):
            out.append(W.batch_member_bubble(nx, ny, m))
    return out


def _systems(members):
    """(op, b) per member: operator assembled on the member's Sol0 and the
    scaled right-hand side of its projection, each on its own handle."""
    from pybella_amd import second_projection as sp
    ops, bs = [], []
    for Sol, Sol0, mpv, elem, node, ud, th in members:
        op = sp.Operator2D(sp.grid_of(node, ud))
        pr = sp._prepare(copy.deepcopy(Sol), copy.deepcopy(mpv), elem, node, ud, th, 0.05, copy.deepcopy(Sol0), op)
        ops.append(op)
        bs.append(pr.b)
    return ops, bs


def test_fp64_batch_equals_single_solves():
    import torch
    from pybella_amd.batch import BatchSolver
    print("case batch f64fast", flush=True)
    members = _members()
    ops, bs = _systems(members)
    for k, op in enumerate(ops):
        op.set_tile((2, 4, 8)[k % 3])          # uneven tiles across members
    singles = [op.solve(b, tol=1e-10, mode="f64fast") for op, b in zip(ops, bs)]
    singles = [(x.clone(), info) for x, info in singles]
    xs, infos = BatchSolver(ops).solve(bs, tol=1e-10, mode="f64fast")
    torch.cuda.synchronize()
    for (xs1, i1), xb, ib in zip(singles, xs, infos):
        assert ib["info"] == 0 and ib["iters"] == i1["iters"], (ib, i1)
        assert torch.equal(xs1, xb)


def test_mixed_batch_converges_and_agrees():
    import torch
    from pybella_amd.batch import BatchSolver
    print("case batch mixed", flush=True)
    members = _members()
    ops, bs = _systems(members)
    xs, infos = BatchSolver(ops).solve(bs, tol=1e-8, mode="mixed")
    for op, b, x, info in zip(ops, bs, xs, infos):
        assert info["info"] == 0 and info["true_rel_resid"] <= 1e-8, info
        xs1, i1 = op.solve(b, tol=1e-8, mode="mixed")
        # both are within 1e-8 of the exact solution in residual; compare modulo the null space
        d = (x - xs1).cpu().numpy()
        from oracle import nn as onn
        d = onn.project_null(d[None])[0]
        ref = onn.project_null(xs1.cpu().numpy()[None])[0]
        assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(ref)
    torch.cuda.synchronize()


def test_batched_projection_matches_single_dropin():
    from pybella_amd import second_projection as sp
    print("case batch impl_part", flush=True)
    members = _members(per_size=1)
    for m in members:
        m[5].tol = 1e-12      # both paths converge far below the comparison bound
    single = [copy.deepcopy(m) for m in members]
    batched = [copy.deepcopy(m) for m in members]
    for Sol, Sol0, mpv, elem, node, ud, th in single:
        sp.euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=Sol0)
    infos = sp.euler_backward_non_advective_impl_part_batch(
        [(Sol, mpv, elem, node, ud, th, Sol0) for Sol, Sol0, mpv, elem, node, ud, th in batched], 0.05,
        mode="f64fast")
    assert all(i["info"] == 0 for i in infos)
    for a, b in zip(single, batched):
        for k in ("rhou", "rhov", "rhoX"):
            assert rel(getattr(b[0], k), getattr(a[0], k)) < 1e-7, k


def test_oracle_residual_per_member():
    print("case batch oracle residual", flush=True)
    from pybella_amd.batch import BatchSolver
    members = [W.batch_member_bubble(33, 17, 0), W.batch_member_bubble(65, 33, 1)]
    ops, bs = _systems(members)
    xs, infos = BatchSolver(ops).solve(bs, tol=1e-9, mode="mixed")
    for (Sol, Sol0, mpv, elem, node, ud, th), b, x, info in zip(members, bs, xs, infos):
        # the oracle operator on the same state the GPU assembled from (Sol0 after the ghost fill)
        from oracle import boundary as obd
        S0 = copy.deepcopy(Sol0)
        m2 = copy.deepcopy(mpv)
        obd.set_explicit_boundary_data(S0, elem, ud, th, m2)
        op_o, _, _ = oracle_op(S0, m2, elem, node, ud, th, 0.05)
        bb = b.cpu().numpy().ravel()
        r = bb - olp.lap2d(x.cpu().numpy().ravel(), op_o)
        assert np.linalg.norm(r) <= 2e-9 * np.linalg.norm(bb), (np.linalg.norm(r) / np.linalg.norm(bb), info)


@pytest.mark.parametrize("ry", [2, 5])
def test_flattened_batch_equals_single_solves(ry):
    """Uniform tile height: the batch runs one launch over every member's tiles
    per phase; each member's iterates are still bitwise those of its own solve."""
    import torch
    from pybella_amd.batch import BatchSolver
    print(f"case flattened batch ry={ry}", flush=True)
    members = _members()
    ops, bs = _systems(members)
    for op in ops:
        op.set_tile(ry)
    singles = [op.solve(b, tol=1e-10, mode="f64fast") for op, b in zip(ops, bs)]
    singles = [(x.clone(), info) for x, info in singles]
    batch = BatchSolver(ops)
    for _ in range(2):   # a second solve replays the captured graph
        xs, infos = batch.solve(bs, tol=1e-10, mode="f64fast")
        torch.cuda.synchronize()
        for (xs1, i1), xb, ib in zip(singles, xs, infos):
            assert ib["info"] == 0 and ib["iters"] == i1["iters"], (ib, i1)
            assert torch.equal(xs1, xb)
    xs, infos = batch.solve(bs, tol=1e-8, mode="mixed")
    for op, b, info in zip(ops, bs, infos):
        assert info["info"] == 0 and info["true_rel_resid"] <= 1e-8, info
