"""CPU oracle for the nodal elliptic pressure solve — TEST INFRASTRUCTURE ONLY.

This package is the checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  ``pybella_amd`` never imports, links or executes anything here.

It is an FP64 NumPy restatement of pyBELLA's second-projection solve, written
to follow the reference's operation order line by line (citations are
file:line in the pyBELLA tree):

  boundary.py        set_explicit_boundary_data, set_ghostnodes_p2
                     (src/dycore/utils/boundary.py:7-132, 296-369)
  projection.py      operator_coefficients_nodes, scale_wall_node_values,
                     multiply_inverse_coriolis, divergence_nodes, grad_nodes,
                     correction_nodes, rhs_from_p_old,
                     euler_backward_non_advective_impl_part
                     (src/dycore/physics/low_mach/second_projection.py:97-652)
  laplacian.py       precon_diag_prepare, lap2D_gather_new (vectorised and a
                     scalar per-node loop), the 27-point operator
                     (src/dycore/physics/low_mach/laplacian.py:496-1026, 1284-1356)
  bicgstab.py        BiCGStab of Barrett et al. (Templates, 1994, sec. 2.3.8) with
                     the reference's call semantics (second_projection.py:205-230,
                     SciPy 1.7.3 legacy tolerance)
  nn.py              NumPy forward/backward of the learned-preconditioner network
                     (no reference counterpart; checker for the GPU kernels)

How each function is pinned (the reference ships no solver-level fixtures):

  * lap2d: scalar per-node loop (a literal restatement of the Numba loop) vs the
    vectorised form; A*1 and A*checkerboard reduce to the centre term for any
    coefficients; the textbook rotated 9-point Laplacian for unit coefficients;
    probed CSR == matvec; manufactured solutions recovered by SciPy splu;
    left null vectors (trapezoid weights, trapezoid*checkerboard) verified.
  * coefficients / diag / divergence / correction: closed forms on uniform and
    analytic states, 2x2 box sums vs SciPy fftconvolve (the reference's own
    routine), discrete divergence theorem, grad of constants/checkerboard = 0.
  * bicgstab: SciPy 1.15's bicgstab(rtol=tol, atol=0) run on the same system
    (same algorithm, independent code); legacy early exit and b = 0 cases.
  * boundary: closed forms (wrap, mirror) and the psinc background image.
  * The 27-point operator has no runnable reference (its branch in
    second_projection.py:183-203 uses names before assignment): its restatement
    is pinned only by invariants (A*1 = centre term, symmetry, CSR probe) and is
    "parity unpinned" against the reference itself.
  * SciPy 1.7.3's Fortran reverse-communication iterate sequence (iteration
    count +-1) is not reproducible without the reference's own environment:
    iteration counts are "parity unpinned" at +-1.
"""
