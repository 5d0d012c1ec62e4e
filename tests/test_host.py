"""CPU checks of host-side logic: training-data generation and option plumbing."""
import numpy as np
import torch

from pybella_amd import workloads as W


def test_device_rhs_is_consistent():
    from pybella_amd.nn_train import fourier_rhs_batch
    g = torch.Generator()
    g.manual_seed(3)
    F = fourier_rhs_batch(g, 3, 33, 17, nmodes=16, kmax=6, device="cpu").numpy()
    assert F.shape == (3, 19, 35)
    Wn = W.left_null_vectors(33, 17)
    for fb in F:
        f = fb.T   # -> host [x][y]
        assert not f[0].any() and not f[-1].any() and not f[:, 0].any() and not f[:, -1].any()
        inner = f[1:-1, 1:-1]
        assert np.abs(inner).max() > 0.1
        for v in Wn:
            assert abs((v * inner).sum()) < 1e-10 * np.abs(inner).sum()
        # the periodic end columns of each mode coincide (integer kx) before
        # projection, so they coincide after it as well
        assert np.allclose(inner[0], inner[-1], atol=1e-12)
    assert not np.allclose(F[0], F[1])


def test_numpy_rhs_is_consistent():
    Sol, mpv, elem, node, ud, th = W.stratified_state(33, 17)
    f = W.stratified_rhs(node, seed=5, nmodes=16, kmax=6)[1:-1, 1:-1]
    for v in W.left_null_vectors(33, 17):
        assert abs((v * f).sum()) < 1e-10 * np.abs(f).sum()


def test_solve_options_layout():
    from pybella_amd import _abi
    o = _abi.SolveOpts(_abi.PB_SOLVE_NN, 1e-8, 300, 1e-4, 12, 32, 0, 6000, 40, 1, 0.6, 0)
    assert o.use_network == 0 and o.krylov_m == 40 and abs(o.omega - 0.6) < 1e-15
