"""Golden fixtures (tests/golden/impl_part_33x17.npz, written by
tests/golden/make_golden.py from the FP64 oracle): the oracle still reproduces
them (CPU), and the GPU drop-in reproduces their outputs (gpu)."""
import copy
import os

import numpy as np
import pytest

from pybella_amd import workloads as W

from _cases import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "impl_part_33x17.npz")
FIELDS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")


def load_case(name):
    g = np.load(GOLD)
    if name == "bubble":
        Sol, Sol0, mpv, elem, node, ud, th = W.rising_bubble_state(33, 17)
    else:
        Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=9, compressible=True)
        Sol0 = None
    for f in FIELDS:
        getattr(Sol, f)[...] = g[f"{name}/in_{f}"]
        if Sol0 is not None:
            getattr(Sol0, f)[...] = g[f"{name}/in0_{f}"]
    mpv.p2_nodes[...] = g[f"{name}/in_p2_nodes"]
    ud.tol = 1e-12
    return g, (Sol, Sol0, mpv, elem, node, ud, th)


@pytest.mark.parametrize("name", ["bubble", "comp"])
def test_oracle_reproduces_golden(name):
    from oracle import projection as opj
    g, (Sol, Sol0, mpv, elem, node, ud, th) = load_case(name)
    res = opj.euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=Sol0)
    assert res.info == 0 and res.niter == int(g[f"{name}/iters"])
    for k in ("rhou", "rhov", "rhoX"):
        assert rel(getattr(Sol, k), g[f"{name}/out_{k}"]) < 1e-12, k
    assert rel(mpv.p2_nodes, g[f"{name}/out_p2_nodes"]) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bubble", "comp"])
def test_gpu_dropin_reproduces_golden(name):
    from pybella_amd import second_projection as sp
    g, (Sol, Sol0, mpv, elem, node, ud, th) = load_case(name)
    print(f"golden {name}", flush=True)
    info = sp.euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=Sol0)
    assert info["info"] == 0
    assert info["true_rel_resid"] < 1e-11
    for k in ("rhou", "rhov", "rhoX"):
        assert rel(getattr(Sol, k), g[f"{name}/out_{k}"]) < 1e-7, k
