"""Writes tests/golden/impl_part_33x17.npz: inputs and outputs of the second
projection (euler_backward_non_advective_impl_part, second_projection.py:97-277)
computed by the FP64 oracle on two seeded 33x17 states.

Run from the repository root:  python tests/golden/make_golden.py

Cases
  bubble  rising-bubble state (psinc, Sol0 given), tol 1e-12
  comp    compressible stratified state with seeded smooth momenta, tol 1e-12
Stored per case (prefix "<case>/"): the input fields (rho, rhou, rhov, rhow,
rhoY, rhoX of Sol and, for bubble, Sol0; p2_nodes), the assembled operator
(wplus0, wcenter, diag_inv), the scaled right-hand side b, and the outputs
(rhou, rhov, rhoX after the projection, p2_nodes after the update, the
oracle's iteration count).
"""
import copy
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import projection as opj  # noqa: E402
from pybella_amd import workloads as W  # noqa: E402

FIELDS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "impl_part_33x17.npz")


def cases():
    Sol, Sol0, mpv, elem, node, ud, th = W.rising_bubble_state(33, 17)
    yield "bubble", Sol, Sol0, mpv, elem, node, ud, th
    Sol, mpv, elem, node, ud, th = W.random_momentum_state(33, 17, seed=9, compressible=True)
    yield "comp", Sol, None, mpv, elem, node, ud, th


def main():
    out = {}
    for name, Sol, Sol0, mpv, elem, node, ud, th in cases():
        ud.tol = 1e-12
        for f in FIELDS:
            out[f"{name}/in_{f}"] = getattr(Sol, f).copy()
            if Sol0 is not None:
                out[f"{name}/in0_{f}"] = getattr(Sol0, f).copy()
        out[f"{name}/in_p2_nodes"] = mpv.p2_nodes.copy()
        S, m = copy.deepcopy(Sol), copy.deepcopy(mpv)
        S0 = copy.deepcopy(Sol0) if Sol0 is not None else None
        op, b, dinv = opj.assemble_system(copy.deepcopy(S), copy.deepcopy(m), elem, node, ud, th, 0.05,
                                          copy.deepcopy(S0))
        res = opj.euler_backward_non_advective_impl_part(S, m, elem, node, ud, th, 0.0, 0.05, 1.0, Sol0=S0)
        assert res.info == 0
        out[f"{name}/b"] = b
        out[f"{name}/diag_inv"] = dinv
        out[f"{name}/out_rhou"] = S.rhou
        out[f"{name}/out_rhov"] = S.rhov
        out[f"{name}/out_rhoX"] = S.rhoX
        out[f"{name}/out_p2_nodes"] = m.p2_nodes
        out[f"{name}/iters"] = np.array(res.niter)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
