"""Ensemble stepping (pybella_amd.ensemble): rising-bubble members on one
grid, stepped in lockstep with their pressure solves batched (pb_batch_solve),
against each member stepped alone by the single-member GPU time step with
the same solver mode (FP64 fast): every field, p2_nodes and dt equal bit for
bit after three steps; and the members' states agree with the time-step
oracle (oracle/timestep.py do()) to the solve tolerance."""
import copy

import numpy as np
import pytest

from oracle import timestep as ots

from _cases import rel, strip_node_null

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "rhou", "rhov", "rhow", "rhoY", "rhoX")


def test_ensemble_equals_single_member_steps():
    from pybella_amd import ensemble as en
    from pybella_amd import second_projection as sp
    from pybella_amd import time_update as tu
    print("case ensemble 4 members 81x41, 3 steps", flush=True)
    members, elem, node, ud, th = en.rising_bubble_members(4, 81, 41)
    host = copy.deepcopy(members)
    E = en.Ensemble(copy.deepcopy(members), elem, node, ud, th, mode="f64fast")
    ts, recs = E.run(0.0, 1e9, [0, 0], 3)
    got = copy.deepcopy(members)
    E.download(got)
    base = dict(sp.solver_options)
    sp.solver_options["mode"] = "f64fast"
    try:
        for m, (Sol, flux, mpv) in enumerate(copy.deepcopy(host)):
            ud_m = copy.deepcopy(ud)
            M = tu.DeviceModel(elem, node, ud_m, th, op=sp.Operator2D(sp.grid_of(node, ud_m)))
            M.upload(Sol, flux, mpv)
            t = 0.0
            for k in range(3):
                dt, _, _, infos = tu.step(M, t, 1e9, k, k)
                assert dt == recs[k][m][0], (m, k)
                assert [i["iters"] for i in infos] == [i["iters"] for i in recs[k][m][1]]
                t += dt
            M.download(Sol, flux, mpv)
            for f in FIELDS:
                assert np.array_equal(getattr(Sol, f), getattr(got[m][0], f)), (m, f)
            assert np.array_equal(mpv.p2_nodes, got[m][2].p2_nodes), m
    finally:
        sp.solver_options.clear()
        sp.solver_options.update(base)


def test_ensemble_members_match_oracle():
    from pybella_amd import ensemble as en
    print("case ensemble 3 members 81x41 vs oracle, 2 steps", flush=True)
    members, elem, node, ud, th = en.rising_bubble_members(3, 81, 41, base_seed=7)
    ud.tol = 1e-12
    host = copy.deepcopy(members)
    E = en.Ensemble(copy.deepcopy(members), elem, node, ud, th, mode="f64fast")
    E.run(0.0, 1e9, [0, 0], 2)
    got = copy.deepcopy(members)
    E.download(got)
    for m, (Sol, flux, mpv) in enumerate(host):
        uo = copy.deepcopy(ud)
        uo.stepmax = 2
        So, fo, mo, _ = ots.do(Sol, flux, mpv, 0.0, 1e9, uo, elem, node, [0, 0], th)
        errs = {f: rel(getattr(got[m][0], f), getattr(So, f)) for f in ("rho", "rhou", "rhov", "rhoY", "rhoX")}
        errs["p2_nodes"] = strip_node_null(got[m][2].p2_nodes, mo.p2_nodes)
        print(m, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert all(v <= 1e-8 for v in errs.values()), (m, errs)
