"""GPU time step (pybella_amd.time_update, C ABI pb_advective_fluxes,
pb_advect_split, pb_advect_rk_update, pb_expl_part, pb_euler_forward,
pb_cfl_maxima) against the time-step oracle (oracle/timestep.py).

* Each explicit function on seeded states with gravity (rising bubble, x
  periodic or walls) and without (doubly periodic with wind): every cell
  field, every face flux, rhs and p2_nodes agree to 1e-13 relative (the
  same operation order; pow and division rounding only).
* dt of dynamic_timestep equal to 1e-15.
* The rising-bubble time step at 161 x 81 (one step, dt = 21.69/t_ref, both
  projections, advection) against the oracle's do(): with the solves run to
  a relative residual of 1e-12, every field to 1e-9 and p2_nodes to 1e-9
  modulo the operator's null space (measured 2.3e-10); five steps at the
  reference tolerance 1e-8 to 1e-7 (measured 1.4e-9).
* The reference's own regression targets: the travelling vortex's 101 steps
  (src/tests/test_travelling_vortex.py, initial projection included) and the
  internal long wave's 31 steps (src/tests/test_internal_long_wave.py:
  gravity, stratification, Coriolis, compressible) on the GPU reproduce the
  float32 sums of src/tests/test_targets.yml under the numpy.isclose check of
  src/tests/diagnostics.py:42-58; the long wave's GPU state is also within
  1e-7 of the oracle's, field by field (measured <= 3.0e-8).
"""
import copy
import json
import os

import numpy as np
import pytest

from oracle import projection as opj
from oracle import timestep as ots
from pybella_amd import workloads as W

from _cases import rel, strip_node_null, timestep_state

pytestmark = pytest.mark.gpu

FIELDS = ots.FIELDS
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_targets.json")


def _tu():
    from pybella_amd import time_update as tu
    return tu


def compare(M, Sol, flux, mpv, tol, what, nodes=True):
    """Download M and compare it with the oracle objects (max-norm relative)."""
    S2, f2, m2 = copy.deepcopy(Sol), copy.deepcopy(flux), copy.deepcopy(mpv)
    M.download(S2, f2, m2, halves=False)
    errs = {}
    for k in FIELDS:
        errs[k] = rel(getattr(S2, k), getattr(Sol, k))
        for d in (0, 1):
            ref = getattr(flux[d], k)
            if np.max(np.abs(ref)) > 0:
                errs[f"f{d}.{k}"] = rel(getattr(f2[d], k), ref)
    if nodes:
        errs["p2_nodes"] = rel(m2.p2_nodes, mpv.p2_nodes)
        if np.max(np.abs(mpv.rhs)) > 0:
            errs["rhs"] = rel(m2.rhs, mpv.rhs)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, (what, bad)


@pytest.mark.parametrize("kind", ["bubble", "xwall", "periodic"])
def test_explicit_functions_match_oracle(kind):
    tu = _tu()
    print(f"case explicit {kind}", flush=True)
    Sol, flux, mpv, elem, node, ud, th = timestep_state(kind, 33, 17 if kind != "periodic" else 33)
    ud.nonhydrostasy = 1.0
    M = tu.DeviceModel(elem, node, ud, th)
    M.upload(Sol, flux, mpv)
    tol = 1e-13
    ots.recompute_advective_fluxes(flux, Sol)
    tu.recompute_advective_fluxes(M)
    compare(M, Sol, flux, mpv, tol, "advective fluxes")
    for split in (0, 1):
        for tag in ("rk", None):
            lam = 0.03 / elem.dxyz[split]
            ots.explicit_step_and_flux(Sol, flux[split], lam, elem, split, ud, th, mpv, tag=tag)
            tu.explicit_step_and_flux(M, split, lam, tag=tag)
            compare(M, Sol, flux, mpv, tol, f"split {split} {tag}")
    ots.recompute_advective_fluxes(flux, Sol)
    tu.recompute_advective_fluxes(M)
    ots.advect_rk(Sol, flux, 0.02, elem, 0, ud, th, mpv)
    tu.advect_rk(M, 0.02, 0)
    compare(M, Sol, flux, mpv, tol, "advect_rk")
    for odd in (1, 0):
        ots.recompute_advective_fluxes(flux, Sol)
        tu.recompute_advective_fluxes(M)
        ots.advect(Sol, flux, 0.03, elem, odd, ud, th, mpv)
        tu.advect(M, 0.03, odd)
        compare(M, Sol, flux, mpv, tol, f"advect odd={odd}")
    ots.euler_backward_non_advective_expl_part(Sol, mpv, elem, 0.02, ud, th)
    tu.euler_backward_non_advective_expl_part(M, 0.02)
    compare(M, Sol, flux, mpv, tol, "expl_part")
    ots.euler_forward_non_advective(Sol, mpv, elem, node, 0.02, ud, th)
    tu.euler_forward_non_advective(M, 0.02)
    compare(M, Sol, flux, mpv, tol, "euler_forward")


@pytest.mark.parametrize("kind", ["bubble", "periodic"])
def test_dynamic_timestep_matches_oracle(kind):
    tu = _tu()
    Sol, flux, mpv, elem, node, ud, th = timestep_state(kind, 65, 33 if kind != "periodic" else 65)
    M = tu.DeviceModel(elem, node, ud, th)
    M.upload(Sol, flux, mpv)
    for step in (0, 3):
        want = ots.dynamic_timestep(Sol, 0.0, 10.0, elem, ud, th, step)
        got = tu.dynamic_timestep(M, 0.0, 10.0, step)
        assert abs(got[0] - want[0]) <= 1e-15 * want[0], (got, want)
        assert abs(got[2] - want[2]) <= 1e-14 * want[2], (got, want)


def _bubble_pair(tol):
    ud = W.rising_bubble_ud(161, 81)
    ud.tol = tol
    host = W.rising_bubble_init(ud)
    return host, copy.deepcopy(host)


@pytest.mark.parametrize("tol,nsteps,bound", [(1e-12, 1, 1e-9), (1e-8, 5, 1e-7)])
def test_rising_bubble_steps_match_oracle(tol, nsteps, bound):
    tu = _tu()
    print(f"case rising bubble 161x81, {nsteps} steps, tol {tol}", flush=True)
    (Sol, flux, mpv, elem, node, ud, th), (So, fo, mo, eo, no, uo, to) = _bubble_pair(tol)
    ud.stepmax = uo.stepmax = nsteps
    rec = []
    So, fo, mo, so = ots.do(So, fo, mo, 0.0, 10.0, uo, eo, no, [0, 0], to, record=rec)
    Sol, flux, mpv, st = tu.do(Sol, flux, mpv, 0.0, 10.0, ud, elem, node, [0, 0], th)
    assert st == so == [nsteps, nsteps]
    M = tu.model_for(elem, node, ud, th)
    assert abs(M.infos[0]["dt"] - 21.69 / ud.t_ref) == 0.0
    # later steps' dt follow the CFL maxima of the state, which the solves
    # fix only to their tolerance
    for r, o in zip(M.infos, rec):
        assert abs(r["dt"] - o.dt) <= bound * o.dt, (r["dt"], o.dt)
    for r in M.infos:
        for s in r["solves"]:
            assert s["info"] == 0, s
    errs = {k: rel(getattr(Sol, k), getattr(So, k)) for k in ("rho", "rhou", "rhov", "rhoY", "rhoX")}
    errs["p2_nodes"] = strip_node_null(mpv.p2_nodes, mo.p2_nodes)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(v <= bound for v in errs.values()), errs
    assert np.array_equal(Sol.rhow, So.rhow)


def test_travelling_vortex_reproduces_reference_targets():
    from pybella_amd import second_projection as sp
    tu = _tu()
    print("case travelling vortex 64x64, 101 steps", flush=True)
    Sol, flux, mpv, elem, node, ud, th = W.travelling_vortex_init(sp.euler_backward_non_advective_impl_part)
    Sol, flux, mpv, st = tu.do(Sol, flux, mpv, 0.0, ud.tout[0], ud, elem, node, [0, 0], th)
    assert st == [101, 101]
    with open(GOLD) as f:
        targets = json.load(f)["targets"]["test_travelling_vortex"]
    bad = {}
    for key, ref in targets.items():
        arr = mpv.p2_nodes if key == "p2_nodes" else getattr(Sol, key)
        test = arr.astype("float32").sum()
        if not np.isclose(ref, test):
            bad[key] = (ref, float(test))
    assert bad == {}


def test_internal_long_wave_reproduces_reference_targets():
    from test_oracle_timestep import internal_long_wave_state
    tu = _tu()
    print("case internal long wave 301x10, 31 steps", flush=True)
    host = internal_long_wave_state()
    Sol, flux, mpv, elem, node, ud, th = copy.deepcopy(host)
    Sol, flux, mpv, st = tu.do(Sol, flux, mpv, 0.0, ud.tout[0], ud, elem, node, [0, 0], th)
    assert st == [31, 31]
    with open(GOLD) as f:
        targets = json.load(f)["targets"]["test_internal_long_wave"]
    bad = {}
    for key, ref in targets.items():
        arr = mpv.p2_nodes if key == "p2_nodes" else getattr(Sol, key)
        test = arr.astype("float32").sum()
        if not np.isclose(ref, test):
            bad[key] = (ref, float(test))
    assert bad == {}
    So, fo, mo, eo, no, uo, to = copy.deepcopy(host)
    So, fo, mo, _ = ots.do(So, fo, mo, 0.0, uo.tout[0], uo, eo, no, [0, 0], to)
    errs = {k: rel(getattr(Sol, k), getattr(So, k)) for k in ("rho", "rhov", "rhow", "rhoY", "rhoX")}
    errs["p2_nodes"] = rel(mpv.p2_nodes, mo.p2_nodes)
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    # rhov is small here (sum 4.5e-4) and its relative error carries the
    # solves' breakdown-level noise: measured 3.0e-8, the others <= 1.7e-11
    assert all(v <= 1e-7 for v in errs.values()), errs
