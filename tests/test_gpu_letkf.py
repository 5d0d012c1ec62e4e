"""GPU LETKF local analysis (pb_letkf_analyse, pybella_amd.letkf.LETKF)
against the LETKF oracle (oracle/letkf.py analyse_grid, itself pinned to the
observation-space Kalman update): cell attributes (rho, rhou, rhov, rhoY)
and node attributes (p2_nodes) of k seeded members, sparse noisy
observations of a further member (10 % of the locations, the reference's
obs_frac), the reference's gaussian localisation over 11 x 11 windows,
periodic x and walls in y; k = 6, 12 and 24 (the 8-, 16- and 32-member
kernels).  Every analysed inner value agrees to 1e-10 relative; the ghost
layers are zero as the reference leaves them."""
import numpy as np
import pytest

from oracle import letkf as ol

from _cases import timestep_state

pytestmark = pytest.mark.gpu

CELL = ("rho", "rhou", "rhov", "rhoY")


@pytest.mark.parametrize("k,inflation", [(6, 1.0), (12, 1.05), (24, 1.0)])
def test_letkf_matches_oracle(k, inflation):
    from pybella_amd import time_update as tu
    from pybella_amd import letkf as lk
    from pybella_amd import second_projection as sp
    print(f"case letkf k={k} inflation={inflation} 81x41", flush=True)
    nx, ny = 81, 41
    members = [timestep_state("bubble", nx, ny, seed=100 + m) for m in range(k)]
    truth = timestep_state("bubble", nx, ny, seed=999)
    Sol0, flux0, mpv0, elem, node, ud, th = members[0]
    rng = np.random.default_rng(k)
    obs, mask, covar = {}, {}, {}
    for a in CELL + ("p2_nodes",):
        f = (truth[2].p2_nodes if a == "p2_nodes" else getattr(truth[0], a))[2:-2, 2:-2]
        var = 0.05 * ((f - f.mean()) ** 2).mean() + 1e-12
        obs[a] = f + np.sqrt(var) * rng.standard_normal(f.shape)
        mask[a] = (rng.random(f.shape) >= 0.10).astype(float)
        covar[a] = var
    models = []
    for (Sol, flux, mpv, *_rest) in members:
        M = tu.DeviceModel(elem, node, ud, th, op=sp.Operator2D(sp.grid_of(node, ud)))
        M.upload(Sol, flux, mpv)
        models.append(M)
    L = lk.LETKF(ud, attributes=CELL + ("p2_nodes",), obs_X=11, obs_Y=11, inflation=inflation)
    L.analyse(models, obs, mask, covar)
    loc = ol.loc_matrix(11, 11)
    for attrs, get_h, get_d, shape in (
            (CELL, lambda S, m, a: getattr(S, a), lambda M, a: M.sol.f[a], (nx - 1, ny - 1)),
            (("p2_nodes",), lambda S, m, a: m.p2_nodes, lambda M, a: M.p2n, (nx, ny))):
        state = np.array([[get_h(S, m, a)[2:-2, 2:-2] for a in attrs] for (S, _, m, *_r) in members])
        bc = ol.boundary_mask(shape, (True, False), 5, 5)
        ref = ol.analyse_grid(state, np.array([obs[a] for a in attrs]), np.array([mask[a] for a in attrs]),
                              np.array([covar[a] for a in attrs]), loc, bc, 11, 11, inflation)
        for m, M in enumerate(models):
            for ai, a in enumerate(attrs):
                got = sp.from_dev(get_d(M, a))
                r = ref[m, ai]
                err = np.max(np.abs(got[2:-2, 2:-2] - r)) / max(np.max(np.abs(r)), 1e-300)
                assert err <= 1e-10, (m, a, err)
                g2 = got.copy()
                g2[2:-2, 2:-2] = 0.0
                assert not np.any(g2), (m, a)
        # the analysis moved the members (the observations carry information)
        assert np.max(np.abs(ref - state)) > 1e-6


def test_rloc_dropin_on_reference_containers():
    """RLoc.analyse with the reference's argument structure (members as
    [Sol, flux, mpv, steps], obs/mask per DA time with ghost layers, covar as
    per-time (cells, nodes) variance arrays) equals the oracle."""
    import types
    from pybella_amd import letkf as lk
    print("case rloc drop-in k=8 65x33", flush=True)
    nx, ny, k = 65, 33, 8
    members = [timestep_state("bubble", nx, ny, seed=300 + m) for m in range(k)]
    Sol0, flux0, mpv0, elem, node, ud, th = members[0]
    results = [[S, f, m, [0, 0]] for (S, f, m, *_r) in members]
    rng = np.random.default_rng(11)
    attrs = ("rho", "rhov", "p2_nodes")
    obs_t, mask_t = {}, {}
    for a in attrs:
        full = mpv0.p2_nodes if a == "p2_nodes" else getattr(Sol0, a)
        obs_t[a] = full + 0.01 * rng.standard_normal(full.shape)
        mask_t[a] = (rng.random(full.shape) >= 0.2).astype(float)
    covar = (np.array([[1e-4, 2e-4]]), np.array([[3e-4]]))
    loc = ol.loc_matrix(7, 7)
    dap = types.SimpleNamespace(da_times=[0.5], obs_attributes=list(attrs), obs_X=7, obs_Y=7,
                                localisation_matrix=loc, inflation_factor=1.0,
                                loc={"rho": 0, "rhov": 0, "p2_nodes": 2})
    before = {a: np.array([(m[2].p2_nodes if a == "p2_nodes" else getattr(m[0], a))[2:-2, 2:-2] for m in results])
              for a in attrs}
    out = lk.RLoc(ud, elem, node, dap, k).analyse(results, [obs_t], covar, [mask_t], k, 0.5)
    for grid_attrs, var, shape in ((("rho", "rhov"), covar[0][0], (nx - 1, ny - 1)), (("p2_nodes",), covar[1][0], (nx, ny))):
        state = np.stack([before[a] for a in grid_attrs], axis=1)
        ref = ol.analyse_grid(state, np.array([obs_t[a][2:-2, 2:-2] for a in grid_attrs]),
                              np.array([mask_t[a][2:-2, 2:-2] for a in grid_attrs]), var, loc,
                              ol.boundary_mask(shape, (True, False), 3, 3), 7, 7, 1.0)
        for m, mem in enumerate(out):
            for ai, a in enumerate(grid_attrs):
                got = mem[2].p2_nodes if a == "p2_nodes" else getattr(mem[0], a)
                r = ref[m, ai]
                assert np.max(np.abs(got[2:-2, 2:-2] - r)) <= 1e-10 * np.max(np.abs(r)), (m, a)
                g2 = got.copy()
                g2[2:-2, 2:-2] = 0.0
                assert not np.any(g2)
