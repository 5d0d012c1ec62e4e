#!/usr/bin/env python3
"""Benchmark: solves/s of the 1025x513-node nodal elliptic solve on one MI355X.

Workload (the metric's configuration): isolated 2D nodal 9-point pressure
solve at 1025 x 513 nodes (x in [-10, 10] km periodic, z in [0, 10] km walls),
coefficients of a hydrostatically stratified atmosphere (theta 300 -> 330 K
linear in z, Exner from hydrostatic balance), pseudo-incompressible
(wcenter = 0), half-step dt = 0.05 (5 s), seeded smooth random RHS (64
Fourier modes, |k| <= 16, seed 20261202) projected onto the operator's range.

One step = FP64 operator assembly on the GPU (coefficients, Coriolis factors,
diagonal, folded stencil weights) + D^-1 scaling of the RHS + the solve to
true relative residual ||D^-1 (b - A x)|| / ||D^-1 b|| <= 1e-8.  Inputs are
resident in HBM before the timed region; every timed step does all of that
work.

Printed: ONE JSON line with the contract fields plus ``roofline`` (dominant
kernel, HBM-bound) and ``cpu_baseline`` (the C restatement of the reference's
serial matvec + BiCGStab, oracle/c/pb_oracle.c, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, public spec

# Algorithmic bytes per node of each solver kernel (zero Coriolis, zero centre
# term; cell arrays counted as one word per node): every distinct array read
# once and written once, in words of the Krylov precision, plus FP64 x bytes.
#   0 matvec:                read p, Wxx, Wyy, dinv; write y                         -> 5 words
#   1 A launch (fast):       read r, p, v, t, rt, Wxx, Wyy, dinv; write r, p, v      -> 11 words
#                            + x read and written in FP64                            -> +16 B
#   2 B launch:              read r, v, rt, Wxx, Wyy, dinv; write t                  -> 7 words
#   3 FP64 true residual:    read x, b, Wxx, Wyy, dinv; write r                      -> 6 words
KERNEL_WORDS = {0: 5, 1: 11, 2: 7, 3: 6}
KERNEL_XBYTES = {1: 16}
KERNEL_NAMES = {0: "matvec", 1: "bicg_fast_A_pupdate_matvec", 2: "bicg_B_s_matvec_dots", 3: "true_residual_f64"}


SOLVER_NAMES = {
    "mixed": "plain BiCGStab, FP32 Krylov (FP64 iterate) in FP64 iterative refinement",
    "f64": "plain BiCGStab FP64 (reference semantics, 3 launches/iteration)",
    "f64fast": "plain BiCGStab FP64 (rho, ||r|| from s/t dots, 2 launches/iteration)",
    "f64fused": "plain BiCGStab FP64 (as f64fast; r~.v as u.p with u = A^T r~, 1 launch/iteration)",
    "nn": "learned preconditioner: FP64 flexible GMRES preconditioned by the FP32 two-branch multigrid cycle "
          "whose coarse level (65x33 nodes) the BF16 network solves; plain BiCGStab FP64 only past the budget",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(nx=1025, ny=513, seed=20261202):
    from pybella_amd import workloads as W
    Sol, mpv, elem, node, ud, th = W.stratified_state(nx, ny)
    rhs = W.stratified_rhs(node, seed=seed)
    return (Sol, mpv, elem, node, ud, th), rhs


def _strip_null(v2d):
    """v minus its least-squares constant and checkerboard components."""
    ny, nx = v2d.shape
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    B = np.stack([np.ones(ny * nx), np.where((ii + jj) % 2 == 0, 1.0, -1.0).ravel()], axis=1)
    c, *_ = np.linalg.lstsq(B, v2d.ravel(), rcond=None)
    return (v2d.ravel() - B @ c).reshape(ny, nx)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_checks(state, rhs, full_iters, sample_iters, gpu_x, dt=0.05):
    """The checker leg (oracle only): assemble the system with the NumPy
    oracle, solve it with the serial C restatement of the reference's solve
    (oracle/c/pb_oracle.c, FP64 BiCGStab with the reference's semantics) to
    relative residual 1e-8, and report every GPU solution's relative L2
    distance to it modulo the null space {1, checkerboard} and its FP64 true
    residual under the oracle's operator.  The same solve's per-iteration time
    is the CPU baseline: one core (the reference's serial loop) on a bounded
    sample of iterations, and all host cores (NumPy's vectorised
    lap2D_gather_new restatement, BLAS threads = the cores available)."""
    import copy
    from oracle import bicgstab as obs
    from oracle import cport
    from oracle import laplacian as olp
    from oracle import projection as opj
    Sol, mpv, elem, node, ud, th = [copy.deepcopy(o) for o in state]
    t0 = time.perf_counter()
    opj.operator_coefficients_nodes(elem, node, Sol, mpv, ud, th, dt)
    mpv.wcenter[...] *= ud.compressibility
    cor = opj.multiply_inverse_coriolis(mpv, Sol, mpv, ud, elem, node, dt, attrs=("u", "v", "w"), get_coeffs=True)
    dinv = olp.precon_diag_prepare(mpv, elem, node, ud)
    op = olp.op2d_arrays(mpv, node, cor, dinv, ud)
    b = (rhs * dinv)[1:-1, 1:-1].T.ravel().copy()
    t_asm = time.perf_counter() - t0
    c = cport.COp(op)
    t0 = time.perf_counter()
    xo, info_o, it_o, _, _ = c.bicgstab(b, tol=1e-8, maxiter=6000)
    t_oracle = time.perf_counter() - t0
    A = lambda v: olp.lap2d(v, op)
    ref = _strip_null(xo.reshape(op["ny"], op["nx"]))
    parity = {"oracle": "oracle/c/pb_oracle.c serial FP64 BiCGStab (reference semantics), tol 1e-8",
              "oracle_iters": it_o, "oracle_info": info_o,
              "oracle_true_rel_resid": obs.true_rel_resid(A, b, xo), "oracle_seconds": round(t_oracle, 2)}
    for name, xg in gpu_x.items():
        xs = _strip_null(xg)
        parity[f"rel_l2_{name}"] = float(np.linalg.norm(xs - ref) / np.linalg.norm(ref))
        parity[f"true_rel_resid_{name}_under_oracle_operator"] = obs.true_rel_resid(A, b, xg.ravel())
    # CPU baseline, one core: the C restatement on a bounded sample
    t0 = time.perf_counter()
    _, _, it, _, _ = c.bicgstab(b, tol=1e-30, maxiter=sample_iters)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    t_solve = t_asm + t_it * full_iters
    cores = len(os.sched_getaffinity(0))
    # all cores: the vectorised NumPy restatement's iteration (matvec + BLAS-1
    # work of one BiCGStab iteration), threads = the cores available
    t_np = None
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=cores):
            t0 = time.perf_counter()
            res = obs.bicgstab(A, b, tol=1e-30, maxiter=20)
            t_np = (time.perf_counter() - t0) / max(res.niter, 1)
    except ImportError:
        pass
    cpu = {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{it} BiCGStab iterations (serial C restatement of lap2D_gather_new + Templates "
                   f"BiCGStab, FP64) of the same 1025x513 solve, {t_it * 1e3:.2f} ms/iteration, scaled to "
                   f"the {full_iters} iterations the FP64 solve needs, plus {t_asm:.2f} s NumPy assembly; "
                   f"CPU: {_cpu_model()}"),
    }
    if t_np is not None:
        cpu["all_cores"] = {"value": 1.0 / (t_asm + t_np * full_iters), "unit": "solves/s", "cores": cores,
                            "kind": "port",
                            "sample": (f"20 iterations of the vectorised NumPy restatement (BLAS threads {cores}), "
                                       f"{t_np * 1e3:.2f} ms/iteration, scaled to {full_iters} iterations")}
    return parity, cpu


def extra_workloads(args):
    """The 3D and batched workloads' summaries for the default line, each
    from its own bench.py process (--workload 3d / batch) run here."""
    import subprocess
    out = {}
    for wl, key, extra in (("timestep", "time_steps", []), ("ensemble", "ensemble_steps", []), ("3d", "workload_3d", []),
                           ("batch", "batch", ["--members-per-size", "16"])):
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", wl, "--steps", "2", *extra]
        log("running", " ".join(cmd[1:]))
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            out[key] = {"error": r.stderr[-500:]}
            continue
        line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
        rec = json.loads(line)
        rec["wall_s"] = round(time.perf_counter() - t0, 1)
        out[key] = rec
    return out


def _node_null_rel(p, q):
    """Max-norm relative difference of node arrays (reference layout) on the
    inner nodes, modulo the constant and the checkerboard."""
    a, b = p[2:-2, 2:-2], q[2:-2, 2:-2]
    i, j = np.meshgrid(np.arange(a.shape[0]), np.arange(a.shape[1]), indexing="ij")
    B = np.stack([np.ones(a.size), np.where((i + j) % 2 == 0, 1.0, -1.0).ravel()], axis=1)
    d = (a - b).ravel()
    c, *_ = np.linalg.lstsq(B, d, rcond=None)
    return float(np.max(np.abs(d - B @ c)) / np.max(np.abs(b)))


def bench_ensemble(args, count=16, nx=161, ny=81, nsteps=6):
    """Member time steps per second of a rising-bubble ensemble on one grid:
    the members stepped in lockstep with their solves batched (one graph per
    chunk of iterations over all members), beside the same members stepped
    one after the other as the reference's member loop does
    (src/__main__.py:223-254), both with the FP64 fast solver."""
    import copy
    import torch
    from pybella_amd import ensemble as en
    from pybella_amd import second_projection as sp
    from pybella_amd import time_update as tu
    members, elem, node, ud, th = en.rising_bubble_members(count, nx, ny)
    E = en.Ensemble(copy.deepcopy(members), elem, node, ud, th, mode="f64fast")
    E.run(0.0, 1e9, [0, 0], 1)                      # warm-up (graphs, workspaces)
    for M, (Sol, flux, mpv) in zip(E.models, members):
        M.upload(Sol, flux, mpv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, recs = E.run(0.0, 1e9, [0, 0], nsteps)
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    base = dict(sp.solver_options)
    sp.solver_options["mode"] = "f64fast"
    try:
        models = [tu.DeviceModel(elem, node, copy.deepcopy(ud), th, op=sp.Operator2D(sp.grid_of(node, ud)))
                  for _ in members]
        for M, (Sol, flux, mpv) in zip(models, members):
            M.upload(Sol, flux, mpv)
            tu.step(M, 0.0, 1e9, 0, 0)
            M.upload(Sol, flux, mpv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for M in models:
            t = 0.0
            for k in range(nsteps):
                dt, _, _, _ = tu.step(M, t, 1e9, k, k)
                t += dt
        torch.cuda.synchronize()
        el_s = time.perf_counter() - t0
    finally:
        sp.solver_options.clear()
        sp.solver_options.update(base)
    its = [sum(i["iters"] for i in r[1]) for rr in recs for r in rr]
    out = {"workload": f"rising-bubble ensemble, {count} members on {nx}x{ny} nodes (delth, xc, r0 per member from "
                       f"default_rng(20261204 + m)), {nsteps} steps, FP64 fast solves",
           "member_steps_per_s_batched": round(count * nsteps / el_b, 2),
           "member_steps_per_s_sequential": round(count * nsteps / el_s, 2),
           "batched_over_sequential": round(el_s / el_b, 2),
           "mean_iterations_per_member_step": round(float(np.mean(its)), 1)}
    out["letkf"] = bench_letkf(args, E, count, nx, ny)
    log("letkf:", out["letkf"])
    log("ensemble:", out)
    print(json.dumps(out), flush=True)


def bench_letkf(args, E, count, nx, ny, obs_X=11):
    """One LETKF analysis (r-localisation, the reference's defaults: rho,
    rhou, rhov, rhoY on cells and p2_nodes on nodes, 11 x 11 windows,
    gaussian weights, 10 % of the locations observed, VarCov noise of 5 %)
    of the ensemble's members, observed from the ensemble mean plus noise;
    timed on the GPU, and on one host core by the oracle on a bounded
    sample of grid points."""
    import copy
    import torch
    from pybella_amd import letkf as lk
    from pybella_amd import second_projection as sp
    cell = ("rho", "rhou", "rhov", "rhoY")
    rng = np.random.default_rng(20261206)
    obs, mask, covar = {}, {}, {}
    for a in cell + ("p2_nodes",):
        f = np.mean([sp.from_dev(M.p2n if a == "p2_nodes" else M.sol.f[a]) for M in E.models], axis=0)[2:-2, 2:-2]
        var = 0.05 * ((f - f.mean()) ** 2).mean() + 1e-12
        obs[a] = f + np.sqrt(var) * rng.standard_normal(f.shape)
        mask[a] = (rng.random(f.shape) >= 0.10).astype(float)
        covar[a] = var
    L = lk.LETKF(E.ud, attributes=cell + ("p2_nodes",), obs_X=obs_X, obs_Y=obs_X)
    saved = [[t.clone() for t in list(M.sol.f.values()) + [M.p2n]] for M in E.models]
    L.analyse(E.models, obs, mask, covar)   # warm-up
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for M, sv in zip(E.models, saved):
            for t, v in zip(list(M.sol.f.values()) + [M.p2n], sv):
                t.copy_(v)
        L.analyse(E.models, obs, mask, covar)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    out = {"members": count, "grid": f"{nx - 1}x{ny - 1} cells + {nx}x{ny} nodes", "window": f"{obs_X}x{obs_X}",
           "analysis_ms": round(el * 1e3, 3), "grid_points": (nx - 1) * (ny - 1) + nx * ny}
    if not args.no_cpu_baseline:
        from oracle import letkf as ol
        names = list(E.models[0].sol.f)
        cols = obs_X + 1                            # a bounded sample: obs_X + 1 columns of each grid
        ms, parts = 0.0, []
        for attrs, pick, gx, gy in ((cell, lambda sv, a: sv[names.index(a)], nx - 1, ny - 1),
                                    (("p2_nodes",), lambda sv, a: sv[6], nx, ny)):
            state = np.array([[sp.from_dev(pick(sv, a))[2:-2, 2:-2][:cols] for a in attrs] for sv in saved])
            t0 = time.perf_counter()
            ol.analyse_grid(state, np.array([obs[a][:cols] for a in attrs]), np.array([mask[a][:cols] for a in attrs]),
                            np.array([covar[a] for a in attrs]), ol.loc_matrix(obs_X, obs_X),
                            ol.boundary_mask((cols, gy), (True, False), (obs_X - 1) // 2, (obs_X - 1) // 2),
                            obs_X, obs_X, 1.0)
            per_point = (time.perf_counter() - t0) / (cols * gy)
            ms += 1e3 * per_point * gx * gy
            parts.append(f"{per_point * 1e3:.3f} ms per {len(attrs)}-attribute point")
        out["cpu_baseline"] = {"value": round(ms, 1), "unit": "ms per analysis", "cores": 1, "kind": "port",
                               "sample": f"the oracle's per-point analysis (oracle/letkf.py, NumPy + SciPy eigh, "
                                         f"the same window) on {cols} columns of the cell and of the node grid "
                                         f"({', '.join(parts)}), scaled to every grid point"}
        out["gpu_over_cpu"] = round(ms / out["analysis_ms"], 1)
    return out


class _SkipSmall(Exception):
    pass


def bench_timestep(args):
    """Time steps per second of the rising bubble (time_update.do's loop
    body: advection, explicit parts, both projections; the state resident on
    the GPU, dt from the device CFL maxima): 161x81 nodes with the reference's
    solver semantics (the CPU-runnable configuration), and 1025x513 nodes
    with the plain FP64 solve and with the learned mode.  The first 161x81
    steps are checked against the oracle's do() and timed on the CPU."""
    import copy
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import time_update as tu
    from pybella_amd import workloads as W
    out = {"workload": "rising bubble (inputs/rising_bubble.py sol_init), psinc, dt = 21.69/t_ref on steps 0-1 "
                       "then CFL 0.5; one step = both projections + RK and Strang advection + explicit parts"}

    def run(nx, ny, nsteps, opts):
        sp.solver_options.update(opts)
        ud = W.rising_bubble_ud(nx, ny)
        host = W.rising_bubble_init(ud)
        Sol, flux, mpv, elem, node, ud, th = copy.deepcopy(host)
        M = tu.model_for(elem, node, ud, th)
        M.upload(Sol, flux, mpv)
        tu.step(M, 0.0, 1e9, 0, 0)          # warm-up (workspaces, graphs)
        M.upload(Sol, flux, mpv)
        rate, t, recs = tu.steps_per_second(M, 0.0, 1e9, nsteps)
        S2, f2, m2 = copy.deepcopy((Sol, flux, mpv))
        M.download(S2, f2, m2, halves=False)
        return host, rate, t, recs, (S2, m2)

    base = dict(sp.solver_options)
    want = (lambda k: args.ts_only in (None, k))
    try:
        # 161 x 81, reference semantics
        n1 = max(args.steps, 10)
        if not want("161x81"):
            raise _SkipSmall
        host, rate, t, recs, _ = run(161, 81, n1, {"mode": "f64"})
        out["161x81"] = {"steps_per_s": round(rate, 2), "ms_per_step": round(1e3 / rate, 3), "steps": n1,
                         "t_end": round(t, 4), "solver": SOLVER_NAMES["f64"],
                         "iterations_per_step": [r["iters"] for r in recs]}
        log("time steps 161x81:", out["161x81"])
        # parity and the CPU leg on the first two steps
        _, _, _, _, (S2, m2) = run(161, 81, 2, {"mode": "f64"})
        if not args.no_cpu_baseline:
            from oracle import timestep as ots
            Sol, flux, mpv, elem, node, ud, th = copy.deepcopy(host)
            ud.stepmax = 2
            t0 = time.perf_counter()
            Sol, flux, mpv, _ = ots.do(Sol, flux, mpv, 0.0, 1e9, ud, elem, node, [0, 0], th)
            el = time.perf_counter() - t0
            errs = {k: float(np.max(np.abs(getattr(S2, k) - getattr(Sol, k))) / np.max(np.abs(getattr(Sol, k))))
                    for k in ("rho", "rhou", "rhov", "rhoY", "rhoX")}
            errs["p2_nodes_mod_null"] = _node_null_rel(m2.p2_nodes, mpv.p2_nodes)
            out["161x81"]["parity_vs_oracle_after_2_steps"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
            out["161x81"]["cpu_baseline"] = {
                "value": round(2 / el, 4), "unit": "time steps/s", "cores": 1, "kind": "port",
                "sample": f"the first 2 steps of the same run by the NumPy oracle (oracle/timestep.py do(), "
                          f"vectorised restatement, FP64 BiCGStab to 1e-8), {el:.2f} s; CPU: {_cpu_model()}"}
            out["161x81"]["gpu_over_cpu"] = round(rate / (2 / el), 1)
    except _SkipSmall:
        pass
    try:
        # 1025 x 513: plain FP64 solve and the learned mode
        for name, opts in (("f64fast", {"mode": "f64fast"}), ("nn", {"mode": "nn", "nn_budget": 300})):
            if not want(f"1025x513_{name}"):
                continue
            if name == "nn":
                from pybella_amd.nn import Network, default_weights_path
                net = Network(seed=0)
                net.load(default_weights_path())
                ud = W.rising_bubble_ud(1025, 513)
                sp.operator_for(W.rising_bubble_init(ud)[4], ud).attach_nn(net)
            _, rate, t, recs, _ = run(1025, 513, 4, opts)
            out[f"1025x513_{name}"] = {"steps_per_s": round(rate, 3), "ms_per_step": round(1e3 / rate, 2),
                                       "steps": 4, "t_end": round(t, 4), "solver": SOLVER_NAMES[name],
                                       "iterations_per_step": [r["iters"] for r in recs]}
            log(f"time steps 1025x513 {name}:", out[f"1025x513_{name}"])
    finally:
        sp.solver_options.clear()
        sp.solver_options.update(base)
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


# Algorithmic words per node of the 3D kernels (cell weights ~ one word per
# node; compressible, so the centre term is read):
#   0 matvec: read p, h, hc, dinv; write y                                 -> 5
#   1 launch 1: read r, p, v, t, rt, x, h, hc, dinv; write r, p, v, x       -> 13
#   2 launch 2: read r, v, rt, h, hc, dinv; write t                         -> 7
KERNEL3_WORDS = {0: 5, 1: 13, 2: 7}
KERNEL3_NAMES = {0: "matvec27", 1: "bicg3_launch1_update_pupdate_matvec", 2: "bicg3_launch2_s_matvec_dots"}


def bench_3d(args, nx=257, ny=129, nz=129):
    """Isolated 3D solve: stratified compressible atmosphere, seeded 64-mode
    zero-mean RHS (|k| <= 12), mixed FP32 Krylov in FP64 refinement to true
    relative residual 1e-8; per-kernel device times and algorithmic GB/s."""
    import torch
    from pybella_amd import workloads as W
    from pybella_amd.op3d import Operator3D, grid3_of, phys3_of
    log(f"building 3D workload {nx}x{ny}x{nz} ...")
    st = W.stratified_state_3d(nx, ny, nz)
    rho = torch.from_numpy(st.rho_c).to("cuda")
    rhoY = torch.from_numpy(st.rhoY_c).to("cuda")
    f = torch.from_numpy(W.fourier_rhs_3d(nx, ny, nz, seed=20261203)).to("cuda")
    op = Operator3D(grid3_of(st))
    phys = phys3_of(st)
    di = torch.empty(nz, ny, nx, dtype=torch.float64, device="cuda")
    x = torch.empty_like(di)
    if args.kernel_only is not None:   # one kernel only, for profiler counter runs
        op.assemble(phys, 0.05, rho, rhoY, None, di)
        k = args.kernel_only
        us = op.time_kernel(k, args.precision, args.kernel_reps)
        alg = KERNEL3_WORDS[k] * (4 if args.precision == "f32" else 8) * nx * ny * nz
        print(json.dumps({"kernel": KERNEL3_NAMES[k], "precision": args.precision, "avg_us": us,
                          "alg_bytes": alg, "alg_GBps": alg / (us * 1e-6) / 1e9}), flush=True)
        return

    def step(mode):
        op.assemble(phys, 0.05, rho, rhoY, None, di)
        b = f * di
        _, info = op.solve(b, x, tol=1e-8, mode=mode)
        if info["info"] != 0 or info["true_rel_resid"] > 1e-8:
            raise SystemExit(f"3D solve did not reach the tolerance: {info}")
        return info

    rates, infos = {}, {}
    for mode in ("mixed", "f64"):
        step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            infos[mode] = step(mode)
        torch.cuda.synchronize()
        rates[mode] = args.steps / (time.perf_counter() - t0)
        log(f"3D {mode}: {rates[mode]:.2f} solves/s {infos[mode]}")
    n = nx * ny * nz
    kernels = {}
    for prec, wb in (("f32", 4), ("f64", 8)):
        for k, words in KERNEL3_WORDS.items():
            us = op.time_kernel(k, prec, args.kernel_reps)
            alg = words * wb * n
            kernels[f"{KERNEL3_NAMES[k]}/{prec}"] = {"avg_us": round(us, 3), "alg_bytes": alg,
                                                     "alg_GBps": round(alg / (us * 1e-6) / 1e9, 1),
                                                     "frac": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    print(json.dumps({"workload": f"isolated 3D nodal 27-point solve, {nx}x{ny}x{nz} nodes, stratified theta "
                                  "300->330 K, compressible, seeded 64-mode zero-mean RHS, true rel. resid <= 1e-8",
                      "solves_per_s": {k: round(v, 3) for k, v in rates.items()},
                      "iterations": {k: v["iters"] for k, v in infos.items()},
                      "refinements": {k: v["outer"] for k, v in infos.items()},
                      "true_rel_resid": {k: v["true_rel_resid"] for k, v in infos.items()},
                      "kernels": kernels, "hbm_peak_GBps": HBM_PEAK_GBS}), flush=True)


def projection_step_rate(mode, rows_per_thread, steps=3, nx=1025, ny=513, dt=0.05):
    """Device-resident projection steps/s at a given tile height: the drop-in's
    device work (preparation, solve, update) on the 1025x513 rising bubble of
    ``--workload step``, its state restored from device copies before each
    step, inside the timed region."""
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    Sol, Sol0, mpv, elem, node, ud, th = W.rising_bubble_state(nx, ny, dt=dt)
    op = sp.operator_for(node, ud)
    op.set_tile(rows_per_thread)
    pr = sp._upload(Sol, mpv, elem, node, ud, th, dt, Sol0, op)
    dev0 = [(t, t.clone()) for t in list(pr.sol.f.values()) + list(pr.sol0.f.values()) + [pr.p2n]]

    def step():
        for t, t0 in dev0:
            t.copy_(t0)
        sp._project_prepare(pr)
        x, info = op.solve(pr.b, tol=float(ud.tol), maxiter=int(ud.max_iterations), mode=mode)
        sp._project_update(pr, x)
        if info["info"] != 0:
            raise SystemExit(f"projection step solve did not converge: {info}")
        return info

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        info = step()
    torch.cuda.synchronize()
    return {"workload": f"projection (drop-in impl_part device work), {nx}x{ny} rising bubble, dt {dt}",
            "mode": mode, "rows_per_thread": rows_per_thread,
            "projections_per_s": round(steps / (time.perf_counter() - t0), 3), "iterations": info["iters"]}


def bench_step(args, nx=1025, ny=513, dt=0.05):
    """Projection steps/s: the whole drop-in euler_backward_non_advective_impl_part
    (second_projection.py:97-277) on a 1025x513 rising-bubble state after its
    buoyancy impulse (pseudo-incompressible, Sol0 given): ghost cells, the
    correction with the old p2, divergence, assembly, the solve to the
    reference's tolerance 1e-8, the scatter, the second correction and the
    Exner update.

    device-resident: the state is uploaded once; every timed step first
    restores the six cell fields, Sol0's fields and p2 from device copies of
    the initial state (about 60 MB of device-to-device copies, inside the
    timed region), so each step projects the same divergent state.
    host-boundary: the drop-in call itself, host arrays in and out (PCIe
    transfers of Sol, Sol0, mpv included); the host state is restored before
    each call, outside the timed region.  Its rate is from the median call
    (host-side time varies with the host's load); every call's time is listed."""
    import copy
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    from pybella_amd.autotune import tune_tile
    log(f"building rising-bubble projection step {nx}x{ny} ...")
    Sol, Sol0, mpv, elem, node, ud, th = W.rising_bubble_state(nx, ny, dt=dt)
    host0 = (copy.deepcopy(Sol), copy.deepcopy(Sol0), copy.deepcopy(mpv))
    op = sp.operator_for(node, ud)
    solver_options = dict(sp.solver_options)

    def restore_host():
        for dst, src in ((Sol, host0[0]), (Sol0, host0[1])):
            for k in sp._CELL_FIELDS:
                getattr(dst, k)[...] = getattr(src, k)
        mpv.p2_nodes[...] = host0[2].p2_nodes

    pr = sp._upload(Sol, mpv, elem, node, ud, th, dt, Sol0, op)
    dev0 = [(t, t.clone()) for t in list(pr.sol.f.values()) + list(pr.sol0.f.values()) + [pr.p2n]]

    def dev_step(mode, parts=None):
        for t, t0 in dev0:
            t.copy_(t0)
        if parts is not None:
            torch.cuda.synchronize()
            c0 = time.perf_counter()
        sp._project_prepare(pr)
        if parts is not None:
            torch.cuda.synchronize()
            c1 = time.perf_counter()
        x, info = op.solve(pr.b, tol=float(ud.tol), maxiter=int(ud.max_iterations), mode=mode)
        if parts is not None:
            torch.cuda.synchronize()
            c2 = time.perf_counter()
        sp._project_update(pr, x)
        if parts is not None:
            torch.cuda.synchronize()
            parts.update(prepare_ms=round((c1 - c0) * 1e3, 3), solve_ms=round((c2 - c1) * 1e3, 3),
                         update_ms=round((time.perf_counter() - c2) * 1e3, 3))
        if info["info"] != 0:
            raise SystemExit(f"projection step solve did not converge: {info}")
        return info

    out = {}
    try:
        for mode in ("f64", "f64fast", "mixed"):
            sp.solver_options["mode"] = mode
            best, _ = tune_tile(op, lambda: dev_step(mode), key=(nx, ny, "step", mode))
            parts = {}
            dev_step(mode, parts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                info = dev_step(mode)
            torch.cuda.synchronize()
            dev_rate = args.steps / (time.perf_counter() - t0)
            host_ms = []
            for it in range(args.steps + 1):   # the first call (untimed) allocates the pinned staging buffers
                restore_host()
                t0 = time.perf_counter()
                hinfo = sp.euler_backward_non_advective_impl_part(Sol, mpv, elem, node, ud, th, 0.0, dt, 0.0,
                                                                  Sol0=Sol0)
                if it > 0:
                    host_ms.append((time.perf_counter() - t0) * 1e3)
                if hinfo["info"] != 0:
                    raise SystemExit(f"drop-in projection step did not converge: {hinfo}")
            out[mode] = {"device_resident_steps_per_s": round(dev_rate, 3), "rows_per_thread": best,
                         "host_boundary_steps_per_s": round(1e3 / float(np.median(host_ms)), 3),
                         "host_boundary_call_ms": [round(v, 2) for v in host_ms],
                         "iterations": info["iters"], "true_rel_resid": info["true_rel_resid"],
                         "parts_device_resident": parts}
            log(f"projection step {mode}: {out[mode]}")
    finally:
        sp.solver_options.clear()
        sp.solver_options.update(solver_options)
    print(json.dumps({"workload": f"projection step euler_backward_non_advective_impl_part, {nx}x{ny} nodes, "
                                  "rising bubble (dtheta 2 K, r0 2 km) after its buoyancy impulse, "
                                  f"pseudo-incompressible, dt {dt}, tol {ud.tol}",
                      "steps": args.steps, "modes": out}), flush=True)


def bench_batch(args, sizes=((129, 65), (257, 129), (513, 257), (1025, 513))):
    """64 rising-bubble pressure solves (members_per_size at each grid size,
    seeded bubble amplitude, centre and radius), each member's operator and
    right-hand side from its own projection step; all solved by one batched
    solve (one hipGraph per chunk of iterations over every member) vs the
    same solves one after the other.  Tile heights come from the autotuner,
    one per grid size."""
    import torch
    from pybella_amd import second_projection as sp
    from pybella_amd import workloads as W
    from pybella_amd.autotune import tune_tile
    from pybella_amd.batch import BatchSolver
    log("building batch workload ...")
    members, ops, bs = [], [], []
    for nx, ny in sizes:
        for m in range(args.members_per_size):
            Sol, Sol0, mpv, elem, node, ud, th = W.batch_member_bubble(nx, ny, m)
            op = sp.Operator2D(sp.grid_of(node, ud))
            pr = sp._prepare(Sol, mpv, elem, node, ud, th, 0.05, Sol0, op)
            members.append((nx, ny, m))
            ops.append(op)
            bs.append(pr.b)
    torch.cuda.synchronize()
    tiles = {}
    for nx, ny in sizes:   # autotune the tile height on the first member of each size
        k = members.index((nx, ny, 0))
        best, table = tune_tile(ops[k], lambda: ops[k].solve(bs[k], tol=1e-8, mode="mixed"),
                                key=(nx, ny, "mixed"))
        tiles[f"{nx}x{ny}"] = {"rows_per_thread": best, "seconds": {str(a): round(b, 5) for a, b in table.items()}}
        for (mx, my, _), op in zip(members, ops):
            if (mx, my) == (nx, ny):
                op.set_tile(best)
    batch = BatchSolver(ops)
    # one tile height for the whole batch lets it run one launch over every
    # member's tiles per phase; the autotuner times the batch itself
    per_size = {v["rows_per_thread"] for v in tiles.values()}
    btable = {}
    from pybella_amd.autotune import CANDIDATES
    for ry in CANDIDATES:
        for op in ops:
            op.set_tile(ry)
        batch.solve(bs, tol=1e-8, mode="f64fast")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.solve(bs, tol=1e-8, mode="f64fast")
        torch.cuda.synchronize()
        btable[ry] = time.perf_counter() - t0
    bry = min(btable, key=btable.get)
    for op in ops:
        op.set_tile(bry)
    tiles["batch"] = {"rows_per_thread": bry, "seconds": {str(a): round(b, 5) for a, b in btable.items()},
                      "per_size_choices": sorted(per_size)}
    out = {"workload": f"{len(ops)} rising-bubble pressure solves, {args.members_per_size} each at "
                       + ", ".join(f"{nx}x{ny}" for nx, ny in sizes) + " nodes, true rel. resid <= 1e-8",
           "unknowns": int(sum(op.nx * op.ny for op in ops)), "tiles": tiles}
    for mode in ("mixed", "f64fast"):
        batch.solve(bs, tol=1e-8, mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, infos = batch.solve(bs, tol=1e-8, mode=mode)
        torch.cuda.synchronize()
        t_batch = (time.perf_counter() - t0) / args.steps
        bad = [i for i in infos if i["info"] != 0 or i["true_rel_resid"] > 1e-8]
        if bad:
            raise SystemExit(f"batch member did not converge: {bad[0]}")
        for op, b in zip(ops, bs):
            op.solve(b, tol=1e-8, mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for op, b in zip(ops, bs):
                op.solve(b, tol=1e-8, mode=mode)
        torch.cuda.synchronize()
        t_seq = (time.perf_counter() - t0) / args.steps
        out[mode] = {"batched_solves_per_s": round(len(ops) / t_batch, 2),
                     "sequential_solves_per_s": round(len(ops) / t_seq, 2),
                     "batched_ms": round(1e3 * t_batch, 2), "sequential_ms": round(1e3 * t_seq, 2),
                     "iterations_per_size": {f"{nx}x{ny}": int(np.mean([i["iters"] for (mx, my, _), i in
                                                                       zip(members, infos) if (mx, my) == (nx, ny)]))
                                             for nx, ny in sizes}}
        log(f"batch {mode}: {out[mode]}")
    print(json.dumps(out), flush=True)


def only_run(args, op, phys, dt, rho, rhoY, dS, rhs_d, b, x):
    """Warm-up and K timed solves of one mode (assembly + scaling + solve per
    step, as in the default line), for profiler runs."""
    import torch
    cfg = dict(mode=args.only if args.only in ("f64fast", "f64fused") else "f64fast", maxiter=6000)
    if args.only in ("nn", "mg"):
        from pybella_amd.nn import Network, default_weights_path
        net = Network(seed=0)
        net.load(default_weights_path())
        op.attach_nn(net)
        cfg = dict(mode="nn", maxiter=2000, sweeps=2, krylov_m=30, fallback_maxiter=6000,
                   use_network=True if args.only == "nn" else "mg")
    else:
        op.set_tile(args.only_tile)

    def step():
        op.assemble(phys, dt, rho, rhoY, rho, rhoY, dS)
        op.scale_rhs(rhs_d, b)
        return op.solve(b, x, tol=1e-8, **cfg)[1]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"only": args.only, "solves_per_s": round(args.steps / el, 3),
                      "ms_per_solve": round(1e3 * el / args.steps, 3), "iterations": infos[-1]["iters"],
                      "true_rel_resid": infos[-1]["true_rel_resid"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="auto", choices=["auto", "nn", "f64", "f64fast", "f64fused", "mixed"],
                    help="auto = the learned-preconditioned solve timed beside the fastest plain mode")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the oracle checks and the CPU baseline")
    ap.add_argument("--no-extra-workloads", action="store_true",
                    help="skip the 3D and batch summaries of the default line")
    ap.add_argument("--only", default=None, choices=["nn", "f64fast", "f64fused", "mg"],
                    help="time only this solve of the 2D workload (no calibration; for profiler runs): the "
                         "learned mode, plain BiCGStab (two launches or one per iteration), or the cycle "
                         "continued to the coarsest level")
    ap.add_argument("--only-tile", type=int, default=5, help="rows per thread of the plain solves under --only")
    ap.add_argument("--cpu-sample-iters", type=int, default=300)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--kernel-only", type=int, default=None,
                    help="only time solver kernel K (0 matvec, 1 fast A launch, 2 B launch, 3 FP64 true "
                         "residual) on the workload, for profiler runs")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32"])
    ap.add_argument("--tile", type=int, default=4, help="rows per thread for --kernel-only")
    ap.add_argument("--workload", default="2d", choices=["2d", "3d", "batch", "step", "timestep", "ensemble"],
                    help="3d: the isolated 257x129x129 27-point solve; batch: 64 uneven rising-bubble "
                         "solves in one graph per chunk; step: whole projection steps (the drop-in "
                         "impl_part) at 1025x513; timestep: rising-bubble time steps/s at 161x81 and "
                         "1025x513; ensemble: 16 rising-bubble members stepped with batched solves (each its "
                         "own JSON line, not the metric)")
    ap.add_argument("--ts-only", default=None, choices=["161x81", "1025x513_f64fast", "1025x513_nn"],
                    help="--workload timestep: run only this configuration (for profiler runs)")
    ap.add_argument("--members-per-size", type=int, default=16)
    ap.add_argument("--batch-sizes", default="129x65,257x129,513x257,1025x513",
                    help="grid sizes of the batch workload, NXxNY comma-separated")
    ap.add_argument("--copy-probe-mib", type=int, default=None,
                    help="only run a device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(a solve would shard by domain decomposition into row slabs). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pybella_amd import second_projection as sp

    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 10}))
        return

    if args.workload == "3d":
        bench_3d(args)
        return
    if args.workload == "step":
        bench_step(args)
        return
    if args.workload == "timestep":
        bench_timestep(args)
        return
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "batch":
        bench_batch(args, tuple(tuple(int(v) for v in z.split("x")) for z in args.batch_sizes.split(",")))
        return

    log("building workload 1025x513 ...")
    state, rhs = build_workload()
    Sol, mpv, elem, node, ud, th = state
    dt = 0.05
    grid = sp.grid_of(node, ud)
    phys = sp.phys_of(ud, th)
    op = sp.Operator2D(grid)
    rho, rhoY = sp.to_dev(Sol.rho), sp.to_dev(Sol.rhoY)
    dS = sp.to_dev(sp.dSdy_cells(mpv, elem, node))
    rhs_d = sp.to_dev(rhs)
    b = torch.empty(grid.ny, grid.nx, dtype=torch.float64, device="cuda")
    x = torch.empty_like(b)
    torch.cuda.synchronize()

    if args.kernel_only is not None:
        op.assemble(phys, dt, rho, rhoY, rho, rhoY, dS)
        op.set_tile(args.tile)
        us = op.time_kernel(args.kernel_only, args.precision, args.kernel_reps)
        n = grid.nx * grid.ny
        words = KERNEL_WORDS[args.kernel_only]
        wb = 8 if args.precision == "f64" else 4
        alg = words * wb * n + (KERNEL_XBYTES.get(args.kernel_only, 0) * n)
        print(json.dumps({"kernel": KERNEL_NAMES[args.kernel_only], "precision": args.precision,
                          "tile_rows_per_thread": args.tile, "avg_us": us, "alg_bytes": alg,
                          "alg_GBps": alg / (us * 1e-6) / 1e9}))
        return

    if args.only is not None:
        only_run(args, op, phys, dt, rho, rhoY, dS, rhs_d, b, x)
        return

    nn_net = None
    if args.mode in ("auto", "nn"):
        from pybella_amd.nn import Network, default_weights_path
        nn_net = Network(seed=0)
        nn_net.load(default_weights_path())
        op.attach_nn(nn_net)

    def step(mode, cfg=None):
        op.assemble(phys, dt, rho, rhoY, rho, rhoY, dS)
        op.scale_rhs(rhs_d, b)
        if mode == "nn":
            _, info = op.solve(b, x, tol=1e-8, mode="nn", fallback_maxiter=6000, **cfg)
        else:
            _, info = op.solve(b, x, tol=1e-8, maxiter=6000, mode=mode, inner_tol=1e-3)
        return info

    def rate_of(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return reps / (time.perf_counter() - t0)

    from pybella_amd.autotune import tune_tile
    # plain FP64 reference-semantics solve: the iteration count the learned
    # mode's budget is measured against
    info64 = step("f64")
    log("f64 (exact) solve:", info64)
    rates, tiles, infos_cal = {}, {}, {}
    plain_modes = ["f64", "f64fast", "f64fused", "mixed"] if args.mode in ("auto", "nn") else [args.mode]
    for m in plain_modes:
        best, table = tune_tile(op, lambda: step(m), key=(grid.nx, grid.ny, m))
        infos_cal[m] = step(m)
        rates[m] = rate_of(lambda: step(m))
        tiles[m] = {"rows_per_thread": best, "seconds": {str(k): round(v, 5) for k, v in table.items()}}
        log(f"mode {m}: {rates[m]:.2f} solves/s, tile rows/thread {best}, {infos_cal[m]}")
    plain = max(rates, key=rates.get)
    op.set_tile(tiles[plain]["rows_per_thread"])

    # the learned mode: FGMRES preconditioned by the two-branch multigrid
    # cycle whose coarse level the network solves; plain BiCGStab takes over
    # only past 2x the plain FP64 iteration count (the fallback budget)
    nn_cfg, nn_table, fallback_after = None, {}, None
    if nn_net is not None:
        fallback_after = 2 * int(info64["iters"])
        for sw in (1, 2):
            c = dict(maxiter=fallback_after, sweeps=sw, krylov_m=30)
            nn_table[json.dumps(c, sort_keys=True)] = round(rate_of(lambda: step("nn", c)), 3)
            log(f"mode nn {c}: {nn_table[json.dumps(c, sort_keys=True)]:.2f} solves/s")
        nn_cfg = json.loads(max(nn_table, key=nn_table.get))
    mode = "nn" if nn_cfg is not None else plain

    def timed(m, cfg):
        for _ in range(args.warmup):
            step(m, cfg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infs = [step(m, cfg) for _ in range(args.steps)]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        for inf in infs:
            if inf["info"] != 0 or inf["true_rel_resid"] > 1e-8:
                raise SystemExit(f"solve did not reach the tolerance: {inf}")
        return el, infs

    el, infos = timed(mode, nn_cfg)
    log("timed:", mode, nn_cfg, infos[-1])
    value = args.steps / el
    x_main = x.clone()
    plain_rate, plain_info = None, None
    comparators = {}
    if mode == "nn":   # the plain BiCGStab it is compared with, timed the same way
        el_p, infos_p = timed(plain, None)
        plain_rate, plain_info = args.steps / el_p, infos_p[-1]
        log("timed plain:", plain, plain_info)
        # the same FGMRES and cycle without the network (its level smoothed by
        # Jacobi sweeps instead), and the cycle continued to the coarsest level
        for name, use in (("cycle_without_network", False), ("multigrid_to_coarsest_no_network", "mg")):
            c = dict(nn_cfg, use_network=use)
            r = rate_of(lambda: step("nn", c))
            inf = step("nn", c)
            comparators[name] = {"solves_per_s": round(r, 3), "iterations": inf["iters"],
                                 "true_rel_resid": inf["true_rel_resid"], "fallback": inf["fallback"]}
            log(name, comparators[name])
    x_plain = x.clone()
    op.set_tile(tiles[plain]["rows_per_thread"])

    # roofline of the plain BiCGStab's dominant launch: HIP events around every
    # launch inside a real solve, and on pattern-filled vectors beside it
    n = grid.nx * grid.ny
    kt = {k: op.time_kernel(k, "f64", args.kernel_reps) for k in (0, 1, 2, 3)}
    op.assemble(phys, dt, rho, rhoY, rho, rhoY, dS)
    op.scale_rhs(rhs_d, b)
    in_solve = {}
    for k in (1, 2):
        us, cnt, _ = op.time_in_solve(b, k)
        in_solve[k] = (us, cnt)
    dom = max((1, 2), key=lambda k: in_solve[k][0])
    alg = KERNEL_WORDS[dom] * 8 * n + KERNEL_XBYTES.get(dom, 0) * n
    achieved = alg / (in_solve[dom][0] * 1e-6) / 1e9
    achieved_syn = alg / (kt[dom] * 1e-6) / 1e9
    roof = {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_us": round(in_solve[dom][0], 3), "timed_in_solve_launches": in_solve[dom][1], "alg_bytes": alg,
            "synthetic_state": {"avg_us": round(kt[dom], 3), "achieved": round(achieved_syn, 1),
                                "frac": round(achieved_syn / HBM_PEAK_GBS, 4)},
            "kernels_us_in_solve": {KERNEL_NAMES[k]: round(v[0], 3) for k, v in in_solve.items()},
            "kernels_us_synthetic": {KERNEL_NAMES[k]: round(v, 3) for k, v in kt.items()}}
    try:   # HBM bytes per launch: rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel (profiles/)
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_counters.json")) as f:
            hc = json.load(f)
        rec = hc["kernels"].get(f"{KERNEL_NAMES[dom]}/f64/{tiles[plain]['rows_per_thread']}")
        if rec:
            roof["traffic"] = int(1024 * (rec["fetch_kib"] * hc["fetch_scale"] + rec["write_kib"] * hc["write_scale"]))
            roof["traffic_source"] = "profiles/hbm_counters.json (rocprofv3 --pmc of the same kernel and tile)"
    except (OSError, KeyError, ValueError):
        pass

    # projections (the drop-in's whole device work per call) per second
    proj_rate = None
    if "f64fast" in tiles:
        proj_rate = projection_step_rate("f64fast", tiles["f64fast"]["rows_per_thread"])
        log("projection step:", proj_rate)

    parity = None
    cpu = None
    if not args.no_cpu_baseline:
        log("oracle solve and cpu baseline ...")
        parity, cpu = oracle_checks(state, rhs, info64["iters"], args.cpu_sample_iters,
                                    {"nn": x_main.cpu().numpy(), plain: x_plain.cpu().numpy()})
        log("parity:", parity)
    extra = {}
    if not args.no_extra_workloads:
        extra = extra_workloads(args)

    out = {
        "metric": "solves/s at rel. resid 1e-8, 1024x512 nodal solve: NN-precond vs plain BiCGStab",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32", "nn": "f64+f32+bf16"}.get(mode, "f64"),
        "data": "synthetic",
        "config": {"workload": "isolated 2D nodal 9-point solve, 1025x513 nodes, stratified theta 300->330 K, "
                               "psinc, seeded 64-mode RHS, true rel. resid <= 1e-8",
                   "solver": SOLVER_NAMES[mode],
                   "nn_config": nn_cfg,
                   "nn_fallback_after_iters": fallback_after,
                   "iterations": infos[-1]["iters"], "refinements": infos[-1]["outer"],
                   "fallback_used": infos[-1].get("fallback", 0),
                   "true_rel_resid": infos[-1]["true_rel_resid"],
                   "f64_iterations": info64["iters"],
                   "plain_bicgstab": None if plain_rate is None else {
                       "solver": SOLVER_NAMES[plain], "solves_per_s": round(plain_rate, 3),
                       "iterations": plain_info["iters"], "true_rel_resid": plain_info["true_rel_resid"]},
                   "nn_over_plain": None if plain_rate is None else round(value / plain_rate, 3),
                   "comparators": comparators,
                   "parity_vs_oracle": parity,
                   "calibration_solves_per_s": {**{k: round(v, 3) for k, v in rates.items()},
                                                **{"nn " + k: v for k, v in nn_table.items()}},
                   "tiles": tiles,
                   "projections": proj_rate,
                   **extra},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
