#!/usr/bin/env python3
"""Benchmark: DOF-RHS evaluations per second of the DGSEM right-hand side + RK stage.

Workload (default): 3D compressible Navier-Stokes Taylor-Green vortex, periodic
[0, 2 pi]^3, 32^3 hexahedra, p = 3 (2,097,152 GLL nodes, 5 conserved variables,
FP64), Ma = 0.1, Re = 1600, Pr = 0.71, gamma = 1.4, Chandrashekar EC volume flux,
the reference's Chandrashekar surface flux, BR1 viscous terms, 5-stage low-storage
RK4 (Carpenter-Kennedy 2N) at a fixed dt.

One step = one LSRK4(5) time step = 5 RHS evaluations, each with its stage update
fused into the flux kernel's epilogue.  value = nodes x RHS evaluations / s over the
timed steps (state resident in HBM).  A DOF is a GLL node (NE * (p+1)^3).

Also reported: the bare DGSEMOperator::Mult rate, per-kernel HIP-event timings and
the roofline of the dominant kernel, and the CPU oracle's rate on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic HBM bytes per node per launch (compulsory traffic: every array read or
# written once; the second read of a face trace by the neighbouring element excluded)
# -- FP64, 3D, p=3, 5 conserved variables:
#   k_lift : read u (40); write the viscous volume term dv (4 x 8 = 32) and the face
#            traces Ff (6 faces x 16 points x 4 values x 8 B / 64 nodes = 48)      = 120
#   k_rhs  : Mult epilogue: read u (40) + dv (32) + Ff (48), write dudt (40)        = 160
#            LSRK epilogue: read u, dv, Ff, k (160), write k and u' (80)            = 240
BYTES_PER_NODE = {"lift": 120.0, "rhs_mult": 160.0, "rhs_lsrk": 240.0}
# One RHS evaluation's two-pass minimum (BR1 materialises the gradient information between the
# passes): pass A reads u (40) and writes the gradient information (120); pass B reads u (40) and
# it (120) and writes du (40)
STAGE_BYTES_PER_NODE = 360.0

# HBM bytes per launch of the dominant kernel (k_rhs_node FP64, LSRK stage, TGV 32^3 p=3),
# measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (two runs; counter KiB x 1024, the
# factor a plain copy of known size reads back): 211.1 MB + 167.8 MB.  Recorded in
# profiles/pmc_tgv32_p3_fp64_mixed.md; PMC counters cannot be read inside this process.
RHS_LSRK_TRAFFIC_TGV32_P3 = 378.9e6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# The GPU leaves its idle clock and power state over the first ~25 ms of stepping: a 20-step
# chunk right after 5 warm-up steps runs 5-10 % slower than every later chunk (measured on the
# MI355X: 8.2-8.6 against 9.0-9.1 G DOF-RHS/s, again after a 1 s pause).  Every timed leg
# therefore warms up for its W steps and then for at least this long, in chunks of W steps.
WARMUP_MIN_S = 0.25


def warm_up(run_steps, steps):
    """Run `steps` untimed steps (run_steps(n) queues n steps), then more chunks of `steps` until
    WARMUP_MIN_S has passed.  Returns the number of warm-up steps run."""
    import torch

    steps = max(1, steps)
    t0 = time.perf_counter()
    done = 0
    while True:
        run_steps(steps)
        torch.cuda.synchronize()
        done += steps
        if time.perf_counter() - t0 >= WARMUP_MIN_S:
            return done


def set_warmup_min(seconds: float):
    global WARMUP_MIN_S
    WARMUP_MIN_S = max(0.0, float(seconds))


def physical_cores():
    """Physical cores of the host (distinct (core, socket) pairs of `lscpu -p`), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=core,socket"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {tuple(line.split(",")) for line in out.splitlines() if line and not line.startswith("#")}
    return len(pairs) or None


def cpu_baseline(mesh, u, phys, threads, budget_s=15.0, single_evals=3):
    """The CPU oracle (a restatement of the reference's algorithm) on the host cores, and on one."""
    import oracle

    oracle.build()
    ph = oracle.phys(gamma=phys["gamma"], Pr=phys["Pr"], mu=phys["mu"], viscous=True)

    def timed(n_threads, max_evals, budget):
        used = oracle.set_threads(n_threads)
        oracle.rhs(mesh, ph, u)  # warm-up
        times = []
        t_start = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            oracle.rhs(mesh, ph, u)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start > budget or len(times) >= max_evals:
                break
        return used, times

    used, times = timed(threads, 20, budget_s)
    med = statistics.median(times)
    _, times1 = timed(1, single_evals, budget_s)
    med1 = statistics.median(times1)
    oracle.set_threads(threads)
    return {
        "value": mesh.num_nodes / med,
        "unit": "DOF-RHS evals/s",
        "cores": used,
        "kind": "port",
        "sample": f"{len(times)} oracle RHS evaluations of the same TGV {mesh.n[0]}^3 p={mesh.order} NS state "
                  f"(median {med:.3f} s each, {used} OpenMP threads, {budget_s:.0f} s budget); "
                  f"{len(times1)} more on one thread (median {med1:.3f} s)",
        "value_1core": mesh.num_nodes / med1,
        "physical_cores": physical_cores(),
    }


def ke_terms(u, du, W):
    """Volume-averaged kinetic energy, its rate from the RHS, and the total entropy S = sum W rho s."""
    vel = [u[1 + k] / u[0] for k in range(3)]
    v2 = sum(v * v for v in vel)
    ke = (W * 0.5 * u[0] * v2).sum() / W.sum()
    rate = (W * (sum(vel[k] * du[1 + k] for k in range(3)) - 0.5 * v2 * du[0])).sum() / W.sum()
    p = 0.4 * (u[4] - 0.5 * u[0] * v2)
    ent = (W * u[0] * (p.log() - 1.4 * u[0].log())).sum() / W.sum()
    return float(ke), float(rate), float(ent)


def compare_mixed(args, dev):
    """FP64 and FP32-mixed TGV runs side by side: throughput of each and the mixed run's errors
    against the FP64 run (density L2, kinetic-energy dissipation rate, entropy drift)."""
    import torch

    from prandtl_amd import capi, cases

    mesh = cases.tgv_mesh(n=args.n, order=args.order)
    phys = cases.tgv_physics()
    u0 = torch.from_numpy(cases.tgv_state(mesh)).to(dev)
    W = torch.from_numpy(mesh.quadrature_weights()).to(dev)
    dt = cases.cfl_dt(mesh, 11.0, args.cfl)
    nsteps = int(math.ceil(args.t_final / dt))
    dt = args.t_final / nsteps
    nout = 10
    chunk = [nsteps // nout + (1 if i < nsteps % nout else 0) for i in range(nout)]
    runs = {}
    for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
        op = capi.Operator(mesh, precision=prec, **phys)
        u = u0.clone()
        du = torch.empty_like(u)
        hist = []
        op.mult(u, du)
        hist.append((0.0,) + ke_terms(u, du, W))
        t, ms = 0.0, 0.0
        for c in chunk:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            t = op.rk_run(capi.LSRK45, u, t, dt, c)
            e1.record()
            torch.cuda.synchronize()
            ms += e0.elapsed_time(e1)
            op.mult(u, du)
            hist.append((t,) + ke_terms(u, du, W))
            log(f"{name}: t={t:.4f} K={hist[-1][1]:.8f} -dK/dt={-hist[-1][2]:.6e}")
        runs[name] = dict(u=u, hist=hist, ms=ms, value=mesh.num_nodes * 5 * nsteps / (ms * 1e-3))
        op.close()
    h64, hmx = runs["fp64"]["hist"], runs["mixed"]["hist"]
    eps_err = max(abs(b[2] - a[2]) / abs(a[2]) for a, b in zip(h64, hmx))
    rho_l2 = float(((W * (runs["mixed"]["u"][0] - runs["fp64"]["u"][0]) ** 2).sum() /
                    (W * runs["fp64"]["u"][0] ** 2).sum()).sqrt())
    ent_drift = {k: (r["hist"][-1][3] - r["hist"][0][3]) / abs(r["hist"][0][3]) for k, r in runs.items()}
    out = {
        "metric": "DOF-RHS evals/s, 3D TGV p=3 NS: FP32-mixed vs FP64 (LSRK4(5), stage update fused)",
        "config": {"workload": f"TGV Navier-Stokes {args.n}^3 hexes p={args.order} ({mesh.num_nodes} nodes), "
                               f"Ma=0.1 Re=1600, to t={args.t_final} in {nsteps} steps (dt={dt:.4e})"},
        "fp64_value": runs["fp64"]["value"],
        "mixed_value": runs["mixed"]["value"],
        "mixed_speedup": runs["mixed"]["value"] / runs["fp64"]["value"],
        "ke_dissipation_rate_max_rel_err": eps_err,
        "density_l2_rel_diff_at_t_final": rho_l2,
        "entropy_drift_rel": ent_drift,
        "history": {k: r["hist"] for k, r in runs.items()},
    }
    print(json.dumps(out), flush=True)


def adjoint_case(dev):
    """Gradients of J = integrated kinetic energy after 40 LSRK4(5) steps (200 stages) of the
    3D Euler TGV (16^3 hexes, p=4, Ma=0.1, + 64 seeded Fourier velocity modes of amplitude 1e-3)
    with respect to the Mach number and the 64 mode amplitudes, by the discrete adjoint with a
    checkpoint every 20 steps; checked against central finite differences (second order at
    eps = 1e-6 relative, and sixth order at a larger step); adjoint/forward cost ratio."""
    import torch

    from prandtl_amd import capi, cases
    from prandtl_amd import torch_adjoint as TA

    mesh = cases.tgv_mesh(n=16, order=4)
    modes = cases.fourier_modes(count=64, kmax=4, seed=4242)
    W = torch.from_numpy(mesh.quadrature_weights()).to(dev)
    op = capi.Operator(mesh)
    nsteps, ckpt = 40, 20
    dt = cases.cfl_dt(mesh, 11.0, 0.5)

    def J_of(mach, amps):
        u0 = TA.tgv_modes_state(mesh, mach, amps, modes)
        return TA.kinetic_energy(TA.RKSolve.apply(u0, op, dt, nsteps, ckpt), W)

    mach = torch.tensor(0.1, dtype=torch.float64, device=dev, requires_grad=True)
    amps = torch.full((64,), 1e-3, dtype=torch.float64, device=dev, requires_grad=True)
    # timing: plain forward against the adjoint sweep (taped forward before it, untimed), best of 3
    # after one warm-up sweep (the first sweep allocates the reverse sweep's stage buffers)
    with torch.no_grad():
        u0 = TA.tgv_modes_state(mesh, mach, amps, modes)
    uu = u0.clone()
    op.rk_run(capi.LSRK45, uu, 0.0, dt, 2)
    torch.autograd.grad(J_of(mach, amps), [mach, amps])
    torch.cuda.synchronize()
    fwd, adj = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        uu = u0.clone()
        op.rk_run(capi.LSRK45, uu, 0.0, dt, nsteps)
        torch.cuda.synchronize()
        fwd.append(time.perf_counter() - t0)
        J = J_of(mach, amps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gM, gA = torch.autograd.grad(J, [mach, amps])
        torch.cuda.synchronize()
        adj.append(time.perf_counter() - t1)
    t_fwd, t_adj = min(fwd), min(adj)
    # finite differences: the second-order central difference at eps = 1e-6 (relative to the
    # parameter), and the sixth-order Richardson combination at a step where its truncation and
    # rounding errors are both far below 1e-6 relative (the reference value of the check)
    checks = []
    with torch.no_grad():
        def fd_entry(name, g, f, x, h2, h6):
            fd2 = (f(x + h2) - f(x - h2)) / (2 * h2)
            d4, d42, fd6 = TA.central_fd(f, x, h6)
            checks.append({"param": name, "adjoint": g, "fd6": fd6, "fd4_h": d4, "fd4_2h": d42,
                           "fd2_eps1e-6": fd2, "rel_err": abs(g - fd6) / max(abs(fd6), 1e-300),
                           "rel_err_fd2": abs(g - fd2) / max(abs(fd2), 1e-300)})

        fd_entry("Ma", gM.item(), lambda x: J_of(x, amps).item(), mach, 1e-6 * 0.1, 1e-2 * 0.1)
        for i in (0, 17, 42, 63):
            e = torch.zeros_like(amps)
            e[i] = 1.0
            fd_entry(f"a_{i}", gA[i].item(), lambda x: J_of(mach, amps + (x - amps[i]) * e).item(), amps[i],
                     1e-6 * 1e-3, 1e-4)
    for c in checks:
        log(f"adjoint {c['param']}: {c['adjoint']:.12e} FD6 {c['fd6']:.12e} rel {c['rel_err']:.2e} "
            f"(FD2 eps=1e-6 {c['fd2_eps1e-6']:.12e}, rel {c['rel_err_fd2']:.2e})")
    # checkpoint schedules: the segments of the case (every 20 steps) against binomial (revolve)
    # checkpointing with a few snapshot buffers, on this case and on the NS TGV 64^3 p=3
    sched = [schedule_cost(op, u0, dt, nsteps, capi.TAPE_SEGMENTS, ckpt, t_fwd, mesh.num_nodes)]
    for snaps in (2, 4):
        sched.append(schedule_cost(op, u0, dt, nsteps, capi.TAPE_REVOLVE, snaps, t_fwd, mesh.num_nodes))
    for c in sched:
        log(f"adjoint schedule {c['schedule']}: {c['states']} states ({c['state_gb']:.2f} GB), "
            f"{c['recomputed_steps']} recomputed steps, adjoint/forward {c['adjoint_over_forward']:.2f}")
    op.close()
    del op
    ns = ns_adjoint_schedules(dev)
    out = {
        "metric": "adjoint/forward cost, 3D Euler TGV 16^3 p=4 (200 LSRK4(5) stages, checkpoint every 20 steps)",
        "config": {"workload": f"TGV Euler 16^3 hexes p=4 ({mesh.num_nodes} nodes), Ma=0.1, 64 Fourier modes "
                               f"(|k|<=4, amplitude 1e-3, seed 4242), J = integrated kinetic energy, dt={dt:.4e}"},
        "forward_s": t_fwd,
        "adjoint_s": t_adj,
        "adjoint_over_forward": t_adj / t_fwd,
        "timing": "best of 3 (plain forward run; reverse sweep incl. its segment recomputation) after one warm-up "
                  "sweep",
        "forward_dof_rhs_per_s": mesh.num_nodes * 5 * nsteps / t_fwd,
        "gradient_checks": checks,
        "max_rel_err": max(c["rel_err"] for c in checks),
        "schedules": sched,
        "ns_tgv64_schedules": ns,
    }
    print(json.dumps(out), flush=True)


def schedule_cost(op, u0, dt, nsteps, schedule, param, t_fwd, nodes, scheme=1):
    """One taped run with the given checkpoint schedule and its reverse sweep: the sweep's time over
    the plain forward run's, the device states the sweep holds and the steps it recomputes."""
    import torch

    from prandtl_amd import capi

    u = u0.clone()
    op.rk_run_tape_schedule(u, 0.0, dt, nsteps, schedule, param, scheme=scheme)
    g = torch.ones_like(u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.rk_adjoint(g)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    states, rec = op.tape_info()
    name = ("segments of %d steps" if schedule == capi.TAPE_SEGMENTS else "revolve, %d snapshots") % param
    return {"schedule": name, "states": states, "state_gb": states * u.numel() * 8 / 1e9,
            "recomputed_steps": rec, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd}


def ns_adjoint_schedules(dev, n=64, nsteps=10):
    """The NS TGV n^3 p=3 adjoint over nsteps LSRK4(5) steps: segment and revolve schedules."""
    import torch

    from prandtl_amd import capi, cases

    mesh = cases.tgv_mesh(n=n, order=3)
    op = capi.Operator(mesh, **cases.tgv_physics())
    u0 = torch.from_numpy(cases.tgv_state(mesh)).to(dev)
    dt = cases.cfl_dt(mesh, 11.0, 0.5)
    u = u0.clone()
    op.rk_run(capi.LSRK45, u, 0.0, dt, 1)
    u = u0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.rk_run(capi.LSRK45, u, 0.0, dt, nsteps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    res = [schedule_cost(op, u0, dt, nsteps, capi.TAPE_SEGMENTS, nsteps, t_fwd, mesh.num_nodes)]
    for snaps in (1, 3):
        res.append(schedule_cost(op, u0, dt, nsteps, capi.TAPE_REVOLVE, snaps, t_fwd, mesh.num_nodes))
    for c in res:
        log(f"NS TGV {n}^3 adjoint, {nsteps} steps, {c['schedule']}: {c['states']} states "
            f"({c['state_gb']:.1f} GB), {c['recomputed_steps']} recomputed steps, adjoint/forward "
            f"{c['adjoint_over_forward']:.2f}")
    op.close()
    return {"workload": f"TGV Navier-Stokes {n}^3 hexes p=3 ({mesh.num_nodes} nodes), {nsteps} LSRK4(5) steps",
            "forward_s": t_fwd, "schedules": res}


def khi_case(args, dev):
    """Kelvin-Helmholtz instability (2D Euler, [-1,1]^2, 256^2 quads p=3, 1,048,576 nodes;
    rho = 0.5 + 0.75 B, u = 0.5 (B - 1), v = 0.1 sin(2 pi x) + U(-1e-2, 1e-2) seeded 1234, p = 1,
    B = tanh((y+0.5)/0.1) - tanh((y-0.5)/0.1)), SSP-RK3 with the entropy filter and positivity
    limiter after every stage, CFL 0.5 with the reference's dt formula, lambda_max recomputed
    every 10 steps, to t = 3 (or --t-final)."""
    import numpy as np
    import torch

    from prandtl_amd import capi, cases

    n = 256 if args.n == 32 else args.n
    mesh = cases.khi_mesh(n=n, order=args.order)
    u0 = cases.khi_state(mesh)
    hooks = capi.HOOK_FILTER | capi.HOOK_LIMITER
    op = capi.Operator(mesh, stage_hooks=hooks)
    ud = torch.from_numpy(u0).to(dev)
    W = torch.from_numpy(mesh.quadrature_weights()).to(dev)
    t_final = args.t_final if args.t_final != 2.0 else 3.0
    t, steps, nf_tot, nl_tot = 0.0, 0, 0, 0
    op.hook_counts()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = t0
    while t < t_final - 1e-12:
        lam = op.max_char_speed(ud)
        dt = cases.cfl_dt(mesh, lam, 0.5)
        k = 10
        if t + k * dt > t_final:
            k = max(1, int(math.ceil((t_final - t) / dt)))
            dt = (t_final - t) / k
        op.rk_run(capi.SSPRK3, ud, t, dt, k, k)
        t += k * dt
        steps += k
        if time.perf_counter() - last > 20.0:
            torch.cuda.synchronize()
            last = time.perf_counter()
            log(f"KHI t={t:.3f} steps={steps} ({last - t0:.1f} s)")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nf, nl = op.hook_counts()
    u = ud.cpu().numpy()
    Wn = mesh.quadrature_weights()
    rho, vel, p = cases.conserv_to_prim(u)
    cons = [float(abs((Wn * (u[v] - u0[v])).sum()) / (Wn * np.abs(u0[v])).sum()) for v in range(4)]
    out = {
        "metric": "KHI 2D Euler with entropy filter + positivity limiter, SSP-RK3 to t=%g" % t_final,
        "config": {"workload": f"KHI [-1,1]^2 {n}^2 quads p={args.order} ({mesh.num_nodes} nodes), width 0.1, "
                               "noise 1e-2 seed 1234, CFL 0.5, lambda_max every 10 steps"},
        "steps": steps, "t_final": t, "wall_s": wall,
        "dof_rhs_per_s_incl_hooks": mesh.num_nodes * 3 * steps / wall,
        "elements_filtered": nf, "elements_limited": nl,
        "rho_min": float(rho.min()), "p_min": float(p.min()), "finite": bool(np.isfinite(u).all()),
        "conservation_rel": cons,
    }
    op.close()
    print(json.dumps(out), flush=True)


def blend_case(args, dev):
    """Subcell finite-volume blending (the reference's SUBCELL_FV_BLENDING build):
    (a) the default TGV 32^3 p=3 NS workload with blending on: every RHS evaluation computes the
        Persson-Peraire coefficients; on this smooth state alpha = 0 in every element, so the
        FV work is skipped and the figure is the coefficient kernel's overhead;
    (b) a periodic 2D Sod shock tube, Euler, [0,1] x [0,1/8], 512 x 64 quads p=3 (524,288 nodes):
        rho, p = 1, 1 for |x - 1/2| < 1/4 and 0.125, 0.1 outside, at rest; SSP-RK3 at CFL 0.3 to
        t = 0.1 with blending, the two shocks and rarefactions running outward.  Reported: the
        RHS rate with blending, the troubled-element fraction, positivity and conservation."""
    import numpy as np
    import torch

    from prandtl_amd import capi, cases

    out = {"metric": "DOF-RHS evals/s with subcell finite-volume blending (Persson-Peraire alpha every RHS)"}
    # (a) TGV with blending on
    mesh = cases.tgv_mesh(n=32, order=3)
    u0 = cases.tgv_state(mesh)
    ud = torch.from_numpy(u0).to(dev)
    lam = cases.max_wave_speed(u0)
    dt = cases.cfl_dt(mesh, lam, 0.5)
    rates = {}
    for blend in (False, True):
        op = capi.Operator(mesh, blending=blend, **cases.tgv_physics())
        uu = ud.clone()
        op.rk_run(capi.LSRK45, uu, 0.0, dt, 5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.rk_run(capi.LSRK45, uu, 0.0, dt, 50)
        torch.cuda.synchronize()
        rates[blend] = mesh.num_nodes * 5 * 50 / (time.perf_counter() - t0)
        if blend:
            a, _ = op.blend_coefficients(uu)
            out["tgv_alpha_positive_elements"] = int((a > 0).sum().item())
        op.close()
    out["tgv32_p3_ns_blend_off"] = rates[False]
    out["tgv32_p3_ns_blend_on"] = rates[True]
    out["tgv32_blend_overhead"] = rates[False] / rates[True] - 1.0
    log(f"TGV 32^3 NS: blending off {rates[False] / 1e9:.3f}, on {rates[True] / 1e9:.3f} G DOF-RHS/s")
    # (b) 2D Sod shock tube
    from prandtl_amd import mesh as M

    m = M.periodic_lattice(2, 3, (512, 64), lengths=[1.0, 0.125])
    X = m.node_coords()
    inside = np.abs(X[0] - 0.5) < 0.25
    rho = np.where(inside, 1.0, 0.125)
    p = np.where(inside, 1.0, 0.1)
    u0 = cases.prim_to_conserv(rho, [np.zeros_like(rho), np.zeros_like(rho)], p)
    op = capi.Operator(m, blending=True)
    ud = torch.from_numpy(u0).to(dev)
    t, steps, t_final = 0.0, 0, 0.1
    rhs_time = 0.0
    alpha_pos = []
    while t < t_final - 1e-12:
        lam = op.max_char_speed(ud)
        dt = cases.cfl_dt(m, lam, 0.3)
        k = 10
        if t + k * dt > t_final:
            k = max(1, int(math.ceil((t_final - t) / dt)))
            dt = (t_final - t) / k
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.rk_run(capi.SSPRK3, ud, t, dt, k, k)
        torch.cuda.synchronize()
        rhs_time += time.perf_counter() - t0
        t += k * dt
        steps += k
        a, _ = op.blend_coefficients(ud)
        alpha_pos.append(float((a > 0).float().mean().item()))
    u = ud.cpu().numpy()
    W = m.quadrature_weights()
    rho, vel, p = cases.conserv_to_prim(u)
    mscale = float((W * np.sqrt(u[1] ** 2 + u[2] ** 2)).sum())
    out["sod2d"] = {
        "workload": f"periodic 2D Sod tube [0,1]x[0,1/8], 512x64 quads p=3 ({m.num_nodes} nodes), SSP-RK3 CFL 0.3 "
                    f"to t={t_final}",
        "steps": steps, "dof_rhs_per_s": m.num_nodes * 3 * steps / rhs_time,
        "troubled_fraction_mean": float(np.mean(alpha_pos)), "troubled_fraction_max": float(np.max(alpha_pos)),
        "rho_min": float(rho.min()), "p_min": float(p.min()), "finite": bool(np.isfinite(u).all()),
        # change of total mass, momentum and energy over the run: mass and energy relative to their
        # totals, both momentum components relative to the final |momentum| total (the tube starts
        # at rest, and the y-momentum stays zero by symmetry)
        "conservation_rel": [float(abs((W * (u[v] - u0[v])).sum()) / scale) for v, scale in
                             enumerate([(W * u0[0]).sum(), mscale, mscale, (W * u0[3]).sum()])],
    }
    log(f"Sod 2D: {steps} steps, {out['sod2d']['dof_rhs_per_s'] / 1e9:.3f} G DOF-RHS/s, "
        f"troubled {out['sod2d']['troubled_fraction_mean']:.3f}")
    op.close()
    print(json.dumps(out), flush=True)


def cavity_case(args, dev):
    """Lid-driven cavity (the reference's TestCases/NavierStokes/2D/LidDrivenCavity): [0,2]^2,
    16 x 16 quads p=3, rho = 1 at rest, p = 1/(gamma Ma^2) with Ma = 0.1, mu = 0.02 (Re = 100 on
    the lid speed 1 and the side 2), Pr = 0.72, R = 287.05, no-slip adiabatic walls, the lid
    moving at (1, 0), blending on, SSP-RK3 at dt = 1e-4 to t = 3 (30,000 steps), then to the
    steady state (t = 40, dt = 2e-4) against Ghia, Ghia & Shin (1982) at Re = 100 -- and the same
    physics on 256 x 256 quads (1,048,576 nodes) for throughput, against the periodic mesh of
    that size."""
    import numpy as np
    import torch

    from prandtl_amd import capi, cases
    from prandtl_amd import mesh as M

    gamma, mach, mu, Pr, R = 1.4, 0.1, 0.02, 0.72, 287.05
    kw = dict(viscous=True, mu=mu, Pr=Pr, R_gas=R, gamma=gamma)
    walls = [capi.bc(capi.BC_NOSLIP_ADIABATIC, 0.0, 0.0, 0.0), capi.bc(capi.BC_NOSLIP_ADIABATIC, 1.0, 0.0, 0.0)]

    def make(n, periodic=False):
        if periodic:
            return M.periodic_lattice(2, 3, (n, n), lengths=[2.0, 2.0])
        return M.periodic_lattice(2, 3, (n, n), lengths=[2.0, 2.0], periodic=(False, False),
                                  boundary={(0, 0): 0, (0, 1): 0, (1, 0): 0, (1, 1): 1})

    def rest(m):
        one = np.ones((m.num_elements, m.nodes_per_element))
        return cases.prim_to_conserv(one, [0 * one, 0 * one], one / (gamma * mach * mach), gamma=gamma)

    out = {"metric": "lid-driven cavity, NS with no-slip adiabatic walls (boundary conditions on the GPU)"}
    # (a) the reference case
    m = make(16)
    u0 = rest(m)
    op = capi.Operator(m, bcs=walls, blending=True, **kw)
    ud = torch.from_numpy(u0).to(dev)
    dt, t_final = 1e-4, (args.t_final if args.t_final != 2.0 else 3.0)
    nsteps = int(round(t_final / dt))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < nsteps:
        k = min(1000, nsteps - done)
        op.rk_run(capi.SSPRK3, ud, done * dt, dt, k, k)
        done += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    u = ud.cpu().numpy()
    rho, vel, p = cases.conserv_to_prim(u, gamma=gamma)
    X = m.node_coords()
    W = m.quadrature_weights()
    centre = np.abs(X[0] - 1.0) < 1e-9          # the vertical centreline x = 1 (element faces)
    ys, us = X[1][centre], vel[0][centre]
    out["reference_case"] = {
        "workload": f"[0,2]^2 16x16 quads p=3 ({m.num_nodes} nodes), Ma 0.1, Re 100, blending on, "
                    f"SSP-RK3 dt=1e-4 to t={t_final}",
        "steps": nsteps, "wall_s": wall, "steps_per_s": nsteps / wall,
        "kinetic_energy": float((W * 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2)).sum()),
        "mass_change_rel": float(abs((W * (u[0] - u0[0])).sum()) / (W * u0[0]).sum()),
        "centreline_u_min": float(us.min()), "centreline_y_at_min": float(ys[np.argmin(us)]),
        "finite": bool(np.isfinite(u).all()),
    }
    # steady state (t = 40 at dt = 2e-4) against Ghia, Ghia & Shin (1982), Re = 100
    t_s, dt_s = 0.0, 2e-4
    uds = torch.from_numpy(u0).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(8):
        t_s = op.rk_run(capi.SSPRK3, uds, t_s, dt_s, 25000)
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    rho_s, vel_s, _ = cases.conserv_to_prim(uds.cpu().numpy(), gamma=gamma)
    uc = cases.lattice_line_values(m, vel_s[0], 1.0, 2.0 * cases.GHIA_RE100_Y)
    out["ghia_re100"] = {"t": t_s, "steps": 200000, "wall_s": wall_s,
                         "max_abs_error": float(np.abs(uc - cases.GHIA_RE100_U).max()),
                         "u_min": float(uc.min()), "u_min_ghia": float(cases.GHIA_RE100_U.min())}
    op.close()
    log(f"cavity 16x16: {nsteps} steps in {wall:.2f} s, centreline u_min {us.min():.4f} at y={ys[np.argmin(us)]:.3f}; "
        f"steady t={t_s:.0f}: max |u - Ghia| {out['ghia_re100']['max_abs_error']:.4f}")
    # (b) throughput at 256 x 256, boundaries against periodic
    for periodic in (False, True):
        m = make(256, periodic)
        op = capi.Operator(m, bcs=None if periodic else walls, **kw)
        ud = torch.from_numpy(rest(m)).to(dev)
        op.rk_run(capi.SSPRK3, ud, 0.0, 1e-5, 5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.rk_run(capi.SSPRK3, ud, 0.0, 1e-5, 100)
        torch.cuda.synchronize()
        rate = m.num_nodes * 3 * 100 / (time.perf_counter() - t0)
        out["cavity256_p3_ns" + ("_periodic" if periodic else "_walls") + "_dof_rhs_per_s"] = rate
        op.close()
    out["boundary_overhead"] = (out["cavity256_p3_ns_periodic_dof_rhs_per_s"] /
                                out["cavity256_p3_ns_walls_dof_rhs_per_s"] - 1)
    log(f"256^2 p=3 NS: walls {out['cavity256_p3_ns_walls_dof_rhs_per_s'] / 1e9:.3f}, periodic "
        f"{out['cavity256_p3_ns_periodic_dof_rhs_per_s'] / 1e9:.3f} G DOF-RHS/s")
    print(json.dumps(out), flush=True)


def geometry_case(args, dev):
    """General meshes (per-node metrics, oriented faces):
    (a) the default TGV 32^3 p=3 NS workload on a curved mesh -- every node of the periodic
        lattice of [0, 2 pi]^3 moved by x_c += 0.1 prod_d sin(x_d + phase_{c,d}), so every element
        is a curved cubic map with its own metrics at each node -- against the affine lattice:
        LSRK4(5) DOF-RHS/s and the per-kernel times of both;
    (b) the reference's forward-facing-step mesh (Gmsh, unstructured: faces meet in other
        directions and flipped), uniformly refined twice (54,400 quadrilaterals, p=3: 870,400
        nodes), Euler at Mach 3 with supersonic inflow / outflow and slip walls, blending on,
        SSP-RK3 at CFL 0.3 for 200 steps from the uniform stream."""
    import numpy as np
    import torch

    from prandtl_amd import capi, cases
    from prandtl_amd import mesh as M

    out = {"metric": "DOF-RHS evals/s on general meshes (per-node metrics, oriented faces)"}
    # (a) curved TGV
    lat = cases.tgv_mesh(n=32, order=3)
    X = lat.node_coords()
    Y = X.copy()
    for c in range(3):
        s = np.ones_like(X[0])
        for d in range(3):
            s = s * np.sin(X[d] + 0.7 * c + 0.3 * d)
        Y[c] = X[c] + 0.1 * s
    ng = M.node_geometry(Y, 4)
    curved = M.Mesh(3, 3, None, None, None, lat.neighbors, np.ascontiguousarray(ng[:, 0, :]), ng, None, Y,
                    lat.h_min() * 0.7)   # |grad of the displacement| <= 0.3
    phys = cases.tgv_physics()
    res = {}
    for name, m in (("affine", lat), ("curved", curved)):
        u0 = cases.tgv_state(m)
        op = capi.Operator(m, **phys)
        ud = torch.from_numpy(u0).to(dev)
        lam = op.max_char_speed(ud)
        dt = cases.cfl_dt(m, lam, 0.5)
        op.rk_run(capi.LSRK45, ud, 0.0, dt, 5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.rk_run(capi.LSRK45, ud, 0.0, dt, 50)
        torch.cuda.synchronize()
        rate = m.num_nodes * 5 * 50 / (time.perf_counter() - t0)
        lift_ms, rhs_ms = op.profile(ud, epilogue=1, iters=20)
        res[name] = {"dof_rhs_per_s": rate, "k_lift_ms": lift_ms, "k_rhs_node_lsrk_ms": rhs_ms,
                     "finite": bool(torch.isfinite(ud).all().item())}
        op.close()
        log(f"TGV 32^3 NS {name}: {rate / 1e9:.3f} G DOF-RHS/s, k_lift {lift_ms:.4f} ms, k_rhs {rhs_ms:.4f} ms")
    out["tgv32_p3_ns"] = res
    out["curved_cost"] = res["affine"]["dof_rhs_per_s"] / res["curved"]["dof_rhs_per_s"] - 1.0
    # (b) forward-facing step, refined twice
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "meshes", "ForwardFacingStep.msh")
    t0 = time.perf_counter()
    gm = M.gmsh_mesh(path, 3, {"left": 0, "right": 1, "top": 2, "bottom": 2}, refine=2)
    build_s = time.perf_counter() - t0
    a = math.sqrt(1.4)
    one = np.ones((gm.num_elements, gm.nodes_per_element))
    u0 = cases.prim_to_conserv(one, [3.0 * a * one, 0.0 * one], one)
    inflow = u0[:, 0, 0]
    bcs = [capi.bc(capi.BC_SUPERSONIC_INFLOW, *inflow), capi.bc(capi.BC_SUPERSONIC_OUTFLOW),
           capi.bc(capi.BC_SLIP_WALL)]
    op = capi.Operator(gm, bcs=bcs, blending=True)
    ud = torch.from_numpy(u0).to(dev)
    lam = op.max_char_speed(ud)
    dt = cases.cfl_dt(gm, lam, 0.3)
    op.rk_run(capi.SSPRK3, ud, 0.0, dt, 5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nsteps = 200
    op.rk_run(capi.SSPRK3, ud, 0.0, dt, nsteps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    al, _ = op.blend_coefficients(ud)
    u = ud.cpu().numpy()
    rho, vel, p = cases.conserv_to_prim(u)
    out["step_mach3"] = {
        "workload": f"forward-facing step (Gmsh, 3400 quads refined twice: {gm.num_elements} quads p=3, "
                    f"{gm.num_nodes} nodes), Euler Mach 3, supersonic in/outflow + slip walls, blending, "
                    f"SSP-RK3 CFL 0.3, {nsteps} steps, dt={dt:.3e}",
        "mesh_build_s": build_s,
        "dof_rhs_per_s": gm.num_nodes * 3 * nsteps / wall,
        "troubled_fraction": float((al > 0).float().mean().item()),
        "rho_min": float(rho.min()), "p_min": float(p.min()), "finite": bool(np.isfinite(u).all()),
    }
    op.close()
    log(f"step Mach 3: {out['step_mach3']['dof_rhs_per_s'] / 1e9:.3f} G DOF-RHS/s, troubled "
        f"{out['step_mach3']['troubled_fraction']:.3f}, p_min {out['step_mach3']['p_min']:.3f}")
    print(json.dumps(out), flush=True)


def vortex_case(dev, steps=100, reps=5):
    """The 2D isentropic vortex (16x16 quads, p=3, 4,096 nodes; Chandrashekar volume flux, Rusanov
    surface flux, SSP-RK3 at CFL 0.5, 100 steps): launch-bound, so the 100 steps are also captured
    once in a HIP graph (torch.cuda.CUDAGraph around prandtl_rk_run) and replayed.  Returns the
    DOF-RHS rates of the eager and the replayed run (median of `reps` timed runs each)."""
    import torch

    from prandtl_amd import capi, cases

    mesh = cases.vortex_mesh(n=16, order=3)
    u0 = torch.from_numpy(cases.vortex_state(mesh)).to(dev)
    op = capi.Operator(mesh, surface_flux=capi.SURF_RUSANOV)
    lam = op.max_char_speed(u0)
    dt = cases.cfl_dt(mesh, lam, 0.5)
    u = u0.clone()
    op.rk_run(capi.SSPRK3, u, 0.0, dt, 10)   # warm-up; allocates the stepper's registers
    eager = []
    for _ in range(reps):
        u.copy_(u0)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        op.rk_run(capi.SSPRK3, u, 0.0, dt, steps)
        b.record()
        torch.cuda.synchronize()
        eager.append(a.elapsed_time(b))
    ref = u.clone()
    ug = u0.clone()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            op.rk_run(capi.SSPRK3, ug, 0.0, dt, steps)
    torch.cuda.current_stream().wait_stream(cap)
    graph = []
    for _ in range(reps):
        ug.copy_(u0)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        graph.append(a.elapsed_time(b))
        assert torch.equal(ug, ref), "graph replay differs from the eager run"
    op.close()
    rhs = mesh.num_nodes * 3 * steps
    e_ms, g_ms = statistics.median(eager), statistics.median(graph)
    return {"workload": f"2D Euler isentropic vortex 16x16 quads p=3 ({mesh.num_nodes} nodes), Chandrashekar "
                        f"volume + Rusanov surface flux, SSP-RK3 CFL 0.5 (dt={dt:.4e}), {steps} steps",
            "eager_dof_rhs_per_s": rhs / (e_ms * 1e-3), "graph_dof_rhs_per_s": rhs / (g_ms * 1e-3),
            "eager_ms": e_ms, "graph_ms": g_ms, "launches_per_step": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--warmup-min-s", type=float, default=WARMUP_MIN_S,
                    help="warm-up time before each timed leg (0: exactly --warmup steps; for profiler runs)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--compare-mixed", action="store_true",
                    help="run the FP32-mixed TGV comparison (FP64 and mixed side by side to --t-final) "
                         "instead of the default benchmark")
    ap.add_argument("--t-final", type=float, default=2.0)
    ap.add_argument("--khi", action="store_true",
                    help="run the Kelvin-Helmholtz case (2D Euler 256^2 p=3, entropy filter + positivity limiter, "
                         "SSP-RK3 to t=3) instead of the default benchmark")
    ap.add_argument("--blend", action="store_true",
                    help="run the subcell-blending cases (TGV 32^3 NS with blending on; periodic 2D Sod tube) "
                         "instead of the default benchmark")
    ap.add_argument("--cavity", action="store_true",
                    help="run the lid-driven cavity (NS, no-slip walls; the reference's 16x16 case and a 256^2 "
                         "throughput run) instead of the default benchmark")
    ap.add_argument("--geometry", action="store_true",
                    help="run the general-mesh cases (TGV 32^3 NS on a curved mesh against the affine lattice; "
                         "the Gmsh forward-facing step refined twice, Mach 3 with blending) instead of the default "
                         "benchmark")
    ap.add_argument("--vortex", action="store_true",
                    help="run only the 2D vortex case (16x16 p=3, 100 SSP-RK3 steps, eager and replayed from a "
                         "HIP graph)")
    ap.add_argument("--adjoint", action="store_true",
                    help="run the discrete-adjoint case (TGV Euler 16^3 p=4 + 64 Fourier modes, 200 RK stages, "
                         "checkpoints every 20 steps) instead of the default benchmark")
    args = ap.parse_args()
    set_warmup_min(args.warmup_min_s)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (run with --gpus 1); "
              "the element partitioning it would use is described with its design notes.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from prandtl_amd import capi, cases

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.compare_mixed:
        compare_mixed(args, dev)
        return
    if args.adjoint:
        adjoint_case(dev)
        return
    if args.vortex:
        v = vortex_case(dev)
        print(json.dumps({"metric": "DOF-RHS evals/s, 2D Euler isentropic vortex (SSP-RK3, 3 RHS per step)",
                          "value": v["graph_dof_rhs_per_s"], "unit": "DOF-RHS evals/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64", "data": "synthetic (analytic Shu vortex)",
                          "config": {"workload": v["workload"]}, **v}), flush=True)
        return
    if args.khi:
        khi_case(args, dev)
        return
    if args.blend:
        blend_case(args, dev)
        return
    if args.cavity:
        cavity_case(args, dev)
        return
    if args.geometry:
        geometry_case(args, dev)
        return

    mesh = cases.tgv_mesh(n=args.n, order=args.order)
    phys = cases.tgv_physics()
    u0 = cases.tgv_state(mesh)
    nodes = mesh.num_nodes
    op = capi.Operator(mesh, **phys)
    ud = torch.from_numpy(u0).to(dev)
    lam = op.max_char_speed(ud)
    dt = cases.cfl_dt(mesh, lam, args.cfl)
    log(f"TGV {args.n}^3 p={args.order}: {nodes} nodes, lambda_max={lam:.4f}, dt={dt:.4e}")

    # ---- timed LSRK4(5) steps -------------------------------------------------
    warm_steps = warm_up(lambda n: op.rk_run(capi.LSRK45, ud, 0.0, dt, n), args.warmup)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t = op.rk_run(capi.LSRK45, ud, 0.0, dt, args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    assert torch.isfinite(ud).all().item(), "non-finite state after the timed steps"
    rhs_evals = 5 * args.steps
    value = nodes * rhs_evals / (ms / 1e3)
    log(f"{args.steps} LSRK4(5) steps: {ms:.2f} ms, {value / 1e9:.3f} G DOF-RHS/s")

    # ---- the dominant kernels' durations inside the same loop -------------------
    # every lifting and flux-differencing launch of a further run of the timed steps bracketed by
    # HIP events on the launch stream (prandtl_op_timing)
    op.timing(True)
    op.rk_run(capi.LSRK45, ud, t, dt, min(args.steps, 50))
    (loop_lift_ms, loop_rhs_ms), loop_counts = op.timing_read()
    op.timing(False)
    log(f"in-loop kernel durations: k_lift {loop_lift_ms:.4f} ms, k_rhs_node {loop_rhs_ms:.4f} ms "
        f"({loop_counts[1]} launches)")

    # ---- the same steps with the Rusanov surface flux -----------------------------
    opr = capi.Operator(mesh, surface_flux=capi.SURF_RUSANOV, **phys)
    ur = torch.from_numpy(u0).to(dev)
    warm_up(lambda n: opr.rk_run(capi.LSRK45, ur, 0.0, dt, n), args.warmup)
    e4 = torch.cuda.Event(enable_timing=True)
    e5 = torch.cuda.Event(enable_timing=True)
    e4.record(stream)
    opr.rk_run(capi.LSRK45, ur, 0.0, dt, args.steps)
    e5.record(stream)
    torch.cuda.synchronize()
    rus_ms = e4.elapsed_time(e5)
    rus_value = nodes * rhs_evals / (rus_ms / 1e3)
    log(f"Rusanov surface flux: {rus_ms:.2f} ms, {rus_value / 1e9:.3f} G DOF-RHS/s")
    opr.close()
    del ur

    # ---- the same steps in FP32-mixed precision ---------------------------------
    opm = capi.Operator(mesh, precision=capi.MIXED, **phys)
    um = torch.from_numpy(u0).to(dev)
    warm_up(lambda n: opm.rk_run(capi.LSRK45, um, 0.0, dt, n), args.warmup)
    e2 = torch.cuda.Event(enable_timing=True)
    e3 = torch.cuda.Event(enable_timing=True)
    e2.record(stream)
    opm.rk_run(capi.LSRK45, um, 0.0, dt, args.steps)
    e3.record(stream)
    torch.cuda.synchronize()
    mixed_ms = e2.elapsed_time(e3)
    mixed_value = nodes * rhs_evals / (mixed_ms / 1e3)
    log(f"mixed precision: {mixed_ms:.2f} ms, {mixed_value / 1e9:.3f} G DOF-RHS/s")
    opm.close()
    del um

    # ---- bare Mult (DGSEMOperator::Mult) ---------------------------------------
    du = torch.empty_like(ud)
    for _ in range(5):
        op.mult(ud, du)
    torch.cuda.synchronize()
    mult_ms = []
    for _ in range(20):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        op.mult(ud, du)
        b.record(stream)
        torch.cuda.synchronize()
        mult_ms.append(a.elapsed_time(b))
    mult_med = statistics.median(mult_ms)

    # ---- per-kernel timing back to back (HIP events on the launch stream) ---------
    lift_ms, rhs_ms = op.profile(ud, epilogue=1, iters=20)
    lift_mult_ms, rhs_mult_ms = op.profile(ud, epilogue=0, iters=20)
    # the roofline of the dominant kernel from its average duration inside the stepping loop
    dom = "k_rhs_node" if loop_rhs_ms >= loop_lift_ms else "k_lift"
    dom_ms = max(loop_rhs_ms, loop_lift_ms)
    dom_bytes = nodes * (BYTES_PER_NODE["rhs_lsrk"] if dom == "k_rhs_node" else BYTES_PER_NODE["lift"])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    # the whole stage against the two-pass minimum traffic of the RHS + stage update (360 B/node)
    stage_ms = ms / rhs_evals
    stage_achieved = STAGE_BYTES_PER_NODE * nodes / (stage_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": f"{dom} (LSRK4(5) stage epilogue)" if dom == "k_rhs_node" else dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": (RHS_LSRK_TRAFFIC_TGV32_P3 if (dom == "k_rhs_node" and args.n == 32 and args.order == 3) else None),
        "traffic_source": "rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, profiles/pmc_tgv32_p3_fp64_mixed.md",
        "duration_source": f"mean of {loop_counts[1]} launches inside a run of the timed LSRK4(5) steps, "
                           f"HIP events around each launch on the launch stream",
        "kernel_ms": {"k_lift": round(loop_lift_ms, 4), "k_rhs_lsrk": round(loop_rhs_ms, 4)},
        "kernel_ms_back_to_back": {"k_lift": round(lift_ms, 4), "k_rhs_lsrk": round(rhs_ms, 4),
                                   "k_rhs_mult": round(rhs_mult_ms, 4)},
        "bytes_per_node": BYTES_PER_NODE,
        "stage": {"achieved": round(stage_achieved, 1), "frac": round(stage_achieved / HBM_PEAK_GBS, 4),
                  "bytes_per_node": STAGE_BYTES_PER_NODE, "ms": round(stage_ms, 5),
                  "note": "one LSRK4(5) stage (lifting pass + flux pass with fused update) over its two-pass "
                          "minimum traffic"},
    }
    del du

    # ---- the small 2D case, eager and from a HIP graph ---------------------------
    vortex = vortex_case(dev)
    log(f"vortex 16x16 p=3: eager {vortex['eager_dof_rhs_per_s'] / 1e9:.3f}, graph "
        f"{vortex['graph_dof_rhs_per_s'] / 1e9:.3f} G DOF-RHS/s")

    # ---- CPU baseline ------------------------------------------------------------
    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU oracle baseline on {threads} threads ...")
        cpu = cpu_baseline(mesh, u0, phys, threads, budget_s=args.cpu_budget)
        log(f"CPU oracle: {cpu['value'] / 1e6:.3f} M DOF-RHS/s")

    out = {
        "metric": "DOF-RHS evals/s, 3D TGV p=3 NS FP64 (LSRK4(5) stages, stage update fused)",
        "value": value,
        "unit": "DOF-RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": warm_steps,
        "warmup_note": f"the W warm-up steps, then further chunks of W until {WARMUP_MIN_S} s of stepping "
                       "(the GPU's clock ramp); the timed region is exactly the K steps",
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Taylor-Green initial state)",
        "config": {
            "workload": f"TGV Navier-Stokes {args.n}^3 hexes p={args.order} ({nodes} nodes), Ma=0.1 Re=1600 "
                        f"Pr=0.71, EC volume flux + Chandrashekar surface flux, BR1, LSRK4(5), dt={dt:.4e} "
                        f"(CFL {args.cfl})",
            "nodes": nodes,
            "rhs_evals_per_step": 5,
            "parallelism": "single GPU",
        },
        "mixed_value": mixed_value,
        "mixed_speedup": mixed_value / value,
        "rusanov_value": rus_value,
        "vortex_2d": vortex,
        "mult_dof_rhs_per_s": nodes / (mult_med * 1e-3),
        "mult_ms": mult_med,
        "t_final": t,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    op.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
