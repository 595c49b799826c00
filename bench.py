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

# Algorithmic HBM bytes per node per launch (compulsory traffic; face-neighbour trace
# re-reads excluded) -- FP64, 3D, 5 conserved variables:
#   k_lift : read u (40) + write F~v, 3 directions x 4 values (96)           = 136
#   k_rhs  : Mult epilogue: read u (40) + F~v (96) + write dudt (40)         = 176
#            LSRK epilogue: read u (40) + F~v (96) + k (40), write k, u (80)  = 256
BYTES_PER_NODE = {"lift": 136.0, "rhs_mult": 176.0, "rhs_lsrk": 256.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(mesh, u, phys, threads, budget_s=15.0):
    """The CPU oracle (a restatement of the reference's algorithm) on the host cores."""
    import oracle

    oracle.build()
    used = oracle.set_threads(threads)
    ph = oracle.phys(gamma=phys["gamma"], Pr=phys["Pr"], mu=phys["mu"], viscous=True)
    oracle.rhs(mesh, ph, u)  # warm-up
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        oracle.rhs(mesh, ph, u)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 20:
            break
    med = statistics.median(times)
    return {
        "value": mesh.num_nodes / med,
        "unit": "DOF-RHS evals/s",
        "cores": used,
        "kind": "port",
        "sample": f"{len(times)} oracle RHS evaluations of the same TGV {mesh.n[0]}^3 p={mesh.order} NS state "
                  f"(median {med:.3f} s each, {used} OpenMP threads, {budget_s:.0f} s budget)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1). "
              "The sharding design is described in DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from prandtl_amd import capi, cases

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

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
    op.rk_run(capi.LSRK45, ud, 0.0, dt, args.warmup)
    torch.cuda.synchronize()
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

    # ---- per-kernel timing (HIP events on the launch stream) -------------------
    lift_ms, rhs_ms = op.profile(ud, epilogue=1, iters=20)
    lift_mult_ms, rhs_mult_ms = op.profile(ud, epilogue=0, iters=20)
    dom = "k_rhs" if rhs_ms >= lift_ms else "k_lift"
    dom_ms = max(rhs_ms, lift_ms)
    dom_bytes = nodes * (BYTES_PER_NODE["rhs_lsrk"] if dom == "k_rhs" else BYTES_PER_NODE["lift"])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": f"{dom} (LSRK4(5) stage epilogue)" if dom == "k_rhs" else dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel_ms": {"k_lift": round(lift_ms, 4), "k_rhs_lsrk": round(rhs_ms, 4),
                      "k_rhs_mult": round(rhs_mult_ms, 4)},
        "bytes_per_node": BYTES_PER_NODE,
    }
    del du

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
