"""A reference case run from its configuration on the GPU, with checkpoints in the reference's
format (configuration and checkpoint formats pinned on the CPU in tests/test_config_io.py).

* the lid-driven cavity configuration (Gmsh mesh, NS, no-slip walls, blending, SSP-RK3 at the
  configured dt) steps on the GPU as the oracle does (5 steps, 1e-11 relative);
* a run interrupted by a checkpoint (written, then read back into a fresh operator) continues
  bit for bit as the uninterrupted run;
* the forward-facing-step configuration (refined Gmsh mesh, Euler Mach 3, supersonic inflow and
  outflow, slip walls, blending) against the oracle for 3 steps.
"""
import os

import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, checkpoint, config

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = os.path.join(HERE, "golden", "configs")
MESHES = os.path.join(os.path.dirname(HERE), "data", "meshes")


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _oracle_phys(case):
    return oracle.phys(**{k: v for k, v in case.physics.items() if k in ("gamma", "Pr", "mu", "R_gas", "viscous")})


def test_cavity_from_configuration(device, tmp_path):
    import torch

    _say("lid-driven cavity from its configuration: oracle steps and checkpoint restart")
    case = config.load_case(os.path.join(CONFIGS, "LidDrivenCavity.json"), mesh_dir=MESHES)
    u0 = config.reference_initial_state(case)
    dt = case.dt
    ref, _ = oracle.rk_run_ex(case.mesh, _oracle_phys(case), u0, oracle_scheme(case), 5, dt,
                              [oracle.bc(t, *q) for t, q in case.bc_specs], blend=case.options["blending"])
    op = case.operator()
    d = torch.from_numpy(u0).to(device)
    op.rk_run(case.scheme, d, 0.0, dt, 5)
    out = d.cpu().numpy()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[config] cavity 5 steps: rel.err {rel:.3e}", flush=True)
    assert rel <= 1e-11
    # 40 steps straight, against 20 + checkpoint + 20 in a fresh operator
    d = torch.from_numpy(u0).to(device)
    op.rk_run(case.scheme, d, 0.0, dt, 40)
    straight = d.cpu().numpy()
    op.close()
    op1 = case.operator()
    d = torch.from_numpy(u0).to(device)
    t = op1.rk_run(case.scheme, d, 0.0, dt, 20)
    folder = str(tmp_path / "Checkpoints")
    checkpoint.save_checkpoint(folder, 20, t, d.cpu().numpy(), case.dim, case.order)
    op1.close()
    back, t1, cyc = checkpoint.load_checkpoint(folder, 20, case.dim, case.order, case.mesh.num_elements)
    assert cyc == 20 and abs(t1 - 20 * dt) < 1e-15
    op2 = case.operator()
    d = torch.from_numpy(back).to(device)
    op2.rk_run(case.scheme, d, t1, dt, 20)
    resumed = d.cpu().numpy()
    op2.close()
    assert np.array_equal(resumed, straight)


def oracle_scheme(case):
    assert case.scheme == config.SSPRK3
    return 0


def test_step_from_configuration(device):
    import torch

    _say("forward-facing step from its configuration: oracle steps")
    case = config.load_case(os.path.join(CONFIGS, "ForwardFacingStep.json"), mesh_dir=MESHES)
    u0 = config.reference_initial_state(case)
    X = case.mesh.node_coords()
    u0[0] *= 1.0 + 0.05 * np.sin(7.0 * X[0]) * np.cos(5.0 * X[1])      # a disturbance on the stream
    lam = cases.max_wave_speed(u0)
    dt = cases.cfl_dt(case.mesh, lam, 0.3)
    bcs = [oracle.bc(t, *q) for t, q in case.bc_specs]
    ref, _ = oracle.rk_run_ex(case.mesh, _oracle_phys(case), u0, 0, 3, dt, bcs, blend=True)
    op = case.operator()
    d = torch.from_numpy(u0).to(device)
    op.rk_run(case.scheme, d, 0.0, dt, 3)
    out = d.cpu().numpy()
    op.close()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[config] step 3 steps: rel.err {rel:.3e}", flush=True)
    assert rel <= 1e-11
