"""The GPU operator against the exact discrete operator, at the headline size.

The exact discrete operator of the double-precision inputs is the oracle evaluated in extended
precision (oracle.rhs_truth: long double throughout, 64-bit significand).  On the Taylor-Green
state at Ma = 0.1 (p ~ 71 against an O(1) right-hand side) the double sums cancel by three to
five digits, so no double evaluation reaches 1e-12 relative there: the double oracle's own error
against the exact operator is printed beside the GPU's, and the GPU's is held to at most twice the
oracle's.  On a well-conditioned full-size state (Ma = 0.3 with 1 % seeded jitter) the literal
bar holds: ||gpu - oracle||_inf / ||oracle||_inf <= 1e-12 and the same against the exact operator.
Each case prints a "[parity-summary]" JSON line with the measured errors.
"""
import json

import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _gpu_rhs(mesh, u, phys):
    import torch

    op = capi.Operator(mesh, **phys)
    d = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    du = torch.empty_like(d)
    op.mult(d, du)
    torch.cuda.synchronize()
    out = du.cpu().numpy()
    op.close()
    return out


def _errors(mesh, u, phys):
    oph = oracle.phys(gamma=phys.get("gamma", 1.4), Pr=phys.get("Pr", 0.72), mu=phys.get("mu", 0.0),
                      viscous=phys.get("viscous", False))
    gpu = _gpu_rhs(mesh, u, phys)
    orc, _ = oracle.rhs(mesh, oph, u)
    orc_fma, _ = oracle.rhs(mesh, oph, u, fma=True)
    hi, lo, _ = oracle.rhs_truth(mesh, oph, u)
    return dict(gpu_vs_truth=oracle.truth_error(gpu, hi, lo), oracle_vs_truth=oracle.truth_error(orc, hi, lo),
                oracle_fma_vs_truth=oracle.truth_error(orc_fma, hi, lo),
                gpu_vs_oracle=float(np.abs(gpu - orc).max() / np.abs(orc).max()))


@pytest.mark.parametrize("n", [6, 16, 32])
def test_tgv_against_the_exact_operator(device, n):
    _say(f"TGV {n}^3 p=3 NS Ma=0.1 against the extended-precision operator")
    m = cases.tgv_mesh(n=n, order=3)
    u = cases.tgv_state(m)
    e = _errors(m, u, cases.tgv_physics())
    print("[parity-summary] " + json.dumps(dict(case=f"TGV {n}^3 p=3 NS Ma=0.1 (bench state)", **e)), flush=True)
    assert e["gpu_vs_truth"] <= 2.0 * e["oracle_vs_truth"], e


def test_well_conditioned_full_size(device):
    _say("TGV 32^3 p=3 NS Ma=0.3 with 1% jitter: the literal 1e-12 bar")
    m = cases.tgv_mesh(n=32, order=3)
    rng = np.random.default_rng(5)
    u = cases.tgv_state(m, mach=0.3)
    u = u * (1.0 + 1e-2 * rng.uniform(-1, 1, u.shape))
    e = _errors(m, u, cases.tgv_physics(mach=0.3))
    print("[parity-summary] " + json.dumps(dict(case="TGV 32^3 p=3 NS Ma=0.3, 1% jitter (seed 5)", **e)), flush=True)
    assert e["gpu_vs_oracle"] <= 1e-12, e
    assert e["gpu_vs_truth"] <= 1e-12, e


def _log_mean_branches(mesh, u):
    """Fractions of the volume pairs' density logarithmic means, (a, a+k mod Np) along every line of
    every element, with u = ((y-x)/(y+x))^2 below 1e-4, in [1e-4, 0.05) and from 0.05 up."""
    n = mesh.order + 1
    rho = np.asarray(u[0]).reshape((-1,) + (n,) * mesh.dim)
    us = []
    for ax in range(1, mesh.dim + 1):
        for k in range(1, n):
            b = np.roll(rho, -k, axis=ax)
            us.append((((b - rho) / (b + rho)) ** 2).ravel())
    us = np.concatenate(us)
    return [float((us < 1e-4).mean()), float(((us >= 1e-4) & (us < 0.05)).mean()), float((us >= 0.05).mean())]


@pytest.mark.parametrize("case", ["khi2d", "tgv3d"])
def test_log_mean_ranges_against_the_exact_operator(device, case):
    """States whose node pairs put the logarithmic means on all three of the GPU's branches (the
    short series, the series to u^11 up to u = 0.05, the closed form ln(y/x)/(y-x) beyond), where
    the oracle keeps the closed form from u = 1e-4 on: the GPU against the extended-precision
    operator, at most twice the double oracle's error, and both within 1e-12."""
    rng = np.random.default_rng(11)
    if case == "khi2d":
        _say("KHI 16^2 p=3 Euler, 10% jitter: every log-mean branch, against the extended-precision operator")
        m = cases.khi_mesh(n=16, order=3)
        u = cases.khi_state(m)
        u = u * (1.0 + 0.1 * rng.uniform(-1, 1, u.shape))
        phys = dict(gamma=1.4, viscous=False)
    else:
        _say("TGV 6^3 p=3 NS Ma=0.3, 30% jitter: every log-mean branch, against the extended-precision operator")
        m = cases.tgv_mesh(n=6, order=3)
        u = cases.tgv_state(m, mach=0.3)
        u = u * (1.0 + 0.3 * rng.uniform(-1, 1, u.shape))
        phys = cases.tgv_physics(mach=0.3)
    fr = _log_mean_branches(m, u)
    assert min(fr) > 0.01, fr
    e = _errors(m, u, phys)
    print("[parity-summary] " + json.dumps(dict(case=f"{case}: log-mean branch fractions {fr}", **e)), flush=True)
    assert e["gpu_vs_truth"] <= 2.0 * e["oracle_vs_truth"], e
    assert e["gpu_vs_truth"] <= 1e-12 and e["gpu_vs_oracle"] <= 1e-12, e
