"""prandtl_rk_run under HIP graph capture (torch.cuda.CUDAGraph around the C-ABI call): the stepper
launches only kernels on the caller's stream (no allocation, no copy, no memset node once the
stepper's registers exist), so a captured run replays.  Each replay starts from the initial state
(restored by eager work between replays) and must equal the eager run bit for bit: the 2D vortex
(SSP-RK3, the launch-bound 16x16 case) and a small Navier-Stokes Taylor-Green run (LSRK4(5), lifting
and flux passes)."""
import pytest

from prandtl_amd import capi, cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _replays(op, scheme, u0, dt, steps, reps=3):
    import torch

    u = u0.clone()
    op.rk_run(scheme, u, 0.0, dt, 2)          # allocates the stepper's registers outside the capture
    u.copy_(u0)
    op.rk_run(scheme, u, 0.0, dt, steps)
    torch.cuda.synchronize()
    ref = u.clone()
    ug = u0.clone()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            op.rk_run(scheme, ug, 0.0, dt, steps)
    torch.cuda.current_stream().wait_stream(cap)
    for r in range(reps):
        ug.copy_(u0)
        g.replay()
        torch.cuda.synchronize()
        same = torch.equal(ug, ref)
        print(f"[graph] replay {r}: identical to the eager run: {same}", flush=True)
        assert same
        # eager work on the same operator between replays
        w = u0.clone()
        op.rk_run(scheme, w, 0.0, dt, 1)
        torch.cuda.synchronize()


def test_graph_vortex_sspk3(device):
    import torch

    _say("graph capture: vortex 16x16 p=3, SSP-RK3, 20 steps")
    m = cases.vortex_mesh(n=16, order=3)
    u0 = torch.from_numpy(cases.vortex_state(m)).to(device)
    op = capi.Operator(m, surface_flux=capi.SURF_RUSANOV)
    dt = cases.cfl_dt(m, op.max_char_speed(u0), 0.5)
    _replays(op, capi.SSPRK3, u0, dt, 20)
    op.close()


def test_graph_tgv_ns_lsrk(device):
    import torch

    _say("graph capture: TGV 4^3 p=3 NS, LSRK4(5), 3 steps")
    m = cases.tgv_mesh(n=4, order=3)
    u0 = torch.from_numpy(cases.tgv_state(m)).to(device)
    op = capi.Operator(m, **cases.tgv_physics())
    _replays(op, capi.LSRK45, u0, 1e-3, 3)
    op.close()
