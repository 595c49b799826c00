"""The torch binding of the adjoint (CPU): RKSolve differentiates the run its own forward taped.

The operator handle keeps one tape (prandtl_rk_run_tape / prandtl_rk_adjoint), so a second taped
run on the same operator replaces the first.  RKSolve saves the tape identifier of its forward and
its backward refuses to run against another one.  A stand-in operator with the binding's methods
(an exact linear map u -> 2 u per run, adjoint g -> 2 g) exercises the check without a GPU.
"""
import pytest
import torch

from prandtl_amd import torch_adjoint as TA


class FakeOperator:
    def __init__(self):
        self._next = 0
        self._tape = 0
        self.t0 = []

    def rk_run_tape(self, u, t, dt, nsteps, ckpt_every, scheme=1):
        self._tape = 0
        self.t0.append(t)
        u.mul_(2.0)
        self._next += 1
        self._tape = self._next
        return t + dt * nsteps

    def tape_id(self):
        return self._tape

    def rk_adjoint(self, g):
        if self._tape == 0:
            raise RuntimeError("no taped run")
        g.mul_(2.0)


def test_backward_of_the_last_tape():
    op = FakeOperator()
    u0 = torch.ones(4, dtype=torch.float64, requires_grad=True)
    out = TA.RKSolve.apply(u0, op, 0.1, 3, 1, 1, 0.25)
    out.sum().backward()
    assert torch.equal(u0.grad, torch.full((4,), 2.0, dtype=torch.float64))
    assert op.t0 == [0.25]   # the start time reaches the library


def test_second_tape_refuses_the_first_backward():
    op = FakeOperator()
    u0 = torch.ones(4, dtype=torch.float64, requires_grad=True)
    a = TA.RKSolve.apply(u0, op, 0.1, 3, 1)
    b = TA.RKSolve.apply(u0 * 3.0, op, 0.1, 3, 1)
    with pytest.raises(TA.TapeMismatch):
        a.sum().backward()
    # the second trajectory's backward is still valid
    u1 = torch.ones(4, dtype=torch.float64, requires_grad=True)
    c = TA.RKSolve.apply(u1, op, 0.1, 3, 1)
    c.sum().backward()
    assert torch.equal(u1.grad, torch.full((4,), 2.0, dtype=torch.float64))
    del b


def test_failed_tape_is_no_tape():
    op = FakeOperator()
    op._tape = 0
    with pytest.raises(TA.TapeMismatch):
        TA.check_tape(op, 0)
