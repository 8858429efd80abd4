"""The fitting harness's host side, without a GPU: CpimFitLoss hands the
gradient its forward computed to torch's backward (scaled by the incoming
gradient), and the parameter layout matches the C ABI."""
import numpy as np
import torch

from capacmd_amd import fit, lib


class QuadraticHarness:
    """Stands in for fit.Harness: L = sum w_k theta_k^2, dL/dtheta = 2 w theta."""

    def __init__(self):
        self.w = np.arange(1.0, 7.0)
        self.calls = 0

    def loss_grad(self, theta):
        self.calls += 1
        theta = np.asarray(theta)
        return float(np.sum(self.w * theta ** 2)), 2.0 * self.w * theta


def test_autograd_function_passes_the_gradient():
    h = QuadraticHarness()
    theta = torch.tensor([10.1, 3.8, 3.0, 1.4, 1.2, 34.4], dtype=torch.float64, requires_grad=True)
    loss = fit.CpimFitLoss.apply(theta, h)
    (3.0 * loss).backward()
    expect = 3.0 * 2.0 * h.w * theta.detach().numpy()
    np.testing.assert_allclose(theta.grad.numpy(), expect, rtol=1e-15)
    assert h.calls == 1                      # one evaluation gives loss and gradient


def test_lbfgs_on_the_function_converges():
    h = QuadraticHarness()
    u = torch.tensor(np.full(6, 0.7), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([u], max_iter=50, line_search_fn="strong_wolfe", tolerance_grad=1e-14,
                            tolerance_change=1e-20)

    def closure():
        opt.zero_grad()
        loss = fit.CpimFitLoss.apply(u, h)
        loss.backward()
        return loss

    opt.step(closure)
    assert np.abs(u.detach().numpy()).max() < 1e-6


def test_parameter_layout():
    assert lib.NTHETA == len(fit.THETA_NAMES) == 6
    src = open(lib.__file__.replace("lib.py", "../include/capac_gpu.h")).read()
    assert "#define CAPAC_NTHETA 6" in src
