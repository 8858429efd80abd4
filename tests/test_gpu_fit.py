"""Adjoint gradients of the force-matching loss (capac_fit_frame) against
central finite differences.

The frames are two states of a short FP64 MD run on a small golden deck, with
that run's metal forces and energy as targets.  theta is moved a few percent
off the deck's values, so the loss and its gradient are not zero.  Solves run
at rTr < 1e-22 bTb (the GPU) or exactly (the oracle, dense LU), so the
differences see the loss and not the solver tolerance.  Checks:
  - GPU adjoint gradient vs central differences of the oracle's loss (the FP64
    C restatement with exact solves, h = 1e-5 theta_k): the FP64 reference
    gradient;
  - GPU adjoint gradient vs central differences of the GPU's own loss;
  - the torch.autograd.Function returns that gradient through backward().
Bar, per component k:
  |g_adj,k - g_fd,k| <= 1e-6 max(|g_fd,k|, 1e-3 max_k' |g_fd,k'|).
"""
import os

import numpy as np
import pytest
import torch

from capacmd_amd import fit, lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

PERTURB = np.array([0.03, -0.02, 0.04, -0.03, 0.02, -0.04])
W_E = 1e-2
CRIT = 1e-22


def oracle_loss(deck, frames, theta, w_e):
    o = Oracle(deck)
    o.set_theta(theta)
    N, Nm = deck.n_atoms, deck.n_metal
    q = deck.array("charge", N)
    E_field = np.array(deck.p.efield[:])
    total = 0.0
    for fr in frames:
        r = fr.r
        b = o.vec_ext(r)
        o.relay(r)
        x = np.linalg.solve(o.relay_dense(), b)
        f, pot, _ = o.nonb(r)
        f2, p2, _, _ = o.qsc(r)
        f3, p3, _ = o.cpff_force(r, x)
        f, pot = f + f2 + f3, pot + p2 + p3
        R = r.reshape(3, N)
        for c in range(3):
            f[c * N:(c + 1) * N] += E_field[c] * q
        pot[9] -= np.sum(E_field[0] * q * R[0] + E_field[1] * q * R[1] + E_field[2] * q * R[2])
        fm = f.reshape(3, N)[:, :Nm].reshape(-1)
        e = float(np.sum(pot[1:]))
        total += float(np.sum((fm - fr.f_ref) ** 2)) + w_e * (e - fr.e_ref) ** 2
    return total


def check(g_adj, g_fd, what):
    floor = 1e-3 * np.abs(g_fd).max()
    err = np.abs(g_adj - g_fd) / np.maximum(np.abs(g_fd), floor)
    for k, name in enumerate(fit.THETA_NAMES):
        print(f"  {what} {name:12s} adjoint {g_adj[k]: .10e} fd {g_fd[k]: .10e} rel {err[k]:.2e}")
    assert err.max() <= 1e-6, (what, err)


@pytest.fixture(scope="module", params=["au13_w", "2au13_field"])
def setup(request, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, request.param))
    frames = fit.make_frames(deck, n_frames=2, stride=3)
    h = fit.Harness(deck, frames, w_e=W_E, solve_crit=CRIT)
    theta = h.theta0 * (1.0 + PERTURB)
    loss, g = h.loss_grad(theta)
    print(request.param, "theta0", h.theta0, "loss", loss, "adjoint iters", flush=True)
    yield request.param, deck, frames, h, theta, loss, g
    h.close()


def test_loss_matches_oracle(setup):
    name, deck, frames, h, theta, loss, _ = setup
    lo = oracle_loss(deck, frames, theta, W_E)
    print(name, "loss gpu", loss, "oracle", lo)
    assert abs(loss - lo) <= 1e-8 * abs(lo)


def test_gradient_vs_oracle_fd(setup):
    name, deck, frames, h, theta, _, g = setup
    g_fd = np.zeros(6)
    for k in range(6):
        hk = 1e-5 * theta[k]
        tp = theta.copy(); tp[k] += hk
        tm = theta.copy(); tm[k] -= hk
        g_fd[k] = (oracle_loss(deck, frames, tp, W_E) - oracle_loss(deck, frames, tm, W_E)) / (2 * hk)
    check(g, g_fd, name + " vs oracle FD")


def test_gradient_vs_gpu_fd(setup):
    name, deck, frames, h, theta, _, g = setup
    g_fd = np.zeros(6)
    for k in range(6):
        hk = 1e-5 * theta[k]
        tp = theta.copy(); tp[k] += hk
        tm = theta.copy(); tm[k] -= hk
        g_fd[k] = (h.loss_grad(tp, grad=False)[0] - h.loss_grad(tm, grad=False)[0]) / (2 * hk)
    check(g, g_fd, name + " vs GPU FD")


def test_autograd_function(setup):
    name, deck, frames, h, theta, loss, g = setup
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    out = fit.CpimFitLoss.apply(t, h)
    out.backward()
    assert abs(out.item() - loss) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(t.grad.numpy(), g, rtol=1e-12, atol=0)
