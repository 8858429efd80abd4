"""Force-matching fit of the CPIM and Sutton-Chen parameters on the GPU.

Frames are positions plus the checkpointed induced charges and dipoles
(vec_pq) of an FP64 MD run.  The targets are each frame's metal forces and
potential energy at the starting parameters.  The per-frame loss
  L = sum_{metal i} |F_i - F_ref,i|^2 + w_E (E - E_ref)^2
and its gradient with respect to theta = (alpha, c, R_q, R_p, eps_SC, c_SC)
come from capac_fit_frame: the adjoint method, run in libcapac_gpu.so with
the BiCGSTAB kernels and forward-mode derivative kernels.  torch reaches
them through CpimFitLoss, a torch.autograd.Function whose backward returns
the gradient its forward computed.  This module is the fitting harness; the
MD and force paths do not import it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import lib

THETA_NAMES = ("alpha", "capacitance", "R_q", "R_p", "sc_eps", "sc_c")


@dataclass
class Frame:
    r: np.ndarray          # 3N positions (SoA)
    x: np.ndarray          # n_mat vec_pq solved at r (warm start of the frame's solve)
    f_ref: np.ndarray      # 3 N_m metal forces (SoA)
    e_ref: float           # potential energy, sum of pot[1..14]


def make_frames(deck: lib.Deck, n_frames: int, stride: int, precision: int = lib.PREC_FP64) -> list:
    """Runs MD from the deck and takes a frame every `stride` steps.  The
    targets are the forces and energy of the run itself (the parameters the
    deck starts from)."""
    ctx = lib.Context(deck, precision=precision)
    nm, n = deck.n_metal, deck.n_atoms
    frames = []
    try:
        ctx.force()
        ctx.md_begin()
        step = 1
        for _ in range(n_frames):
            ctx.md_run(step, stride, 10 ** 9)
            step += stride
            r, _, x = ctx.get_state()
            f = ctx.forces().reshape(3, n)[:, :nm].reshape(-1).copy()
            e = float(np.sum(ctx.observables()["pot"][1:]))
            frames.append(Frame(r=r, x=x.copy(), f_ref=f, e_ref=e))
    finally:
        ctx.close()
    return frames


class Harness:
    """Loss and gradient over a list of frames on one device context."""

    def __init__(self, deck: lib.Deck, frames: list, w_e: float = 1e-2, solve_crit: float = 1e-8,
                 precision: int = lib.PREC_FP64):
        self.ctx = lib.Context(deck, precision=precision)
        self.ctx.set_solve_tolerance(solve_crit)
        self.frames = frames
        self.w_e = w_e
        self.theta0 = self.ctx.get_theta()

    def close(self):
        self.ctx.close()

    def loss_grad(self, theta, grad: bool = True):
        self.ctx.set_theta(np.asarray(theta, dtype=np.float64))
        total, g = 0.0, np.zeros(lib.NTHETA)
        for fr in self.frames:
            self.ctx.set_state(r=fr.r, x=fr.x)
            loss, gk = self.ctx.fit_frame(fr.f_ref, fr.e_ref, self.w_e, grad=grad)
            total += loss
            if grad:
                g += gk
        return total, (g if grad else None)


class CpimFitLoss(torch.autograd.Function):
    """loss = CpimFitLoss.apply(theta, harness): theta a float64 tensor of 6."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, harness: Harness):
        loss, g = harness.loss_grad(theta.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(g).to(theta.dtype))
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g.to(grad_out.device), None
