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


def main(argv=None) -> int:
    """Fitting workflow on the Au309 deck: frames from an FP64 run, the
    parameters moved off the deck's values, then L-BFGS on the GPU loss and
    its adjoint gradient.  Prints one JSON line."""
    import argparse
    import json
    import sys
    import tempfile
    import time

    from . import gen

    ap = argparse.ArgumentParser(description=main.__doc__.splitlines()[0])
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--stride", type=int, default=50)
    ap.add_argument("--steps", type=int, default=6, help="L-BFGS iterations")
    ap.add_argument("--solve-crit", type=float, default=1e-14)
    ap.add_argument("--w-e", type=float, default=1e-2)
    ap.add_argument("--fd-check", action="store_true",
                    help="central differences of the loss (h = 1e-5 theta_k) at the start")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731

    spec = gen.config(a.config)
    tmp = tempfile.mkdtemp(prefix="capac_fit_")
    gen.write_deck(spec, tmp)
    deck = lib.Deck.from_dir(tmp)
    log(f"deck {spec.name}: N={deck.n_atoms} Nm={deck.n_metal}")
    t0 = time.perf_counter()
    frames = make_frames(deck, a.frames, a.stride)
    t_frames = time.perf_counter() - t0
    log(f"{len(frames)} frames in {t_frames:.1f} s")
    h = Harness(deck, frames, w_e=a.w_e, solve_crit=a.solve_crit)
    perturb = np.array([0.03, -0.02, 0.04, -0.03, 0.02, -0.04])
    theta0 = h.theta0.copy()
    start = theta0 * (1.0 + perturb)
    t0 = time.perf_counter()
    loss0, g0 = h.loss_grad(start)
    t_eval = time.perf_counter() - t0
    out = {"workload": f"{spec.name}: {deck.n_metal} Au, {deck.n_atoms} atoms, {len(frames)} frames "
                       f"every {a.stride} steps", "solve_crit": a.solve_crit,
           "seconds_per_loss_and_gradient": round(t_eval, 4),
           "ms_per_frame": round(1e3 * t_eval / len(frames), 3),
           "loss_start": loss0, "theta_names": list(THETA_NAMES),
           "theta_reference": theta0.tolist(), "theta_start": start.tolist()}
    if a.fd_check:
        fd = np.zeros(len(start))
        for k in range(len(start)):
            hk = 1e-5 * start[k]
            tp, tm = start.copy(), start.copy()
            tp[k] += hk
            tm[k] -= hk
            fd[k] = (h.loss_grad(tp, grad=False)[0] - h.loss_grad(tm, grad=False)[0]) / (2 * hk)
        floor = 1e-3 * np.abs(fd).max()
        out["fd_check"] = {"adjoint": g0.tolist(), "central_difference": fd.tolist(),
                           "max_rel": float((np.abs(g0 - fd) / np.maximum(np.abs(fd), floor)).max())}
    # L-BFGS in scaled coordinates theta = theta_ref_scale * u
    scale = torch.tensor(start, dtype=torch.float64)
    u = torch.ones(len(start), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([u], lr=1.0, max_iter=a.steps, line_search_fn="strong_wolfe")
    history = []

    def closure():
        opt.zero_grad()
        loss = CpimFitLoss.apply(u * scale, h)
        loss.backward()
        history.append(loss.item())
        return loss

    t0 = time.perf_counter()
    opt.step(closure)
    t_fit = time.perf_counter() - t0
    theta = (u.detach() * scale).numpy()
    out.update({"lbfgs_evaluations": len(history), "loss_history": history, "fit_seconds": round(t_fit, 3),
                "theta_fit": theta.tolist(),
                "theta_rel_error_start": (np.abs(start - theta0) / theta0).tolist(),
                "theta_rel_error_fit": (np.abs(theta - theta0) / theta0).tolist()})
    h.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
