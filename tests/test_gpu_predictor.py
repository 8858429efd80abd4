"""Warm-start predictor of the MD solves (capac_set_predictor).

The oracle runs the same MD with the same extrapolation formed in Python
from its own solutions (set_state(x=...) before each oracle step: linear
2 x1 - x2, quadratic 3 (x1 - x2) + x3, the order dropping to what is stored),
so both sides start every solve from the extrapolated point.  Both solve to
rTr < 1e-18 bTb: the predictor changes where a solve starts, not where it
stops, so at that tolerance the two trajectories agree far inside the MD
parity bar (test_gpu_parity.py::test_md_three_steps: positions within
1e-7 nm, velocities, temperature and the chain within 1e-5 relative), which
the test holds.  At the reference's tolerance the stored solutions of the two
sides differ at the solver tolerance (~1e-5 relative), the quadratic
extrapolation multiplies that by up to sqrt(19), and a start that lands the
two sides a few iterations apart moves a step's positions by ~1e-6 nm: a
difference of iteration paths, not of the predictor.  The iteration counts
are held in total, over the six steps, within 20% (plus two) of the
oracle's.

On au55_w the iterations to the same stop test fall, in the oracle, from 19
to 13 (linear) and 7 (quadratic) per step after the first steps; the GPU
must show the same fall.
"""
import os
import tempfile

import numpy as np
import pytest

from capacmd_amd import gen, lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def predict(xs, order):
    order = min(order, len(xs) - 1)
    if order == 1:
        return 2.0 * xs[-1] - xs[-2]
    if order == 2:
        return 3.0 * (xs[-1] - xs[-2]) + xs[-3]
    return None


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_predictor_md_matches_oracle(name, order, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    o = Oracle(deck)
    o.set_solve_crit(1e-18)
    o.force()
    o.md_begin()
    xs = [o.get_state()[2].copy()]
    its_g, its_o = [], []
    ctx = lib.Context(deck)
    assert ctx.predictor == 0
    ctx.set_predictor(order)
    ctx.set_solve_tolerance(1e-18)
    ctx.force()
    ctx.md_begin()
    for step in range(1, 7):
        xp = predict(xs, order)
        if xp is not None:
            o.set_state(x=xp)
        assert o.md_step(step, deck.n_save) == 0
        assert ctx.md_run(step, 1, deck.n_save) == 1
        ro, vo, xo, _ = o.get_state()
        rg, vg, xg = ctx.get_state()
        xs.append(xo.copy())
        it_o, it_g = o.solve_stats()[0], ctx.solve_stats()[0]
        obo, obg = o.observables(), ctx.observables()
        print(name, order, step, "iters", it_g, it_o, "r", np.abs(rg - ro).max(), "v", rel(vg, vo), flush=True)
        its_g.append(it_g)
        its_o.append(it_o)
        assert np.abs(rg - ro).max() < 1e-7
        assert rel(vg, vo) < 1e-5
        assert abs(obg["temp"] - obo["temp"]) < 1e-5 * obo["temp"]
        np.testing.assert_allclose(obg["vq"], obo["vq"], rtol=1e-5, atol=1e-12)
    assert abs(sum(its_g) - sum(its_o)) <= 0.2 * sum(its_o) + 2, (its_g, its_o)
    ctx.close()
    o.close()


@pytest.mark.parametrize("prec", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
def test_predictor_cuts_iterations(prec, tmp_path):
    """12 MD steps of au55_w at each order; mean iterations over steps 4-12.
    The FP32-refined path's GMRES runs with the Jacobi preconditioner here:
    its block preconditioner inverts au55's single 64-atom block outright,
    which leaves a handful of columns at every order (checked last)."""
    gen.write_deck(gen.small("au55_w"), str(tmp_path))
    deck = lib.Deck.from_dir(str(tmp_path))

    def run(order, block):
        ctx = lib.Context(deck, prec)
        ctx.set_predictor(order)
        if prec == lib.PREC_FP32_REFINED and not block:
            ctx.set_block_precond(0)
        ctx.force()
        ctx.md_begin()
        its = []
        for step in range(1, 13):
            assert ctx.md_run(step, 1, 1) == 1
            its.append(ctx.solve_stats()[0])
        ctx.close()
        print(prec, order, "block" if block else "jacobi", its, flush=True)
        return float(np.mean(its[3:]))

    mean = {order: run(order, False) for order in (0, 1, 2)}
    assert mean[1] < 0.85 * mean[0]
    assert mean[2] < 0.6 * mean[0]
    if prec == lib.PREC_FP32_REFINED:
        assert run(2, True) <= run(0, True)


def test_predictor_default_and_errors(golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, "au13_w"))
    c64 = lib.Context(deck, lib.PREC_FP64)
    c32 = lib.Context(deck, lib.PREC_FP32_REFINED)
    assert c64.predictor == 0 and c32.predictor == 2
    with pytest.raises(lib.CapacError, match="predictor order"):
        c64.set_predictor(3)
    c64.close()
    c32.close()
