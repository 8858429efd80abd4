"""Energy minimisation on the device (capac_em_begin / capac_em_step /
capac_em_direction, Context.minimise) against the oracle's restatement of
the reference loop (mpi-main.c:706-865), which test_oracle.py pins bit for
bit to the reference program's own step lines and last frame.

FP64: both sides solve to rTr < 1e-20 bTb, so the comparison does not see
the reference's solver tolerance (at its 1e-8 the oracle's own energy after
18 steps moves by 5e-7 relative and its positions by 4e-5 nm).  Bars: E and
F_max per step to 1e-9 relative, final positions to 1e-9 nm, solved charges
and dipoles to 1e-7 of the largest.

FP32-refined: its refinement floor is rTr ~ 1e-12 bTb, so it runs with the
stop test at 1e-11.  A minimiser trajectory amplifies any difference in the
solutions (a one-ulp change of the summation order moves the final
positions by microns of a nanometre), so the bar is calibrated on the oracle
itself: the FP32-refined trajectory stays within 10 times the distance
between the oracle's trajectories at stop tests 1e-11 and 1e-20 (E, F_max,
positions, charges and dipoles; measured 1.5 times on au13_em).
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib
from test_oracle import oracle_minimise

pytestmark = pytest.mark.gpu

CRIT = 1e-20
CRIT32 = 1e-11


@pytest.fixture(scope="module", params=["au13_em", "2au13_cg"])
def reference(request, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, request.param))
    o, lines, converged = oracle_minimise(deck, CRIT)
    r, v, x, _ = o.get_state()
    o.close()
    # the oracle's own distance between stop tests 1e-11 and 1e-20
    o11, lines11, _ = oracle_minimise(deck, CRIT32)
    r11, _, x11, _ = o11.get_state()
    o11.close()
    scale = dict(
        e=max(abs(a[2] - b[2]) / abs(b[2]) for a, b in zip(lines11, lines)),
        f=max(abs(a[1] - b[1]) / abs(b[1]) for a, b in zip(lines11, lines)),
        r=np.abs(r11 - r).max(), x=np.abs(x11 - x).max() / np.abs(x).max())
    print(request.param, "oracle 1e-11 vs 1e-20", scale, flush=True)
    yield request.param, deck, lines, converged, r, x, scale


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_minimiser_matches_oracle(reference, prec):
    name, deck, lines_o, conv_o, r_o, x_o, scale = reference
    fp32 = prec == "fp32"
    ctx = lib.Context(deck, lib.PREC_FP32_REFINED if fp32 else lib.PREC_FP64)
    ctx.set_solve_tolerance(CRIT32 if fp32 else CRIT)
    ctx.force()
    print(name, prec, "steps", deck.em_steps, flush=True)
    lines, conv = ctx.minimise(deck.run_type == "cg", deck.em_steps, deck.em_length, deck.em_tol)
    r, v, x = ctx.get_state()
    ctx.close()
    if fp32:
        tol_e, tol_f, tol_r, tol_x = (10 * scale[k] for k in ("e", "f", "r", "x"))
    else:
        tol_e, tol_f, tol_r, tol_x = 1e-9, 1e-9, 1e-9, 1e-7
    assert conv == conv_o and len(lines) == len(lines_o)
    de = max(abs(a[2] - b[2]) / abs(b[2]) for a, b in zip(lines, lines_o))
    df = max(abs(a[1] - b[1]) / abs(b[1]) for a, b in zip(lines, lines_o))
    dr = np.abs(r - r_o).max()
    dx = np.abs(x - x_o).max() / np.abs(x_o).max()
    print(name, prec, "E %.2e F_max %.2e r %.2e x %.2e" % (de, df, dr, dx))
    assert [a[0] for a in lines] == [b[0] for b in lines_o]
    assert de <= tol_e and df <= tol_f
    assert dr <= tol_r and dx <= tol_x
    assert not v.any()
    if deck.p.fix_metal == 1:
        nm = deck.n_metal
        r0 = deck.r0.reshape(3, -1)[:, :nm]
        np.testing.assert_array_equal(r.reshape(3, -1)[:, :nm], r0)


def test_minimiser_api_errors(golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, "au13_em"))
    ctx = lib.Context(deck)
    with pytest.raises(lib.CapacError, match="capac_em_begin"):
        ctx.em_step(1e-3)
    with pytest.raises(lib.CapacError, match="capac_em_begin"):
        ctx.em_direction(True)
    ctx.close()
