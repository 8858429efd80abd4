"""Energy minimisation on the device (capac_em_begin / capac_em_step /
capac_em_direction, Context.minimise) against the oracle's restatement of
the reference loop (mpi-main.c:706-865), which test_oracle.py pins bit for
bit to the reference program's own step lines and last frame.

Both sides solve to rTr < 1e-20 bTb, so the comparison does not see the
reference's solver tolerance (at its 1e-8 the energy after 18 steps moves by
5e-7 relative with the solver's iteration path; at 1e-20 and 1e-24 the
oracle agrees with itself to 1e-11).  Bars: FP64, E and F_max per step to
1e-9 relative, final positions to 1e-9 nm and solved charges and dipoles to
1e-7 of the largest; FP32-refined (stop test at its floor, 1e-11), E to 1e-7
relative, positions to 5e-6 nm.  The oracle's final positions move by
4.1e-5 nm between solves to 1e-8 and 1e-20; the residual norm at 1e-11 is
sqrt(1e-3) of that at 1e-8, which scales this to 1.3e-6 nm.
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib
from test_oracle import oracle_minimise

pytestmark = pytest.mark.gpu

CRIT = 1e-20


@pytest.fixture(scope="module", params=["au13_em", "2au13_cg"])
def reference(request, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, request.param))
    o, lines, converged = oracle_minimise(deck, CRIT)
    r, v, x, _ = o.get_state()
    yield request.param, deck, lines, converged, r, x
    o.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_minimiser_matches_oracle(reference, prec):
    name, deck, lines_o, conv_o, r_o, x_o = reference
    fp32 = prec == "fp32"
    ctx = lib.Context(deck, lib.PREC_FP32_REFINED if fp32 else lib.PREC_FP64)
    ctx.set_solve_tolerance(1e-11 if fp32 else CRIT)
    ctx.force()
    print(name, prec, "steps", deck.em_steps, flush=True)
    lines, conv = ctx.minimise(deck.run_type == "cg", deck.em_steps, deck.em_length, deck.em_tol)
    r, v, x = ctx.get_state()
    ctx.close()
    tol_e, tol_r, tol_x = (1e-7, 5e-6, 1e-5) if fp32 else (1e-9, 1e-9, 1e-7)
    assert conv == conv_o and len(lines) == len(lines_o)
    for (s1, f1, e1), (s2, f2, e2) in zip(lines, lines_o):
        assert s1 == s2
        assert abs(e1 - e2) <= tol_e * abs(e2), (s1, e1, e2)
        assert abs(f1 - f2) <= 10 * tol_e * abs(f2), (s1, f1, f2)
    print(name, prec, "max |dr|", np.abs(r - r_o).max(), "max |dx|", np.abs(x - x_o).max())
    assert np.abs(r - r_o).max() <= tol_r
    assert np.abs(x - x_o).max() <= tol_x * np.abs(x_o).max()
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
