"""FP32-refined CPIM path (CAPAC_PREC_FP32_REFINED) against the FP64 oracle.

The relay is stored as 4 FP32 coefficients per metal pair, with its
T-tensors evaluated in FP32, in either layout: ROWS (every ordered pair) or
TILES (each unordered pair once, applied to both of its rows); every test
below runs on both.  GMRES runs on it with FP64 vectors and dots,
right-preconditioned with the reference's Jacobi diagonal or with the
block-Jacobi inverses of the relay's 64-atom diagonal blocks (the default);
the solve tests run with both.  FP64 iterative refinement then applies the
reference's stop test to the true residual, with the FP64 relay recomputed
on the fly.  Bars:
  - relay entries: max |A32 - A64| / max |A64| < 2e-6 (FP32 evaluation);
    the two layouts expand to bitwise the same matrix (same coefficients);
  - solve: the reference criterion rTr < 1e-8 bTb holds for the FP64
    residual b - A64 x (A64 from the oracle, dense); the distance to the
    exact FP64 solution is at most 10x the oracle's own; neutrality holds
    per nanoparticle;
  - everything downstream of the solve is FP64: the force call agrees within
    1e-10 with the oracle evaluated at the GPU's solution;
  - end to end: forces of the FP32-refined path against the oracle's own
    force call (with its own FP64 solution); the relative error is printed
    and bounded by 1e-6, or by 10x the FP64 GPU path's own end-to-end error
    where that is larger (both are then set by the solver tolerance, not by
    FP32);
  - MD: three steps within the FP64 path's metal-deck MD tolerance.
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

METAL_DECKS = ["au13_w", "2au13_field", "au147_vac"]


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


@pytest.fixture(scope="module")
def decks(golden_dir):
    out = {}
    for name in METAL_DECKS:
        d = lib.Deck.from_dir(os.path.join(golden_dir, name))
        print(f"deck {name}: N={d.n_atoms} Nm={d.n_metal} n_mat={d.n_mat}", flush=True)
        out[name] = d
    return out


LAYOUTS = {"rows": lib.LAYOUT_ROWS, "tiles": lib.LAYOUT_TILES}


PRECONDS = {"block": lib.BJ_EVERY_DEFAULT, "jacobi": 0}


def fp32_context(deck, layout, precond="block"):
    c = lib.Context(deck, precision=lib.PREC_FP32_REFINED)
    c.set_fp32_layout(LAYOUTS[layout])
    assert c.fp32_layout == LAYOUTS[layout]
    c.set_block_precond(PRECONDS[precond])
    assert c.block_precond == PRECONDS[precond]
    return c


@pytest.fixture(scope="module", params=[(lay, pc) for lay in LAYOUTS for pc in PRECONDS],
                ids=lambda p: f"{p[0]}-{p[1]}")
def ctxs(decks, request):
    out = {name: fp32_context(d, *request.param) for name, d in decks.items()}
    yield out
    for c in out.values():
        c.close()


@pytest.mark.parametrize("name", METAL_DECKS)
def test_layouts_same_relay(name, decks):
    """Both layouts hold the same FP32 coefficients: their dense expansions
    are bitwise equal.  Their products differ by FP64 summation order only,
    so two solves taken far below the reference's tolerance (rTr < 1e-18
    bTb) agree to round-off; at the reference's own tolerance two valid
    solves may differ by up to about 1e-5."""
    deck = decks[name]
    cr, ct = fp32_context(deck, "rows"), fp32_context(deck, "tiles")
    try:
        ar, _ = cr.cpim_assemble(dense=True)
        at, _ = ct.cpim_assemble(dense=True)
        np.testing.assert_array_equal(ar, at)
        o = Oracle(deck)
        b = o.vec_ext(deck.r0)
        x0 = np.zeros(deck.n_mat)
        for c in (cr, ct):
            c.set_solve_tolerance(1e-18)
        xr, itr, _, _, kr = cr.cpim_solve(b=b, x=x0)
        xt, itt, _, _, kt = ct.cpim_solve(b=b, x=x0)
        print(name, "iterations rows", itr, "tiles", itt, "x rel", rel(xt, xr))
        assert kr == kt
        assert rel(xt, xr) < 1e-8
    finally:
        cr.close()
        ct.close()


def check_refined_solution(deck, o, b, xg, kindg, xo, kindo):
    assert kindg == kindo
    if kindo != 0:
        # early exits (zero right-hand side, zero rho): x is left as it was
        np.testing.assert_array_equal(xg, xo)
        return
    A = o.relay_dense()
    res = b - A @ xg
    assert np.dot(res, res) < 1.0e-8 * np.dot(b, b)
    xe = np.linalg.solve(A, b)
    err_o = rel(xo, xe)
    err_g = rel(xg, xe)
    print("   x rel to exact: fp32-refined", err_g, "oracle", err_o)
    assert err_g <= max(10 * err_o, 1e-9)
    Nm = deck.n_metal
    q = xg[3 * Nm: 4 * Nm]
    first = deck.array("np_first", deck.p.n_nps, np.int64)
    last = deck.array("np_last", deck.p.n_nps, np.int64)
    qtot = deck.array("np_charge", deck.p.n_nps)
    for k in range(deck.p.n_nps):
        assert abs(q[first[k]: last[k] + 1].sum() - qtot[k]) < 1.1e-4 * np.sqrt(np.dot(b, b))


@pytest.mark.parametrize("name", METAL_DECKS)
def test_relay_fp32(name, decks, ctxs):
    deck, ctx = decks[name], ctxs[name]
    o = Oracle(deck)
    o.relay(deck.r0)
    dense_o = o.relay_dense()
    n4 = 4 * deck.n_metal
    ag, diag = ctx.cpim_assemble(dense=True)
    ref = dense_o[:n4, :n4]
    err = np.abs(ag - ref).max() / np.abs(ref).max()
    print(name, "fp32 relay max rel", err)
    assert err < 2e-6
    _, _, _, diag_o = o.relay(deck.r0)
    np.testing.assert_array_equal(diag[:n4], diag_o[:n4])


@pytest.mark.parametrize("name", METAL_DECKS)
def test_bicgstab_fp32_refined(name, decks, ctxs):
    deck, ctx = decks[name], ctxs[name]
    o = Oracle(deck)
    b = o.vec_ext(deck.r0)
    o.relay(deck.r0)
    x0 = np.zeros(deck.n_mat)
    xo, ito, _, _, kindo = o.solve(b, x0)
    ctx.set_state(r=deck.r0)
    xg, itg, rtrg, btbg, kindg = ctx.cpim_solve(b=b, x=x0)
    print(name, "iters fp32-refined", itg, "oracle", ito, "kind", kindg, kindo, "rTr", rtrg, "bTb", btbg)
    assert itg <= ito + max(5, ito // 5)
    check_refined_solution(deck, o, b, xg, kindg, xo, kindo)


def oracle_phases(deck, o, r, x):
    """Oracle force call with the CPIM force at the given solution x."""
    f, pot, vir = o.nonb(r)
    f2, p2, v2, _ = o.qsc(r)
    f3, p3, v3 = o.cpff_force(r, x)
    f, pot, vir = f + f2 + f3, pot + p2 + p3, vir + v2 + v3
    N = deck.n_atoms
    q = deck.array("charge", N)
    E = np.array(deck.p.efield[:])
    R = r.reshape(3, N)
    for c in range(3):
        f[c * N:(c + 1) * N] += E[c] * q
    pot[9] -= np.sum(E[0] * q * R[0] + E[1] * q * R[1] + E[2] * q * R[2])
    return f, pot, vir


@pytest.mark.parametrize("name", METAL_DECKS)
def test_force_call_fp32_refined(name, decks, ctxs):
    deck, ctx = decks[name], ctxs[name]
    r = deck.r0
    ctx.set_state(r=r, x=np.zeros(deck.n_mat))
    fg, pg, vg = ctx.force()
    _, _, xg = ctx.get_state()
    o = Oracle(deck)
    b = o.vec_ext(r)
    o.relay(r)
    xo, _, _, _, kindo = o.solve(b, np.zeros(deck.n_mat))
    check_refined_solution(deck, o, b, xg, ctx.solve_stats()[3], xo, kindo)
    f, pot, vir = oracle_phases(deck, o, r, xg)
    print(name, "force rel at the GPU x", rel(fg, f))
    assert rel(fg, f) < 1e-10
    for k in range(1, 15):
        assert abs(pg[k] - pot[k]) <= 1e-10 * max(1.0, abs(pot[k])), k
    assert rel(vg, vir) < 1e-10
    # end to end against the oracle's own solution, beside the FP64 GPU path
    fo, poto, _ = oracle_phases(deck, o, r, xo)
    c64 = lib.Context(deck)
    try:
        f64, p64, _ = c64.force()
    finally:
        c64.close()
    e32, e64 = rel(fg, fo), rel(f64, fo)
    print(name, "end-to-end force rel: fp32-refined", e32, "fp64", e64,
          "| pot[0] terms fp32", np.sum(pg[1:]), "oracle", np.sum(poto[1:]))
    assert e32 <= max(1e-6, 10 * e64)


MD_TOL = {"au13_w": (1e-7, 1e-5), "2au13_field": (1e-7, 1e-5)}


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_md_three_steps_fp32_refined(name, layout, decks):
    deck = decks[name]
    tol_r, tol_v = MD_TOL[name]
    o = Oracle(deck)
    o.force()
    o.md_begin()
    ctx = fp32_context(deck, layout)
    try:
        ctx.force()
        ctx.md_begin()
        for step in range(1, 4):
            assert o.md_step(step, deck.n_save) == 0
            assert ctx.md_run(step, 1, deck.n_save) == 1
            ro, vo, _, _ = o.get_state()
            rg, vg, _ = ctx.get_state()
            print(name, step, "r", np.abs(rg - ro).max(), "v rel", rel(vg, vo))
            assert np.abs(rg - ro).max() < tol_r
            assert rel(vg, vo) < tol_v
    finally:
        ctx.close()
