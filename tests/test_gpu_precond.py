"""Block-Jacobi preconditioner of the FP32-refined path's GMRES
(capac_set_block_precond): the inverses of the relay's diagonal blocks of 64
metal atoms, formed on the device by Gauss-Jordan elimination in FP32.

Bars:
  - each stored inverse against the block of the dense FP32 relay (its own
    expansion, diagonal included): max |M A_bb - 1| < 1e-3 and within 1e-4
    (relative, Frobenius) of numpy's FP64 inverse of the same block; the
    identity rows of unknowns past Nm are exact;
  - the coarse space: the basis is orthonormal over each block's charges
    (FP32 entries, 1e-5), spans 1 and the atoms' x, y, z there, and the
    stored Ac^-1 inverts Z^T A Z formed here from the dense FP32 relay
    (max |Ac^-1 Ac - 1| < 1e-3);
  - the refresh schedule: with an interval of 2, the second assembly keeps
    the first one's inverses and the third forms them from its own
    positions, bitwise equal to a fresh context's at those positions;
  - solves: the reference's stop test on the FP64 residual, as with the
    Jacobi diagonal, in no more GMRES products than the Jacobi diagonal
    needs from the same start;
  - the knob's argument checks.
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

DECKS = ["au13_w", "2au13_field", "au147_vac"]


def rel(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a, float) - b) / (den if den > 0 else 1.0)


def block_unknowns(nm, b):
    """Dense-relay indices of block b's 256 unknowns (-1 past Nm)."""
    idx = np.full(256, -1, dtype=np.int64)
    for rho in range(256):
        atom = 64 * b + (rho // 3 if rho < 192 else rho - 192)
        if atom < nm:
            idx[rho] = 3 * atom + rho % 3 if rho < 192 else 3 * nm + atom
    return idx


def context(deck, layout, every):
    c = lib.Context(deck, precision=lib.PREC_FP32_REFINED)
    c.set_fp32_layout(layout)
    c.set_block_precond(every)
    return c


@pytest.mark.parametrize("layout", [lib.LAYOUT_ROWS, lib.LAYOUT_TILES], ids=["rows", "tiles"])
@pytest.mark.parametrize("name", DECKS)
def test_block_inverses(name, layout, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    nm = deck.n_metal
    ctx = context(deck, layout, 1)
    try:
        A, _ = ctx.cpim_assemble(dense=True)
        for b in range((nm + 63) // 64):
            M = ctx.block_precond_inverse(b).astype(np.float64)
            idx = block_unknowns(nm, b)
            ok = idx >= 0
            Abb = A[np.ix_(idx[ok], idx[ok])]
            Mv = M[np.ix_(ok, ok)]
            err = np.abs(Mv @ Abb - np.eye(ok.sum())).max()
            ref = np.linalg.inv(Abb)
            print(name, layout, "block", b, "unknowns", ok.sum(), "|MA - 1|", err, "rel to FP64 inverse",
                  rel(Mv, ref), flush=True)
            assert err < 1e-3
            assert rel(Mv, ref) < 1e-4
            # unknowns past Nm: identity rows and columns, exactly
            pad = ~ok
            np.testing.assert_array_equal(M[np.ix_(pad, pad)], np.eye(pad.sum()))
            assert not M[np.ix_(pad, ok)].any() and not M[np.ix_(ok, pad)].any()
    finally:
        ctx.close()


@pytest.mark.parametrize("name", DECKS)
def test_coarse_space(name, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    nm = deck.n_metal
    nb = (nm + 63) // 64
    ctx = context(deck, lib.LAYOUT_AUTO, 1)
    try:
        A, _ = ctx.cpim_assemble(dense=True)
        ainv, basis = ctx.block_precond_coarse()
    finally:
        ctx.close()
    nc = 4 * nb
    assert ainv.shape == (nc, nc)
    Z = np.zeros((4 * nm, nc))
    R = deck.r0.reshape(3, -1)[:, :nm]
    for b in range(nb):
        atoms = np.arange(64 * b, min(64 * b + 64, nm))
        W = basis[atoms].astype(np.float64)
        Z[3 * nm + atoms, 4 * b: 4 * b + 4] = W
        act = np.abs(W).sum(axis=0) > 0
        G = W[:, act].T @ W[:, act]
        np.testing.assert_allclose(G, np.eye(act.sum()), atol=1e-5)
        # the basis spans 1, x, y, z over the block (up to dropped directions)
        F = np.column_stack([np.ones(len(atoms)), (R[:, atoms] - R[:, atoms[:1]]).T])
        resid = F - W[:, act] @ (W[:, act].T @ F)
        assert np.abs(resid).max() <= 1e-4 * np.abs(F).max(), b
        for a in np.flatnonzero(~act):
            assert ainv[4 * b + a, 4 * b + a] == 1.0
    Ac = Z.T @ A @ Z
    act = np.abs(Z).sum(axis=0) > 0
    Ac[~act, ~act] = 1.0
    err = np.abs(ainv @ Ac - np.eye(nc)).max()
    print(name, "coarse unknowns", nc, "active", act.sum(), "|Ainv Ac - 1|", err, flush=True)
    assert err < 1e-3


def test_refresh_schedule(golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, "au147_vac"))
    rng = np.random.default_rng(3)
    r1 = deck.r0 + 2e-3 * rng.standard_normal(deck.r0.shape)
    ctx = context(deck, lib.LAYOUT_TILES, 2)
    fresh = context(deck, lib.LAYOUT_TILES, 2)
    try:
        ctx.cpim_assemble(dense=False)                      # forms them at r0
        m0 = ctx.block_precond_inverse(1)
        ctx.set_state(r=r1)
        ctx.cpim_assemble(dense=False)                      # kept
        np.testing.assert_array_equal(ctx.block_precond_inverse(1), m0)
        ctx.cpim_assemble(dense=False)                      # formed again, at r1
        m1 = ctx.block_precond_inverse(1)
        assert not np.array_equal(m1, m0)
        fresh.set_state(r=r1)
        fresh.cpim_assemble(dense=False)
        np.testing.assert_array_equal(m1, fresh.block_precond_inverse(1))
    finally:
        ctx.close()
        fresh.close()


@pytest.mark.parametrize("name", DECKS)
def test_solve_block_vs_jacobi(name, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    o = Oracle(deck)
    b = o.vec_ext(deck.r0)
    if not np.any(b):
        # a deck without external charges (au147_vac): a random right-hand side
        b = 1e-2 * np.random.default_rng(5).standard_normal(deck.n_mat)
    o.relay(deck.r0)
    A = o.relay_dense()
    x0 = np.zeros(deck.n_mat)
    its = {}
    for every in (0, 1):
        ctx = context(deck, lib.LAYOUT_AUTO, every)
        try:
            x, it, rtr, btb, kind = ctx.cpim_solve(b=b, x=x0)
        finally:
            ctx.close()
        res = b - A @ x
        print(name, "precond every", every, "iterations", it, "kind", kind, "true rTr/bTb",
              np.dot(res, res) / np.dot(b, b), flush=True)
        assert kind == 0
        assert np.dot(res, res) < 1e-8 * np.dot(b, b)
        its[every] = it
    assert its[1] <= its[0]


def test_knob_checks(golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, "au13_w"))
    c32 = lib.Context(deck, precision=lib.PREC_FP32_REFINED)
    c64 = lib.Context(deck)
    try:
        assert c32.block_precond == lib.BJ_EVERY_DEFAULT
        assert c64.block_precond == -1
        with pytest.raises(RuntimeError):
            c32.set_block_precond(-1)
        with pytest.raises(RuntimeError):
            c64.set_block_precond(4)
        c32.set_block_precond(0)
        with pytest.raises(RuntimeError):
            c32.block_precond_inverse(0)                    # none formed with the preconditioner off
        c32.set_block_precond(3)
        c32.cpim_assemble(dense=False)
        with pytest.raises(RuntimeError):
            c32.block_precond_inverse(1)                    # au13_w has one block
        assert c32.block_precond_inverse(0).shape == (256, 256)
    finally:
        c32.close()
        c64.close()
