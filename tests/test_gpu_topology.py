"""General topologies on the device against the oracle: bonded terms (bonds,
1-4 pairs, angles, dihedrals), virtual sites, Buckingham and Morse pair
types, and the whole force call and MD step with them, on the topo_au13 deck
(tests/golden/topo_au13, capacmd_amd.gen.write_topology_deck: Au13 with CPIM
and a field, eight flexible six-atom chains carrying every bonded term type,
forty TIP5P-type waters with two virtual sites each).

The oracle reproduces the reference bit for bit on this deck
(tests/test_oracle.py), so it is the checker here.  Bars:
  - bonded terms, virtual-site construction and spreading: the same
    operations in the same order per molecule; relative 1e-12 (the device's
    acos, sin, exp and pow are its own library's);
  - nonbonded with Buckingham and Morse pairs: 1e-10, as the other pair tests;
  - the force call: 1e-10 against the oracle's phases at the GPU's solution;
  - MD: 1e-7 nm and 1e-5 relative in velocity over three steps, the bar of
    the other CPIM decks (the solution differs within the solver tolerance).
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NAME = "topo_au13"


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


@pytest.fixture(scope="module")
def deck(golden_dir):
    d = lib.Deck.from_dir(os.path.join(golden_dir, NAME))
    p = d.p
    print(f"deck {NAME}: N={d.n_atoms} Nm={d.n_metal} bonds={p.n_bonds} pairs14={p.n_pairs14} "
          f"angles={p.n_angles} dihedrals={p.n_dihedrals} vsites={p.n_vsites}", flush=True)
    return d


@pytest.fixture(scope="module")
def ref(golden_dir):
    return np.load(os.path.join(golden_dir, NAME, "ref_dump.npz"))


def test_bonded_phase(deck, ref):
    ctx = lib.Context(deck)
    try:
        f, pot, vir = ctx.bonded()
        print("bonded f rel", rel(f, ref["bonded_f"]))
        assert rel(f, ref["bonded_f"]) < 1e-12
        for k in (2, 3, 4, 7, 8):
            assert abs(pot[k] - ref["bonded_pot"][k]) <= 1e-12 * max(1.0, abs(ref["bonded_pot"][k])), k
        assert rel(vir, ref["bonded_vir"]) < 1e-12
    finally:
        ctx.close()


def test_vsite_build_and_spread(deck, ref):
    ctx = lib.Context(deck)
    try:
        r = ctx.vsite_build()
        print("vsite pos max diff", np.abs(r - ref["vsite_pos"]).max())
        # products and sums only, contraction off: the reference's bits
        assert np.array_equal(r, ref["vsite_pos"])
        # the spread uses the constructing atoms' positions, which the build
        # leaves as they were
        fs = ctx.vsite_spread(ref["nonb_f"])
        assert np.array_equal(fs, ref["vsite_f"])
    finally:
        ctx.close()


def test_nonbonded_with_buckingham_and_morse(deck, ref):
    ctx = lib.Context(deck)
    try:
        f, pot, vir = ctx.nonbonded()
        print("nonb f rel", rel(f, ref["nonb_f"]))
        assert rel(f, ref["nonb_f"]) < 1e-10
        for k in (7, 8):
            assert abs(pot[k] - ref["nonb_pot"][k]) <= 1e-10 * max(1.0, abs(ref["nonb_pot"][k])), k
        assert rel(vir, ref["nonb_vir"]) < 1e-10
    finally:
        ctx.close()


def test_force_call(deck):
    """capac_force with bonded terms and virtual sites: the reference's order
    (sites built, bonded, pairs, Sutton-Chen, CPIM, field, sites spread) at
    1e-10 against the oracle's phases at the GPU's solution."""
    ctx = lib.Context(deck)
    try:
        r0 = deck.r0
        ctx.set_state(r=r0, x=np.zeros(deck.n_mat))
        fg, pg, vg = ctx.force()
        rg, _, xg = ctx.get_state()
        o = Oracle(deck)
        r = o.vsite_build(r0)
        assert np.array_equal(rg, r)
        f, pot, vir = o.bonded(r)
        f2, p2, v2 = o.nonb(r)
        f3, p3, v3, _ = o.qsc(r)
        b = o.vec_ext(r)
        o.relay(r)
        A = o.relay_dense()
        res = A @ xg - b
        print("GPU solution relative residual", np.sqrt(res @ res / (b @ b)))
        assert res @ res < 1.2e-8 * (b @ b)
        f4, p4, v4 = o.cpff_force(r, xg)
        f = f + f2 + f3 + f4
        pot = pot + p2 + p3 + p4
        vir = vir + v2 + v3 + v4
        N = deck.n_atoms
        q = deck.array("charge", N)
        E = np.array(deck.p.efield[:])
        R = r.reshape(3, N)
        for c in range(3):
            f[c * N:(c + 1) * N] += E[c] * q
        pot[9] -= np.sum(E[0] * q * R[0] + E[1] * q * R[1] + E[2] * q * R[2])
        f = o.vsite_spread(r, f)
        print("force rel", rel(fg, f))
        assert rel(fg, f) < 1e-10
        for k in range(1, 15):
            assert abs(pg[k] - pot[k]) <= 1e-10 * max(1.0, abs(pot[k])), k
        assert rel(vg, vir) < 1e-10
        # the virtual sites keep no force
        vs = deck.array("vsite_atoms", 4 * deck.p.n_vsites, np.int64).reshape(-1, 4)[:, 3]
        assert not fg.reshape(3, N)[:, vs].any()
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
def test_md_three_steps(deck, precision):
    o = Oracle(deck)
    o.force()
    o.md_begin()
    ctx = lib.Context(deck, precision)
    try:
        ctx.set_predictor(0)            # the reference's warm start, as the oracle
        ctx.force()
        ctx.md_begin()
        for step in range(1, 4):
            assert o.md_step(step, deck.n_save) == 0
            assert ctx.md_run(step, 1, deck.n_save) == 1
            ro, vo, xo, fo = o.get_state()
            rg, vg, xg = ctx.get_state()
            obo, obg = o.observables(), ctx.observables()
            print(precision, step, "r", np.abs(rg - ro).max(), "v rel", rel(vg, vo), "T", obg["temp"],
                  obo["temp"], flush=True)
            assert np.abs(rg - ro).max() < 1e-7
            assert rel(vg, vo) < 1e-5
            assert abs(obg["temp"] - obo["temp"]) < 1e-5 * obo["temp"]
            for k in range(1, 15):
                assert abs(obg["pot"][k] - obo["pot"][k]) <= 1e-2 * max(1.0, abs(obo["pot"][k])), k
    finally:
        ctx.close()

