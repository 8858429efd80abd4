"""The host reader's side of general topologies (capacmd_amd/host/deck.c
against file.c:697-1000 and compute_dihedral_9, force.c:1236-1347): terms
expanded to every molecule with absolute indices and per-molecule ranges,
periodic dihedrals converted to Ryckaert-Bellemans coefficients, 1-4 pair
parameters resolved, and the reference's load-time errors."""
import math
import os

import numpy as np
import pytest

from capacmd_amd import lib

NAME = "topo_au13"


@pytest.fixture(scope="module")
def deck(golden_dir):
    return lib.Deck.from_dir(os.path.join(golden_dir, NAME))


def test_terms_expanded_per_molecule(deck):
    p = deck.p
    n_flx, n_w = 8, 40
    assert (p.n_bonds, p.n_pairs14, p.n_angles, p.n_dihedrals, p.n_vsites) == (
        5 * n_flx, 2 * n_flx, 4 * n_flx, 6 * n_flx, 2 * n_w)
    assert p.ndf == 3 * (deck.n_atoms - p.n_vsites) - 3 - 3 * n_w
    M = p.n_mols
    rng = deck.array("mol_terms", 5 * (M + 1), np.int64).reshape(5, M + 1)
    first = deck.array("mol_first", M, np.int64)
    last = deck.array("mol_last", M, np.int64)
    bonds = deck.array("bond_atoms", 2 * p.n_bonds, np.int64).reshape(-1, 2)
    vs = deck.array("vsite_atoms", 4 * p.n_vsites, np.int64).reshape(-1, 4)
    for m in range(M):
        for t in range(rng[0, m], rng[0, m + 1]):
            assert first[m] <= bonds[t].min() and bonds[t].max() <= last[m]
        for t in range(rng[4, m], rng[4, m + 1]):
            assert first[m] <= vs[t].min() and vs[t].max() <= last[m]
    # the first chain molecule's bonds, absolute: 0-1, 1-2 (Morse), 2-3, 3-4, 0-5
    a0 = deck.n_metal
    np.testing.assert_array_equal(bonds[:5], np.array([[0, 1], [1, 2], [2, 3], [3, 4], [0, 5]]) + a0)
    np.testing.assert_array_equal(deck.array("bond_funct", p.n_bonds, np.int64)[:5], [1, 3, 1, 1, 1])
    # sites of the first water: i, j, k = O, H1, H2; s = the lone pairs
    w0 = a0 + 6 * n_flx
    np.testing.assert_array_equal(vs[:2], [[w0, w0 + 1, w0 + 2, w0 + 3], [w0, w0 + 1, w0 + 2, w0 + 4]])
    assert deck.array("mass", deck.n_atoms)[w0 + 3] == 0.0


def rb_from_periodic(phi0, k, n):
    """k (1 + cos(n phi - phi0)) as a polynomial in cos psi, psi = phi - 180
    (cos phi = -cos psi), through the Chebyshev expansion of cos(n phi); the
    coefficients the reference's RB form takes, written out independently."""
    cheb = {1: [0, 1], 2: [-1, 0, 2], 3: [0, -3, 0, 4], 4: [1, 0, -8, 0, 8]}[n]
    sign = 1.0 if phi0 == 0.0 else -1.0              # cos(n phi - 180) = -cos(n phi)
    c = [0.0] * 6
    c[0] += k
    for m, a in enumerate(cheb):
        c[m] += sign * k * a * (-1.0) ** m          # (cos phi)^m = (-cos psi)^m
    return c


def test_periodic_dihedrals_as_rb(deck):
    """The converted coefficients give the same energy as k (1 + cos(n phi -
    phi0)) at every angle (funct 9 n=3, funct 1 n=2 phi0=180, funct 9 n=4,
    funct 4 n=1 phi0=180, funct 9 n=4 phi0=180)."""
    par = deck.array("dihedral_par", 6 * deck.p.n_dihedrals).reshape(-1, 6)
    cases = [(1, 0.0, 2.0, 3), (2, 180.0, 1.5, 2), (3, 0.0, 0.5, 4), (4, 180.0, 0.8, 1), (5, 180.0, 0.6, 4)]
    for row, phi0, k, n in cases:
        np.testing.assert_allclose(par[row], rb_from_periodic(phi0, k, n), rtol=0, atol=1e-12)
        for phi in np.linspace(-math.pi, math.pi, 13):
            cpsi = -math.cos(phi)
            e_rb = sum(par[row][m] * cpsi ** m for m in range(6))
            assert abs(e_rb - k * (1 + math.cos(n * phi - math.radians(phi0)))) < 1e-12


def test_pair14_parameters(deck):
    p = deck.p
    pp = deck.array("pair14_par", 6 * p.n_pairs14).reshape(-1, 6)
    a0 = deck.n_metal
    # funct 1 (0, 3): C-C table entry, the atoms' charges, SCALE_LJ_QQ
    c6, c12 = 4 * 0.3 * 0.35 ** 6, 4 * 0.3 * 0.35 ** 12
    np.testing.assert_allclose(pp[0], [c6, c12, -0.2, -0.2, 0.5, 0.83333333], rtol=1e-7)
    # funct 2 (5, 2): sigma 0.25, epsilon 0.1, q 0.1 / 0.1, scaleQQ 0.5, scaleLJ 1
    s6 = 0.25 ** 6
    np.testing.assert_allclose(pp[1], [0.4 * s6, 0.4 * s6 * s6, 0.1, 0.1, 1.0, 0.5], rtol=1e-15)
    np.testing.assert_array_equal(deck.array("pair14_atoms", 2 * p.n_pairs14, np.int64)[:4],
                                  [a0, a0 + 3, a0 + 5, a0 + 2])


def test_deck_rejects_what_the_reference_rejects(tmp_path, golden_dir):
    """The reference's reader errors for unknown term types, and
    compute_dihedral_9's for n and phi0, arrive as load errors."""
    src = os.path.join(golden_dir, NAME)
    par = open(os.path.join(src, "param.txt")).read()
    cases = [
        (par.replace("1 2 3 4 9  0.0  2.0  3", "1 2 3 4 9  0.0  2.0  5"), "n is not equal to 1, 2, 3 or 4"),
        (par.replace("1 2 3 4 9  0.0  2.0  3", "1 2 3 4 9  90.0  2.0  3"), "phi0 is not equal to 0 or pi"),
        (par.replace("0 1 2 1  111.0  460.0", "0 1 2 2  111.0  460.0"), "unsupported angle function type"),
        (par.replace("0 1 1  0.153  250000.0", "0 1 2  0.153  250000.0"), "unsupported bond function type"),
        (par.replace("5 2 2  0.5", "5 2 3  0.5"), "unsupported pair function type"),
    ]
    for n, (text, msg) in enumerate(cases):
        assert text != par
        d = tmp_path / ("case%d" % n)
        d.mkdir()
        for fn in ("init.gro", "mdset.txt"):
            (d / fn).write_text(open(os.path.join(src, fn)).read())
        (d / "param.txt").write_text(text)
        with pytest.raises(ValueError, match=msg):
            lib.Deck.from_dir(str(d))
