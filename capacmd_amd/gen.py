"""Seeded generator of CapacMD input decks (init.gro, param.txt, mdset.txt).

Systems are gold Mackay icosahedra solvated in rigid SPC/E water.  The files
follow the grammar the reference reads:
  * mdset.txt  key/value lines             (reference file.c:74-130)
  * param.txt  sectioned force field       (reference file.c:286-562, 569-1023)
  * init.gro   title, count, %5d%-5s%5s%5d + free-format numbers from column 20,
               box line                    (reference file.c:138-203)
Metal atoms come first (the reference's metal-metal CPIM force loop assumes
the metal block starts at atom 0, cpff.c:762-764).  Atom and residue numbers
wrap modulo 100000 so that the fixed-width columns stay aligned for systems
of more than 99,999 atoms (the %5d fields of file.c:188-191,221).

Units are the reference's MD units (nm, ps, kJ/mol, e, amu); CPIM parameters
are in atomic units (bohr^3 and bohr), as param.txt expects.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Sequence, Tuple

import numpy as np

EV2KJMOL = 96.4853363969351          # kJ/mol per eV (reference typedef.h:70)
NM2BOHR = 18.8972612456506           # bohr per nm (reference typedef.h:73)
K_BOLTZ = 0.00831446214546895        # kJ/mol/K (reference typedef.h:77)

# SPC/E water (Berendsen et al. 1987; GROMACS spce.itp values)
SPCE_QO = -0.8476
SPCE_QH = 0.4238
SPCE_ROH = 0.1
SPCE_HOH_DEG = 109.47
SPCE_C6 = 2.6171e-3
SPCE_C12 = 2.6331e-6
M_O = 15.9994
M_H = 1.008
M_AU = 196.97

# Au-O / Au-H erf-damped vdW (funct 4), the only metal-water pair terms that
# ship with the reference (example_01/param.txt:17-18).
AU_O_ERFVDW = (0.002522219544, 3.624859037184e-06, 2.25)
AU_H_ERFVDW = (0.001028707002, 1.45541995368e-06, 2.25)

# Sutton-Chen gold: n=10, m=8, eps=1.2793e-2 eV, c=34.408, a=4.08 A
SC_N, SC_M = 10, 8
SC_EPS = 1.2793e-2 * EV2KJMOL
SC_C = 34.408
SC_A = 0.408

# CPIM gold: polarizability 1.5 A^3, capacitance 2.0 A, in atomic units
CPIM_POLAR = 1.5 * (NM2BOHR / 10.0) ** 3
CPIM_CAPAC = 2.0 * (NM2BOHR / 10.0)

AU_NN = 0.288          # nm, nearest-neighbour distance along icosahedron edges
CONTACTS = (0.18, 0.16)  # nm, smallest intermolecular H-H and O-H distances of the benchmark decks
WATER_DENSITY = 1.0    # g/cm^3
AVOGADRO = 6.02214129e23


def icosahedron_vertices() -> np.ndarray:
    """Unit-edge icosahedron vertices (12 x 3) centred at the origin."""
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = []
    for s1 in (-1.0, 1.0):
        for s2 in (-1.0, 1.0):
            v.append((0.0, s1, s2 * phi))
            v.append((s1, s2 * phi, 0.0))
            v.append((s2 * phi, 0.0, s1))
    v = np.array(v) / 2.0   # edge length 1
    return v


def icosahedron_faces(v: np.ndarray) -> List[Tuple[int, int, int]]:
    n = len(v)
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=-1)
    faces = []
    for i in range(n):
        for j in range(i + 1, n):
            if abs(d[i, j] - 1.0) > 1e-9:
                continue
            for k in range(j + 1, n):
                if abs(d[i, k] - 1.0) < 1e-9 and abs(d[j, k] - 1.0) < 1e-9:
                    faces.append((i, j, k))
    assert len(faces) == 20
    return faces


def mackay_icosahedron(shells: int, nn: float = AU_NN) -> np.ndarray:
    """Mackay icosahedron with `shells` shells counting the central atom.

    Shell k (k >= 1) is the surface of an icosahedron of edge k*nn, with its
    edges and faces filled on a triangular grid of spacing nn: 10k^2+2 atoms.
    Totals: 1, 13, 55, 147, 309, 561, 923, 1415, 2057, ...
    """
    v = icosahedron_vertices()
    faces = icosahedron_faces(v)
    pts = [np.zeros(3)]
    for k in range(1, shells):
        shell = {}
        for (a, b, c) in faces:
            A, B, C = v[a] * k * nn, v[b] * k * nn, v[c] * k * nn
            for i in range(k + 1):
                for j in range(k + 1 - i):
                    p = A + (B - A) * (i / k) + (C - A) * (j / k)
                    key = tuple(np.round(p / nn * 1e6).astype(np.int64))
                    shell[key] = p
        assert len(shell) == 10 * k * k + 2, (k, len(shell))
        pts.extend(shell[key] for key in sorted(shell))
    return np.array(pts)


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from random unit quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_template() -> np.ndarray:
    """O, H1, H2 positions of a rigid SPC/E water with O at the origin."""
    half = math.radians(SPCE_HOH_DEG) / 2.0
    h1 = np.array([SPCE_ROH * math.sin(half), 0.0, SPCE_ROH * math.cos(half)])
    h2 = np.array([-SPCE_ROH * math.sin(half), 0.0, SPCE_ROH * math.cos(half)])
    return np.array([np.zeros(3), h1, h2])


def spce_hh() -> float:
    return 2.0 * SPCE_ROH * math.sin(math.radians(SPCE_HOH_DEG) / 2.0)


@dataclasses.dataclass
class Deck:
    """A generated system: coordinates/velocities plus everything the files carry."""
    name: str
    box: np.ndarray                 # (3,) nm
    metal: np.ndarray               # (n_metal, 3)
    np_ranges: List[Tuple[int, int, float]]   # (first, last, total charge), metal indices
    water: np.ndarray               # (n_water, 3, 3)
    vel: np.ndarray                 # (n_atoms, 3)
    n_steps: int
    dt: float = 0.001
    save_step: int = 10
    efield: Tuple[float, float, float] = (0.0, 0.0, 0.0)   # V/nm
    ref_temp: float = 300.0
    tau_t: float = 0.1
    use_cpff: int = 1
    fix_metal: int = 0
    ensemble: str = "nvt"
    coulomb_alpha: float = 2.0
    rcut: float = 1.2
    run_type: str = "md"            # md, em (steepest descent) or cg
    em_steps: int = 100
    em_length: float = 0.0001       # nm
    em_tol: float = 1000.0          # kJ/mol/nm

    @property
    def n_metal(self) -> int:
        return len(self.metal)

    @property
    def n_water(self) -> int:
        return len(self.water)

    @property
    def n_atoms(self) -> int:
        return self.n_metal + 3 * self.n_water

    def positions(self) -> np.ndarray:
        return np.concatenate([self.metal, self.water.reshape(-1, 3)], axis=0)

    def masses(self) -> np.ndarray:
        return np.concatenate([np.full(self.n_metal, M_AU),
                               np.tile([M_O, M_H, M_H], self.n_water)])


def _untangle(centres: np.ndarray, tmpl: np.ndarray, R: np.ndarray, box: np.ndarray,
              rng: np.random.Generator, hh_min: float, oh_min: float,
              max_rounds: int = 400) -> np.ndarray:
    """Redraws the orientation of molecules whose atoms come closer to another
    molecule's than hh_min (H-H) or oh_min (O-H), periodic images included,
    until no such contact is left.  Of each close pair the molecule with the
    larger index turns, so the result depends on the seed only."""
    from scipy.spatial import cKDTree
    n = len(centres)
    mol = np.repeat(np.arange(n), 3)
    is_h = np.tile([False, True, True], n)
    reach = max(hh_min, oh_min)
    moved = None                        # atoms to check: all at first, then the turned molecules'
    for _ in range(max_rounds):
        mols = centres[:, None, :] + np.einsum("nij,aj->nai", R, tmpl)
        w = np.mod(mols, box).reshape(-1, 3)
        tree = cKDTree(w, boxsize=box)
        if moved is None:
            pairs = tree.query_pairs(reach, output_type="ndarray")
            a, b = pairs[:, 0], pairs[:, 1]
        else:
            hits = tree.query_ball_point(w[moved], reach)
            a = np.repeat(moved, [len(h) for h in hits])
            b = np.fromiter((j for h in hits for j in h), dtype=np.int64, count=len(a))
        d = w[a] - w[b]
        d -= box * np.round(d / box)
        r = np.sqrt(np.einsum("ij,ij->i", d, d))
        hh = is_h[a] & is_h[b]
        oh = is_h[a] != is_h[b]
        close = (mol[a] != mol[b]) & ((hh & (r < hh_min)) | (oh & (r < oh_min)))
        if not close.any():
            return mols
        idx = np.unique(np.maximum(mol[a[close]], mol[b[close]]))
        R[idx] = random_rotations(rng, len(idx))
        moved = (3 * idx[:, None] + np.arange(3)[None, :]).ravel()
    raise RuntimeError("water orientations: close contacts remain after %d rounds" % max_rounds)


def fill_water(box: Sequence[float], metal: np.ndarray, rng: np.random.Generator,
               exclusion: float = 0.25, hh_min: float = 0.0, oh_min: float = 0.0) -> np.ndarray:
    """SPC/E waters on a cubic lattice at 1.0 g/cm^3 with random orientations.

    With hh_min / oh_min > 0 the orientations are redrawn until no two
    molecules have H-H closer than hh_min or O-H closer than oh_min (nm):
    lattice water with independent random orientations puts some hydrogens
    of neighbouring molecules ~0.11 nm apart, and those contacts heat the
    box far above the thermostat's temperature in the first steps.
    Molecules with any atom within `exclusion` nm of a metal atom are removed.
    """
    box = np.asarray(box, dtype=float)
    mw = M_O + 2 * M_H
    n_density = WATER_DENSITY / mw * AVOGADRO * 1e-21      # molecules per nm^3
    spacing = n_density ** (-1.0 / 3.0)
    # a lattice at least as dense as the target, thinned by seeded vacancies
    # so that the box holds the 1.0 g/cm^3 count before the metal is carved out
    counts = np.maximum(1, np.ceil(box / spacing).astype(int))
    grid = [(np.arange(c) + 0.5) * (b / c) for c, b in zip(counts, box)]
    gx, gy, gz = np.meshgrid(*grid, indexing="ij")
    centres = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    target = int(round(n_density * float(np.prod(box))))
    if target < len(centres):
        pick = np.sort(rng.choice(len(centres), size=target, replace=False))
        centres = centres[pick]
    tmpl = water_template()
    com = (tmpl * np.array([M_O, M_H, M_H])[:, None]).sum(0) / mw
    tmpl = tmpl - com
    R = random_rotations(rng, len(centres))
    if hh_min > 0.0 or oh_min > 0.0:
        mols = _untangle(centres, tmpl, R, box, rng, hh_min, oh_min)
    else:
        mols = centres[:, None, :] + np.einsum("nij,aj->nai", R, tmpl)
    if len(metal):
        from scipy.spatial import cKDTree
        flat = np.mod(mols.reshape(-1, 3), box)
        dist, _ = cKDTree(np.mod(metal, box), boxsize=box).query(flat, distance_upper_bound=exclusion)
        close = dist < exclusion
        keep = ~close.reshape(-1, 3).any(axis=1)
        mols = mols[keep]
    return mols


def maxwell_boltzmann(masses: np.ndarray, temp: float, rng: np.random.Generator) -> np.ndarray:
    sigma = np.sqrt(K_BOLTZ * temp / masses)
    v = rng.normal(size=(len(masses), 3)) * sigma[:, None]
    vcom = (v * masses[:, None]).sum(0) / masses.sum()
    return v - vcom


def make_deck(name: str, box, clusters: Sequence[Tuple[int, Sequence[float]]],
              n_steps: int, seed: int, save_step: int = 10,
              efield=(0.0, 0.0, 0.0), with_water: bool = True,
              contacts: Tuple[float, float] = (0.0, 0.0), **kw) -> Deck:
    """clusters: list of (shells, centre) for each gold nanoparticle;
    contacts: the (H-H, O-H) distances below which water orientations are
    redrawn (fill_water), 0 for none."""
    box = np.asarray(box, dtype=float)
    metal_blocks, ranges = [], []
    first = 0
    for shells, centre in clusters:
        pts = mackay_icosahedron(shells) + np.asarray(centre, dtype=float)
        metal_blocks.append(pts)
        ranges.append((first, first + len(pts) - 1, 0.0))
        first += len(pts)
    metal = np.concatenate(metal_blocks) if metal_blocks else np.zeros((0, 3))
    rng = np.random.default_rng(1000 + seed)
    water = (fill_water(box, metal, rng, hh_min=contacts[0], oh_min=contacts[1])
             if with_water else np.zeros((0, 3, 3)))
    deck = Deck(name=name, box=box, metal=metal, np_ranges=ranges, water=water,
                vel=np.zeros((0, 3)), n_steps=n_steps, save_step=save_step,
                efield=tuple(efield), **kw)
    vrng = np.random.default_rng(2000 + seed)
    deck.vel = maxwell_boltzmann(deck.masses(), deck.ref_temp, vrng)
    return deck


def config(index: int) -> Deck:
    """The five benchmark systems.

    0: Au147 in a 3.6 nm box, ~1,500 waters, 100 steps
    1: Au561 in a 6 nm box, ~7,000 waters, 500 steps
    2: two Au309 on x with a 1.0 nm surface gap, 8x5x5 nm box, ~6,500 waters,
       0.5 V/nm field along x, 500 steps
    3: Au2057 in a 10 nm box, ~33,000 waters, 1,000 steps
    4: Au309 in a 4.8 nm box, ~4,000 waters, 3,200 steps (frames every 50)

    Water orientations are redrawn until no intermolecular H-H contact is
    below 0.18 nm and no O-H contact below 0.16 nm (CONTACTS).
    """
    if index == 0:
        return make_deck("cfg0_au147", [3.6] * 3, [(4, [1.8] * 3)], 100, 0, contacts=CONTACTS)
    if index == 1:
        return make_deck("cfg1_au561", [6.0] * 3, [(6, [3.0] * 3)], 500, 1, contacts=CONTACTS)
    if index == 2:
        # Au309 vertex radius: 4 * 0.288 * 0.9511 nm; surface gap 1.0 nm
        rad = 4 * AU_NN * 0.9510565162951535
        half = rad + 0.5
        return make_deck("cfg2_2au309_field", [8.0, 5.0, 5.0],
                         [(5, [4.0 - half, 2.5, 2.5]), (5, [4.0 + half, 2.5, 2.5])],
                         500, 2, efield=(0.5, 0.0, 0.0), contacts=CONTACTS)
    if index == 3:
        return make_deck("cfg3_au2057", [10.0] * 3, [(9, [5.0] * 3)], 1000, 3,
                         save_step=100, contacts=CONTACTS)
    if index == 4:
        return make_deck("cfg4_au309", [4.8] * 3, [(5, [2.4] * 3)], 3200, 4,
                         save_step=50, contacts=CONTACTS)
    raise ValueError(f"unknown config {index}")


def small(kind: str) -> Deck:
    """Small systems that the CPU oracle finishes in seconds."""
    if kind == "au13_w":
        return make_deck("au13_w", [2.6] * 3, [(2, [1.3] * 3)], 10, 101, save_step=1)
    if kind == "au55_w":
        return make_deck("au55_w", [3.0] * 3, [(3, [1.5] * 3)], 10, 102, save_step=1)
    if kind == "2au13_field":
        return make_deck("2au13_field", [4.4, 2.6, 2.6],
                         [(2, [1.3, 1.3, 1.3]), (2, [3.1, 1.3, 1.3])], 10, 103,
                         save_step=1, efield=(0.5, 0.0, 0.0))
    if kind == "au147_vac":
        return make_deck("au147_vac", [3.6] * 3, [(4, [1.8] * 3)], 10, 104,
                         save_step=1, with_water=False)
    if kind == "water_only":
        return make_deck("water_only", [2.5] * 3, [], 10, 105, save_step=1)
    if kind == "au13_em":
        return make_deck("au13_em", [2.6] * 3, [(2, [1.3] * 3)], 10, 101, save_step=1,
                         run_type="em", em_steps=30, em_length=0.002, em_tol=2000.0)
    if kind == "2au13_cg":
        return make_deck("2au13_cg", [4.4, 2.6, 2.6],
                         [(2, [1.3, 1.3, 1.3]), (2, [3.1, 1.3, 1.3])], 10, 103,
                         save_step=1, efield=(0.5, 0.0, 0.0), fix_metal=1,
                         run_type="cg", em_steps=15, em_length=0.002, em_tol=2000.0)
    raise ValueError(kind)


# ---------------------------------------------------------------- writers

def _wrap5(n: int) -> int:
    return n % 100000


def write_gro(deck: Deck, path: str) -> None:
    pos = deck.positions()
    vel = deck.vel
    lines = ["step  0  vQ  0.000000  vP  0.000000", "%5d" % deck.n_atoms]
    res = 0
    for i in range(deck.n_metal):
        res += 1
        lines.append("%5d%-5s%5s%5d%14.9f%14.9f%14.9f%14.9f%14.9f%14.9f" % (
            _wrap5(res), "AU", "Au", _wrap5(i + 1), *pos[i], *vel[i]))
    names = ("OW", "HW1", "HW2")
    for w in range(deck.n_water):
        res += 1
        for a in range(3):
            i = deck.n_metal + 3 * w + a
            lines.append("%5d%-5s%5s%5d%14.9f%14.9f%14.9f%14.9f%14.9f%14.9f" % (
                _wrap5(res), "SOL", names[a], _wrap5(i + 1), *pos[i], *vel[i]))
    lines.append("%10.5f%10.5f%10.5f" % tuple(deck.box))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_param(deck: Deck, path: str) -> None:
    out = ["SCALE_LJ_QQ    0.50000000     0.83333333", ""]
    mols = []
    if deck.n_metal:
        mols.append(("gold", deck.n_metal))
    if deck.n_water:
        mols.append(("SOL", deck.n_water))
    out.append("MOLECULES  %d" % len(mols))
    for name, n in mols:
        out.append("%-10s %d" % (name, n))
    out.append("")
    for name, _ in mols:
        if name == "gold":
            out += ["gold_ATOMS  1", "  0.000000  %.6f  0" % M_AU, ""]
        else:
            out += ["SOL_ATOMS  3",
                    "  %.6f  %.6f  1" % (SPCE_QO, M_O),
                    "  %.6f  %.6f  2" % (SPCE_QH, M_H),
                    "  %.6f  %.6f  2" % (SPCE_QH, M_H), ""]
    out.append("NONBONDED  3  6")
    out.append(" 1  0.0  0.0")
    out.append(" 4  %.12g  %.12g  %.12g" % AU_O_ERFVDW)
    out.append(" 4  %.12g  %.12g  %.12g" % AU_H_ERFVDW)
    out.append(" 1  %.8e  %.8e" % (SPCE_C6, SPCE_C12))
    out.append(" 1  0.0  0.0")
    out.append(" 1  0.0  0.0")
    out.append("")
    for name, _ in mols:
        for sec in ("BONDS", "PAIRS", "ANGLES", "DIHEDRALS", "VSITES"):
            out.append("%s_%s  0" % (name, sec))
        if name == "gold":
            out.append("gold_CSTRS  0")
        else:
            out += ["SOL_CSTRS  3", "0 1  %.10f" % SPCE_ROH, "0 2  %.10f" % SPCE_ROH,
                    "1 2  %.10f" % spce_hh()]
        out.append("")
    if deck.n_metal:
        out.append("MM_METAL  0  %d" % (deck.n_metal - 1))
    else:
        out.append("MM_METAL  -1  -1")
    out.append("QSC_METAL  %d  %d  %.10f  %.6f  %.6f" % (SC_N, SC_M, SC_EPS, SC_C, SC_A))
    out.append("CPFF_METAL  %.10f  %.10f   ;  atomic units" % (CPIM_POLAR, CPIM_CAPAC))
    out.append("")
    out.append("NANOPARTICLES  %d" % len(deck.np_ranges))
    for first, last, q in deck.np_ranges:
        out.append("START_END_CHG  %d  %d  %.6f" % (first, last, q))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_mdset(deck: Deck, path: str) -> None:
    out = [
        "run_type          %s" % deck.run_type,
        "ensemble          %s" % deck.ensemble,
        "",
        "em_steps          %d" % deck.em_steps,
        "em_length_in_nm   %g" % deck.em_length,
        "em_tolerance      %r" % deck.em_tol,
        "",
        "number_of_steps   %d" % deck.n_steps,
        "time_step_in_ps   %g" % deck.dt,
        "save_step         %d" % deck.save_step,
        "",
        "vdw_type          shifted",
        "coulomb_type      wolf_sum",
        "cut_off_radius    %g" % deck.rcut,
        "coulomb_alpha     %g" % deck.coulomb_alpha,
        "",
        "heating_steps     0",
        "ref_T_in_K        %g" % deck.ref_temp,
        "tau_T_in_ps       %g" % deck.tau_t,
        "",
        "fix_metal         %d" % deck.fix_metal,
        "use_cpff          %d" % deck.use_cpff,
        "",
        "external_Ex       %g" % deck.efield[0],
        "external_Ey       %g" % deck.efield[1],
        "external_Ez       %g" % deck.efield[2],
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_deck(deck: Deck, directory: str) -> Tuple[str, str, str]:
    os.makedirs(directory, exist_ok=True)
    gro = os.path.join(directory, "init.gro")
    par = os.path.join(directory, "param.txt")
    mds = os.path.join(directory, "mdset.txt")
    write_gro(deck, gro)
    write_param(deck, par)
    write_mdset(deck, mds)
    return gro, par, mds


# ------------------------------------------------- general topologies
#
# A deck that exercises every bonded term, virtual sites and every pair
# type the reference reads (file.c:614-1000): a flexible six-atom chain
# molecule with harmonic and Morse bonds, 1-4 pairs of both kinds, harmonic and Urey-Bradley angles (one with kb_ub = 0), a
# Ryckaert-Bellemans dihedral and periodic dihedrals (funct 1, 4, 9; n = 1-4;
# phi0 = 0 and 180); TIP5P-type water with two virtual sites; Buckingham and
# Morse pair types beside LJ and erf-damped vdW; an Au13 cluster with CPIM
# and an external field.

TIP5P_ROH, TIP5P_HOH_DEG = 0.09572, 104.52
TIP5P_A, TIP5P_C = -0.344908262, 6.4437903      # vsite a = b, +-c (nm^-1)
TIP5P_QH = 0.241
FLX_Q = (-0.2, 0.1, 0.1, -0.2, 0.1, 0.1)
FLX_TYPE = (1, 1, 1, 1, 2, 2)
FLX_M = (12.011, 12.011, 12.011, 12.011, 1.008, 1.008)


def _place(parent: np.ndarray, grand: np.ndarray, great: np.ndarray, bond: float,
           angle_deg: float, dihedral_deg: float) -> np.ndarray:
    """z-matrix placement: the new atom at `bond` from parent, `angle` with
    grand, `dihedral` with great."""
    b = parent - grand
    b /= np.linalg.norm(b)
    n = np.cross(grand - great, b)
    n /= np.linalg.norm(n)
    m = np.cross(n, b)
    a, d = math.radians(angle_deg), math.radians(dihedral_deg)
    local = np.array([-bond * math.cos(a), bond * math.sin(a) * math.cos(d), bond * math.sin(a) * math.sin(d)])
    return parent + local[0] * b + local[1] * m + local[2] * n


def _flx_template() -> np.ndarray:
    c0 = np.zeros(3)
    c1 = np.array([0.153, 0.0, 0.0])
    c2 = c1 + 0.153 * np.array([math.cos(math.radians(69.0)), math.sin(math.radians(69.0)), 0.0])
    c3 = _place(c2, c1, c0, 0.153, 112.0, 180.0)
    h4 = _place(c3, c2, c1, 0.109, 109.5, 60.0)
    h5 = _place(c0, c1, c2, 0.109, 109.5, -60.0)
    x = np.array([c0, c1, c2, c3, h4, h5])
    return x - x.mean(axis=0)


def _tip5p(o: np.ndarray, rot: np.ndarray) -> np.ndarray:
    half = math.radians(TIP5P_HOH_DEG) / 2.0
    h1 = o + rot @ np.array([TIP5P_ROH * math.sin(half), TIP5P_ROH * math.cos(half), 0.0])
    h2 = o + rot @ np.array([-TIP5P_ROH * math.sin(half), TIP5P_ROH * math.cos(half), 0.0])
    rij, rik = h1 - o, h2 - o
    cr = np.cross(rij, rik)
    l1 = o + TIP5P_A * rij + TIP5P_A * rik - TIP5P_C * cr
    l2 = o + TIP5P_A * rij + TIP5P_A * rik + TIP5P_C * cr
    return np.array([o, h1, h2, l1, l2])


def write_topology_deck(directory: str, seed: int = 106, n_steps: int = 10) -> Tuple[str, str, str]:
    """The general-topology deck above (3.0 nm box, 13 Au, 8 chain molecules,
    40 TIP5P-type waters: 261 atoms)."""
    rng = np.random.default_rng(3000 + seed)
    L = 3.0
    box = np.array([L, L, L])
    metal = mackay_icosahedron(2) + L / 2
    sites = []
    g = np.arange(4) * (L / 4) + L / 8
    for x in g:
        for y in g:
            for z in g:
                p = np.array([x, y, z])
                d = p - L / 2
                if np.linalg.norm(d) > 0.62:
                    sites.append(p)
    rng.shuffle(sites)
    n_flx, n_w = 8, 40
    assert len(sites) >= n_flx + n_w
    tmpl = _flx_template()
    rots = random_rotations(rng, n_flx + n_w)
    # bond lengths and angles off their minima, so every term pulls at step 0
    flx = [sites[k] + (tmpl + rng.normal(0.0, 0.006, tmpl.shape)) @ rots[k].T for k in range(n_flx)]
    wat = [_tip5p(sites[n_flx + k] + rng.normal(0.0, 0.01, 3), rots[n_flx + k]) for k in range(n_w)]
    pos = np.concatenate([metal] + flx + wat)
    masses = np.concatenate([np.full(len(metal), M_AU)] + [np.array(FLX_M)] * n_flx +
                            [np.array([M_O, M_H, M_H, 0.0, 0.0])] * n_w)
    vel = np.zeros_like(pos)
    real = masses > 0
    vel[real] = maxwell_boltzmann(masses[real], 300.0, np.random.default_rng(4000 + seed))
    os.makedirs(directory, exist_ok=True)
    gro = os.path.join(directory, "init.gro")
    lines = ["step  0  vQ  0.000000  vP  0.000000", "%5d" % len(pos)]
    res, i = 0, 0
    for m in range(len(metal)):
        res += 1
        lines.append("%5d%-5s%5s%5d%14.9f%14.9f%14.9f%14.9f%14.9f%14.9f" % (res, "AU", "Au", i + 1, *pos[i], *vel[i]))
        i += 1
    for names, count, rn in ((("C0", "C1", "C2", "C3", "H4", "H5"), n_flx, "FLX"),
                             (("OW", "HW1", "HW2", "LP1", "LP2"), n_w, "T5P")):
        for _ in range(count):
            res += 1
            for nm in names:
                lines.append("%5d%-5s%5s%5d%14.9f%14.9f%14.9f%14.9f%14.9f%14.9f" % (res, rn, nm, i + 1, *pos[i], *vel[i]))
                i += 1
    lines.append("%10.5f%10.5f%10.5f" % tuple(box))
    with open(gro, "w") as fh:
        fh.write("\n".join(lines) + "\n")

    def lj(sig, eps):
        return 4.0 * eps * sig ** 6, 4.0 * eps * sig ** 12

    out = ["SCALE_LJ_QQ    0.50000000     0.83333333", "",
           "MOLECULES  3", "gold       %d" % len(metal), "FLX        %d" % n_flx, "T5P        %d" % n_w, "",
           "gold_ATOMS  1", "  0.000000  %.6f  0" % M_AU, "",
           "FLX_ATOMS  6"]
    out += ["  %.6f  %.6f  %d" % (FLX_Q[a], FLX_M[a], FLX_TYPE[a]) for a in range(6)]
    out += ["", "T5P_ATOMS  5", "  0.000000  %.6f  3" % M_O, "  %.6f  %.6f  4" % (TIP5P_QH, M_H),
            "  %.6f  %.6f  4" % (TIP5P_QH, M_H), "  %.6f  0.000000  4" % -TIP5P_QH,
            "  %.6f  0.000000  4" % -TIP5P_QH, ""]
    cc, cw, hh, ow = lj(0.35, 0.3), lj(0.33, 0.4), lj(0.25, 0.1), lj(0.312, 0.669)
    out += ["NONBONDED  5  15",
            " 1  0.0  0.0",                                           # Au-Au (Sutton-Chen)
            " 4  %.12g  %.12g  %.12g" % AU_O_ERFVDW,                  # Au-C
            " 2  2500.0  32.0  1.0e-4",                               # Au-H  Buckingham
            " 4  %.12g  %.12g  %.12g" % AU_O_ERFVDW,                  # Au-O
            " 1  0.0  0.0",                                           # Au-HW/LP
            " 1  %.8e  %.8e" % cc,                                    # C-C
            " 3  0.5  15.0  0.30",                                    # C-H   Morse
            " 1  %.8e  %.8e" % cw,                                    # C-O
            " 1  0.0  0.0",                                           # C-HW/LP
            " 1  %.8e  %.8e" % hh,                                    # H-H
            " 3  0.8  20.0  0.32",                                    # H-O   Morse
            " 1  0.0  0.0",                                           # H-HW/LP
            " 1  %.8e  %.8e" % ow,                                    # O-O
            " 1  0.0  0.0",                                           # O-HW/LP
            " 1  0.0  0.0", ""]                                       # HW/LP
    out += ["gold_BONDS  0", "gold_PAIRS  0", "gold_ANGLES  0", "gold_DIHEDRALS  0", "gold_VSITES  0",
            "gold_CSTRS  0", ""]
    # no exclusion-only bonds (funct 5): the reference's mpi_bonded adds an
    # uninitialised force pair for them (force.c:1710-1742), which the
    # reader leaves without atoms (file.c:752-756)
    out += ["FLX_BONDS  5",
            "0 1 1  0.153  250000.0",
            "1 2 3  0.153  350.0  19.0",
            "2 3 1  0.153  250000.0",
            "3 4 1  0.109  300000.0",
            "0 5 1  0.109  300000.0",
            "FLX_PAIRS  2",
            "0 3 1",
            "5 2 2  0.5  0.1  0.1  0.25  0.1",
            "FLX_ANGLES  4",
            "0 1 2 1  111.0  460.0",
            "1 2 3 5  112.0  400.0  0.252  20000.0",
            "2 3 4 1  109.5  300.0",
            "5 0 1 5  109.5  300.0  0.25  0.0",
            "FLX_DIHEDRALS  6",
            "0 1 2 3 3  9.28  12.16  -13.12  -3.06  26.24  -31.5",
            "1 2 3 4 9  0.0  2.0  3",
            "5 0 1 2 1  180.0  1.5  2",
            "0 1 2 3 9  0.0  0.5  4",
            "4 3 2 1 4  180.0  0.8  1",
            "5 0 1 2 9  180.0  0.6  4",
            "FLX_VSITES  0", "FLX_CSTRS  0", ""]
    out += ["T5P_BONDS  0", "T5P_PAIRS  0", "T5P_ANGLES  0", "T5P_DIHEDRALS  0",
            "T5P_VSITES  2",
            "0 1 2 3 4  %.9f  %.9f  %.7f" % (TIP5P_A, TIP5P_A, -TIP5P_C),
            "0 1 2 4 4  %.9f  %.9f  %.7f" % (TIP5P_A, TIP5P_A, TIP5P_C),
            "T5P_CSTRS  3",
            "0 1  %.10f" % TIP5P_ROH, "0 2  %.10f" % TIP5P_ROH,
            "1 2  %.10f" % (2.0 * TIP5P_ROH * math.sin(math.radians(TIP5P_HOH_DEG) / 2.0)), ""]
    out += ["MM_METAL  0  %d" % (len(metal) - 1),
            "QSC_METAL  %d  %d  %.10f  %.6f  %.6f" % (SC_N, SC_M, SC_EPS, SC_C, SC_A),
            "CPFF_METAL  %.10f  %.10f   ;  atomic units" % (CPIM_POLAR, CPIM_CAPAC), "",
            "NANOPARTICLES  1", "START_END_CHG  0  %d  0.000000" % (len(metal) - 1)]
    par = os.path.join(directory, "param.txt")
    with open(par, "w") as fh:
        fh.write("\n".join(out) + "\n")
    deck = Deck(name="topo_au13", box=box, metal=metal, np_ranges=[(0, len(metal) - 1, 0.0)],
                water=np.zeros((0, 3, 3)), vel=vel, n_steps=n_steps, save_step=1, dt=0.0005,
                efield=(0.0, 0.0, 0.3))
    mds = os.path.join(directory, "mdset.txt")
    write_mdset(deck, mds)
    return gro, par, mds


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("which", help="config index 0-4 or a small-system name")
    ap.add_argument("outdir")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--save", type=int, default=None)
    a = ap.parse_args(argv)
    deck = config(int(a.which)) if a.which.isdigit() else small(a.which)
    if a.steps is not None:
        deck.n_steps = a.steps
    if a.save is not None:
        deck.save_step = a.save
    write_deck(deck, a.outdir)
    print("%s: %d atoms, %d metal, %d waters, n_mat %d" % (
        deck.name, deck.n_atoms, deck.n_metal, deck.n_water,
        4 * deck.n_metal + len(deck.np_ranges)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
