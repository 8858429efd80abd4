"""Writes the golden fixtures under tests/golden/ (run in the build container).

For each small deck it
  1. writes the deck (init.gro, param.txt, mdset.txt) with capacmd_amd.gen,
  2. runs oracle/_ref/ref_dump -- the reference's own compiled phase functions
     driven by oracle/ref_dump.cpp -- and packs the full-precision arrays into
     ref_dump.npz,
  3. for the MD and minimiser decks, runs the reference program itself
     (oracle/_ref/capacmd_ref, built from the reference sources by
     oracle/Makefile) and keeps its stdout, vec_pq.txt, the first traj.gro
     frame and the last traj.dat frame (and, for the minimiser, the last
     traj.gro frame) as fixtures.  The minimiser decks skip step 2.
Needs the reference sources (make -C oracle ref); the GPU box never runs it.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from capacmd_amd import gen  # noqa: E402

REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
REF_PROG = os.path.join(ROOT, "oracle", "_ref", "capacmd_ref")

# deck name -> (MD steps dumped by ref_dump or None for no dump, keep the COO
# matrix, run the program)
DECKS = {
    "au13_w": (3, True, True),
    "2au13_field": (2, True, True),
    "au147_vac": (0, False, False),
    "water_only": (2, False, False),
    "au13_em": (None, False, True),
    "2au13_cg": (None, False, True),
    "topo_au13": (3, False, True),
}


def build():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True,
                   stdout=subprocess.DEVNULL)


def dump_phases(gro, par, mds, out, nsteps, keep_coo):
    with tempfile.TemporaryDirectory() as tmp:
        cmd = [REF_DUMP, gro, par, mds, tmp] + ([str(nsteps)] if nsteps else [])
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
        arrays = {}
        for fn in sorted(os.listdir(tmp)):
            key = fn[:-4]
            if key.startswith("coo_") and not keep_coo:
                continue
            if key in ("pos", "vel"):
                continue            # the deck itself holds them
            arrays[key] = np.load(os.path.join(tmp, fn))
        np.savez_compressed(os.path.join(out, "ref_dump.npz"), **arrays)


def run_program(gro, par, mds, out, minimiser):
    with tempfile.TemporaryDirectory() as tmp:
        for f in (gro, par, mds):
            shutil.copy(f, tmp)
        env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="0")
        res = subprocess.run([REF_PROG, "-gro", "init.gro", "-par", "param.txt",
                              "-mds", "mdset.txt"], cwd=tmp, env=env, check=True,
                             capture_output=True, text=True)
        lines = res.stdout.splitlines()
        start = next(i for i, l in enumerate(lines) if "start MD calculation" in l)
        end = next(i for i, l in enumerate(lines) if "Calculation ended" in l)
        with open(os.path.join(out, "ref_stdout.txt"), "w") as fh:
            fh.write("\n".join(lines[start:end]) + "\n")
        shutil.copy(os.path.join(tmp, "vec_pq.txt"), os.path.join(out, "ref_vec_pq.txt"))
        with open(os.path.join(tmp, "traj.gro")) as fh:
            tg = fh.read().splitlines()
        n = int(tg[1])
        # first frame: the format fixture
        with open(os.path.join(out, "ref_traj_frame0.gro"), "w") as fh:
            fh.write("\n".join(tg[: n + 3]) + "\n")
        if minimiser:
            with open(os.path.join(out, "ref_traj_last.gro"), "w") as fh:
                fh.write("\n".join(tg[-(n + 3):]) + "\n")
        raw = open(os.path.join(tmp, "traj.dat"), "rb").read()
        rec = 8 + 24 + 24 * n
        with open(os.path.join(out, "ref_traj_last.dat"), "wb") as fh:
            fh.write(raw[-rec:])


def main() -> int:
    build()
    for name, (nsteps, keep_coo, run_prog) in DECKS.items():
        out = os.path.join(HERE, name)
        os.makedirs(out, exist_ok=True)
        if name == "topo_au13":         # general topologies (bonded terms, virtual sites)
            gro, par, mds = gen.write_topology_deck(out)
            minimiser = False
        else:
            deck = gen.small(name)
            gro, par, mds = gen.write_deck(deck, out)
            minimiser = deck.run_type != "md"
        if nsteps is not None:
            dump_phases(gro, par, mds, out, nsteps, keep_coo)
        if run_prog:
            run_program(gro, par, mds, out, minimiser)
        print(name, "->", out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
