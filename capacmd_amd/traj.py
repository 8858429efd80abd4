"""Readers for the reference's output files, for analysis and comparisons.

    traj.dat    binary frames: int N, int step, double box[3], double rx[N],
                ry[N], rz[N] (write_binary, file.c:269-278)
    vec_pq.txt  per frame "step S  lambda l_1 .. l_nNP", "%5d" N_m, then N_m
                lines "%15.8f" x 4: the dipole p (e bohr) and the charge q
                (write_vec_pq, file.c:236-260)
    traj.gro    per frame "step S  vQ x  vP y", "%5d" N, N atom lines and
                the box (write_gro, file.c:212-229)

traj.dat is memory-mapped, so frames of large runs are read on demand.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Iterator, List

import numpy as np


@dataclass
class DatFrame:
    step: int
    box: np.ndarray          # (3,) nm
    r: np.ndarray            # (3, N) nm, rx|ry|rz


class Dat:
    """traj.dat as a sequence of frames (len, indexing, iteration)."""

    def __init__(self, path: str):
        size = os.path.getsize(path)
        if size < 8:
            raise ValueError(f"{path}: not a traj.dat file ({size} bytes)")
        n = int(np.fromfile(path, dtype=np.int32, count=1)[0])
        if n <= 0:
            raise ValueError(f"{path}: atom count {n}")
        self.n_atoms = n
        self.dtype = np.dtype([("n", "<i4"), ("step", "<i4"), ("box", "<f8", (3,)),
                               ("r", "<f8", (3, n))])
        if size % self.dtype.itemsize:
            raise ValueError(f"{path}: {size} bytes is not a whole number of "
                             f"{self.dtype.itemsize}-byte frames of {n} atoms")
        self._rec = np.memmap(path, dtype=self.dtype, mode="r")
        if np.any(self._rec["n"] != n):
            raise ValueError(f"{path}: frames with different atom counts")

    def __len__(self) -> int:
        return len(self._rec)

    def __getitem__(self, k: int) -> DatFrame:
        rec = self._rec[k]
        return DatFrame(int(rec["step"]), np.array(rec["box"]), np.array(rec["r"]))

    def __iter__(self) -> Iterator[DatFrame]:
        for k in range(len(self)):
            yield self[k]

    @property
    def steps(self) -> np.ndarray:
        return np.array(self._rec["step"])


def read_dat(path: str) -> Dat:
    return Dat(path)


@dataclass
class PqFrame:
    step: int
    lam: np.ndarray          # (n_NP,) Lagrange multipliers
    p: np.ndarray            # (N_m, 3) dipoles, e bohr
    q: np.ndarray            # (N_m,) charges, e

    def vec_pq(self) -> np.ndarray:
        """The solver's vector [p_x p_y p_z interleaved per atom | q | lambda]."""
        return np.concatenate([self.p.reshape(-1), self.q, self.lam])


def read_vec_pq(path: str) -> List[PqFrame]:
    frames = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = 0
    while k < len(lines):
        head = lines[k].split()
        if len(head) < 3 or head[0] != "step" or head[2] != "lambda":
            raise ValueError(f"{path}:{k + 1}: expected 'step S  lambda ...', got {lines[k]!r}")
        step = int(head[1])
        lam = np.array([float(t) for t in head[3:]])
        nm = int(lines[k + 1])
        rows = np.array([[float(t) for t in l.split()] for l in lines[k + 2:k + 2 + nm]]).reshape(nm, 4)
        frames.append(PqFrame(step, lam, rows[:, :3].copy(), rows[:, 3].copy()))
        k += 2 + nm
    return frames


@dataclass
class GroFrame:
    step: int
    vq: float
    vp: float
    res_id: np.ndarray       # (N,)
    res_name: List[str]
    atom_name: List[str]
    r: np.ndarray            # (N, 3) nm
    v: np.ndarray            # (N, 3) nm/ps
    box: np.ndarray          # (3,) nm


def read_gro(path: str) -> List[GroFrame]:
    """Every frame of a traj.gro (or an init.gro with the reference's title)."""
    frames = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = 0
    while k < len(lines):
        t = lines[k].split()
        if len(t) < 6 or t[0] != "step" or t[2] != "vQ" or t[4] != "vP":
            raise ValueError(f"{path}:{k + 1}: expected 'step S  vQ x  vP y', got {lines[k]!r}")
        n = int(lines[k + 1])
        body = lines[k + 2:k + 2 + n]
        vals = np.array([[float(x) for x in l[20:].split()] for l in body])
        frames.append(GroFrame(
            step=int(t[1]), vq=float(t[3]), vp=float(t[5]),
            res_id=np.array([int(l[0:5]) for l in body]),
            res_name=[l[5:10].strip() for l in body],
            atom_name=[l[10:15].strip() for l in body],
            r=vals[:, :3].copy(), v=vals[:, 3:6].copy() if vals.shape[1] >= 6 else np.zeros((n, 3)),
            box=np.array([float(x) for x in lines[k + 2 + n].split()[:3]])))
        k += n + 3
    return frames
