"""The C host (capacmd_amd/capacmd) against the reference program's own
outputs for the same deck (tests/golden/*/ref_*, written by the reference
built from its sources): the stdout step lines and final energy table
(mpi-main.c:993-995, force.c:337-354), traj.gro (file.c:212-229),
vec_pq.txt (file.c:236-260) and the last traj.dat frame (file.c:269-278).

traj.gro's first frame is compared byte for byte.  The rest passes through
the CPIM solve, whose results agree within the solver tolerance (see
test_gpu_parity.py), so numbers are compared with tolerances stated below,
and formats line by line.
"""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "capacmd_amd", "capacmd")
STEP = re.compile(r"^\s+([0-9.]+)\s+Fmax=(\S+)\s+E=(\S+)\s+T=(\S+)")


def run_cli(golden_dir, name, tmp_path, prec):
    for f in ("init.gro", "param.txt", "mdset.txt"):
        shutil.copy(os.path.join(golden_dir, name, f), tmp_path)
    print("running", CLI, name, prec, flush=True)
    res = subprocess.run([CLI, "-gro", "init.gro", "-par", "param.txt", "-mds", "mdset.txt", "-prec", prec],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return res.stdout


def step_lines(text):
    return [tuple(float(v) for v in m.groups()) for m in map(STEP.match, text.splitlines()) if m]


def energy_table(text):
    lines = text.splitlines()
    k = next(i for i, l in enumerate(lines) if "Potential energies:" in l)
    out = []
    for l in lines[k + 1:k + 14]:
        v, label = l.split(None, 1)
        out.append((float(v), label.strip()))
    return out


def read_dat_frame(raw, n):
    N, step = struct.unpack("ii", raw[:8])
    assert N == n
    box = np.frombuffer(raw[8:32], dtype=np.float64)
    r = np.frombuffer(raw[32:32 + 24 * n], dtype=np.float64)
    return step, box, r


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_cli_matches_reference_program(name, prec, golden_dir, tmp_path):
    g = os.path.join(golden_dir, name)
    out = run_cli(golden_dir, name, tmp_path, prec)
    ref = open(os.path.join(g, "ref_stdout.txt")).read()
    # step lines: time exact; E and T to 1e-5 (7 printed digits, solver
    # tolerance over 10 steps); Fmax is printed with 3 digits
    a, b = step_lines(out), step_lines(ref)
    assert len(a) == len(b) > 0
    for (t1, f1, e1, T1), (t2, f2, e2, T2) in zip(a, b):
        assert t1 == t2
        assert abs(e1 - e2) <= 1e-5 * abs(e2), (t1, e1, e2)
        assert abs(T1 - T2) <= 1e-5 * abs(T2), (t1, T1, T2)
        assert abs(f1 - f2) <= 0.011 * abs(f2), (t1, f1, f2)
    # final energy table: same rows and labels; values to 1e-5 (1e-6 absolute
    # floor), except the CPIM rows, functions of the solved charges and
    # dipoles.  The reference's stop test bounds the relative residual by
    # 1e-4, and on the field deck the reference's own solution is 3.8e-4 from
    # the exact one (test_gpu_parity.py::test_bicgstab); the error that
    # leaves in each CPIM term scales with the CPIM energy as a whole, not
    # with the (sometimes small) term, so each CPIM row is held to 1e-4 of
    # the summed magnitude of the CPIM rows.
    ta, tb = energy_table(out), energy_table(ref)
    assert [l for _, l in ta] == [l for _, l in tb]
    cpim_scale = sum(abs(v) for v, l in tb if l.startswith("CPIM"))
    for (va, la), (vb, _) in zip(ta, tb):
        tol = 1e-4 * cpim_scale if la.startswith("CPIM") else max(1e-5 * abs(vb), 1e-6)
        assert abs(va - vb) <= tol, (la, va, vb)
    # traj.gro: the first frame byte for byte; the file has the same line count
    tg = open(os.path.join(tmp_path, "traj.gro")).read().splitlines()
    n = int(tg[1])
    ref0 = open(os.path.join(g, "ref_traj_frame0.gro")).read().splitlines()
    assert tg[:n + 3] == ref0
    # vec_pq.txt: same layout; values to 1e-4 of the largest entry (solver tolerance)
    va = open(os.path.join(tmp_path, "vec_pq.txt")).read().splitlines()
    vb = open(os.path.join(g, "ref_vec_pq.txt")).read().splitlines()
    assert len(va) == len(vb)
    for la, lb in zip(va, vb):
        fa, fb = la.split(), lb.split()
        assert len(fa) == len(fb)
        for xa, xb in zip(fa, fb):
            try:
                xa, xb = float(xa), float(xb)
            except ValueError:
                assert xa == xb
                continue
            assert abs(xa - xb) <= 2e-4 * max(1.0, abs(xb)), (la, lb)
    # traj.dat: the last frame, positions to 1e-6 nm
    raw = open(os.path.join(tmp_path, "traj.dat"), "rb").read()
    rec = 8 + 24 + 24 * n
    assert len(raw) % rec == 0
    sa, ba, ra = read_dat_frame(raw[-rec:], n)
    sb, bb, rb = read_dat_frame(open(os.path.join(g, "ref_traj_last.dat"), "rb").read(), n)
    assert sa == sb
    np.testing.assert_array_equal(ba, bb)
    assert np.abs(ra - rb).max() < 1e-6
