"""Restart checkpoints (capac_checkpoint_save / capac_checkpoint_load and
the CLI's -cpt / -rst): a run continued from a checkpoint repeats the
uninterrupted run's later steps bit for bit, on both precisions.

The reference restarts from traj.gro, which keeps vQ[0] and vP of the
ten-link Nose-Hoover chain and no CPIM warm start (file.c:152-160,
215-229); the checkpoint carries the whole chain, vec_pq, forces, the
heating ramp and the positions the neighbour rows were built at, so the
rows (and with them the order of every pair sum) are rebuilt as they were.
au55_w (3.0 nm box) runs with the 0.1 nm Verlet skin, so the rows in use at
the checkpoint are reused after it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from capacmd_amd import gen, lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "capacmd_amd", "capacmd")


@pytest.fixture(scope="module")
def au55(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("au55_w"))
    gen.write_deck(gen.small("au55_w"), d)
    return lib.Deck.from_dir(d)


def run(deck, prec, n_first, n_second=0, blob=None):
    ctx = lib.Context(deck, prec)
    if blob is None:
        ctx.force()
        ctx.md_begin()
        assert ctx.md_run(1, n_first, 1) == n_first
        start = n_first
    else:
        start = ctx.restore(blob)
    if n_second:
        assert ctx.md_run(start + 1, n_second, 1) == n_second
    return ctx


@pytest.mark.parametrize("prec", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
@pytest.mark.parametrize("name", ["au55_w", "2au13_field"])
def test_checkpoint_continues_bitwise(name, prec, au55, golden_dir):
    deck = au55 if name == "au55_w" else lib.Deck.from_dir(os.path.join(golden_dir, name))
    print(name, prec, flush=True)
    a = run(deck, prec, 12)
    ra, va, xa = a.get_state()
    oa = a.observables()
    builds = a.nb_builds()
    skin = a.skin
    a.close()
    b = run(deck, prec, 5)
    blob = b.checkpoint(5)
    b.close()
    c = run(deck, prec, 0, 7, blob=blob)
    rc_, vc, xc = c.get_state()
    oc = c.observables()
    c.close()
    print(name, prec, "skin", skin, "builds in 12 steps", builds, flush=True)
    if name == "au55_w":
        assert skin > 0 and builds < 13
    np.testing.assert_array_equal(rc_, ra)
    np.testing.assert_array_equal(vc, va)
    np.testing.assert_array_equal(xc, xa)
    for k in ("eksum", "temp", "fmax", "frms"):
        assert oc[k] == oa[k], k
    np.testing.assert_array_equal(oc["vq"], oa["vq"])
    np.testing.assert_array_equal(oc["pot"], oa["pot"])
    np.testing.assert_array_equal(oc["vir"], oa["vir"])


def test_checkpoint_rejects_other_systems(au55, golden_dir):
    other = lib.Deck.from_dir(os.path.join(golden_dir, "au13_w"))
    a = run(other, lib.PREC_FP64, 1)
    blob = a.checkpoint(1)
    a.close()
    ctx = lib.Context(au55)
    with pytest.raises(lib.CapacError, match="another system"):
        ctx.restore(blob)
    with pytest.raises(lib.CapacError, match="not a capac checkpoint"):
        ctx.restore(b"X" * len(blob))
    with pytest.raises(lib.CapacError, match="truncated"):
        ctx.restore(blob[:100])
    ctx.close()
    ctx = lib.Context(other)
    with pytest.raises(lib.CapacError, match="truncated"):
        ctx.restore(blob[:-8])
    assert ctx.restore(blob) == 1
    ctx.close()


def cli(args, cwd):
    res = subprocess.run([CLI, "-gro", "init.gro", "-par", "param.txt", "-mds", "mdset.txt"] + args,
                         cwd=cwd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return res.stdout


def set_steps(path, n):
    lines = open(path).read().splitlines()
    lines = ["number_of_steps   %d" % n if l.startswith("number_of_steps") else l for l in lines]
    open(path, "w").write("\n".join(lines) + "\n")


def test_cli_restart_repeats_the_run(golden_dir, tmp_path):
    """10 steps straight against 4 steps with -cpt and then -rst up to 10:
    the step lines of steps 5-10, the final energy table and the last
    traj.dat frame are the same bytes."""
    straight, first = tmp_path / "straight", tmp_path / "split"
    for d in (straight, first):
        d.mkdir()
        for f in ("init.gro", "param.txt", "mdset.txt"):
            shutil.copy(os.path.join(golden_dir, "au13_w", f), d)
    out_a = cli([], straight)
    set_steps(first / "mdset.txt", 4)
    cli(["-cpt", "run.cpt"], first)
    set_steps(first / "mdset.txt", 10)
    out_b = cli(["-rst", "run.cpt"], first)

    def body(text, t0):
        lines = text.splitlines()
        steps = [l for l in lines if l.startswith("  ") and "Fmax=" in l and float(l.split()[0]) >= t0]
        k = next(i for i, l in enumerate(lines) if "Potential energies:" in l)
        return steps, lines[k:k + 14]

    sa, ea = body(out_a, 0.005 - 1e-9)
    sb, eb = body(out_b, 0.005 - 1e-9)
    assert len(sa) == 6 and sa == sb
    assert ea == eb
    n = int(open(straight / "traj.gro").read().splitlines()[1])
    rec = 8 + 24 + 24 * n
    da, db = open(straight / "traj.dat", "rb").read(), open(first / "traj.dat", "rb").read()
    assert len(db) == 6 * rec
    assert da[-rec:] == db[-rec:]
