"""Readers of traj.dat, vec_pq.txt and traj.gro (capacmd_amd/traj.py)
against the reference program's own files and against the host library's
writers (file.c:212-278)."""
import os

import numpy as np
import pytest

from capacmd_amd import lib, traj


def test_read_reference_files(golden_dir):
    """The reference program's files: the last traj.dat frame of the 10-step
    MD run (step 10) and of the minimiser (step 18), vec_pq.txt frame 0
    equal to the reference's first solution at the printed precision, and
    the frame-0 traj.gro equal to the deck it started from."""
    g = os.path.join(golden_dir, "au13_w")
    deck = lib.Deck.from_dir(g)
    d = traj.read_dat(os.path.join(g, "ref_traj_last.dat"))
    assert len(d) == 1 and d.n_atoms == deck.n_atoms and d[0].step == 10
    np.testing.assert_array_equal(d[0].box, deck.box[:3])
    assert traj.read_dat(os.path.join(golden_dir, "au13_em", "ref_traj_last.dat"))[0].step == 18
    z = np.load(os.path.join(g, "ref_dump.npz"))
    pq = traj.read_vec_pq(os.path.join(g, "ref_vec_pq.txt"))
    assert [f.step for f in pq] == [0, 10]
    assert pq[0].p.shape == (deck.n_metal, 3) and pq[0].lam.shape == (deck.p.n_nps,)
    assert np.abs(pq[0].vec_pq() - z["force_vec_pq"]).max() <= 5.0000001e-9
    gro = traj.read_gro(os.path.join(g, "ref_traj_frame0.gro"))
    assert len(gro) == 1 and gro[0].step == 0
    r0 = deck.r0.reshape(3, -1).T
    assert np.abs(gro[0].r - r0).max() <= 5.0000001e-6
    assert gro[0].atom_name[:2] == ["Au", "Au"] and gro[0].res_name[-1] == "SOL"


def test_round_trip_through_the_writers(golden_dir, tmp_path):
    """Three frames appended by the host library's writers come back: traj.dat
    bit for bit, vec_pq.txt and traj.gro to their printed precision."""
    deck = lib.Deck.from_dir(os.path.join(golden_dir, "2au13_field"))
    h = lib.host_lib()
    rng = np.random.default_rng(7)
    n, nm, nmat = deck.n_atoms, deck.n_metal, deck.n_mat
    dat, pqf, grof = (str(tmp_path / f) for f in ("t.dat", "pq.txt", "t.gro"))
    rs, xs, vs = [], [], []
    for k, step in enumerate((0, 5, 10)):
        r = rng.uniform(0, 2.5, 3 * n)
        v = rng.normal(0, 0.5, 3 * n)
        x = rng.normal(0, 0.1, nmat)
        mode = b"w" if k == 0 else b"a"
        assert h.capac_write_binary_path(dat.encode(), mode, deck._h, lib.dptr(r), step) == 0
        assert h.capac_write_vec_pq_path(pqf.encode(), mode, deck._h, lib.dptr(x), step) == 0
        assert h.capac_write_gro_path(grof.encode(), mode, deck._h, lib.dptr(r), lib.dptr(v),
                                      0.25 * k, -0.5 * k, step) == 0
        rs.append(r); xs.append(x); vs.append(v)
    d = traj.read_dat(dat)
    assert list(d.steps) == [0, 5, 10]
    for f, r in zip(d, rs):
        np.testing.assert_array_equal(f.r.reshape(-1), r)
    for f, x in zip(traj.read_vec_pq(pqf), xs):
        assert np.abs(f.vec_pq() - x).max() <= 5.0000001e-9
    frames = traj.read_gro(grof)
    assert [f.step for f in frames] == [0, 5, 10]
    for k, (f, r, v) in enumerate(zip(frames, rs, vs)):
        assert np.abs(f.r - r.reshape(3, -1).T).max() <= 5.0000001e-6
        assert np.abs(f.v - v.reshape(3, -1).T).max() <= 5.0000001e-7
        assert f.vq == 0.25 * k and f.vp == -0.5 * k


def test_read_dat_rejects_partial_frames(tmp_path, golden_dir):
    raw = open(os.path.join(golden_dir, "au13_w", "ref_traj_last.dat"), "rb").read()
    p = tmp_path / "cut.dat"
    p.write_bytes(raw[:-8])
    with pytest.raises(ValueError, match="whole number"):
        traj.read_dat(str(p))
