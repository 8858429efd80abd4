"""C-ABI and host-library checks that need no GPU.

  - libcapac_gpu.so loads and exports every function include/capac_gpu.h
    declares, and the Python binding lists exactly those (no compute call is
    made: there is no GPU here);
  - libcapac_host.so exports every function include/capac_host.h declares;
  - struct layouts agree between C and ctypes;
  - the deck reader rejects malformed decks with the reference's messages
    (file.c:74-203, 286-1023);
  - the writers: traj.gro and vec_pq.txt at step 0 are byte-identical to the
    reference program's own output files for the same deck (file.c:212-260).
"""
import ctypes as C
import os
import re
import shutil

import numpy as np
import pytest

from capacmd_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str):
    """Function names declared in a header (comments and macros stripped)."""
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = "\n".join(l for l in txt.splitlines() if not l.lstrip().startswith("#"))
    return set(re.findall(r"\b(capac_\w+)\s*\(", txt))


def test_gpu_library_exports_every_declared_function():
    names = declared("capac_gpu.h")
    assert len(names) > 20
    so = C.CDLL(lib.GPU_SO)
    missing = [n for n in sorted(names) if not hasattr(so, n)]
    assert not missing, missing
    assert set(lib.GPU_EXPORTS) == names
    assert so.capac_version is not None
    so.capac_version.restype = C.c_char_p
    assert so.capac_version().decode().startswith("capac_gpu")


def test_host_library_exports_every_declared_function():
    names = declared("capac_host.h")
    so = C.CDLL(lib.HOST_SO)
    missing = [n for n in sorted(names) if not hasattr(so, n)]
    assert not missing, missing


def test_struct_layouts():
    h = lib.host_lib()
    assert h.capac_params_size() == C.sizeof(lib.CapacParams)
    # capac_timing: 15 doubles + 3 long long
    assert C.sizeof(lib.CapacTiming) == 15 * 8 + 4 * 8


def copy_deck(golden_dir, name, dst):
    for f in ("init.gro", "param.txt", "mdset.txt"):
        shutil.copy(os.path.join(golden_dir, name, f), dst)
    return str(dst)


def test_deck_missing_file(tmp_path):
    with pytest.raises(ValueError, match="Cannot open"):
        lib.Deck(str(tmp_path / "nope.gro"), str(tmp_path / "p.txt"), str(tmp_path / "m.txt"))


def test_deck_bad_param_section(golden_dir, tmp_path):
    d = copy_deck(golden_dir, "au13_w", tmp_path)
    par = os.path.join(d, "param.txt")
    txt = open(par).read()
    assert "NONBONDED" in txt
    open(par, "w").write(txt.replace("NONBONDED", "NONBONDEX", 1))
    with pytest.raises(ValueError, match="expected NONBONDED"):
        lib.Deck.from_dir(d)


def test_deck_atom_count_mismatch(golden_dir, tmp_path):
    d = copy_deck(golden_dir, "au13_w", tmp_path)
    gro = os.path.join(d, "init.gro")
    lines = open(gro).read().splitlines(True)
    n = int(lines[1])
    lines[1] = f" {n - 1}\n"
    del lines[2 + n - 1]
    open(gro, "w").write("".join(lines))
    with pytest.raises(ValueError, match=r"groNAtoms\(%d\) not equal to nAtoms\(%d\)" % (n - 1, n)):
        lib.Deck.from_dir(d)


def test_deck_unknown_coulomb_type(golden_dir, tmp_path):
    d = copy_deck(golden_dir, "au13_w", tmp_path)
    mds = os.path.join(d, "mdset.txt")
    txt = open(mds).read()
    txt = re.sub(r"(coulomb_type\s+)\S+", r"\1ewald", txt)
    open(mds, "w").write(txt)
    with pytest.raises(ValueError, match="unknown coulomb_type ewald"):
        lib.Deck.from_dir(d)


@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_deck_sizes_match_reference_dump(name, golden_dir):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    z = np.load(os.path.join(golden_dir, name, "ref_dump.npz"))
    assert deck.n_mat == z["vec_pq"].size
    assert z["nonb_f"].size == 3 * deck.n_atoms
    np.testing.assert_array_equal(deck.box, z["box"][:3])


@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_write_gro_matches_reference_frame(name, golden_dir, tmp_path):
    """The reference writes frame 0 after removing the centre-of-mass
    velocity (mpi-main.c:630-670); the oracle's md_begin does the same."""
    from oracle.oracle import Oracle
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    o = Oracle(deck)
    o.force()
    o.md_begin()
    r, v, _, _ = o.get_state()
    out = str(tmp_path / "traj.gro")
    h = lib.host_lib()
    rc = h.capac_write_gro_path(out.encode(), b"w", deck._h, lib.dptr(r), lib.dptr(v),
                                deck.vq0, deck.vp, 0)
    assert rc == 0
    got = open(out, "rb").read()
    want = open(os.path.join(golden_dir, name, "ref_traj_frame0.gro"), "rb").read()
    assert got == want


@pytest.mark.parametrize("name", ["au13_w", "2au13_field"])
def test_write_vec_pq_matches_reference_frame(name, golden_dir, tmp_path):
    """vec_pq.txt at step 0 from the reference's own first solution (the
    dump of the first force call) equals the program's first frame."""
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    z = np.load(os.path.join(golden_dir, name, "ref_dump.npz"))
    out = str(tmp_path / "vec_pq.txt")
    h = lib.host_lib()
    x = np.ascontiguousarray(z["force_vec_pq"], dtype=np.float64)
    assert h.capac_write_vec_pq_path(out.encode(), b"w", deck._h, lib.dptr(x), 0) == 0
    got = open(out).read()
    want = open(os.path.join(golden_dir, name, "ref_vec_pq.txt")).read()
    assert want.startswith(got)
    assert got.count("\n") == deck.n_metal + 2 or got.count("\n") > deck.n_metal
