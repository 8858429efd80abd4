"""capac_md_run ends a step without a host synchronisation when nothing reads
its results (not a save step, not the call's last).  The device work is the
same either way, so eight steps in one call must leave the state bit for bit
where eight one-step calls (each ending with a synchronisation) leave it, in
both precisions, and the last step's observables must agree exactly.  Every
force call is counted, also those whose results stay on the device, and a
RATTLE failure inside such a step comes back as the reference's RATTLE error
(at that step or the next), never as another error.  The unobserved steps
also run the water pair kernel without its energies and virial (au13_w,
2au13_field): its forces must be the same bits."""
import os
import re
import shutil

import numpy as np
import pytest

from capacmd_amd import lib

pytestmark = pytest.mark.gpu

STEPS = 8


def run(deck, precision, per_call):
    ctx = lib.Context(deck, precision)
    try:
        ctx.force()
        ctx.md_begin()
        ctx.reset_timing()
        if per_call:
            for step in range(1, STEPS + 1):
                assert ctx.md_run(step, 1, 10 ** 6) == 1
        else:
            assert ctx.md_run(1, STEPS, 10 ** 6) == STEPS
        r, v, x = ctx.get_state()
        assert ctx.timing()[1] == STEPS          # one force call per step
        return r, v, x, ctx.observables()
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
@pytest.mark.parametrize("name", ["au13_w", "2au13_field", "topo_au13"])
def test_one_call_equals_step_by_step(golden_dir, name, precision):
    deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
    print(name, precision, flush=True)
    a = run(deck, precision, per_call=False)
    b = run(deck, precision, per_call=True)
    for k in range(3):
        assert np.array_equal(a[k], b[k]), k
    for key in ("temp", "eksum"):
        assert a[3][key] == b[3][key], key
    assert np.array_equal(a[3]["pot"], b[3]["pot"])


@pytest.mark.parametrize("precision", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
def test_rattle_failure_reported(golden_dir, tmp_path, precision):
    """A water whose three constraint lengths violate the triangle inequality
    (H-H 0.21 nm for O-H 0.1 nm) cannot be satisfied, so RATTLE_1 fails at
    the first step; inside one multi-step call that step ends unobserved.
    (The oracle's RATTLE_1 on this deck stops after its 1000 sweeps with
    every coordinate finite.)"""
    for f in ("init.gro", "param.txt", "mdset.txt"):
        shutil.copy(os.path.join(golden_dir, "au13_w", f), tmp_path / f)
    par = (tmp_path / "param.txt").read_text()
    par, n = re.subn(r"(SOL_CSTRS\s+3\n0 1\s+\S+\n0 2\s+\S+\n1 2\s+)\S+", r"\g<1>0.2100000000", par)
    assert n == 1
    (tmp_path / "param.txt").write_text(par)
    deck = lib.Deck.from_dir(str(tmp_path))
    ctx = lib.Context(deck, precision)
    try:
        ctx.force()
        ctx.md_begin()
        with pytest.raises(lib.CapacError, match="RATTLE_1 not converged"):
            ctx.md_run(1, 6, 10 ** 6)
    finally:
        ctx.close()
