"""capac_md_run ends a step without a host synchronisation when nothing reads
its results (not a save step, not the call's last).  The device work is the
same either way, so eight steps in one call must leave the state bit for bit
where eight one-step calls (each ending with a synchronisation) leave it, in
both precisions, and the last step's observables must agree exactly."""
import os

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
        if per_call:
            for step in range(1, STEPS + 1):
                assert ctx.md_run(step, 1, 10 ** 6) == 1
        else:
            assert ctx.md_run(1, STEPS, 10 ** 6) == STEPS
        r, v, x = ctx.get_state()
        return r, v, x, ctx.observables()
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", [lib.PREC_FP64, lib.PREC_FP32_REFINED])
@pytest.mark.parametrize("name", ["2au13_field", "topo_au13"])
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
