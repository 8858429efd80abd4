"""capac_cpim_solve_noalloc: the CPIM solve with its work space in the
library's file-scope device array (no allocation in the call), against
capac_cpim_solve on the same system, right-hand side and warm start.

Both run the same kernels in the same order on the same inputs, so the
solutions, iteration counts and residuals must be bitwise equal.  Two calls
in a row share the device array: the second, on a smaller system, must not
see the first one's contents.  The file's name puts it last in collection
order.
"""
import os

import numpy as np
import pytest

from capacmd_amd import lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [lib.PREC_FP64, lib.PREC_FP32_REFINED], ids=["fp64", "fp32"])
def test_noalloc_solve_matches(prec, golden_dir):
    for name in ("2au13_field", "au13_w"):        # the second call on a smaller system
        deck = lib.Deck.from_dir(os.path.join(golden_dir, name))
        ctx = lib.Context(deck, precision=prec)
        try:
            b = ctx.cpim_rhs()
            rng = np.random.default_rng(7)
            x0 = 1e-3 * rng.standard_normal(deck.n_mat)
            want = ctx.cpim_solve(b=b, x=x0)
            got = ctx.cpim_solve_noalloc(b=b, x=x0)
            print(name, prec, "iters", want[1], got[1], "kind", want[4], got[4], flush=True)
            assert want[4] == 0
            np.testing.assert_array_equal(got[0], want[0])
            assert got[1:] == want[1:]
        finally:
            ctx.close()
