"""Learned autogrid end to end with the committed weights: the fallback
guarantee against the grid-free sum rule, on held-out molecules."""
import os

import numpy as np
import pytest

from inca_parallel_amd import capi, learned_grid as LG

pytestmark = pytest.mark.gpu


def test_weights_file_is_complete():
    p = np.load(LG.WEIGHTS, allow_pickle=False)
    assert p.dtype == np.float32 and p.size == capi.MLP_NPARAM and np.isfinite(p).all()


def test_fallback_guarantee_on_held_out_molecules():
    net = capi.MLP(seed=1)
    net.set_params(np.load(LG.WEIGHTS, allow_pickle=False))
    accepted = 0
    try:
        for i in range(6):
            m = LG.Molecule(LG.molecule(200000 + i))
            try:
                r = LG.learned_grid_evaluate(net, m)
                err_p = r.get("err_plain")
                if err_p is None:
                    err_p, _, _ = m.uniform(1.0, *LG.PLAIN)
                final = abs(r["integral"] - m.analytic) / abs(m.analytic)
                print("molecule %d: %s, learned err %.2e (%d pts), plain err %.2e, final %.2e" % (
                    i, r["grid"], r["err_learned"], r["points_learned"], err_p, final))
                # never further from the sum rule than the plain grid by more than 1e-8
                assert final <= max(LG.FALLBACK_TOL, err_p + LG.FALLBACK_TOL)
                accepted += r["grid"] == "learned"
            finally:
                m.close()
    finally:
        net.close()
    assert accepted >= 1
