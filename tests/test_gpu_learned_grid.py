"""Learned autogrid end to end with the committed weights, on held-out
molecules (seeds 200000+, never used in training): the fallback guarantee
against the grid-free sum rule, and what the network gains over the plain
autogrid at the same cost.

The gain is judged two ways on 24 molecules: by the exact integral of the
prepared terms (the sum rule, grid-free) and by the FP64 dense grid (nrad
200, nang 974) on the same centres.  The committed network, trained on
20,000 molecules, was measured on 1,000 held-out ones (the bars below sit
well inside what it reached there)."""
import numpy as np
import pytest

from inca_parallel_amd import capi, learned_grid as LG

pytestmark = pytest.mark.gpu


def test_weights_file_is_complete():
    p = np.load(LG.WEIGHTS, allow_pickle=False)
    assert p.dtype == np.float32 and p.size == capi.MLP_NPARAM and np.isfinite(p).all()


def test_learned_grid_gains_on_held_out_molecules():
    net = capi.MLP(seed=1)
    net.set_params(np.load(LG.WEIGHTS, allow_pickle=False))
    rows = []
    try:
        for i in range(24):
            m = LG.Molecule(LG.molecule(200000 + i))
            try:
                r = LG.learned_grid_evaluate(net, m)
                err_p, tot_p, g_p = m.uniform(1.0, *LG.PLAIN)
                err_d, tot_d, _ = m.uniform(1.0, *LG.DENSE)
                final = abs(r["integral"] - m.analytic) / abs(m.analytic)
                rows.append(dict(grid=r["grid"], final=final, plain=err_p,
                                 points=r["points_learned"] / g_p,
                                 dense_final=abs(r["integral"] - tot_d) / abs(tot_d),
                                 dense_plain=abs(tot_p - tot_d) / abs(tot_d)))
                print("molecule %d: %s, final err %.2e (plain %.2e), points %.3f of plain, "
                      "dense-judged %.2e (plain %.2e)" % (
                          i, r["grid"], final, err_p, rows[-1]["points"],
                          rows[-1]["dense_final"], rows[-1]["dense_plain"]))
            finally:
                m.close()
    finally:
        net.close()
    # the guarantee: never further from the sum rule than the plain grid by
    # more than 1e-8
    for r in rows:
        assert r["final"] <= max(LG.FALLBACK_TOL, r["plain"] + LG.FALLBACK_TOL)
    # the gain at the plain grid's cost
    accepted = sum(r["grid"] == "learned" for r in rows)
    ratio = np.exp(np.mean(np.log([max(r["final"], 1e-16) / max(r["plain"], 1e-16)
                                   for r in rows])))
    dense_ratio = np.exp(np.mean(np.log([max(r["dense_final"], 1e-16) /
                                         max(r["dense_plain"], 1e-16) for r in rows])))
    print("accepted %d of %d; geometric-mean error ratio to the plain grid %.3f (sum rule), "
          "%.3f (dense grid)" % (accepted, len(rows), ratio, dense_ratio))
    assert accepted >= len(rows) // 4
    assert max(r["points"] for r in rows) <= 1.06
    assert ratio < 0.8
    assert dense_ratio < 0.9
