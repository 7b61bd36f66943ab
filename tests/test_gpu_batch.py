"""Batched persistent launch against unbatched per-molecule runs.

Several seeded molecules of different sizes (s/p/d primitives, multicentre
autogrids with TFVC, nosym with symmetry-unique records) are prepared in
their own contexts; the batch (class-sorted quartets, cost-ordered work
items, persistent kernel, hipGraph replay) must give each molecule's Ivec
and integral as its own inca_evaluate does, apart from summation order:
class sorting reorders the quartets, so |dI| is bounded by 1e-13 max|I|
(measured about 1e-14: the d-function fits have terms far larger than I
that cancel).  Replays of the captured graph are bitwise identical."""
import numpy as np
import pytest

from inca_parallel_amd import capi, synth

pytestmark = pytest.mark.gpu
THRESH = 1e-12


def _molecule(i):
    nat = 3 + i % 4
    s = synth.config_cluster(seed=500 + i, natoms=nat, nheavy=max(1, nat - 2), nocc=2 + i % 3,
                             shells_heavy={0: 2, 1: 1, 2: i % 2},
                             shells_h={0: 2})
    return s


@pytest.fixture(scope="module")
def molecules():
    out = []
    for i in range(6):
        s = _molecule(i)
        p = s.prims()
        ijkl, d = s.dm2_unique()
        e = capi.Engine(0)
        e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
        cent, Ps, sf = capi.autogrid(p["an"], p["cartes"], nohydro=True)
        nq = len(Ps)
        e.grid_multicenter(cent, [30 + 10 * (i % 3)] * nq, [110 if i % 2 else 146] * nq, sf, Ps,
                           nosym=True)
        e.set_quartets(ijkl=ijkl, d=d)
        iv, cnt = e.evaluate(capi.INCA_FP64, THRESH)
        tot = e.integrate()[:2]
        out.append(dict(s=s, e=e, ivec=iv, tot=tot, cnt=cnt))
        print("molecule %d: %d prims, %d points, %d quartets" % (i, len(p["alpha"]), e.G, cnt[2]))
    yield out
    for m in out:
        m["e"].close()


def _check(batch, mols, tag):
    ti, tv = batch.evaluate()
    worst = 0.0
    for m, mol in enumerate(mols):
        iv = batch.ivec(m)
        ref = mol["ivec"]
        scale = np.abs(ref).max()
        err = np.abs(iv - ref).max() / scale
        worst = max(worst, err)
        assert err < 1e-13, (tag, m, err)
        assert abs(ti[m] - mol["tot"][0]) <= 1e-13 * abs(mol["tot"][0]), (tag, m)
        assert abs(tv[m] - mol["tot"][1]) <= 1e-13 * abs(mol["tot"][1]), (tag, m)
    return worst


def test_batch_matches_unbatched_runs(molecules):
    for mol in molecules:
        mol["e"].prepare(capi.INCA_FP64, THRESH)
    b = capi.Batch([m["e"] for m in molecules])
    try:
        info = b.info()
        print("batch: %s" % info)
        assert info["items"] > 0 and info["blocks"] > 0
        w = _check(b, molecules, "first replay")
        first = [b.ivec(m) for m in range(len(molecules))]
        # replays with eager work of the contexts in between
        for r in range(3):
            molecules[r]["e"].evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
            w = max(w, _check(b, molecules, "replay %d" % (r + 2)))
            for m in range(len(molecules)):
                assert np.array_equal(b.ivec(m), first[m]), ("replay not bitwise", r, m)
        assert b.info()["replays"] == 4
        print("batch vs unbatched: max |dI| / max|I| = %.2e over 4 replays" % w)
    finally:
        b.close()


def test_batch_autotune_recaptures(molecules):
    for mol in molecules:
        mol["e"].prepare(capi.INCA_FP64, THRESH)
    b = capi.Batch([m["e"] for m in molecules])
    try:
        best = b.autotune()
        assert best in (1, 2, 4, 8)
        assert b.info()["points_per_lane"] == best
        _check(b, molecules, "autotuned")
        _check(b, molecules, "autotuned replay")
        print("autotuned points per lane %d, %s" % (best, b.info()))
        for P in (1, 8, best):
            b.set_points(P)
            assert b.info()["points_per_lane"] == P
            _check(b, molecules, "points per lane %d" % P)
        with pytest.raises(capi.IncaError, match="1, 2, 4 or 8"):
            b.set_points(3)
    finally:
        b.close()


def test_batch_mixed_matches_unbatched_mixed(molecules):
    """Molecules prepared for INCA_MIXED: the batch evaluates each one's FP64
    range with the FP64 body and its FP32 range with the mixed body, so it
    must match the molecule's own mixed evaluation apart from the FP64
    summation order, and the FP64 result within the mixed tolerance."""
    ref = []
    try:
        for mol in molecules:
            e = mol["e"]
            e.set_mixed_tail(0.05)
            iv, _ = e.evaluate(capi.INCA_MIXED, THRESH)
            n64, n32 = e.mixed_split()[:2]
            assert n64 > 0 and n32 > 0, (n64, n32)
            ref.append((iv, e.integrate()[0]))
        b = capi.Batch([m["e"] for m in molecules])
        try:
            ti, _ = b.evaluate()
            worst = worst64 = 0.0
            for m, mol in enumerate(molecules):
                iv = b.ivec(m)
                err = np.abs(iv - ref[m][0]).max() / np.abs(ref[m][0]).max()
                worst = max(worst, err)
                assert err < 1e-13, (m, err)
                assert abs(ti[m] - ref[m][1]) <= 1e-13 * abs(ref[m][1])
                e64 = abs(ti[m] - mol["tot"][0]) / abs(mol["tot"][0])
                worst64 = max(worst64, e64)
                assert e64 < 1e-8, (m, e64)
            print("mixed batch vs unbatched mixed %.2e; integral vs FP64 %.2e" % (worst, worst64))
        finally:
            b.close()
    finally:
        for mol in molecules:
            mol["e"].set_mixed_tail(0.02)


def test_batch_outlives_its_contexts():
    """The batch owns copies: destroying the contexts leaves it usable."""
    s = _molecule(1)
    p = s.prims()
    ijkl, d = s.dm2_unique()
    e = capi.Engine(0)
    e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    cent, Ps, sf = capi.autogrid(p["an"], p["cartes"])
    e.grid_multicenter(cent, [30] * len(Ps), [110] * len(Ps), sf, Ps, nosym=True)
    e.set_quartets(ijkl=ijkl, d=d)
    ref, _ = e.evaluate(capi.INCA_FP64, THRESH)
    b = capi.Batch([e])
    e.close()
    try:
        b.evaluate()
        iv = b.ivec(0)
        assert np.abs(iv - ref).max() <= 1e-13 * np.abs(ref).max()
    finally:
        b.close()


def test_batch_rejects_unprepared_context():
    e = capi.Engine(0)
    try:
        with pytest.raises(capi.IncaError):
            capi.Batch([e])
    finally:
        e.close()
