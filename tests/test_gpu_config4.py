"""The batched-molecules workload: seeded molecules of 4-20 atoms with
30-300 s/p/d primitives (synth.config4_molecule, seeds 1000-1255), each on
its own autogrid (nrad 50-150, nang 110-590, nosym), streamed through the
persistent hipGraph-captured batch in arena-sized groups.

All 256 molecules together are about 6e10 records and 3e16 reference
evaluations: about an hour of the device in FP64 and mixed precision
(`python -m inca_parallel_amd.batch_stream --rank lo:hi` runs slices of
them by cost rank).  The checks here run a stratified subset at the
molecules' stated sizes: the cheapest molecule of every atom-count band
plus the cheapest overall.  They are checked through properties that
do not depend on size:
  * the batch against the same molecules' prepared quartets evaluated one
    context at a time (FP64, 1e-12 relative on the integral);
  * the mixed batch against the FP64 one (1e-8 relative on the integral);
  * replays bitwise identical, and results independent of the grouping;
  * with exact polynomials, the analytic integral of the prepared terms
    equals the pair count N(N-1)/2 (up to the screened terms);
  * the grid sum-rule error per molecule is reported.
"""
import numpy as np
import pytest

from inca_parallel_amd import batch_stream as BS
from inca_parallel_amd import capi, synth

pytestmark = pytest.mark.gpu
BANDS = [(4, 6), (7, 9), (10, 12), (13, 15), (16, 18), (19, 20)]


def _cost(seed):
    s, nrad, nang = synth.config4_molecule(seed)
    cent, _, _ = capi.autogrid(s.an, s.coords, nohydro=True)
    npair = s.nprim * (s.nprim + 1) // 2
    return npair * (npair + 1) // 2 * len(cent) * nrad * nang


@pytest.fixture(scope="module")
def subset():
    """Seeds of the stratified subset: the cheapest molecule (records x
    grid points) of each atom-count band, and the six cheapest overall."""
    seeds = list(synth.CONFIG4_SEEDS)
    rows = []
    for seed in seeds:
        rng = np.random.default_rng(seed)
        nat = int(rng.integers(4, 21))  # the first draw of config4_molecule
        rows.append((seed, nat))
    pick = set()
    cost = {}
    for lo, hi in BANDS:
        band = [sd for sd, nat in rows if lo <= nat <= hi]
        # cheapest few of the band by primitives first, then exact cost
        cand = sorted(band, key=lambda sd: synth.config4_sizes(sd)[3])[:3]
        for sd in cand:
            cost[sd] = _cost(sd)
        pick.add(min(cand, key=lambda sd: cost[sd]))
    small = sorted(seeds, key=lambda sd: synth.config4_sizes(sd)[3])[:6]
    pick.update(small)
    return sorted(pick)


def _molecules(seeds):
    out = []
    for sd in seeds:
        s, nrad, nang = synth.config4_molecule(sd)
        out.append(("seed%d" % sd, s, nrad, nang))
    return out


def test_generator_spans_the_stated_ranges():
    sizes = [synth.config4_sizes(sd) for sd in list(synth.CONFIG4_SEEDS)[:64]]
    atoms = [a for a, *_ in sizes]
    prims = [p for _, p, *_ in sizes]
    assert min(atoms) >= 4 and max(atoms) <= 20 and max(atoms) - min(atoms) >= 12
    assert min(prims) >= 30 and max(prims) <= 300 and max(prims) >= 250


def test_config4_subset_batched(subset):
    mols = _molecules(subset)
    names = [m[0] for m in mols]
    atoms = sorted(len(m[1].an) for m in mols)
    print("subset: %d molecules, atoms %s" % (len(mols), atoms))
    assert atoms[0] <= 6 and atoms[-1] >= 19
    # FP64 in several groups (a small arena budget), with single-context checks
    r64, g64 = BS.evaluate_stream(mols, capi.INCA_FP64, arena_bytes=1 << 24,
                                  single=set(names), log=print, solo_bytes=1 << 40)
    assert len(g64) >= 2 and all(g["replays_bitwise_equal"] for g in g64)
    assert not any(g.get("single_context") for g in g64)
    for m in r64:
        rel = abs(m["integral"] - m["single_fp64"]) / abs(m["single_fp64"])
        print("%s: %d atoms, %d prims, %d points, %.3g evals, sum-rule error %.3e, "
              "batch vs single %.1e" % (m["name"], m["atoms"], m["nprim"], m["points"],
                                        m["reference_evals"], m["sum_rule_err"], rel))
        assert rel < 1e-12
    # the same molecules in one group give the same integrals
    r1, g1 = BS.evaluate_stream(mols, capi.INCA_FP64, arena_bytes=1 << 40)
    assert len(g1) == 1
    for a, b in zip(r64, r1):
        assert abs(a["integral"] - b["integral"]) <= 1e-13 * abs(b["integral"])
    # molecules larger than the arena go alone through their own context
    rs, gs = BS.evaluate_stream(mols, capi.INCA_FP64, arena_bytes=1 << 40,
                                solo_bytes=sorted(m["arena_bytes"] for m in r64)[len(mols) // 2])
    assert any(g.get("single_context") for g in gs) and not all(g.get("single_context") for g in gs)
    for a, b in zip(rs, r64):
        assert a["name"] == b["name"]
        assert abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"])
    # mixed precision, batched, against the FP64 batch
    rmx, gmx = BS.evaluate_stream(mols, capi.INCA_MIXED, arena_bytes=1 << 40)
    for a, b in zip(rmx, r64):
        rel = abs(a["integral"] - b["integral"]) / abs(b["integral"])
        print("%s mixed vs FP64: %.2e" % (a["name"], rel))
        assert rel < 1e-8
    print("groups FP64 %s, mixed %s" % (g1, gmx))


def test_config4_subset_analytic_sum_rule(subset):
    """With exact polynomials the prepared terms integrate (analytically) to
    N(N-1)/2 for every molecule of the subset."""
    for name, s, nrad, nang in _molecules(subset):
        e, cnt = BS.prepare_molecule(s, 50, 110, capi.INCA_FP64, exact=True)
        try:
            exact = e.analytic_integral()
        finally:
            e.close()
        npairs = s.nelec * (s.nelec - 1) / 2.0
        print("%s: analytic %.10f vs N(N-1)/2 = %g" % (name, exact, npairs))
        assert abs(exact - npairs) < 1e-6 * npairs
