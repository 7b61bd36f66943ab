"""Evaluation by shell-quartet groups (the default) against the oracle and
against the record-by-record loop on the same prepared quartets.

The records of one shell quartet share sqe and the offset sqe (P_ik - P_jl),
so their terms add into exp(-|rp|^2) times one three-variable polynomial;
the results may differ from the record loop only by summation order.  FP64
bar against the oracle: 1e-10 relative wherever |I| >= 1e-6 max|I| and
1e-13 of max|I| everywhere (the parity bar of test_gpu_parity.py); against
the record loop: 1e-12 relative on the same points.  Mixed: 1e-8 on the
integral against FP64, the target of the mixed path."""
import numpy as np
import pytest

from inca_parallel_amd import capi, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THRESH = 1e-12


def _close(a, b):
    scale = np.abs(b).max()
    big = np.abs(b) >= 1e-6 * scale
    err_rel = (np.abs(a - b)[big] / np.abs(b[big])).max() if big.any() else 0.0
    return err_rel, np.abs(a - b).max() / scale


@pytest.fixture(scope="module")
def dsys():
    """4 atoms (3 heavy) with s, p and d shells: groups of every total degree
    0..8."""
    s = synth.config_cluster(seed=21, natoms=4, nheavy=3, nocc=4,
                             shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
    ijkl, d = s.dm2_unique()
    return s, ijkl, d


def _setup(e, s, ijkl, d):
    p = s.prims()
    e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    e.set_quartets(ijkl=ijkl, d=d)
    return p


def test_groups_match_oracle_and_records(dsys):
    s, ijkl, d = dsys
    with capi.Engine(0) as e:
        p = _setup(e, s, ijkl, d)
        cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
        nq = len(Ps)
        e.grid_multicenter(cent, [40] * nq, [110] * nq, sf, Ps, nosym=True)
        iv, cnt = e.evaluate(capi.INCA_FP64, THRESH)
        ng, ng64, ncoef = e.groups()
        flops = e.eval_flops()
        g = e.get_grid()
        ref, rcnt = O.intracule(p, ijkl, d, g["r"], THRESH)
        assert list(cnt[:3]) == list(rcnt[:3])
        # fewer groups than records; every group at least one coefficient,
        # at most 165 (total degree 8)
        assert 0 < ng < cnt[2] and ng64 == ng and ng <= ncoef <= 165 * ng
        assert flops > 0.0
        er, ea = _close(iv, ref)
        print("groups (%d for %d records, %d coefficients) vs oracle: rel %.2e abs %.2e"
              % (ng, cnt[2], ncoef, er, ea))
        assert er < 1e-10 and ea < 1e-13
        e.set_grouping(0)
        iv0, cnt0 = e.evaluate(capi.INCA_FP64, THRESH)
        assert e.groups()[0] == 0 and e.eval_flops() == 0.0
        assert list(cnt0) == list(cnt)
        er0, ea0 = _close(iv, iv0)
        print("groups vs records: rel %.2e abs %.2e" % (er0, ea0))
        assert er0 < 1e-12 and ea0 < 1e-14
        # the integrals through the grid's weights
        assert abs(np.sum(g["w"] * iv) - np.sum(g["w"] * ref)) < 1e-10 * abs(np.sum(g["w"] * ref))


def test_groups_on_radial_scan_and_points(dsys, ch4):
    s, ijkl, d = dsys
    with capi.Engine(0) as e:
        p = _setup(e, s, ijkl, d)
        e.grid_radial_plot([[0.0, 4.0]], [0.1], [110])
        iv, _ = e.evaluate(capi.INCA_FP64, THRESH)
        assert e.eval_flops() > 0.0
        g = e.get_grid()
        ref, _ = O.intracule(p, ijkl, d, g["r"], THRESH)
        er, ea = _close(iv, ref)
        print("radial scan by groups vs oracle: rel %.2e abs %.2e" % (er, ea))
        assert er < 1e-10 and ea < 1e-13
        # the reference's CH4 values at two points, by groups
        e.set_primitives(ch4["alpha"], ch4["tmn"], ch4["x"], ch4["y"], ch4["z"])
        e.set_quartets(ijkl=ch4["ijkl"], d=ch4["d"])
        e.grid_points(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]))
        two, _ = e.evaluate(capi.INCA_FP64, THRESH)
        assert e.groups()[0] > 0
        assert abs(two[0] - 15.540540526960413) < 1e-10 * 15.54
        assert abs(two[1] - 1.2854504822876072) < 1e-10 * 1.29


def test_ch4_integral_by_groups(ch4):
    with capi.Engine(0) as e:
        e.set_primitives(ch4["alpha"], ch4["tmn"], ch4["x"], ch4["y"], ch4["z"])
        e.set_quartets(ijkl=ch4["ijkl"], d=ch4["d"])
        cent, Ps, sf = O.centercalc(ch4["an"], ch4["cartes"], True, False)
        e.grid_multicenter(cent, [50] * len(Ps), [110] * len(Ps), sf, Ps)
        e.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
        ng, _, _ = e.groups()
        tot = e.integrate()[0]
        print("CH4 by %d groups: integral %.17g" % (ng, tot))
        assert ng > 0
        assert abs(tot - 45.013363086638655) < 1e-10 * 45.0


def test_mixed_groups_within_target(dsys):
    s, ijkl, d = dsys
    with capi.Engine(0) as e:
        p = _setup(e, s, ijkl, d)
        cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
        nq = len(Ps)
        e.grid_multicenter(cent, [40] * nq, [110] * nq, sf, Ps, nosym=True)
        ref, _ = e.evaluate(capi.INCA_FP64, THRESH)
        tot_ref = e.integrate()[0]
        iv, _ = e.evaluate(capi.INCA_MIXED, THRESH)
        tot = e.integrate()[0]
        ng, ng64, _ = e.groups()
        print("mixed groups: %d FP64 of %d; integral rel err %.2e" % (
            ng64, ng, abs(tot - tot_ref) / abs(tot_ref)))
        assert 0 < ng64 < ng
        assert abs(tot - tot_ref) <= 1e-8 * abs(tot_ref)
        assert not np.array_equal(iv, ref)  # the FP32 range ran in FP32
        _, ea = _close(iv, ref)
        assert ea < 1e-8


def test_groups_repeat_bitwise_and_tiles(dsys):
    """Two preparations and evaluations give the same bits (stable sort,
    members summed in record order, fixed-order reductions); every points-
    per-lane choice agrees to summation order."""
    s, ijkl, d = dsys
    with capi.Engine(0) as e:
        p = _setup(e, s, ijkl, d)
        cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
        nq = len(Ps)
        e.grid_multicenter(cent, [30] * nq, [110] * nq, sf, Ps, nosym=True)
        a, _ = e.evaluate(capi.INCA_FP64, THRESH)
        b, _ = e.evaluate(capi.INCA_FP64, THRESH)
        assert np.array_equal(a, b)
        best = e.autotune()
        assert best[0] in (1, 2, 4, 6, 8)
        try:
            for P in (1, 2, 4, 6, 8):
                e.set_tiles(P, 4)
                c, _ = e.evaluate(capi.INCA_FP64, THRESH)
                er, ea = _close(c, a)
                print("groups, %d points per lane: rel %.2e abs %.2e" % (P, er, ea))
                assert er < 1e-13 and ea < 1e-15
        finally:
            e.set_tiles(4, 4)


def test_groups_built_chunk_by_chunk_without_records(dsys):
    """inca_set_keep_records(ctx, 0): the groups are built per pre-pass chunk
    (here 20,000 records) and no prepared record is kept; a shell quartet
    split between chunks becomes two groups.  Same counters, same values up
    to summation order; every record-by-record use reports the release."""
    s, ijkl, d = dsys
    with capi.Engine(0) as a, capi.Engine(0) as b:
        for e in (a, b):
            p = _setup(e, s, ijkl, d)
            cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
            nq = len(Ps)
            e.grid_multicenter(cent, [40] * nq, [110] * nq, sf, Ps, nosym=True)
        b.set_keep_records(0)
        b.set_prepare_chunk(20000)
        for prec in (capi.INCA_FP64, capi.INCA_MIXED):
            ref, cnt = a.evaluate(prec, THRESH)
            iv, cnt2 = b.evaluate(prec, THRESH)
            assert list(cnt2) == list(cnt)
            assert b.groups()[0] >= a.groups()[0]
            assert b.analytic_integral() == pytest.approx(a.analytic_integral(), rel=1e-13)
            er, ea = _close(iv, ref)
            print("chunked groups (%d vs %d) prec %d: rel %.2e abs %.2e" % (
                b.groups()[0], a.groups()[0], prec, er, ea))
            if prec == capi.INCA_FP64:
                assert er < 1e-12 and ea < 1e-14
            else:
                assert ea < 1e-8
        with pytest.raises(capi.IncaError):
            b.prepared(0, 1)
        b.set_grouping(0)
        with pytest.raises(capi.IncaError):
            b.evaluate_prepared(capi.INCA_FP64)


def test_group_rays_match_point_by_point():
    """Radial scans: the groups on ray tiles (P >= 2 points per lane on one
    direction, exp(-s) by the two-product recurrence) against the groups
    point by point (P = 1), two scan blocks with different steps and orders,
    exponents up to 2e4 so that tight groups take the point-by-point branch
    inside ray tiles."""
    s = synth.config_cluster(seed=31, natoms=4, nheavy=3, nocc=4, exp_lo=0.05, exp_hi=2e4,
                             shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
    ijkl, d = s.dm2_unique()
    with capi.Engine(0) as e:
        _setup(e, s, ijkl, d)
        e.grid_radial_plot([[0.0, 3.0], [3.2, 6.0]], [0.1, 0.2], [110, 302])
        out = {}
        for P in (1, 2, 4, 6, 8):
            e.set_tiles(P, 4)
            iv, _ = e.evaluate(capi.INCA_FP64, THRESH)
            assert e.groups()[0] > 0
            out[P] = (iv.copy(), e.radial_values().copy())
        for P in (2, 4, 6, 8):
            er, ea = _close(out[P][0], out[1][0])
            rr, ra = _close(out[P][1], out[1][1])
            print("groups, P=%d rays vs point by point: Ivec rel %.2e abs %.2e; I(s) rel %.2e "
                  "abs %.2e" % (P, er, ea, rr, ra))
            assert not np.array_equal(out[P][0], out[1][0])  # the ray arithmetic ran
            assert er < 1e-11 and ea < 1e-13
            assert rr < 1e-11 and ra < 1e-13


def _used(rec):
    """the header and the coefficient words of each record"""
    cls = rec[:, 4].view(np.int64)
    n = 8 + (cls & 0xff) + ((cls >> 8) & 0xff) + ((cls >> 16) & 0xff) + 3
    mask = np.arange(rec.shape[1])[None, :] < n[:, None]
    return np.where(mask, rec, 0.0)


def _sorted_rows(rec):
    u = _used(rec)
    return u[np.lexsort(u.T[::-1])]


def test_grouped_records_compact_on_first_use(dsys):
    """A grouped preparation (one chunk) keeps its pre-pass slots as they are,
    screen-2 failures marked, and compacts them at the first record-by-record
    use.  The compacted records are the file-order preparation's records,
    bit for bit, in group order; the no-allocation entry point refuses the
    uncompacted state and afterwards matches the file-order records' result
    up to summation order."""
    import torch
    s, ijkl, d = dsys
    with capi.Engine(0) as a, capi.Engine(0) as b:
        for e in (a, b):
            p = _setup(e, s, ijkl, d)
            cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
            e.grid_multicenter(cent, [30] * len(Ps), [110] * len(Ps), sf, Ps, nosym=True)
            e.set_point_cutoff(0)
        a.set_grouping(0)
        ca = a.prepare(capi.INCA_FP64, THRESH)
        cb = b.prepare(capi.INCA_FP64, THRESH)
        assert list(ca) == list(cb) and b.groups()[0] > 0
        G = a.G
        out = torch.full((G,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with pytest.raises(capi.IncaError):
            b.evaluate_prepared_noalloc(capi.INCA_FP64, out.data_ptr(), None)
        ra, rb = a.prepared(), b.prepared()  # b compacts here
        assert ra.shape == rb.shape == (int(ca[2]), ra.shape[1])
        assert not np.array_equal(_used(ra), _used(rb))  # group order, not file order
        assert np.array_equal(_sorted_rows(ra), _sorted_rows(rb))
        res = []
        for e in (a, b):
            o = torch.full((G,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            e.evaluate_prepared_noalloc(capi.INCA_FP64, o.data_ptr(), None)
            torch.cuda.synchronize()
            res.append(o.cpu().numpy())
        er, ea = _close(res[1], res[0])
        print("no-allocation entry point, group-order records vs file order: rel %.2e abs %.2e"
              % (er, ea))
        assert er < 1e-12 and ea < 1e-14
        # the groups are untouched by the compaction
        iv = b.evaluate_prepared(capi.INCA_FP64)
        er, ea = _close(iv, res[0])
        assert er < 1e-12 and ea < 1e-14


def test_group_rays_on_cube_lines(dsys):
    """Cube grids by groups (the per-axis contraction off): eight points of
    one z line per lane, exp(-s) by the recurrence and the polynomial's
    later values from its difference table, against the groups point by
    point (P = 1) and against the oracle."""
    s, ijkl, d = dsys
    with capi.Engine(0) as e:
        p = _setup(e, s, ijkl, d)
        e.grid_cube([0.0, 0.0, 0.0], [0.2, 0.25, 0.15], [12, 10, 33])
        e.set_cube_contraction(0)
        try:
            out = {}
            for P in (1, 8):
                e.set_tiles(P, 4)
                iv, _ = e.evaluate(capi.INCA_FP64, THRESH)
                assert e.groups()[0] > 0
                out[P] = iv.copy()
            g = e.get_grid()
            ref, _ = O.intracule(p, ijkl, d, g["r"], THRESH)
        finally:
            e.set_tiles(4, 4)
            e.set_cube_contraction(2)
        er, ea = _close(out[8], out[1])
        print("cube lines, P=8 rays vs point by point: rel %.2e abs %.2e" % (er, ea))
        assert not np.array_equal(out[8], out[1])  # the ray arithmetic ran
        assert er < 1e-11 and ea < 1e-13
        er, ea = _close(out[8], ref)
        print("cube lines by groups vs oracle: rel %.2e abs %.2e" % (er, ea))
        assert er < 1e-10 and ea < 1e-13


def test_contexts_prepared_from_threads_match_sequential():
    """Contexts are independent: several prepared and evaluated from host
    threads at once give the results of the same contexts run one after the
    other, bit for bit (each context has its own stream and buffers)."""
    from concurrent.futures import ThreadPoolExecutor
    systems = [synth.config_cluster(seed=40 + k, natoms=3, nheavy=2, nocc=3,
                                    shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
               for k in range(6)]

    def run(s):
        ijkl, d = s.dm2_unique()
        with capi.Engine(0) as e:
            p = _setup(e, s, ijkl, d)
            cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
            e.grid_multicenter(cent, [30] * len(Ps), [110] * len(Ps), sf, Ps, nosym=True)
            iv, cnt = e.evaluate(capi.INCA_MIXED, THRESH)
            return iv.copy(), list(cnt), e.integrate()[0]

    seq = [run(s) for s in systems]
    with ThreadPoolExecutor(max_workers=6) as pool:
        par = list(pool.map(run, systems))
    for a, b in zip(seq, par):
        assert a[1] == b[1]
        assert np.array_equal(a[0], b[0]) and a[2] == b[2]
