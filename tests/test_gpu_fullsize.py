"""The radial-plot and cube workloads at their full size, through properties
that do not depend on size.

The workload: 10-atom C/N/O/H cluster, 151 s/p/d primitives, every
symmetry-unique DM2 record (65,855,026), and s = 0..10 in steps of 0.05 on
LD0302 (32,502 points).  The oracle cannot finish this size in seconds, so
the checks are internal:
  * the record count is the closed form npair (npair + 1) / 2 with
    npair = nprim (nprim + 1) / 2, and the screens only remove records;
  * the evaluation by shell-quartet groups (the default) against the record
    kernels on the same prepared quartets: the ray tiles (8 points per
    lane, exp(-s) by recurrence) and the point-by-point kernel;
  * the per-point cutoff against the full quartet x point loop;
  * mixed precision against FP64 (1e-8 on the integral, the precision
    target of the mixed path);
  * two evaluations bitwise identical (fixed-order reductions);
  * the 80^3 cube of the same records by the per-axis contraction against
    the term-by-term evaluation by groups.
The FP64 bar is the one of the oracle parity tests: 1e-10 relative wherever
|I| >= 1e-6 max|I|, and 1e-13 of max|I| everywhere.  The oracle-based parity
of each kernel is in test_gpu_parity.py and test_gpu_ch4.py at sizes the
oracle finishes."""
import numpy as np
import pytest

from inca_parallel_amd import capi, synth

pytestmark = pytest.mark.gpu
THRESH = 1e-12


def _close(a, b):
    """(max relative error where |b| >= 1e-6 max|b|, max |a-b| / max|b|)"""
    scale = np.abs(b).max()
    big = np.abs(b) >= 1e-6 * scale
    err_rel = (np.abs(a - b)[big] / np.abs(b[big])).max() if big.any() else 0.0
    return err_rel, np.abs(a - b).max() / scale


def _integral(rad, radi):
    """The integral of I over all space from the radial values on equally
    spaced radii, up to the constant step: sum_s rintra(s) s^2."""
    return float(np.sum(rad * radi * radi))


@pytest.fixture(scope="module")
def full():
    import torch
    s = synth.config_cluster(seed=2, natoms=10)
    recs = s.dm2_unique_torch(device="cuda")
    e = capi.Engine(0)
    p = s.prims()
    e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    nradi, G = e.grid_radial_plot([[0.0, 10.0]], [0.05], [302])
    e.set_quartets(device_ptr=recs.data_ptr(), n=recs.shape[0])
    n = recs.shape[0]
    del recs
    torch.cuda.empty_cache()
    e.set_tiles(8, 8)
    iv, cnt = e.evaluate(capi.INCA_FP64, THRESH)
    print("full size: %d records, counts %s, %d radii, %d points" % (n, list(cnt), nradi, G),
          flush=True)
    radi, _ = e.radial_layout()
    rad = e.radial_values().copy()
    yield dict(s=s, e=e, n=n, nradi=nradi, G=G, iv=iv.copy(), cnt=cnt.copy(), rad=rad,
               radi=radi, tot=_integral(rad, radi))
    e.close()


def test_record_count_and_screens(full):
    nprim = full["s"].nprim
    npair = nprim * (nprim + 1) // 2
    assert nprim == 151
    assert full["n"] == npair * (npair + 1) // 2 == 65855026
    cnt = full["cnt"]
    assert cnt[0] == full["n"]
    assert 0 < cnt[2] <= cnt[1] <= cnt[0]
    assert (full["nradi"], full["G"]) == (201, 32502)


def test_rays_match_point_by_point(full):
    """Record by record: the ray tiles (8 points per lane) and the
    point-by-point kernel, each against the fixture's evaluation by
    shell-quartet groups."""
    e = full["e"]
    out = {}
    try:
        e.set_grouping(0)
        for P in (8, 1):
            e.set_tiles(P, 8)
            iv, cnt = e.evaluate(capi.INCA_FP64, THRESH)
            assert list(cnt) == list(full["cnt"])
            assert e.eval_flops() == 0.0  # record by record
            out[P] = (iv.copy(), e.radial_values().copy())
    finally:
        e.set_tiles(8, 8)
        e.set_grouping(1)
    assert not np.array_equal(out[1][0], out[8][0])  # two different kernels ran
    for P in (8, 1):
        er, ea = _close(out[P][0], full["iv"])
        rr, ra = _close(out[P][1], full["rad"])
        print("records P=%d vs groups: Ivec rel %.2e abs %.2e; I(s) rel %.2e abs %.2e"
              % (P, er, ea, rr, ra))
        assert er < 1e-10 and ea < 1e-13
        assert rr < 1e-10 and ra < 1e-13


def test_cutoff_matches_full_loop(full):
    e = full["e"]
    try:
        e.set_point_cutoff(0)
        iv0, _ = e.evaluate(capi.INCA_FP64, THRESH)
        ref0, ex0 = e.eval_counts()
    finally:
        e.set_point_cutoff(100)
    assert ref0 == int(full["cnt"][2]) * full["G"]
    # the full loop executes every (group, point) pair
    assert ex0 == e.groups()[0] * full["G"]
    er, ea = _close(full["iv"], iv0)
    print("cutoff vs full loop: rel %.2e abs %.2e" % (er, ea))
    assert er < 1e-10 and ea < 1e-13


def test_mixed_within_target_of_fp64(full):
    e = full["e"]
    e.evaluate(capi.INCA_MIXED, THRESH, want_ivec=False)
    rad = e.radial_values()
    tot = _integral(rad, full["radi"])
    n64, n32, _ = e.mixed_split()
    print("mixed: %d FP64 + %d FP32 quartets, integral rel err %.2e" % (
        n64, n32, abs(tot - full["tot"]) / abs(full["tot"])))
    assert n32 > 0 and n64 + n32 == full["cnt"][2]
    assert abs(tot - full["tot"]) <= 1e-8 * abs(full["tot"])
    rr, _ = _close(rad, full["rad"])
    assert rr < 1e-8


def test_repeat_is_bitwise_identical(full):
    iv, cnt = full["e"].evaluate(capi.INCA_FP64, THRESH)
    assert list(cnt) == list(full["cnt"])
    assert np.array_equal(iv, full["iv"])


def test_cube_contraction_matches_ray_tiles(full):
    """The 80^3 cube (0.15 bohr steps, centred at the origin) on the same
    records: the per-axis contraction (every quartet at every voxel) on the
    vector units and on the matrix cores, against the term-by-term
    evaluation by shell-quartet groups (cutoff on).  Runs last: it replaces
    the grid."""
    e = full["e"]
    n = e.grid_cube([0.0, 0.0, 0.0], [0.15, 0.15, 0.15], [80, 80, 80])
    assert n == 512000
    got = {}
    try:
        e.set_cube_contraction(0)
        ray, cnt = e.evaluate(capi.INCA_FP64, THRESH)
        ray = ray.copy()
        for mode in (1, 2):
            e.set_cube_contraction(mode)
            iv, cnt2 = e.evaluate(capi.INCA_FP64, THRESH)
            ref_n, ex = e.eval_counts()
            assert list(cnt2) == list(full["cnt"])
            assert ex == ref_n == int(cnt[2]) * n
            got[mode] = iv.copy()
    finally:
        e.set_cube_contraction(2)
    assert list(cnt) == list(full["cnt"])
    for mode, iv in got.items():
        er, ea = _close(iv, ray)
        print("80^3 cube, contraction mode %d vs groups: rel %.2e abs %.2e" % (mode, er, ea))
        assert er < 1e-10 and ea < 1e-13
    full["cube_fp64"] = got[1]


def test_cube_mixed_per_voxel(full):
    """The 80^3 cube evaluated mixed: the FP32 range on FP32 factors and the
    FP32 matrix cores, folded into FP64, against the FP64 contraction at
    every voxel (runs after the cube test above, on its grid)."""
    e = full["e"]
    if "cube_fp64" not in full:
        pytest.skip("needs the FP64 cube of test_cube_contraction_matches_ray_tiles")
    ref = full["cube_fp64"]
    mixed, cnt = e.evaluate(capi.INCA_MIXED, THRESH)
    n64, n32, _ = e.mixed_split()
    er, ea = _close(mixed, ref)
    es = abs(mixed.sum() - ref.sum()) / abs(ref.sum())
    print("80^3 cube mixed (%d FP64 + %d FP32 quartets) vs FP64: rel %.2e abs %.2e sum %.2e"
          % (n64, n32, er, ea, es))
    assert n32 > n64 > 0
    assert ea < 1e-9 and es < 1e-9
    assert not np.array_equal(mixed, ref)
