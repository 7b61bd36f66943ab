"""CH4 end-to-end parity through the C-ABI on the GPU.

The reference's own CH4 goldens (tests/run_tests.py:19-28, ch4_ref.rad,
ch4_ref.cube) and the oracle on the same inputs are the checkers.  FP64
bar: 1e-10 relative (the .rad golden carries 8 printed digits, so it is
compared at its printing precision).  Mixed-precision bar: 1e-8 relative on
the integral, 1e-6 relative per point.
"""
import numpy as np
import pytest

import ch4_golden as GD
from inca_parallel_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THRESH = 1e-12


def _setup(engine, ch4):
    engine.set_primitives(ch4["alpha"], ch4["tmn"], ch4["x"], ch4["y"], ch4["z"])
    engine.set_quartets(ijkl=ch4["ijkl"], d=ch4["d"])


def _multicentre(engine, ch4):
    cent, Ps, sf = O.centercalc(ch4["an"], ch4["cartes"], True, False)
    nq = len(Ps)
    return engine.grid_multicenter(cent, [50] * nq, [110] * nq, sf, Ps), (cent, Ps, sf)


def test_radial_integral_fp64(engine, ch4):
    _setup(engine, ch4)
    (maxgrid, rgrid, rrgrid), _ = _multicentre(engine, ch4)
    assert (maxgrid, rgrid, rrgrid) == (5500, 3050, 3050)
    iv, cnt = engine.evaluate(capi.INCA_FP64, THRESH)
    tot, vee, ipc, _ = engine.integrate()
    print("integral %.17g counts %s" % (tot, cnt))
    assert cnt[0] == 142884
    assert abs(tot - GD.INTEGRAL) / GD.INTEGRAL < 1e-10
    assert abs((tot - GD.NELEC * (GD.NELEC - 1) / 2) - GD.INTEGRAL_ERR) < 1e-8
    assert abs(ipc.sum() - tot) < 1e-12 * tot


def test_radial_integral_matches_oracle_pointwise(engine, ch4):
    _setup(engine, ch4)
    _multicentre(engine, ch4)
    iv, cnt = engine.evaluate(capi.INCA_FP64, THRESH)
    grid = engine.get_grid()
    ref, rcnt = O.intracule(ch4, ch4["ijkl"], ch4["d"], grid["r"], THRESH)
    assert list(cnt[:3]) == list(rcnt[:3])
    rel = np.abs(iv - ref) / np.maximum(np.abs(ref), 1e-300)
    print("max rel pointwise", rel.max())
    assert rel.max() < 1e-10


def test_zero_and_two_points(engine, ch4):
    _setup(engine, ch4)
    engine.grid_points(np.zeros((1, 3)))
    iv, _ = engine.evaluate(capi.INCA_FP64, THRESH)
    assert abs(iv[0] / 2 - GD.ZERO_HALF) / GD.ZERO_HALF < 1e-10
    engine.grid_points(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]))
    iv, _ = engine.evaluate(capi.INCA_FP64, THRESH)
    for got, exp in zip(iv, GD.TWO_POINTS):
        assert abs(got - exp) / exp < 1e-10


def test_radial_plot(engine, ch4):
    _setup(engine, ch4)
    nradi, ng = engine.grid_radial_plot([[0.0, 5.0]], [0.05], [110])
    assert nradi == 101
    engine.evaluate(capi.INCA_FP64, THRESH)
    rint = engine.radial_values()
    ref = GD.rad_table()
    radi, smn = engine.radial_layout()
    assert np.abs(radi - ref[:, 0]).max() < 1e-12
    # golden printed with 8 significant digits (ES20.8)
    assert (np.abs(rint - ref[:, 1]) / np.abs(ref[:, 1])).max() < 1e-8
    # and the oracle on the same grid at FP64 parity
    grid = engine.get_grid()
    iv = engine.ivec()
    oiv, _ = O.intracule(ch4, ch4["ijkl"], ch4["d"], grid["r"], THRESH)
    assert (np.abs(iv - oiv) / np.abs(oiv)).max() < 1e-10


def test_cube(engine, ch4):
    _setup(engine, ch4)
    n = engine.grid_cube([0.0, 0.0, 0.0], [0.1, 0.1, 0.1], [20, 20, 20])
    assert n == 8000
    iv, _ = engine.evaluate(capi.INCA_FP64, THRESH)
    ref = GD.cube_values()
    assert (np.abs(iv - ref) / np.abs(ref)).max() < 1e-10


def test_mixed_precision_integral(engine, ch4):
    _setup(engine, ch4)
    _multicentre(engine, ch4)
    iv64, _ = engine.evaluate(capi.INCA_FP64, THRESH)
    t64 = engine.integrate()[0]
    ivm, _ = engine.evaluate(capi.INCA_MIXED, THRESH)
    tm = engine.integrate()[0]
    rel = abs(tm - t64) / abs(t64)
    big = np.abs(iv64) > 1e-12 * np.abs(iv64).max()
    pt = (np.abs(ivm - iv64)[big] / np.abs(iv64[big])).max()
    print("mixed integral rel err %.3e, max pointwise %.3e" % (rel, pt))
    assert rel < 1e-8
    assert pt < 1e-5


def test_mixed_tail_sweep(engine, ch4):
    """Error of the mixed path against FP64 as the FP32 share grows."""
    _setup(engine, ch4)
    _multicentre(engine, ch4)
    engine.evaluate(capi.INCA_FP64, THRESH)
    t64 = engine.integrate()[0]
    errs = {}
    for tail in (1.0, 0.2, 0.05, 0.02, 0.005):
        engine.set_mixed_tail(tail)
        engine.evaluate(capi.INCA_MIXED, THRESH)
        tm = engine.integrate()[0]
        n64, n32, tau = engine.mixed_split()
        errs[tail] = abs(tm - t64) / abs(t64)
        print("tail %-6g fp64 %7d fp32 %7d tau %.3e  integral rel err %.3e" % (
            tail, n64, n32, tau, errs[tail]))
    engine.set_mixed_tail(0.02)
    assert errs[0.02] < 1e-8


def test_config_ch4_flat_stream(tmp_path, engine):
    """The CPU-runnable configuration as specified: CH4 with 27 seeded
    primitives and all 27^4 = 531,441 records written as a flat .dm2p
    stream (12.8 MB), read by the engine's reader, evaluated on the
    50 x LD0110 autogrid.  Checks: the records obey the pair sum rule
    sum D S_ik S_jl = N(N-1) = 90; Ivec matches the oracle to 1e-10; the
    analytic integral of the prepared terms is N(N-1)/2 = 45 up to the
    screened terms; the grid integral equals the oracle's (the 50 x 110
    single-centre grid itself misses 45 by its quadrature error)."""
    from inca_parallel_amd import synth
    s = synth.config_ch4(seed=1)
    ijkl, d = s.dm2_full()
    assert len(d) == 27 ** 4
    assert abs(s.sum_rule(ijkl, d) - 90.0) < 1e-9
    rec = np.empty(len(d), dtype=[("ijkl", "<i4", 4), ("d", "<f8")])
    rec["ijkl"], rec["d"] = ijkl, d
    path = tmp_path / "ch4_flat.dm2p"
    rec.tofile(path)
    assert path.stat().st_size == 24 * 531441
    p = s.prims()
    engine.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    nrec, fmt = engine.load_dm2p(str(path))
    assert (nrec, fmt) == (531441, 0)
    cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
    engine.grid_multicenter(cent, [50] * len(Ps), [110] * len(Ps), sf, Ps)
    iv, cnt = engine.evaluate(capi.INCA_FP64, THRESH)
    g = engine.get_grid()
    ref, rcnt = O.intracule(p, ijkl, d, g["r"], THRESH)
    assert list(cnt) == list(rcnt)
    # far points underflow to zero: relative where |I| >= 1e-6 max|I|
    big = np.abs(ref) >= 1e-6 * np.abs(ref).max()
    assert np.max(np.abs(iv - ref)[big] / np.abs(ref[big])) < 1e-10
    assert np.max(np.abs(iv - ref)) < 1e-13 * np.abs(ref).max()
    tot = engine.integrate()[0]
    assert abs(tot - np.sum(g["w"] * ref)) < 1e-10 * abs(tot)
    exact = engine.analytic_integral()
    print("config 1: grid integral %.12f (error %.3e), analytic %.12f, counts %s"
          % (tot, tot - 45.0, exact, cnt))
    assert abs(exact - 45.0) < 1e-6
