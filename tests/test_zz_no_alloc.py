"""inca_evaluate_prepared_noalloc against inca_evaluate_prepared.

The no-allocation entry point evaluates record by record, point by point,
with the cutoff off, its partial sums in a static per-device array; the
allocating entry point, run with the cutoff off and without shell-quartet
groups, is the checker.  Two calls in a row, the
second on a smaller input, so that leftovers of the first in the shared
array would show; then a grid too large for the array must be refused.
FP64 bar: 1e-12 of max|I| (the two differ only in summation order and, on
radial scans, in the allocating path's ray recurrence); mixed: 1e-12 of
max|I| against the allocating mixed path with the same split.
"""
import numpy as np
import pytest

from inca_parallel_amd import capi, synth

pytestmark = pytest.mark.gpu
THRESH = 1e-12


def _device_out(G):
    import torch
    return torch.full((max(G, 1),), float("nan"), dtype=torch.float64, device="cuda")


def _noalloc(eng, precision, G, stream=None):
    import torch
    out = _device_out(G)
    torch.cuda.synchronize()
    eng.evaluate_prepared_noalloc(precision, out.data_ptr(), stream)
    torch.cuda.synchronize()
    return out[:G].cpu().numpy()


def _err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_noalloc_matches_allocating_entry_point(ch4):
    import torch
    s = synth.config_cluster(seed=21, natoms=4, nheavy=3, nocc=4,
                             shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
    ijkl, d = s.dm2_unique()
    p = s.prims()
    with capi.Engine(0) as big, capi.Engine(0) as small:
        # call 1: d functions on a radial scan (about 16,000 points), FP64
        big.set_point_cutoff(0)
        big.set_grouping(0)
        big.set_tiles(4, 4)
        big.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
        big.set_quartets(ijkl=ijkl, d=d)
        big.grid_radial_plot(np.array([[0.0, 5.0]]), np.array([0.05]), np.array([302], np.int32))
        big.prepare(capi.INCA_FP64, THRESH)
        ref1 = big.evaluate_prepared(capi.INCA_FP64)
        got1 = _noalloc(big, capi.INCA_FP64, big.G)
        print("call 1: %d points, rel err %.2e" % (big.G, _err(got1, ref1)))
        assert big.G > 10000
        assert np.all(np.isfinite(got1)) and _err(got1, ref1) < 1e-12

        # call 2, a smaller input: CH4 at three points, on another context
        # and on torch's current stream
        small.set_point_cutoff(0)
        small.set_grouping(0)
        small.set_tiles(4, 4)
        small.set_primitives(ch4["alpha"], ch4["tmn"], ch4["x"], ch4["y"], ch4["z"])
        small.set_quartets(ijkl=ch4["ijkl"], d=ch4["d"])
        small.grid_points(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.3, -0.4, 0.8]]))
        small.prepare(capi.INCA_FP64, THRESH)
        ref2 = small.evaluate_prepared(capi.INCA_FP64)
        got2 = _noalloc(small, capi.INCA_FP64, 3, torch.cuda.current_stream().cuda_stream)
        print("call 2: rel err %.2e" % _err(got2, ref2))
        assert _err(got2, ref2) < 1e-12
        # the reference's printed values at (0,0,0) and (1,0,0)
        assert abs(got2[0] - 15.540540526960413) < 1e-10 * 15.54
        assert abs(got2[1] - 1.2854504822876072) < 1e-10 * 1.29

        # mixed precision on the multicentre autogrid of the cluster, twice
        # (the second call reuses the array the first filled)
        from oracle import oracle as O
        cent, Ps, sf = O.centercalc(p["an"], p["cartes"], True, False)
        nq = len(Ps)
        big.grid_multicenter(cent, [40] * nq, [110] * nq, sf, Ps)
        big.prepare(capi.INCA_MIXED, THRESH)
        n64, n32, _ = big.mixed_split()
        assert n64 > 0 and n32 > 0
        ref3 = big.evaluate_prepared(capi.INCA_MIXED)
        got3 = _noalloc(big, capi.INCA_MIXED, big.G)
        got3b = _noalloc(big, capi.INCA_MIXED, big.G)
        print("mixed: %d points, rel err %.2e" % (big.G, _err(got3, ref3)))
        assert _err(got3, ref3) < 1e-12
        assert np.array_equal(got3, got3b)


def test_noalloc_refuses_a_grid_beyond_its_scratch(ch4):
    with capi.Engine(0) as eng:
        eng.set_point_cutoff(0)
        eng.set_grouping(0)
        eng.set_primitives(ch4["alpha"], ch4["tmn"], ch4["x"], ch4["y"], ch4["z"])
        eng.set_quartets(ijkl=ch4["ijkl"], d=ch4["d"])
        eng.prepare(capi.INCA_FP64, THRESH)
        n = 4_500_000  # one FP64 partial row alone is 36 MB > 32 MiB
        r = np.zeros((n, 3))
        r[:, 0] = np.linspace(0.0, 5.0, n)
        eng.grid_points(r)
        out = _device_out(1)
        with pytest.raises(capi.IncaError) as e:
            eng.evaluate_prepared_noalloc(capi.INCA_FP64, out.data_ptr())
        assert "scratch" in str(e.value)
        assert np.isnan(out.cpu().numpy()[0])  # nothing was enqueued
