"""One-electron cube files of &cube_job (mod_cubefile.f90:16-128 with the
functions of mod_functions.f90) on the device, against the oracle's
restatement (orc_onee, pinned in test_oracle.py by finite differences and
integrals): through the C-ABI at 1e-12 relative, and through the Fortran
executable on a .wfx input at the files' printed precision (E16.6: six
significant digits, so 1e-5 relative)."""
import os
import subprocess

import numpy as np
import pytest

from inca_parallel_amd import capi, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "inca_parallel_amd", "bin", "inca")
CENTER, STEP, NP = [0.1, 0.2, -0.1], [0.25, 0.3, 0.2], [14, 11, 16]


@pytest.fixture(scope="module")
def system():
    rng = np.random.default_rng(41)
    elements = ["O", "H", "H"]
    coords = np.array([[0.0, 0.0, 0.0], [1.43, 1.1, 0.0], [-1.43, 1.1, 0.3]])
    alpha, tmn, cen, ra = synth.make_basis(rng, elements, coords, {0: 2, 1: 1, 2: 1}, {0: 2},
                                           0.2, 3.0)
    alpha = np.array([float("%.6g" % a) for a in alpha])
    C = synth.orbitals(rng, synth.overlap_matrix(alpha, tmn, cen), 3)
    s = synth.System(elements, coords, alpha, tmn, cen, ra, C)
    norm = np.array([synth.normprim(a, t) for a, t in zip(alpha, tmn)])
    return s, (s.C * norm[:, None]).T


def _cmp(a, b, rel, absr):
    scale = np.abs(b).max()
    big = np.abs(b) >= 1e-6 * scale
    er = (np.abs(a - b)[big] / np.abs(b[big])).max()
    ea = np.abs(a - b).max() / scale
    return er < rel and ea < absr, (er, ea)


def test_onee_matches_oracle(system):
    s, T = system
    p = s.prims()
    with capi.Engine(0) as e:
        e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"], normalize=False)
        n = e.grid_cube(CENTER, STEP, NP)
        r = e.get_grid()["r"]
        assert n == np.prod(NP)
        cases = [(1, q, None, None) for q in (1, 5, s.nprim)]
        cases += [(3, i, T, None) for i in (1, 2, 3)]
        cases += [(5, 0, T, [2.0] * 3), (6, 0, T, [2.0] * 3), (5, 0, T[:2], [1.0] * 2)]
        for kind, idx, TT, w in cases:
            got = e.onee(kind, idx, TT, w)
            ref = O.onee(kind, idx, p, r, TT, w)
            ok, err = _cmp(got, ref, 1e-12, 1e-14)
            print("kind %d index %d: rel %.2e abs %.2e" % (kind, idx, *err))
            assert ok
        with pytest.raises(capi.IncaError, match="kind 2"):
            e.onee(2, 1, T, [1.0] * 3)
        with pytest.raises(capi.IncaError, match="outside"):
            e.onee(1, s.nprim + 1)


def test_cube_job_through_the_executable(system, tmp_path):
    s, T = system
    synth.write_wfx(s, str(tmp_path / "w.wfx"))
    inp = ("&files\n  wfxfilename = 'w.wfx'\n/\n&intracule_job\n/\n&cube_job\n"
           "  cube = .true., primcube = .true., mocube = .true., denscube = .true.,\n"
           "  laplacian = .true., npr = 4, mo = 2,\n"
           "  center = %s,\n  step = %s,\n  np = %s\n/\n" % (
               ", ".join(map(str, CENTER)), ", ".join(map(str, STEP)),
               ", ".join(map(str, NP))))
    (tmp_path / "job.inp").write_text(inp)
    r = subprocess.run([EXE, "job.inp"], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout
    p = s.prims()
    pts = O.gridpoints3(CENTER, STEP, NP)
    refs = {"job_prim.cube": O.onee(1, 4, p, pts),
            "job_mo.cube": O.onee(3, 2, p, pts, T),
            "job_dens.cube": O.onee(5, 0, p, pts, T, [2.0] * 3),
            "job_lapl.cube": O.onee(6, 0, p, pts, T, [2.0] * 3)}
    for name, ref in refs.items():
        lines = open(tmp_path / name).read().splitlines()
        head = [l.split() for l in lines[:6 + len(s.an)]]
        assert head[0] == ["CUBE", "FILE"]
        assert [int(head[3][0]), int(head[4][0]), int(head[5][0])] == NP
        assert abs(float(head[2][1]) - CENTER[0]) < 1e-12
        vals = np.array([float(l) for l in lines[6 + len(s.an):]])
        assert len(vals) == np.prod(NP)
        ok, err = _cmp(vals, ref, 1e-5, 1e-5)
        print("%s: rel %.2e abs %.2e" % (name, *err))
        assert ok
