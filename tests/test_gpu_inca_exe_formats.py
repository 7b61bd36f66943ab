"""Boundary formats of the drop-in, through the Fortran executable `inca`:
the .wfx reader (mod_read_files.f90:12-287) against the .fchk + .bas pair,
the automatic centres (centercalc, mod_autogrid.f90:12-131: pair centres,
the 0.5 bohr z snap, the 0.1 bohr merge) on a molecule with three heavy
atoms, and manual_grid (mod_input_reader.f90:196-206).  The oracle's
centres, grid and Ivec on the same seeded inputs are the checkers (FP64
bar 1e-10 relative on the integral).

The molecule: C at the origin, N at (2.5, 0, 0.3), O at (5, 0.05, 0.55), H
at (0, 2, 0) bohr.  The C-N and N-O displacements have |z| < 0.5 (snapped to
0) and lie within 0.1 bohr of each other (merged), so the autogrid keeps
five centres.  Coordinates and exponents have at most 8 significant digits,
so the .fchk (8 digits), .bas and .wfx (16) files all hold the same doubles.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from inca_parallel_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "inca_parallel_amd", "bin", "inca")
THRESH = 1e-12


@pytest.fixture(scope="module")
def molecule(tmp_path_factory):
    rng = np.random.default_rng(77)
    elements = ["C", "N", "O", "H"]
    coords = np.array([[0.0, 0.0, 0.0], [2.5, 0.0, 0.3], [5.0, 0.05, 0.55], [0.0, 2.0, 0.0]])
    alpha, tmn, cen, ra = synth.make_basis(rng, elements, coords, {0: 2, 1: 1}, {0: 1}, 0.1, 10.0)
    alpha = np.array([float("%.6g" % a) for a in alpha])
    C = synth.orbitals(rng, synth.overlap_matrix(alpha, tmn, cen), 4)
    s = synth.System(elements, coords, alpha, tmn, cen, ra, C)
    d = tmp_path_factory.mktemp("formats")
    s.write_fchk(str(d / "mol.fchk"), title="formats test")
    s.write_bas(str(d / "mol.bas"))
    synth.write_wfx(s, str(d / "mol.wfx"))
    ijkl, dm = s.dm2_unique()
    synth.write_dm2p(str(d / "mol.dm2p"), ijkl, dm)
    return s, ijkl, dm, d


def _run(d, name, files, job):
    inp = "&files\n%s,\n  dm2name = 'mol.dm2p'\n/\n&intracule_job\n%s\n/\n&cube_job\n/\n" % (
        files, job)
    (d / name).write_text(inp)
    r = subprocess.run([EXE, name], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout
    return open(d / name.replace(".inp", ".out")).read()


def _floats(text):
    return [float(v.replace("D", "E")) for v in text.split()]


def _total(out):
    return float(re.search(r"TOTAL VALUE OF INTRACULE\s*=\s*(\S+)", out).group(1))


def _centres(out):
    n = int(re.search(r"Number of centres:\s+(\d+)", out).group(1))
    cent = [_floats(m) for m in re.findall(r"^\s*\d+\s+:::\s+(.*)$", out, re.M)]
    # list-directed output may wrap a long list over several lines
    ps = _floats(re.search(r"Weights for centres:(.*?)Alpha parameter:", out, re.S).group(1))
    sf = _floats(re.search(r"Alpha parameter:(.*?)Gauss-Legendre", out, re.S).group(1))
    return n, np.array(cent), np.array(ps), np.array(sf)


FCHK = "  fchkfilename = 'mol.fchk',\n  basfilename = 'mol.bas'"
WFX = "  wfxfilename = 'mol.wfx'"
AUTO = ("  calc_type = 'radial_integral',\n  thresh = 1.0d-12,\n  nosym = .true.,\n"
        "  multicenter = .true.,\n  nrad = 30,\n  nang = 110")


def _oracle_integral(s, ijkl, dm, cent, nradc, nangc, sf, Ps):
    g = O.gridpoints(cent, nradc, nangc, sf, Ps, nosym=True)
    iv, cnt = O.intracule(s.prims(), ijkl, dm, g["r"], THRESH)
    return float(np.sum(g["w"] * iv)), cnt


def test_autogrid_centres_and_integral(molecule):
    s, ijkl, dm, d = molecule
    out = _run(d, "auto_fchk.inp", FCHK, AUTO)
    n, cent, ps, sf = _centres(out)
    oc, ops, osf = O.centercalc(s.an, s.coords, True, False)
    assert n == len(ops) == 5
    assert np.abs(cent - oc).max() < 1e-12
    assert np.abs(ps - ops).max() < 1e-12 and np.abs(sf - osf).max() < 1e-12
    # the snapped and merged pair centres
    assert np.any(np.all(np.abs(oc - [2.5, 0.0, 0.0]) < 1e-12, axis=1))
    tot, cnt = _oracle_integral(s, ijkl, dm, oc, [30] * n, [110] * n, osf, ops)
    print("autogrid integral %.15f oracle %.15f" % (_total(out), tot))
    assert abs(_total(out) - tot) < 1e-10 * abs(tot)
    assert int(re.search(r"Quartets actually computed:\s+(\d+)", out).group(1)) == cnt[2]


def test_wfx_reader_gives_the_fchk_job(molecule):
    s, ijkl, dm, d = molecule
    a = _run(d, "auto_fchk2.inp", FCHK, AUTO)
    b = _run(d, "auto_wfx.inp", WFX, AUTO)
    keys = ["TOTAL VALUE OF INTRACULE", "TOTAL VALUE OF Vee", "Quartets actually computed",
            "Total surviving grid points", "Number of centres"]
    for k in keys:
        la = [l for l in a.splitlines() if k in l]
        lb = [l for l in b.splitlines() if k in l]
        assert la and la == lb, (k, la, lb)
    assert _centres(a)[1].tolist() == _centres(b)[1].tolist()


def test_manual_grid(molecule):
    s, ijkl, dm, d = molecule
    cent = np.array([[0.0, 0.0, 0.0], [2.5, 0.0, 0.0], [-2.5, 0.0, 0.0]])
    nradc, nangc = [30, 40, 25], [110, 146, 110]
    sf, Ps = [1.2, 0.9, 1.5], [10.0, 5.0, 5.0]
    job = ("  calc_type = 'radial_integral',\n  thresh = 1.0d-12,\n  nosym = .true.,\n"
           "  multicenter = .true.,\n  manual_grid = .true.,\n  nquad_in = 3,\n"
           "  cent_in = %s,\n  nradc_in = %s,\n  nangc_in = %s,\n  sfalpha_in = %s,\n"
           "  Ps_in = %s" % (", ".join("%.1f" % v for v in cent.reshape(-1)),
                              ", ".join(map(str, nradc)), ", ".join(map(str, nangc)),
                              ", ".join(map(str, sf)), ", ".join(map(str, Ps))))
    out = _run(d, "manual.inp", FCHK, job)
    n, c, ps, a = _centres(out)
    assert n == 3 and np.array_equal(c, cent)
    assert "Gauss-Legendre nodes:" in out
    assert _floats(re.search(r"Gauss-Legendre nodes:(.*?)Gauss-Lebedev", out, re.S).group(1)) == nradc
    assert _floats(re.search(r"Gauss-Lebedev nodes:(.*?)-{5}", out, re.S).group(1)) == nangc
    tot, _ = _oracle_integral(s, ijkl, dm, cent, nradc, nangc, sf, Ps)
    print("manual grid integral %.15f oracle %.15f" % (_total(out), tot))
    assert abs(_total(out) - tot) < 1e-10 * abs(tot)
