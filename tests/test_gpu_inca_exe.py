"""End-to-end drop-in test: the Fortran host `inca` on the reference's own
CH4 test inputs (tests/inputs/*.inp and tests/data/ch4.* of
markelgtx/INCA_parallel), checked like the reference's harness
(tests/run_tests.py:96-254) but at the goldens' printed precision:
1e-10 relative for ES25.16 values and cube voxels, 8 digits for the .rad.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import ch4_golden as GD

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "inca_parallel_amd", "bin", "inca")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")


def run_case(tmp_path, template, extra=""):
    for f in ("ch4.fchk", "ch4_ops.bas", "ch4.dm2p"):
        shutil.copy(os.path.join(GD.G, f), tmp_path / f)
    txt = open(os.path.join(INPUTS, template)).read()
    txt = (txt.replace("{fchk_file}", "ch4.fchk").replace("{bas_file}", "ch4_ops.bas")
           .replace("{dm2p_file}", "ch4.dm2p").replace("{prefix}", "ch4").replace("{name}", "Methane"))
    if extra:
        txt = txt.replace("&intracule_job", "&intracule_job\n  " + extra + ",", 1)
    name = template.replace("test_", "ch4_")
    (tmp_path / name).write_text(txt)
    r = subprocess.run([EXE, name], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout
    return tmp_path / name.replace(".inp", ".out")


# The reference's .out layout (mod_intracule.f90:410-503), line by line as
# regular expressions: a parser of the reference's output reads the same
# fields on the same lines.  Engine lines may only follow the closing rule.
_RULE = r"=+$"
_HEAD = [_RULE, r"INCA INTRACULE CALCULATION$", _RULE, r"--- GRID AND ACCURACY SETTINGS ---$",
         r"Original grid points:\s+\d+$", r"Symmetry reduced grid points:\s+\d+$",
         r"Total surviving grid points:\s+\d+$", r"Quartets in the dm2p file:\s+\d+$",
         r"Quartets after 1st screening:\s+\d+$", r"Quartets actually computed:\s+\d+$",
         r"Screening limit \(lim\):\s+\S+$", r"Thresholds used for DM2prim:\s+\S+\s+\S+$",
         r"$", r"--- RESULTS ---$"]
_TAIL = [r"$", r"--- COMPUTATIONAL TIME ---$", r"Total CPU time:\s+\S+ seconds$",
         r"Grid generation time:\s+\S+ seconds$", r"Integration loop time:\s+\S+ seconds$",
         _RULE]
_NUM = r"\s+\S+"
_BODY = {
    "integral": [r"Calculation Type: Radial Integral$", r"Number of centres:\s+1$",
                 r"1\s+:::" + _NUM * 3 + "$", r"Weights for centres:" + _NUM + "$",
                 r"Alpha parameter:" + _NUM + "$", r"Gauss-Legendre nodes:\s+50$",
                 r"Gauss-Lebedev nodes:\s+110$", r"-+$", r"TOTAL VALUE OF INTRACULE = " + _NUM + "$",
                 r"TOTAL VALUE OF Vee       = " + _NUM + "$",
                 r"Radial_integral error    = " + _NUM + "$", r"-+$",
                 r"-------LOCAL DECOMPOSITION \(TFVC\)-------$",
                 r"Center\s+Coordinates\s+Pairs\(N_pairs\)\s+Vee\(Hartree\)$",
                 r"1" + _NUM * 5 + r"\s+\(INTRA-ATOMIC\)$", r"-+$"],
    "radial": [r"Calculation Type: Radial Plot, I\(s\) vs s$",
               r"Radius evaluated from" + _NUM + r"\s+to" + _NUM + r"\s+bohr$",
               r"Data written to:\s+ch4_scan.rad$"],
    "cube": [r"Calculation Type: Vectorial Cubefile$", r"Data written to:\s+ch4_vector.cube$"],
    "zero": [r"INTRACULE AT ZERO = " + _NUM + "$"],
    "twopoints": [r"Point 1:" + _NUM * 3 + r"\s+Intracule:" + _NUM + "$",
                  r"Point 2:" + _NUM * 3 + r"\s+Intracule:" + _NUM + "$"],
}


def check_layout(out, kind):
    lines = [l.strip() for l in open(out).read().splitlines()]
    want = _HEAD + _BODY[kind] + _TAIL
    assert len(lines) >= len(want)
    for k, (line, pat) in enumerate(zip(lines, want)):
        assert re.match(pat, line), "line %d: %r does not match %r" % (k + 1, line, pat)


def grab(path, pat):
    m = re.search(pat, open(path).read())
    assert m, pat
    return float(m.group(1).replace("D", "E").replace("d", "e"))


def test_integral(tmp_path):
    out = run_case(tmp_path, "test_integral.inp")
    val = grab(out, r"TOTAL VALUE OF INTRACULE\s*=\s*(\S+)")
    err = grab(out, r"Radial_integral error\s*=\s*(\S+)")
    assert abs(val - GD.INTEGRAL) / GD.INTEGRAL < 1e-10
    assert abs(err - GD.INTEGRAL_ERR) < 1e-7
    assert "Quartets actually computed" in open(out).read()
    check_layout(out, "integral")
    sm = tmp_path / "ch4_integral_smearing.dat"
    rows = np.loadtxt(sm, comments="#")
    assert rows.shape == (1000, 3)
    # values: radial_smearing (mod_intracule.f90:576-649) of the oracle's Ivec
    # on the oracle's grid, u_max = max |centre| + 5 = 5 (CH4: the origin only)
    u, P, Ia = _oracle_smearing()
    assert np.abs(rows[:, 0] - u).max() < 1e-8 * 5.0
    big = np.abs(P) > 1e-6 * np.abs(P).max()
    assert (np.abs(rows[big, 1] - P[big]) / np.abs(P[big])).max() < 1e-7
    assert np.abs(rows[:, 1] - P).max() < 1e-8 * np.abs(P).max()
    bigi = np.abs(Ia) > 1e-6 * np.abs(Ia).max()
    assert (np.abs(rows[bigi, 2] - Ia[bigi]) / np.abs(Ia[bigi])).max() < 1e-7
    assert rows[0, 2] == 0.0


def _oracle_smearing():
    from oracle import oracle as O
    from inca_inputs import read_bas, read_dm2p, read_fchk
    mol = read_bas(os.path.join(GD.G, "ch4_ops.bas"), read_fchk(os.path.join(GD.G, "ch4.fchk")))
    ijkl, d = read_dm2p(os.path.join(GD.G, "ch4.dm2p"))
    cent, Ps, sf = O.centercalc(mol["an"], mol["cartes"], True, False)
    g = O.gridpoints(cent, [50] * len(Ps), [110] * len(Ps), sf, Ps)
    iv, _ = O.intracule(mol, ijkl, d, g["r"], 1e-12)
    u_max = float(np.max(np.linalg.norm(cent, axis=1))) + 5.0
    return O.smearing(g["r"], g["w"], iv, u_max, 1000, 0.15)


def test_zero(tmp_path):
    out = run_case(tmp_path, "test_zero.inp")
    v = grab(out, r"INTRACULE AT ZERO\s*=\s*(\S+)")
    assert abs(v - GD.ZERO_HALF) / GD.ZERO_HALF < 1e-10
    check_layout(out, "zero")


def test_two_points(tmp_path):
    out = run_case(tmp_path, "test_twopoints.inp")
    p1 = grab(out, r"Point 1:.*Intracule:\s*(\S+)")
    p2 = grab(out, r"Point 2:.*Intracule:\s*(\S+)")
    assert abs(p1 - GD.TWO_POINTS[0]) / GD.TWO_POINTS[0] < 1e-10
    assert abs(p2 - GD.TWO_POINTS[1]) / GD.TWO_POINTS[1] < 1e-10
    check_layout(out, "twopoints")


def test_radial_scan(tmp_path):
    check_layout(run_case(tmp_path, "test_radial.inp"), "radial")
    gen = [l.split() for l in open(tmp_path / "ch4_scan.rad")
           if l.strip() and not l.strip().startswith("#")]
    ref = GD.rad_table()
    assert len(gen) == len(ref)
    g = np.array([[float(v) for v in r] for r in gen])
    assert g.shape == ref.shape == (101, 4)
    assert np.abs(g[:, 0] - ref[:, 0]).max() < 1e-9
    # columns 2-4 (mod_intracule.f90:529): I(s), 2 pi s^2 I(s) and the
    # reference's 2 pi s I(s) under its "s * I(s)" header, to the 8 digits
    # the golden file prints
    for c in (1, 2, 3):
        nz = ref[:, c] != 0.0
        assert (np.abs(g[nz, c] - ref[nz, c]) / np.abs(ref[nz, c])).max() < 1e-8
        assert np.all(g[~nz, c] == 0.0)
    assert np.allclose(g[:, 3], 2 * np.pi * g[:, 0] * g[:, 1], rtol=1e-7, atol=0.0)


def test_cube(tmp_path):
    check_layout(run_case(tmp_path, "test_vectorial.inp"), "cube")
    gen = open(tmp_path / "ch4_vector.cube").read().splitlines()
    ref = open(os.path.join(GD.G, "ch4_ref.cube")).read().splitlines()
    assert len(gen) == len(ref)
    for lg, lr in zip(gen[:11], ref[:11]):
        tg, tr = lg.split(), lr.split()
        assert len(tg) == len(tr)
        for a, b in zip(tg, tr):
            try:
                assert abs(float(a) - float(b)) < 1e-12
            except ValueError:
                assert a == b
    vg = np.array([float(v) for v in gen[11:]])
    vr = GD.cube_values()
    assert (np.abs(vg - vr) / np.abs(vr)).max() < 1e-10


def test_integral_mixed_precision(tmp_path):
    out = run_case(tmp_path, "test_integral.inp", "engine_precision = 'mixed'")
    val = grab(out, r"TOTAL VALUE OF INTRACULE\s*=\s*(\S+)")
    print("mixed integral rel err vs golden %.3e" % (abs(val - GD.INTEGRAL) / GD.INTEGRAL))
    assert abs(val - GD.INTEGRAL) / GD.INTEGRAL < 1e-8


def test_integral_learned_grid(tmp_path):
    """learned_grid_weights: the network's radii and orders on the autogrid
    centres, with the sum-rule fallback.  The final integral may not be
    further from the analytic integral of the prepared terms than the plain
    autogrid's (the golden integral, same centres and orders) by more than
    1e-8 relative; a rejected learned grid reproduces the golden integral."""
    w = os.path.join(ROOT, "inca_parallel_amd", "data", "autogrid_mlp.npy")
    out = run_case(tmp_path, "test_integral.inp", "learned_grid_weights = '%s'" % w)
    txt = open(out).read()
    note = re.search(r"Learned autogrid: (accepted|rejected).*analytic integral\s*(\S+)", txt)
    assert note, txt
    exact = float(note.group(2))
    val = grab(out, r"TOTAL VALUE OF INTRACULE\s*=\s*(\S+)")
    err_final = abs(val - exact) / abs(exact)
    err_plain = abs(GD.INTEGRAL - exact) / abs(exact)
    print("learned grid: %s, err final %.3e, err plain %.3e" % (note.group(1), err_final, err_plain))
    assert err_final <= err_plain + 1e-8
    if note.group(1) == "rejected":
        assert abs(val - GD.INTEGRAL) / GD.INTEGRAL < 1e-10


def test_learned_grid_bad_weights(tmp_path):
    """A weights file that is not a (9993,) float32 .npy stops the job with
    the loader's message."""
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"not a numpy file")
    with pytest.raises(AssertionError) as e:
        run_case(tmp_path, "test_integral.inp", "learned_grid_weights = '%s'" % bad)
    assert "inca_mlp_load" in str(e.value)
