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
    sm = tmp_path / "ch4_integral_smearing.dat"
    rows = np.loadtxt(sm, comments="#")
    assert rows.shape == (1000, 3)


def test_zero(tmp_path):
    out = run_case(tmp_path, "test_zero.inp")
    v = grab(out, r"INTRACULE AT ZERO\s*=\s*(\S+)")
    assert abs(v - GD.ZERO_HALF) / GD.ZERO_HALF < 1e-10


def test_two_points(tmp_path):
    out = run_case(tmp_path, "test_twopoints.inp")
    p1 = grab(out, r"Point 1:.*Intracule:\s*(\S+)")
    p2 = grab(out, r"Point 2:.*Intracule:\s*(\S+)")
    assert abs(p1 - GD.TWO_POINTS[0]) / GD.TWO_POINTS[0] < 1e-10
    assert abs(p2 - GD.TWO_POINTS[1]) / GD.TWO_POINTS[1] < 1e-10


def test_radial_scan(tmp_path):
    run_case(tmp_path, "test_radial.inp")
    gen = [l.split() for l in open(tmp_path / "ch4_scan.rad")
           if l.strip() and not l.strip().startswith("#")]
    ref = GD.rad_table()
    assert len(gen) == len(ref)
    g = np.array([[float(v) for v in r] for r in gen])
    assert np.abs(g[:, 0] - ref[:, 0]).max() < 1e-9
    assert (np.abs(g[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])).max() < 1e-8


def test_cube(tmp_path):
    run_case(tmp_path, "test_vectorial.inp")
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
