"""Generate the polycoef roundoff-audit fixtures (tests/golden/polycoef_audit.*).

    python tests/golden/gen_polycoef_audit.py            # about 5 minutes, 8 cores

Inputs (all seeded, generated here):
  A  the 4-atom s/p/d cluster of the GPU parity tests (synth.config_cluster
     seed 21, 35 primitives), every symmetry-unique record;
  B  the radial-plot benchmark cluster (synth.config_cluster seed 2, 151
     primitives): a seeded subset of 2^20 of its 65,855,026 unique records,
     and a census of the whole stream: every fit that some arithmetic
     variant zeroes or fits with max |C| >= 100.

Each fit is classified by polycoef_audit (keep / zero / roundoff /
unbounded: whether the mathematics or the rounding decides the
reference's |C| > 1e3 zeroing, mod_intrastuff.f90:169-173), and fitted by
four FP64 implementations of the reference's algorithm:
  V0  oracle mode 0 (repeated multiplication, no FMA: the device kernels'
      operation sequence),
  V1  oracle mode 1 (the reference's libm pow forms and decimal roots),
  V2  mode 1 built with FMA contraction (the reference's -O3 -march=native
      on an FMA host),
  V3  mode 0 with FMA contraction.
The Ivec spread between V0, V1 and V2 (whole oracle runs) is measured on A's
multicentre grid and on B's subset over the benchmark's radial scan.

Writes polycoef_audit.json (summary) and polycoef_audit_rows.npz (the
census rows that are not plain "keep", with their classes and decisions).
"""
from __future__ import annotations

import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import polycoef_audit as PA  # noqa: E402
from inca_parallel_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

THRESH = 1e-12
SUBSET = 1 << 20
SUBSET_SEED = 20
VALIDATE = 2000


def system_a():
    s = synth.config_cluster(seed=21, natoms=4, nheavy=3, nocc=4,
                             shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
    ijkl, d = s.dm2_unique()
    return s, ijkl, d


def system_b_subset(ijkl, d):
    rng = np.random.default_rng(SUBSET_SEED)
    idx = np.sort(rng.choice(len(d), size=SUBSET, replace=False))
    return np.ascontiguousarray(ijkl[idx]), np.ascontiguousarray(d[idx]), idx


def variants(rows):
    """Zeroing status of every row under V0..V3."""
    out = []
    for L, mode in ((O.lib(), 0), (O.lib(), 1), (O.lib_fma(), 1), (O.lib_fma(), 0)):
        with O.arith(mode, L):
            C, st = O.polycoef_rows(rows, L)
        out.append((C, st))
    return out


def _exact(row):
    Cs, d, valid, cls = PA.exact_fit(row)
    return Cs, d, valid, cls


def exact_all(rows, procs=8):
    with Pool(procs) as p:
        res = p.map(_exact, list(rows), chunksize=64)
    return res


def classify(rows, name):
    """Classes of every row (FP64 tier + 60-digit tier) and checks."""
    V = variants(rows)
    C0 = V[0][0]
    cls, dmax, cond = PA.fp64_classify(rows, C0, V[0][1])
    todo = np.nonzero(cls < 0)[0]
    t = time.time()
    res = exact_all(rows[todo])
    print("%s: %d fits, %d solved at 60 digits (%.0f s)" % (name, len(rows), len(todo),
                                                              time.time() - t), flush=True)
    cmax_exact = np.full(len(rows), np.nan)
    for i, (Cs, d, valid, c) in zip(todo, res):
        cls[i] = c
        dmax[i] = d.max() if valid else np.inf
        cmax_exact[i] = np.abs(Cs).max()
    # validation of the FP64 tier at 60 digits
    rng = np.random.default_rng(7)
    done = np.nonzero(np.isnan(cmax_exact))[0]
    samp = np.sort(rng.choice(done, size=min(VALIDATE, len(done)), replace=False))
    vres = exact_all(rows[samp])
    agree = sum(int(c == cls[i]) for i, (_, _, _, c) in zip(samp, vres))
    within = 0
    for i, (Cs, d, valid, c) in zip(samp, vres):
        npt = int(rows[i, 13])
        within += int(np.all(np.abs(C0[i, :npt] - Cs) <= d))
    # every variant against the classes
    checks = {}
    for vname, (C, st) in zip(("V0", "V1", "V2", "V3"), V):
        checks[vname] = dict(
            zeroed=int(st.sum()),
            zeroed_in_keep=int(np.sum(st[cls == PA.KEEP])),
            kept_in_zero=int(np.sum(st[cls == PA.ZERO] == 0)),
        )
    counts = {k: int(np.sum(cls == v)) for k, v in
              (("keep", PA.KEEP), ("zero", PA.ZERO), ("roundoff", PA.ROUNDOFF),
               ("unbounded", PA.UNBOUNDED))}
    summary = dict(fits=int(len(rows)), solved_60_digits=int(len(todo)), classes=counts,
                   variants=checks, fp64_tier_validation=dict(
                       sample=int(len(samp)), class_agrees=int(agree),
                       v0_within_bound=int(within)))
    print(name, json.dumps(summary), flush=True)
    return cls, dmax, cmax_exact, V, summary


def ivec_spread(prims, ijkl, d, r):
    out = {}
    iv = {}
    for vname, (L, mode) in (("V0", (O.lib(), 0)), ("V1", (O.lib(), 1)),
                             ("V2", (O.lib_fma(), 1))):
        with O.arith(mode, L):
            iv[vname], cnt = O.intracule(prims, ijkl, d, r, THRESH, L=L, nthreads=8)
        out["counts_" + vname] = [int(c) for c in cnt]
    ex, _ = O.intracule(prims, ijkl, d, r, THRESH, exact=True, nthreads=8)
    scale = np.abs(iv["V0"]).max()
    big = np.abs(iv["V0"]) >= 1e-6 * scale
    for a, b in (("V0", "V1"), ("V0", "V2"), ("V1", "V2")):
        dlt = np.abs(iv[a] - iv[b])
        out["%s_%s" % (a, b)] = dict(max_rel=float((dlt[big] / np.abs(iv[b][big])).max()),
                                     max_abs_over_max=float(dlt.max() / scale))
    dlt = np.abs(iv["V0"] - ex)
    out["V0_exact_polynomials"] = dict(max_rel=float((dlt[big] / np.abs(ex[big])).max()),
                                       max_abs_over_max=float(dlt.max() / scale))
    out["points"] = int(r.shape[0])
    return out


def main():
    t0 = time.time()
    report = {"thresh": THRESH, "subset": SUBSET, "subset_seed": SUBSET_SEED}
    # ---- A
    sA, ijA, dA = system_a()
    pA = sA.prims()
    rowsA = O.fit_rows(pA, ijA, dA, THRESH)
    clsA, _, _, _, report["A"] = classify(rowsA, "A")
    report["A"]["records"] = int(len(dA))
    report["A"]["sum_d"] = float(dA.sum())
    cent, Ps, sf = O.centercalc(pA["an"], pA["cartes"], True, False)
    g = O.gridpoints(cent, [40] * len(Ps), [110] * len(Ps), sf, Ps, nosym=True)
    report["A"]["ivec"] = ivec_spread(pA, ijA, dA, g["r"])
    print("A ivec", report["A"]["ivec"], flush=True)
    # ---- B
    sB = synth.config_cluster(seed=2)
    pB = sB.prims()
    ijB, dB = sB.dm2_unique()
    report["B_records"] = int(len(dB))
    ijS, dS, _ = system_b_subset(ijB, dB)
    rowsS = O.fit_rows(pB, ijS, dS, THRESH)
    clsS, _, _, _, report["B_subset"] = classify(rowsS, "B_subset")
    report["B_subset"]["sum_d"] = float(dS.sum())
    g2 = O.gridpoints2([0.0, 10.0], [0.05], [302])
    report["B_subset"]["ivec"] = ivec_spread(pB, ijS, dS, g2["r"])
    print("B ivec", report["B_subset"]["ivec"], flush=True)
    # ---- B census over the whole stream
    parts = [O.fit_census(pB, ijB, dB, THRESH, 100.0, L=L) for L in (O.lib(), O.lib_fma())]
    keys = {}
    for part in parts:
        for r in part:
            keys.setdefault((int(r[0]), int(r[1])), r)
    # inputs of the plain build for every (record, axis)
    recs = np.array(sorted({k[0] for k in keys}), dtype=np.int64)
    plain = O.fit_rows(pB, ijB[recs], dB[recs], THRESH)
    plain[:, 0] = recs[plain[:, 0].astype(int)]
    want = set(keys)
    rowsC = plain[[(int(r[0]), int(r[1])) in want for r in plain]]
    clsC, dmaxC, cmaxC, VC, report["B_census"] = classify(rowsC, "B_census")
    report["B_census"]["cmin"] = 100.0
    # fixture rows: every census row that is not a settled "keep" with
    # max |C*| < 500, plus a seeded sample of the rest
    cmax0 = np.abs(VC[0][0]).max(1)
    sel = (clsC != PA.KEEP) | (np.nan_to_num(cmaxC, nan=0.0) >= 500.0) | (cmax0 >= 500.0)
    rng = np.random.default_rng(11)
    rest = np.nonzero(~sel)[0]
    extra = rng.choice(rest, size=min(200, len(rest)), replace=False)
    sel[extra] = True
    idx = np.nonzero(sel)[0]
    np.savez_compressed(
        os.path.join(HERE, "polycoef_audit_rows.npz"), rows=rowsC[idx], cls=clsC[idx],
        dmax=dmaxC[idx], cmax_exact=cmaxC[idx],
        status=np.stack([v[1][idx] for v in VC], axis=1).astype(np.int8))
    report["B_census"]["fixture_rows"] = int(len(idx))
    report["seconds"] = round(time.time() - t0, 1)
    with open(os.path.join(HERE, "polycoef_audit.json"), "w") as f:
        json.dump(report, f, indent=1)
    print("wrote", report["seconds"], "s")


if __name__ == "__main__":
    main()
