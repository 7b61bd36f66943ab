"""Learned autogrid: labels, training and use (drives the C-ABI on one GPU).

No reference counterpart.  The reference builds its multicentre grid from
Bragg-Slater radii and one nrad/nang for every centre (mod_autogrid.f90:
12-185, mod_input_reader.f90:185-221).  Here a per-atom network
(inca_mlp_*; HIP kernels for inference and training) predicts a factor on
each atom's Bragg-Slater radius and per-centre nrad/nang classes, and the
centres keep the reference's formulas (inca_learned_autogrid).

Accuracy is judged without a reference grid: the pre-pass gives the
analytic integral of every prepared term over all space (the pair sum rule
of the terms, inca_get_analytic_integral), which any grid integral
approximates.  err(grid) = |I_grid - I_analytic| / |I_analytic|.

Labels (per training molecule): over the label candidates (radius factor,
nrad, nang; LABEL_CANDIDATES, the choices that win on a sweep of 5 factors x
4 x 4 orders), the most accurate grid whose point count stays within 1.05
times the plain autogrid's (Bragg-Slater radii, the reference defaults
nrad 50 / nang 590): accuracy at the reference's cost.  Every atom of the
molecule gets that label.  The errors come from the engine itself
(shell-quartet groups, mixed precision: within 1e-8 of FP64, far below the
grid errors it ranks).

Fallback: the learned grid is accepted when its error is at most 1e-8, or
else when it moves the sum rule by no more than 1e-8 beyond the plain
autogrid's error; otherwise the plain autogrid's result is returned.

Judge: besides the grid-free sum rule, the FP64 dense grid (nrad 200, nang
974) on the same centres.

    python -m inca_parallel_amd.learned_grid label --seeds 100000:120000 --out labels.npz
    python -m inca_parallel_amd.learned_grid train --labels labels.npz --out weights.npy

The shipped network (data/autogrid_mlp.npy) was trained by `train` on the
labels of seeds 100000-119999 (data/autogrid_labels.npz: seed, label, and the
plain and label grids' errors and points) for 1,500 full-batch Adam steps.
    python -m inca_parallel_amd.learned_grid ablate --weights weights.npy --ntest 1000 --dense
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from inca_parallel_amd import capi, synth

PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG, "data", "autogrid_mlp.npy")
SCALES = (0.5, 0.7, 1.0, 1.4, 2.0)
PLAIN = (50, 590)          # reference defaults (mod_input_reader.f90:55-66)
# (radius factor, nrad, nang) tried for each training molecule
LABEL_CANDIDATES = ((1.0, 50, 590), (1.4, 50, 590), (2.0, 50, 590), (0.7, 50, 590),
                    (0.5, 50, 302), (0.7, 50, 302), (1.4, 50, 302), (2.0, 50, 302),
                    (0.5, 50, 194), (1.0, 50, 110), (0.7, 75, 302), (0.7, 75, 194))
DENSE = (200, 974)         # judge grid
FALLBACK_TOL = 1e-8
THRESH = 1e-12


def bragg_slater(an):
    """The library's Bragg-Slater radii in bohr (get_bs_radius)."""
    return capi.bragg_slater(an)


def molecule(seed):
    """Molecules like the batch workload: 3-5 atoms, s/p/d primitives."""
    nat = 3 + seed % 3
    return synth.config_cluster(seed=seed, natoms=nat, nheavy=max(1, nat - 1), nocc=2 + seed % 3,
                                shells_heavy={0: 2, 1: 1, 2: seed % 2}, shells_h={0: 2})


class Molecule:
    """One molecule's context with its quartets prepared once.  precision:
    INCA_FP64 (the reference's arithmetic) or INCA_MIXED; device_records:
    generate the DM2 records on the device (torch) instead of the host."""

    def __init__(self, s, nohydro=True, precision=capi.INCA_FP64, device_records=False):
        self.s = s
        p = s.prims()
        self.p = p
        self.nohydro = nohydro
        self.precision = precision
        if device_records:
            # torch's HIP runtime starts before the library's first context
            # (it finds no device when it comes second)
            import torch
            torch.cuda.init()
        self.e = capi.Engine(0)
        self.e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
        if device_records:
            recs = s.dm2_unique_torch(device="cuda")
            torch.cuda.synchronize()
            self.e.set_quartets(device_ptr=recs.data_ptr(), n=recs.shape[0])
            del recs
        else:
            ijkl, d = s.dm2_unique()
            self.e.set_quartets(ijkl=ijkl, d=d)
        self.e.prepare(precision, THRESH)
        self.e.release_scratch()
        self.analytic = self.e.analytic_integral()

    def evaluate(self, cent, Ps, sf, nradc, nangc):
        """(relative sum-rule error, integral, points) on the given centres."""
        self.e.grid_multicenter(cent, nradc, nangc, sf, Ps, nosym=True)
        self.e.evaluate_prepared(self.precision, want_ivec=False)
        tot = self.e.integrate()[0]
        return abs(tot - self.analytic) / abs(self.analytic), tot, self.e.G

    def uniform(self, scale, nrad, nang):
        p = self.p
        rad = bragg_slater(p["an"]) * scale
        cent, Ps, sf = capi.autogrid(p["an"], p["cartes"], self.nohydro, False, radius=rad)
        nq = len(Ps)
        return self.evaluate(cent, Ps, sf, [nrad] * nq, [nang] * nq)

    def features(self):
        p = self.p
        return capi.atom_features(p["an"], p["cartes"], p["alpha"], p["tmn"], self.s.ra)

    def close(self):
        self.e.close()


def label(m, budget=1.05):
    """Per-molecule label [log factor, nrad class, nang class] and the plain
    error: the most accurate of LABEL_CANDIDATES within `budget` times the
    plain grid's points (ties: fewer points)."""
    e_plain, _, g_plain = m.uniform(1.0, *PLAIN)
    best = None
    for sc, nr, na in LABEL_CANDIDATES:
        if (sc, nr, na) == (1.0,) + PLAIN:
            err, g = e_plain, g_plain
        else:
            err, _, g = m.uniform(sc, nr, na)
        if g > budget * g_plain:
            continue
        key = (err, g)
        if best is None or key < best[0]:
            best = (key, sc, nr, na, err, g)
    _, sc, nr, na, err, g = best
    lab = np.array([np.log(sc), capi.NRAD_CLASSES.index(nr), capi.NANG_CLASSES.index(na)],
                   np.float32)
    return lab, dict(err_plain=e_plain, g_plain=g_plain, err_label=err, g_label=g)


def learned_grid_evaluate(net, m, tol=FALLBACK_TOL):
    """Evaluate on the learned grid with the fallback rule.  Returns a dict
    with the integral used and which grid produced it."""
    p = m.p
    cent, Ps, sf, nr, na, _ = net.learned_autogrid(p["an"], p["cartes"], p["alpha"], p["tmn"],
                                                   m.s.ra, nohydro=m.nohydro)
    err_l, tot_l, g_l = m.evaluate(cent, Ps, sf, nr, na)
    out = dict(err_learned=err_l, points_learned=g_l, integral=tot_l, grid="learned")
    if err_l <= tol:
        return out
    err_p, tot_p, g_p = m.uniform(1.0, *PLAIN)
    out.update(err_plain=err_p, points_plain=g_p)
    if err_l > err_p + tol:
        out.update(integral=tot_p, grid="plain")
    return out


def train(net, feats, targets, epochs=600, lr=2e-3, log=None):
    x = np.concatenate(feats).astype(np.float32)
    t = np.concatenate(targets).astype(np.float32)
    hist = []
    for ep in range(epochs):
        loss, _ = net.loss_grad(x, t)
        net.adam_step(lr if ep < epochs // 2 else lr * 0.3)
        if ep % 100 == 0 or ep == epochs - 1:
            hist.append((ep, loss))
            if log:
                log("epoch %d loss %.5f" % (ep, loss))
    return hist


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _seeds(spec):
    a, b = (int(v) for v in spec.split(":"))
    return range(a, b)


def cmd_label(a):
    """Label molecules on the GPU (records generated on the device, mixed
    precision) and save per molecule: seed, label, plain and label errors
    and points, and the per-atom features."""
    t0 = time.time()
    seeds = list(_seeds(a.seeds))
    rows, feats = [], []

    def save():
        np.savez(a.out, seeds=np.array([r[0] for r in rows], np.int64),
                 labels=np.array([r[1] for r in rows], np.float32),
                 info=np.array([r[2] for r in rows], np.float64),
                 natoms=np.array([len(f) for f in feats], np.int32),
                 features=np.concatenate(feats).astype(np.float32) if feats else
                 np.zeros((0, capi.MLP_IN), np.float32))

    for i, seed in enumerate(seeds):
        m = Molecule(molecule(seed), precision=capi.INCA_MIXED, device_records=True)
        lab, info = label(m)
        rows.append((seed, lab, [info["err_plain"], info["g_plain"], info["err_label"],
                                 info["g_label"]]))
        feats.append(m.features())
        m.close()
        if (i + 1) % 500 == 0 or i + 1 == len(seeds):
            save()
            _log("labelled %d/%d (%.0f s)" % (i + 1, len(seeds), time.time() - t0))
    print(json.dumps(dict(labelled=len(rows), seconds=time.time() - t0)))


def cmd_train(a):
    """Train on the GPU from one or more label files; save the weights."""
    t0 = time.time()
    feats, targets = [], []
    nmol = 0
    for path in a.labels:
        z = np.load(path, allow_pickle=False)
        lab = z["labels"]
        if "features" in z:
            f, nat = z["features"], z["natoms"]
        else:  # a compact label file (data/autogrid_labels.npz): features from the seeds
            fs = []
            for seed in z["seeds"]:
                s = molecule(int(seed))
                p = s.prims()
                fs.append(capi.atom_features(p["an"], p["cartes"], p["alpha"], p["tmn"], s.ra))
            f, nat = np.concatenate(fs), np.array([len(x) for x in fs])
        feats.append(f)
        targets.append(np.repeat(lab, nat, axis=0))
        nmol += len(lab)
    net = capi.MLP(seed=a.seed)
    hist = train(net, feats, targets, epochs=a.epochs, log=_log)
    np.save(a.out, net.params())
    net.close()
    print(json.dumps(dict(molecules=nmol, atoms=int(sum(len(f) for f in feats)), epochs=a.epochs,
                          loss=hist, seconds=time.time() - t0)))


def cmd_ablate(a):
    """Held-out molecules: the learned grid with the fallback rule against the
    plain autogrid, judged by the grid-free sum rule and (--dense) by the FP64
    dense grid on the same centres."""
    t0 = time.time()
    import torch
    torch.cuda.init()  # before the library's first context (see Molecule)
    net = capi.MLP(seed=1)
    net.set_params(np.load(a.weights, allow_pickle=False))
    rows = []
    for i in range(a.ntest):
        m = Molecule(molecule(a.first + i), device_records=True)
        r = learned_grid_evaluate(net, m)
        r["points_evaluated"] = r["points_learned"] + r.get("points_plain", 0)
        err_p, tot_p, g_p = m.uniform(1.0, *PLAIN)
        r.update(err_plain=err_p, points_plain=g_p, integral_plain=tot_p)
        if a.dense:
            r["err_dense"], tot_d, r["points_dense"] = m.uniform(1.0, *DENSE)
            r["dense_judged_final"] = abs(r["integral"] - tot_d) / abs(tot_d)
            r["dense_judged_plain"] = abs(tot_p - tot_d) / abs(tot_d)
        rows.append(r)
        m.close()
        if (i + 1) % 100 == 0:
            _log("held out %d/%d (%.0f s)" % (i + 1, a.ntest, time.time() - t0))
    net.close()
    final = [r["err_learned"] if r["grid"] == "learned" else r["err_plain"] for r in rows]

    def geo(v):
        return float(np.exp(np.mean(np.log(np.maximum(v, 1e-16)))))

    ratio = [max(f, 1e-16) / max(r["err_plain"], 1e-16) for f, r in zip(final, rows)]
    summary = dict(
        molecules=len(rows), accepted=int(sum(r["grid"] == "learned" for r in rows)),
        better_than_plain=int(sum(f < r["err_plain"] for f, r in zip(final, rows))),
        never_worse_than_plain_plus_1e8=bool(all(f <= r["err_plain"] + FALLBACK_TOL
                                                 for f, r in zip(final, rows))),
        geomean_err_ratio_final_vs_plain=geo(ratio),
        median_err_ratio_final_vs_plain=float(np.median(ratio)),
        median_points_ratio_learned_vs_plain=float(np.median(
            [r["points_learned"] / r["points_plain"] for r in rows])),
        median_err_plain=float(np.median([r["err_plain"] for r in rows])),
        median_err_learned=float(np.median([r["err_learned"] for r in rows])),
        median_err_final=float(np.median(final)),
        seconds=time.time() - t0,
    )
    if a.dense:
        summary["median_err_dense_sum_rule"] = float(np.median([r["err_dense"] for r in rows]))
        summary["geomean_dense_judged_final"] = geo([r["dense_judged_final"] for r in rows])
        summary["geomean_dense_judged_plain"] = geo([r["dense_judged_plain"] for r in rows])
        summary["median_dense_judged_final"] = float(np.median([r["dense_judged_final"]
                                                                for r in rows]))
        summary["median_dense_judged_plain"] = float(np.median([r["dense_judged_plain"]
                                                                for r in rows]))
    out = dict(summary=summary, rows=rows)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f)
    print(json.dumps(summary))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    lb = sub.add_parser("label")
    lb.add_argument("--seeds", default="100000:100128")
    lb.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--labels", nargs="+", required=True)
    t.add_argument("--epochs", type=int, default=1500)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=WEIGHTS)
    b = sub.add_parser("ablate")
    b.add_argument("--weights", default=WEIGHTS)
    b.add_argument("--ntest", type=int, default=32)
    b.add_argument("--first", type=int, default=200000)
    b.add_argument("--dense", action="store_true")
    b.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    return {"label": cmd_label, "train": cmd_train, "ablate": cmd_ablate}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
