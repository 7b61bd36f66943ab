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

Labels (per training molecule): over radius factors SCALES and the class
sets nrad {50,75,100,150} x nang {110,194,302,590} (uniform over centres),
the most accurate grid whose point count stays within 1.05 times the plain
autogrid's (Bragg-Slater radii, the reference defaults nrad 50 / nang 590):
accuracy at the reference's cost.  Every atom of the molecule gets that
label.

Fallback: the learned grid is accepted when its error is at most 1e-8, or
else when it moves the sum rule by no more than 1e-8 beyond the plain
autogrid's error; otherwise the plain autogrid's result is returned.

    python -m inca_parallel_amd.learned_grid train --out weights.npy
    python -m inca_parallel_amd.learned_grid ablate --weights weights.npy
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
SCALES = (0.7, 1.0, 1.4)
PLAIN = (50, 590)          # reference defaults (mod_input_reader.f90:55-66)
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
    """One molecule's context with its quartets prepared once."""

    def __init__(self, s, nohydro=True):
        self.s = s
        p = s.prims()
        self.p = p
        self.nohydro = nohydro
        ijkl, d = s.dm2_unique()
        self.e = capi.Engine(0)
        self.e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
        self.e.set_quartets(ijkl=ijkl, d=d)
        self.e.prepare(capi.INCA_FP64, THRESH)
        self.analytic = self.e.analytic_integral()

    def evaluate(self, cent, Ps, sf, nradc, nangc):
        """(relative sum-rule error, integral, points) on the given centres."""
        self.e.grid_multicenter(cent, nradc, nangc, sf, Ps, nosym=True)
        self.e.evaluate_prepared(capi.INCA_FP64, want_ivec=False)
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
    """Per-molecule label [log factor, nrad class, nang class] and the plain error."""
    e_plain, _, g_plain = m.uniform(1.0, *PLAIN)
    best = None
    for sc in SCALES:
        for ir, nr in enumerate(capi.NRAD_CLASSES):
            for ia, na in enumerate(capi.NANG_CLASSES):
                err, _, g = m.uniform(sc, nr, na)
                if g > budget * g_plain:
                    continue
                key = (err, g)
                if best is None or key < best[0]:
                    best = (key, sc, ir, ia, err, g)
    _, sc, ir, ia, err, g = best
    return np.array([np.log(sc), ir, ia], np.float32), dict(err_plain=e_plain, g_plain=g_plain,
                                                             err_label=err, g_label=g)


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


def cmd_train(a):
    t0 = time.time()
    feats, targets, meta = [], [], []
    for i in range(a.ntrain):
        m = Molecule(molecule(100000 + i))
        lab, info = label(m)
        f = m.features()
        feats.append(f)
        targets.append(np.repeat(lab[None, :], len(f), axis=0))
        meta.append(info)
        m.close()
        if i % 16 == 0:
            _log("labelled %d/%d (%.0f s): %s %s" % (i + 1, a.ntrain, time.time() - t0, lab, info))
    net = capi.MLP(seed=a.seed)
    hist = train(net, feats, targets, epochs=a.epochs, log=_log)
    np.save(a.out, net.params())
    net.close()
    print(json.dumps(dict(ntrain=a.ntrain, epochs=a.epochs, loss=hist, seconds=time.time() - t0)))


def cmd_ablate(a):
    net = capi.MLP(seed=1)
    net.set_params(np.load(a.weights, allow_pickle=False))
    rows = []
    for i in range(a.ntest):
        m = Molecule(molecule(200000 + i))
        r = learned_grid_evaluate(net, m)
        if "err_plain" not in r:
            r["err_plain"], _, r["points_plain"] = m.uniform(1.0, *PLAIN)
        if a.dense:
            r["err_dense"], _, r["points_dense"] = m.uniform(1.0, *DENSE)
        rows.append(r)
        m.close()
    net.close()
    acc = [r for r in rows if r["grid"] == "learned"]
    final = [r["err_learned"] if r["grid"] == "learned" else r["err_plain"] for r in rows]
    summary = dict(
        better_than_plain=int(sum(f < r["err_plain"] for f, r in zip(final, rows))),
        geomean_err_ratio_final_vs_plain=float(np.exp(np.mean(
            [np.log(max(f, 1e-16) / max(r["err_plain"], 1e-16)) for f, r in zip(final, rows)]))),
        molecules=len(rows), accepted=len(acc),
        median_points_ratio=float(np.median([r["points_learned"] / r["points_plain"] for r in rows])),
        median_err_plain=float(np.median([r["err_plain"] for r in rows])),
        median_err_learned=float(np.median([r["err_learned"] for r in rows])),
        median_err_final=float(np.median([r["err_learned"] if r["grid"] == "learned"
                                          else r["err_plain"] for r in rows])),
    )
    if a.dense:
        summary["median_err_dense"] = float(np.median([r["err_dense"] for r in rows]))
    print(json.dumps(dict(summary=summary, rows=rows)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--ntrain", type=int, default=128)
    t.add_argument("--epochs", type=int, default=600)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=WEIGHTS)
    b = sub.add_parser("ablate")
    b.add_argument("--weights", default=WEIGHTS)
    b.add_argument("--ntest", type=int, default=32)
    b.add_argument("--dense", action="store_true")
    a = ap.parse_args(argv)
    return {"train": cmd_train, "ablate": cmd_ablate}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
