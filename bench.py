"""Benchmark of the intracule hot path on one MI355X.

Workload (the radial-plot configuration): a seeded 10-atom C/N/O/H cluster
(bond lengths 1.8-2.9 bohr), 151 Cartesian s/p/d primitives with exponents
log-uniform in [0.05, 500], an HF-like DM2 from 20 seeded occupied orbitals
written as symmetry-unique records (about 6.6e7, 1.6 GB), and I(s) for
s = 0..10 bohr in steps of 0.05 on 302-point Lebedev spheres, thresh 1e-12.
The records are generated directly in HBM before timing.

One step = one pass of the hot path over the resident records:
inca_evaluate = decode + normalise + screen 1 + prefactors + polycoef +
screen 2 + compaction + shell-quartet grouping + the quartet x grid-point
contraction (by groups) + reduction.  `value` is reference-equivalent
evals/s of the FP64 path (evals = quartets passing both screens x grid
points, the reference's count); the mixed-precision path is timed the same
way and reported beside it with its error against FP64, and the same step
record by record (the reference's loop order) beside both.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_VALU = 78.6e12   # MI355X FP64 vector, FLOP/s
PEAK_FP32_VALU = 157.3e12  # MI355X FP32 vector (packed), FLOP/s
THRESH = 1e-12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_eval(hist):
    L = np.arange(len(hist))
    return float(np.sum(hist * (16 + 2 * L))) / max(1, int(hist.sum()))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sysm, rec_host, grid_r, seconds_target=15.0):
    """Oracle (FP64 CPU restatement, OpenMP over quartets, all host cores) on
    a seeded random subset of the records over the full grid, built here
    with -O3 -march=native -fopenmp for this host."""
    import tempfile
    from oracle import oracle as O
    rng = np.random.default_rng(12345)
    cores = os.cpu_count() or 1
    prims = sysm.prims()
    n_total = len(rec_host)
    tmp = tempfile.mkdtemp(prefix="orc_native_")
    L = O.lib_native(os.path.join(tmp, "liborc_native.so"))

    def run(n):
        idx = np.sort(rng.choice(n_total, size=n, replace=False))
        sub = rec_host[idx]
        ijkl = np.ascontiguousarray(sub[:, :4])
        d = np.ascontiguousarray(sub[:, 4:6]).view(np.float64).reshape(-1)
        t0 = time.perf_counter()
        _, cnt = O.intracule(prims, ijkl, d, grid_r, THRESH, nthreads=cores, dynamic=True, L=L)
        return time.perf_counter() - t0, cnt

    n = 4000
    t, cnt = run(n)
    n2 = int(min(n_total, max(n, n * seconds_target / max(t, 1e-3))))
    t, cnt = run(n2)
    evals = float(cnt[2]) * grid_r.shape[0]
    return dict(value=evals / t, unit="evals/s", cores=cores, kind="port",
                cpu=cpu_model(), build="gcc -O3 -march=native -fopenmp (built on this host)",
                sample="%d seeded random records of %d (%d pass both screens) x %d grid "
                       "points, FP64 oracle, OpenMP dynamic, %.1f s" % (
                           n2, n_total, int(cnt[2]), grid_r.shape[0], t))


def record_loop(capi, eng, ref, evals):
    """The same FP64 step record by record (the reference's loop order: one
    exponential and three axis polynomials per record and point; ray tiles
    on a radial scan), for comparison with the grouped step."""
    eng.set_grouping(0)
    eng.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
    rec_tiles = eng.autotune()
    t0 = time.perf_counter()
    eng.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
    rec_step = time.perf_counter() - t0
    rec_kern = eng.timings()
    rec_rintra = eng.radial_values()
    hist = eng.class_histogram(64)
    eng.set_grouping(1)
    eng.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
    big = np.abs(ref) >= 1e-6 * np.abs(ref).max()
    return {
        "note": "the same step record by record (the reference's loop order; ray tiles)",
        "step_s": rec_step, "evals_per_s": evals / rec_step,
        "kernel_ms": {"prep": float(rec_kern[0]), "eval": float(rec_kern[1])},
        "points_per_lane": rec_tiles[0], "flop_per_eval": flops_per_eval(hist),
        "max_rel_err_rintra_vs_groups": float(np.max(np.abs(rec_rintra - ref)[big] /
                                                     np.abs(ref)[big])),
    }


def batch_bench(capi, synth, torch, count=32):
    """Batched molecules (the batched workload: seeded molecules of 4-20
    atoms, 30-300 s/p/d primitives, each on its own autogrid with nrad
    50-150 and nang 110-590, nosym, symmetry-unique records generated in
    HBM): the `count` cheapest of its 256 molecules by records x grid points,
    streamed through the persistent batch in arena-sized groups, in FP64 and
    mixed; against the same molecules evaluated one context at a time."""
    from inca_parallel_amd import batch_stream as BS
    seeds = BS.config4_by_cost(count)
    mols = []
    for sd in seeds:
        s, nrad, nang = synth.config4_molecule(sd)
        mols.append(("seed%d" % sd, s, nrad, nang))
    # every molecule also one context at a time, or every 4th beyond 32
    names = {m[0] for k, m in enumerate(mols) if len(mols) <= 32 or k % 4 == 0}
    r64, g64 = BS.evaluate_stream(mols, capi.INCA_FP64, single=names, log=log)
    rmx, gmx = BS.evaluate_stream(mols, capi.INCA_MIXED, log=log)
    evals = sum(g["evals_per_replay"] for g in g64)
    ms64 = sum(g["replay_ms"] for g in g64)
    msmx = sum(g["replay_ms"] for g in gmx)
    singles = [m for m in r64 if "single_fp64" in m]
    single_ms = sum(m["single_kernel_ms"] for m in singles)
    single_evals = sum(m["reference_evals"] for m in singles)
    err_single = max(abs(m["integral"] - m["single_fp64"]) / abs(m["single_fp64"])
                     for m in singles)
    err_mixed = max(abs(a["integral"] - b["integral"]) / abs(b["integral"])
                    for a, b in zip(rmx, r64))
    return {
        "workload": "batched molecules: the %d cheapest of the 256 seeded molecules (seeds "
                    "1000-1255, 4-20 atoms, 30-300 s/p/d primitives, nrad 50-150 x nang "
                    "110-590 autogrids, nosym)" % len(mols),
        "molecules": len(mols), "groups": len(g64),
        "atoms": [min(m["atoms"] for m in r64), max(m["atoms"] for m in r64)],
        "primitives": [min(m["nprim"] for m in r64), max(m["nprim"] for m in r64)],
        "evals_per_replay": evals,
        "value": evals / (ms64 * 1e-3), "unit": "evals/s",
        "replay_ms": ms64,
        "mixed_value": evals / (msmx * 1e-3), "mixed_replay_ms": msmx,
        "unbatched_kernel_evals_per_s": single_evals / (single_ms * 1e-3),
        "unbatched_molecules": len(singles),
        "max_rel_err_batch_vs_single_fp64": err_single,
        "max_rel_err_mixed_vs_fp64": err_mixed,
        "replays_bitwise_equal": all(g["replays_bitwise_equal"] for g in g64 + gmx),
        "median_abs_sum_rule_err": float(np.median([abs(m["sum_rule_err"]) for m in r64])),
        "points_per_lane": [g["points_per_lane"] for g in g64],
        "note": "replay = one hipGraph replay of a group's persistent class-sorted launch; "
                "unbatched = the same prepared quartets evaluated one context at a time "
                "(eval + reduce kernel time)",
    }


def target_bench(capi, torch, nmol=32):
    """Radial intracule integrals of many small molecules (3-5 atoms, s/p/d,
    the learned grid's held-out family), from records and grids already
    resident on the device: the plain FP64 GPU kernel (the reference's loop:
    record by record, every term at every point, one inca_evaluate per
    molecule) against the engine with everything on (shell-quartet groups,
    mixed precision, per-point cutoff, learned autogrid, one batched
    hipGraph replay, the sum-rule fallback).  The plain kernel is timed on
    the plain autogrid (nrad 50 / nang 590 per centre, Bragg-Slater radii)
    and on the learned grid; the engine's integrals are checked against the
    FP64 full loop on the grid it used."""
    from inca_parallel_amd import learned_grid as LG
    net = capi.MLP(seed=1)
    net.set_params(np.load(LG.WEIGHTS, allow_pickle=False))
    mols = []
    for i in range(nmol):
        s = LG.molecule(300000 + i)
        p = s.prims()
        recs = s.dm2_unique_torch(device="cuda")
        mols.append(dict(s=s, p=p, recs=recs))
    torch.cuda.synchronize()

    def engine(m, cent, Ps, sf, nr, na):
        p = m["p"]
        e = capi.Engine(0)
        e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
        e.grid_multicenter(cent, nr, na, sf, Ps, nosym=True)
        e.set_quartets(device_ptr=m["recs"].data_ptr(), n=m["recs"].shape[0])
        return e

    def plain_grid(m):
        p = m["p"]
        cent, Ps, sf = capi.autogrid(p["an"], p["cartes"], True, False)
        nq = len(Ps)
        return cent, Ps, sf, [LG.PLAIN[0]] * nq, [LG.PLAIN[1]] * nq

    def learned(m):
        p = m["p"]
        cent, Ps, sf, nr, na, _ = net.learned_autogrid(p["an"], p["cartes"], p["alpha"], p["tmn"],
                                                       m["s"].ra, nohydro=True)
        return cent, Ps, sf, list(nr), list(na)

    # the plain FP64 kernel: record by record, every term at every point, one
    # molecule at a time (contexts, grids and records set up beforehand)
    t_plain = {"plain_grid": 0.0, "learned_grid": 0.0}
    evals = {"plain_grid": 0.0, "learned_grid": 0.0}
    for key, make in (("plain_grid", plain_grid), ("learned_grid", learned)):
        engines = [engine(m, *make(m)) for m in mols]
        for e in engines:
            e.set_grouping(0)
            e.set_point_cutoff(0)
        engines[0].evaluate(capi.INCA_FP64, THRESH, want_ivec=False)  # warm-up
        torch.cuda.synchronize()
        for m, e in zip(mols, engines):
            t0 = time.perf_counter()
            _, cnt = e.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
            tot = e.integrate()[0]
            t_plain[key] += time.perf_counter() - t0
            evals[key] += float(cnt[2]) * e.G
            m[key] = dict(fp64=tot, analytic=e.analytic_integral(), points=e.G)
        for e in engines:
            e.close()
    # the engine: learned grid, groups, mixed, cutoff, batched (contexts,
    # grids and records set up beforehand)
    engines = [engine(m, *learned(m)) for m in mols]
    torch.cuda.synchronize()
    # the contexts prepare on host threads at once (each context has its own
    # stream; a small molecule's pre-pass is mostly launch and wait latency)
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(max_workers=8)
    t0 = time.perf_counter()
    list(pool.map(lambda e: e.prepare(capi.INCA_MIXED, THRESH), engines))
    t_prep = time.perf_counter() - t0
    b = capi.Batch(engines)
    t_create = time.perf_counter() - t0 - t_prep
    t1 = time.perf_counter()
    ti, _ = b.evaluate()
    t_first = time.perf_counter() - t1
    replay = []
    for _ in range(3):
        b.evaluate()
        replay.append(b.info()["last_ms"] * 1e-3)
    t_engine = t_prep + t_create + t_first
    # the sum-rule fallback: the plain autogrid is evaluated (mixed, on the
    # prepared quartets) for every molecule whose learned-grid integral is
    # more than 1e-8 from the analytic integral
    t1 = time.perf_counter()
    err_l = [abs(ti[i] - m["learned_grid"]["analytic"]) / abs(m["learned_grid"]["analytic"])
             for i, m in enumerate(mols)]
    check = [i for i in range(nmol) if err_l[i] > LG.FALLBACK_TOL]
    tp = {}

    def plain_on(i):  # the plain autogrid on the context that holds the prepared quartets
        cent, Ps, sf, nr, na = plain_grid(mols[i])
        engines[i].grid_multicenter(cent, nr, na, sf, Ps, nosym=True)

    list(pool.map(plain_on, check))
    t_grids = time.perf_counter() - t1
    if check:  # one batch over them (groups, mixed), as the learned grids
        bp = capi.Batch([engines[i] for i in check])
        tpv, _ = bp.evaluate()
        bp.close()
        tp = {i: float(tpv[k]) for k, i in enumerate(check)}
    final, used = [], []
    for i, m in enumerate(mols):
        ia = m["learned_grid"]["analytic"]
        val, grid = ti[i], "learned_grid"
        if i in tp and err_l[i] > abs(tp[i] - ia) / abs(ia) + LG.FALLBACK_TOL:
            val, grid = tp[i], "plain_grid"
        final.append(val)
        used.append(grid)
    t_fallback = time.perf_counter() - t1
    pool.shutdown()
    for e in engines:
        e.close()
    b.close()
    net.close()
    # accuracy: the engine against the FP64 full loop on the grid it used
    err_fast = []
    sum_rule = {"plain_grid": [], "learned_grid": [], "engine": []}
    for m, v, grid in zip(mols, final, used):
        r = m[grid]["fp64"]
        err_fast.append(abs(v - r) / abs(r))
        for key in ("plain_grid", "learned_grid"):
            sum_rule[key].append(abs(m[key]["fp64"] - m[key]["analytic"]) / abs(m[key]["analytic"]))
        sum_rule["engine"].append(abs(v - m["learned_grid"]["analytic"]) /
                                  abs(m["learned_grid"]["analytic"]))
    return {
        "molecules": nmol,
        "plain_fp64_s": {k: round(v, 4) for k, v in t_plain.items()},
        "plain_fp64_evals": evals,
        "engine_s": t_engine,
        "engine_parts_s": {"prepare_mixed": t_prep, "batch_create": t_create,
                           "first_replay": t_first, "replay": float(np.mean(replay))},
        "fallback_check_s": t_fallback,
        "fallback_grid_build_s": t_grids,
        "speedup_vs_plain_same_grid": t_plain["learned_grid"] / t_engine,
        "speedup_vs_plain_same_grid_replay": t_plain["learned_grid"] / float(np.mean(replay)),
        "speedup_vs_plain_with_fallback": t_plain["plain_grid"] / (t_engine + t_fallback),
        "max_rel_err_engine_vs_fp64_same_grid": float(max(err_fast)),
        "fallbacks": used.count("plain_grid"),
        "median_sum_rule_err": {k: float(np.median(v)) for k, v in sum_rule.items()},
        "note": "from records and grids resident on the device (context set-up excluded on both "
                "sides); plain: inca_evaluate FP64 record by record with the cutoff off (the "
                "reference's full loop) plus the integral, one molecule at a time, on the learned "
                "grid (same grid) and on the plain autogrid; engine: inca_prepare(INCA_MIXED) per "
                "molecule on 8 host threads (pre-pass and shell-quartet groups, each context on "
                "its own stream), batch create and the first replay "
                "on the learned grid (replay: later replays); fallback check: the plain autogrids "
                "of the molecules whose learned-grid integral misses the analytic integral by more "
                "than 1e-8, built on their contexts and evaluated in one batch (mixed, by groups, "
                "on the prepared quartets)",
    }


PROFILE = os.path.join(ROOT, "profiles", "geval_fp64_radial_plot_cluster151.md")


def profiled_traffic():
    """HBM bytes of one FP64 step's group kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated on a 1 GiB
    copy)."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_fp64_step_eval"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def dm2p_file_bench(capi, eng, recs, step_s, rintra):
    """The drop-in path from a file: the radial-plot workload's records written
    as a stream-format .dm2p (24-byte records) in the host's temporary
    directory, flushed and dropped from the page cache, then read by
    inca_load_dm2p (pinned double-buffered copies to the device) and one FP64
    step on them; a second read from the page cache beside it."""
    import tempfile
    host = recs.cpu().numpy()
    nbytes = host.nbytes
    fd, path = tempfile.mkstemp(suffix=".dm2p")
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(host.tobytes())
            f.flush()
            os.fsync(f.fileno())
        del host
        with open(path, "rb") as f:
            os.posix_fadvise(f.fileno(), 0, 0, os.POSIX_FADV_DONTNEED)
        t0 = time.perf_counter()
        n, fmt = eng.load_dm2p(path)
        t_cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        eng.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
        t_step = time.perf_counter() - t0
        same = bool(np.array_equal(eng.radial_values(), rintra))
        t0 = time.perf_counter()
        eng.load_dm2p(path)
        t_warm = time.perf_counter() - t0
    finally:
        os.unlink(path)
    return {
        "bytes": nbytes, "records": int(n), "format": int(fmt),
        "load_s": t_cold, "load_GB_per_s": nbytes / t_cold / 1e9,
        "load_page_cache_s": t_warm, "load_page_cache_GB_per_s": nbytes / t_warm / 1e9,
        "step_after_load_s": t_step, "drop_in_s": t_cold + t_step,
        "step_resident_s": step_s, "rintra_equal_to_resident_path": same,
        "note": "stream .dm2p written to the host's temporary directory, fsync'd and dropped from "
                "the page cache (posix_fadvise), read by inca_load_dm2p; drop-in = that read plus "
                "one FP64 step (pre-pass and grid loop); a second read from the page cache beside "
                "it; the bench line's value keeps the records resident",
    }


def cube_bench(capi, eng):
    """The cube workload on the same cluster and DM2: 80^3 voxels, step 0.15
    bohr, centred at the origin (the cubeintra grid).  One FP64 and one mixed
    grid loop over the prepared quartets (the per-axis contraction: every
    term at every voxel), with the mixed error per voxel."""
    n = eng.grid_cube([0.0, 0.0, 0.0], [0.15, 0.15, 0.15], [80, 80, 80])
    out = {"points": n, "method": "per-axis contraction (every quartet at every voxel) on the "
                                  "matrix cores, factor tiles staged in LDS: FP64; mixed: its FP32 "
                                  "range on FP32 factors and FP32 matrix cores, folded into FP64"}
    # the FP64 contraction on the vector units, for comparison
    eng.prepare(capi.INCA_FP64, THRESH)
    eng.set_cube_contraction(1)
    eng.evaluate_prepared(capi.INCA_FP64, want_ivec=False)
    tm = eng.timings()
    out["fp64_vector_units_eval_ms"] = float(tm[1] + tm[2])
    eng.set_cube_contraction(2)
    for prec, name in ((capi.INCA_FP64, "fp64"), (capi.INCA_MIXED, "mixed")):
        eng.prepare(prec, THRESH)
        t0 = time.perf_counter()
        iv = eng.evaluate_prepared(prec, want_ivec=True)
        wall = time.perf_counter() - t0
        tm = eng.timings()
        ref, ex = eng.eval_counts()
        out[name] = {"eval_ms": float(tm[1] + tm[2]), "wall_s": wall,
                     "evals_per_s": ref / ((tm[1] + tm[2]) * 1e-3)}
        out["reference_evals"], out["executed_evals"] = ref, ex
        if name == "fp64":
            iv64 = iv
    big = np.abs(iv64) >= 1e-6 * np.abs(iv64).max()
    out["mixed"]["max_rel_err_voxels_vs_fp64"] = float(
        np.max(np.abs(iv - iv64)[big] / np.abs(iv64[big])))
    out["mixed"]["rel_err_sum_vs_fp64"] = float(abs(iv.sum() - iv64.sum()) / abs(iv64.sum()))
    # FP64 by shell-quartet groups on the z lines of the cube (ray tiles,
    # per-point cutoff), against the contraction's every term at every voxel
    eng.set_cube_contraction(0)
    try:
        eng.prepare(capi.INCA_FP64, THRESH)
        eng.evaluate_prepared(capi.INCA_FP64, want_ivec=False)
        t0 = time.perf_counter()
        ivg = eng.evaluate_prepared(capi.INCA_FP64, want_ivec=True)
        wall = time.perf_counter() - t0
        tm = eng.timings()
        ref, ex = eng.eval_counts()
        out["groups_fp64"] = {
            "eval_ms": float(tm[1] + tm[2]), "wall_s": wall,
            "evals_per_s": ref / ((tm[1] + tm[2]) * 1e-3),
            "executed_group_evals": ex, "groups": int(eng.groups()[0]),
            "max_rel_err_voxels_vs_contraction": float(
                np.max(np.abs(ivg - iv64)[big] / np.abs(iv64[big]))),
            "max_abs_err_vs_contraction_over_max": float(
                np.abs(ivg - iv64).max() / np.abs(iv64).max()),
            "note": "shell-quartet groups on ray tiles along z (eight points per lane, the "
                    "polynomial's later values from its difference table), per-point cutoff "
                    "2^-100 of each term's peak"}
    finally:
        eng.set_cube_contraction(2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dm2p-file", action="store_true",
                    help="skip the leg that reads the workload from a .dm2p file")
    ap.add_argument("--natoms", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-molecules", type=int, default=32,
                    help="molecules of the batched workload in the batch leg (cheapest first; "
                         "256 is the whole workload, several hours)")
    ap.add_argument("--no-cube", action="store_true")
    ap.add_argument("--no-record-loop", action="store_true",
                    help="skip the record-by-record comparison step (profiling runs)")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the learned-grid + mixed + batched leg against the plain FP64 kernel")
    ap.add_argument("--tiles", default=None,
                    help="points per lane 'fp64,fp32' instead of autotuning (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one device copy of 1 GiB and exit: the reference run for the "
                         "rocprofv3 byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        return 2

    import torch
    from inca_parallel_amd import capi, synth

    if a.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d bytes" % (src.numel() * 8))
        return 0

    t0 = time.perf_counter()
    sysm = synth.config_cluster(seed=2, natoms=a.natoms)
    recs = sysm.dm2_unique_torch(device="cuda")
    torch.cuda.synchronize()
    Q = recs.shape[0]
    log("generated %d records (%d primitives) in %.1f s" % (Q, sysm.nprim, time.perf_counter() - t0))

    eng = capi.Engine(0)
    p = sysm.prims()
    eng.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    nradi, G = eng.grid_radial_plot([[0.0, 10.0]], [0.05], [302])
    eng.set_quartets(device_ptr=recs.data_ptr(), n=Q)
    log("grid: %d radii, %d points" % (nradi, G))

    # tile autotuning on this grid and these quartets (outside the timed steps)
    if a.tiles:
        tiles = tuple(int(v) for v in a.tiles.split(","))
        eng.set_tiles(*tiles)
        log("points per lane set: fp64 %d, fp32 %d" % tiles)
    else:
        eng.evaluate(capi.INCA_FP64, THRESH, want_ivec=False)
        tiles = eng.autotune()
        log("autotuned points per lane: fp64 %d, fp32 %d" % tiles)

    results = {}
    for prec, name in ((capi.INCA_FP64, "fp64"), (capi.INCA_MIXED, "mixed")):
        for _ in range(a.warmup):
            eng.evaluate(prec, THRESH, want_ivec=False)
        torch.cuda.synchronize()
        kern = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            _, cnt = eng.evaluate(prec, THRESH, want_ivec=False)
            kern.append(eng.timings())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        kern = np.array(kern)
        results[name] = dict(step_s=dt, counts=cnt, kern_ms=kern.mean(axis=0),
                             rintra=eng.radial_values(), groups=eng.groups(),
                             evals=eng.eval_counts(), flops=eng.eval_flops())
        log("%s: %.3f s/step, kernels (prep, eval, reduce) ms = %s, counts %s" % (
            name, dt, np.round(kern.mean(axis=0), 2), cnt))

    cnt = results["fp64"]["counts"]
    q2 = int(cnt[2])
    evals = float(q2) * G
    r64 = results["fp64"]
    rm = results["mixed"]
    eval_s = r64["kern_ms"][1] * 1e-3
    executed = float(r64["evals"][1])
    ngroups, _, ncoef = eng.groups()
    # the roofline counts the work the group kernel executed: 14 + 2 M FP64
    # operations per (group, point) pair it evaluated (M coefficients; the
    # cutoff skips (group, tile) pairs whose terms are below 2^-100 of their
    # peak)
    achieved = r64["flops"] / eval_s
    W = r64["flops"] / max(executed, 1.0)
    ref = r64["rintra"]
    mixed_err = float(np.max(np.abs(rm["rintra"] - ref) / np.maximum(np.abs(ref), 1e-300)))
    # the record loop (the reference's order: one exponential and three axis
    # polynomials per record and point; ray tiles on this scan) for comparison
    rec = None
    if not a.no_record_loop:
        rec = record_loop(capi, eng, ref, evals)

    out = {
        "metric": "quartet x gridpoint evals/s; time to 1e-8 rel integral, mixed vs FP64 GPU",
        "value": evals / r64["step_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r64["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cluster, basis and HF-like DM2 generated in HBM)",
        "config": {
            "workload": "radial_plot: 10-atom C/N/O/H cluster, %d s/p/d primitives, "
                        "%d symmetry-unique DM2 records, s=0..10 step 0.05 x LD0302 "
                        "(%d points), thresh 1e-12" % (sysm.nprim, Q, G),
            "records": Q, "quartets_after_screen1": int(cnt[1]),
            "quartets_computed": q2, "grid_points": G,
            "shell_quartet_groups": ngroups, "group_coefficients": ncoef,
            "flop_per_executed_group_eval": W,
            "reference_evals": evals, "executed_group_evals": executed,
            "point_cutoff": "2^-100 of each term's peak",
            "points_per_lane": {"fp64": tiles[0], "fp32": tiles[1]},
        },
        "roofline": {
            "bound": "valu", "kernel": "k_geval (shell-quartet groups, FP64)",
            "achieved": achieved / 1e12, "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_VALU, "traffic": profiled_traffic(),
            "note": "FLOP = 14 + 2 M per executed (group, point) pair (M coefficients of the "
                    "group's polynomial; FMA = 2; exp not counted), summed by the skip-mask pass "
                    "(inca_get_eval_flops); grid loop %.1f ms per step (HIP events on the engine "
                    "stream around the tile bounds, skip mask and group kernel); traffic = HBM "
                    "bytes of the group kernel per step from rocprofv3 FETCH_SIZE + WRITE_SIZE (%s)"
                    % (eval_s * 1e3, os.path.relpath(PROFILE, ROOT)),
        },
        "record_loop": rec,
        "mixed": {
            "value": evals / rm["step_s"], "unit": "evals/s",
            "speedup_vs_fp64": r64["step_s"] / rm["step_s"],
            "max_rel_err_rintra_vs_fp64": mixed_err,
            "eval_kernel_ms": float(rm["kern_ms"][1]),
            "groups_fp64": int(rm["groups"][1]),
            "groups_fp32": int(rm["groups"][0] - rm["groups"][1]),
        },
        "dm2_stream": {"bytes": 24 * Q, "prep_ms": float(r64["kern_ms"][0]),
                       "GB_per_s": 24 * Q / (float(r64["kern_ms"][0]) * 1e-3) / 1e9,
                       "note": "24-byte .dm2p records decoded, screened, prepared and grouped "
                               "per step"},
        "kernel_ms": {"prep": float(r64["kern_ms"][0]), "eval": float(r64["kern_ms"][1]),
                      "reduce": float(r64["kern_ms"][2])},
    }
    grid_r = eng.get_grid()["r"]  # the radial-plot grid, for the CPU baseline
    if not a.no_dm2p_file:
        import shutil
        import tempfile
        need = 2 * recs.numel() * recs.element_size()
        try:
            if shutil.disk_usage(tempfile.gettempdir()).free < need:
                out["dm2p_file"] = {"skipped": "less than %d bytes free in %s" % (
                    need, tempfile.gettempdir())}
            else:
                out["dm2p_file"] = dm2p_file_bench(capi, eng, recs, r64["step_s"], r64["rintra"])
        except OSError as err:  # the file leg only: the measured step is done
            out["dm2p_file"] = {"skipped": "file I/O: %s" % err}
        log("dm2p_file: %s" % out["dm2p_file"])
    if not a.no_cube:
        out["cube"] = cube_bench(capi, eng)
        log("cube: %s" % out["cube"])
    if not a.no_batch:
        out["batch"] = batch_bench(capi, synth, torch, a.batch_molecules)
        log("batch: %s" % out["batch"])
    if not a.no_target:
        out["target"] = target_bench(capi, torch)
        log("target: %s" % out["target"])
    if not a.no_cpu:
        rec_host = recs.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(sysm, rec_host, grid_r)
    print(json.dumps(out))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
