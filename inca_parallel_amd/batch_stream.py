"""Many molecules through the persistent batch, streamed in arena-sized groups.

The reference runs one molecule per process (inca_main.f90:66-77) and reads
each molecule's DM2 in batches of 40,960 records (mod_intracule.f90:66,
229-262).  Here a list of molecules is streamed through the engine:

  for each molecule: its symmetry-unique DM2 records are generated in device
      memory, prepared by the chunked pre-pass (inca_prepare; the records are
      freed afterwards, only the prepared quartets stay), and its autogrid
      (nosym: the records are symmetry-unique) is built;
  when the pending molecules (prepared quartets, grids, and the batch's
      partial rows and skip masks) would exceed the arena budget, the group
      is handed to inca_batch_create (class-sorted work queue, one
      hipGraph-captured persistent launch), whose copies let the contexts be
      closed at once; it is autotuned (or captured with the first group's
      choice), replayed, and released before the next group starts;
  a molecule larger than the whole arena is evaluated alone by its own
      context's grid loop.

Results do not depend on the grouping: every molecule's work items write
their own partial rows.  Used by the batch tests and bench.py.
"""
from __future__ import annotations

import time

import numpy as np

from inca_parallel_amd import capi


def prepared_bytes(engine, counts):
    """Device bytes a molecule adds to a batch arena: its prepared quartets
    (shell-quartet group headers of 96 bytes and their coefficients, FP64
    and FP32, when the context grouped them; else the records), its grid
    (points, weights and maps), and the batch's per-molecule scratch.  The
    scratch grows with items x points: one partial row of G doubles per
    chunk of at most 4096 items, and one 64-bit skip word per 64 items and
    64-point wave tile at one point per lane (the autotune's smallest tile),
    together 8 / 2048 bytes per (item, point)."""
    ng, _, ncoef = engine.groups()
    if ng > 0:
        items, stored = ng, ng * 96 + ncoef * 12
    else:
        stride = engine.prepared(0, 0).shape[1] if counts[2] else 0
        items, stored = int(counts[2]), int(counts[2]) * stride * 8
    return stored + engine.G * 8 * 8 + int(float(items) * engine.G * 8 / 2048)


def prepare_molecule(s, nrad, nang, precision, thresh=1e-12, exact=False):
    """An engine holding the molecule's autogrid and prepared quartets; the
    DM2 records are generated on the device and released after the
    pre-pass."""
    import torch
    e = capi.Engine(0)
    p = s.prims()
    e.set_primitives(p["alpha"], p["tmn"], p["x"], p["y"], p["z"])
    if exact:
        e.set_polynomials(True)
    # the groups alone, built chunk by chunk: no molecule holds all its
    # prepared records at once (the batch and the solo path take groups)
    e.set_keep_records(0)
    cent, Ps, sf = capi.autogrid(p["an"], p["cartes"], nohydro=True)
    nq = len(Ps)
    e.grid_multicenter(cent, [nrad] * nq, [nang] * nq, sf, Ps, nosym=True)
    recs = s.dm2_unique_torch(device="cuda")
    e.set_quartets(device_ptr=recs.data_ptr(), n=recs.shape[0])
    del recs
    torch.cuda.empty_cache()
    cnt = e.prepare(precision, thresh)
    return e, cnt


def evaluate_stream(molecules, precision=capi.INCA_FP64, arena_bytes=64 << 30, thresh=1e-12,
                    replays=2, autotune="first", single=(), log=None, solo_bytes=None):
    """Evaluate `molecules` (name, System, nrad, nang) in arena-sized groups.

    autotune: "first" times points-per-lane 1/2/4/8 on the first group and
    captures the later groups with its choice, "each" times every group,
    an int uses that value throughout.
    single: names of molecules also evaluated one context at a time (FP64,
    the same grid and prepared quartets) before batching, for comparison.
    solo_bytes: molecules needing more than this (default: the arena) are
    evaluated alone by their own context.
    Returns (per-molecule results, per-group results)."""
    mols, groups = [], []
    pending = []
    used = 0
    chosen = autotune if isinstance(autotune, int) and not isinstance(autotune, bool) else None

    def flush():
        nonlocal pending, used, chosen
        if not pending:
            return
        b = capi.Batch([m["engine"] for m in pending])
        # the batch holds its own copies: the contexts go before the replays
        for m in pending:
            m.pop("engine").close()
        if chosen is None or autotune == "each":
            best = b.autotune()
            if autotune == "first":
                chosen = best
        else:
            b.set_points(chosen)
            best = chosen
        outs = []
        t0 = time.perf_counter()
        for _ in range(replays):
            outs.append(b.evaluate())
        wall = (time.perf_counter() - t0) / max(1, replays)
        info = b.info()
        evals = float(sum(m["reference_evals"] for m in pending))
        same = all(np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])
                   for o in outs)
        groups.append(dict(molecules=len(pending), evals_per_replay=evals,
                           replay_ms=info["last_ms"], replay_wall_ms=wall * 1e3,
                           evals_per_s=evals / (info["last_ms"] * 1e-3),
                           points_per_lane=best, work_items=info["items"],
                           replays_bitwise_equal=same))
        for k, m in enumerate(pending):
            m["integral"], m["vee"] = float(outs[-1][0][k]), float(outs[-1][1][k])
            m["group"] = len(groups) - 1
        b.close()
        if log:
            log("group %d: %d molecules, %.3g evals, replay %.1f ms (%.3g evals/s), bitwise %s"
                % (len(groups) - 1, len(pending), evals, info["last_ms"],
                   groups[-1]["evals_per_s"], same))
        mols.extend(pending)
        pending, used = [], 0

    def solo(m):
        e = m.pop("engine")
        outs, ms = [], 0.0
        try:
            for _ in range(replays):
                e.evaluate_prepared(precision, want_ivec=False)
                outs.append(e.integrate()[:2])
                tm = e.timings()
                ms = float(tm[1] + tm[2])
        finally:
            e.close()
        same = all(o == outs[0] for o in outs)
        m["integral"], m["vee"] = float(outs[-1][0]), float(outs[-1][1])
        groups.append(dict(molecules=1, evals_per_replay=m["reference_evals"], replay_ms=ms,
                           replay_wall_ms=ms, evals_per_s=m["reference_evals"] / (ms * 1e-3),
                           points_per_lane=None, work_items=None, replays_bitwise_equal=same,
                           single_context=True))
        m["group"] = len(groups) - 1
        mols.append(m)
        if log:
            log("group %d: %s alone (larger than the arena), %.3g evals, %.1f ms, bitwise %s"
                % (len(groups) - 1, m["name"], m["reference_evals"], ms, same))

    for name, s, nrad, nang in molecules:
        e, cnt = prepare_molecule(s, nrad, nang, precision, thresh)
        need = prepared_bytes(e, cnt)
        npairs = s.nelec * (s.nelec - 1) / 2.0
        m = dict(name=name, engine=e, atoms=len(s.an), nprim=s.nprim, nelec=s.nelec,
                 nrad=nrad, nang=nang, points=e.G, counts=[int(c) for c in cnt],
                 reference_evals=float(cnt[2]) * e.G, npairs=npairs, arena_bytes=need,
                 analytic=e.analytic_integral())
        if name in single:
            e.evaluate_prepared(capi.INCA_FP64, want_ivec=False)
            m["single_fp64"] = e.integrate()[0]
            tm = e.timings()
            m["single_kernel_ms"] = float(tm[1] + tm[2])
        if log:
            log("prepared %s: %d atoms, %d primitives, %d quartets, %d points"
                % (name, m["atoms"], m["nprim"], cnt[2], m["points"]))
        # only the grid and the prepared quartets (their groups, when the
        # context made them) stay on the device
        e.release_scratch()
        if need > (arena_bytes if solo_bytes is None else solo_bytes):
            # larger than a whole arena: the context's own grid loop (rows
            # in bounded chunks) evaluates it as a group of one
            flush()
            solo(m)
            continue
        if pending and used + need > arena_bytes:
            flush()
        pending.append(m)
        used += need
    flush()
    for m in mols:
        # against the pair count N(N-1)/2 (the reference's check), and
        # against the exact integral of the prepared terms (the grid's own
        # quadrature error, free of the fits' roundoff)
        m["sum_rule_err"] = (m["integral"] - m["npairs"]) / m["npairs"]
        m["grid_err"] = (m["integral"] - m["analytic"]) / abs(m["analytic"])
    return mols, groups


def config4_by_cost(count):
    """The `count` cheapest molecules of the batched workload (seeds
    1000-1255) by records x grid points, in seed order."""
    from inca_parallel_amd import synth
    cost = []
    for seed in synth.CONFIG4_SEEDS:
        _, el, coords, alpha, _, _, _, _, nrad, nang = synth._config4_layout(seed)
        an = np.array([synth.ZNUM[x] for x in el], np.int32)
        cent, _, _ = capi.autogrid(an, coords, nohydro=True)
        npair = len(alpha) * (len(alpha) + 1) // 2
        cost.append((npair * (npair + 1) // 2 * len(cent) * nrad * nang, seed))
    return sorted(sd for _, sd in sorted(cost)[:count])


def config4_by_rank(lo, hi):
    """Seeds of the batched workload ranked lo..hi-1 by records x grid
    points (cheapest first), in rank order."""
    from inca_parallel_amd import synth
    cost = []
    for seed in synth.CONFIG4_SEEDS:
        _, el, coords, alpha, _, _, _, _, nrad, nang = synth._config4_layout(seed)
        an = np.array([synth.ZNUM[x] for x in el], np.int32)
        cent, _, _ = capi.autogrid(an, coords, nohydro=True)
        npair = len(alpha) * (len(alpha) + 1) // 2
        cost.append((npair * (npair + 1) // 2 * len(cent) * nrad * nang, seed))
    return [sd for _, sd in sorted(cost)[lo:hi]]


def main(argv=None):
    """Run molecules of the batched workload (ranks lo:hi by cost) through
    the stream in FP64 and mixed precision, every `single`-th of them also
    one context at a time, and write the per-molecule and per-group results
    as JSON."""
    import argparse
    import json
    import sys
    from inca_parallel_amd import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--rank", default="0:32")
    ap.add_argument("--single-every", type=int, default=8)
    ap.add_argument("--out", required=True)
    ap.add_argument("--precision", choices=("fp64", "mixed", "both"), default="both")
    ap.add_argument("--replays", type=int, default=2,
                    help="FP64 replays per group (2: replays compared bit for bit)")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()  # torch's HIP runtime before the library's first context
    lo, hi = (int(v) for v in a.rank.split(":"))
    seeds = config4_by_rank(lo, hi)
    mols = []
    for sd in seeds:
        s, nrad, nang = synth.config4_molecule(sd)
        mols.append(("seed%d" % sd, s, nrad, nang))
    single = {m[0] for k, m in enumerate(mols) if a.single_every and k % a.single_every == 0}

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    keep = ("name", "atoms", "nprim", "nelec", "nrad", "nang", "points", "counts",
            "reference_evals", "npairs", "analytic", "integral", "vee", "sum_rule_err", "grid_err",
            "group",
            "single_fp64", "single_kernel_ms")
    out = dict(rank=[lo, hi])

    def save():
        with open(a.out, "w") as f:
            json.dump(out, f)

    # results are written after each precision's pass
    t0 = time.time()
    if a.precision in ("fp64", "both"):
        r64, g64 = evaluate_stream(mols, capi.INCA_FP64, single=single, log=log,
                                   replays=a.replays)
        out.update(seconds_fp64=time.time() - t0, groups_fp64=g64,
                   fp64=[{k: m[k] for k in keep if k in m} for m in r64])
        save()
    t1 = time.time()
    if a.precision in ("mixed", "both"):
        rmx, gmx = evaluate_stream(mols, capi.INCA_MIXED, replays=1, log=log)
        out.update(seconds_mixed=time.time() - t1, groups_mixed=gmx,
                   mixed=[{k: m[k] for k in keep if k in m} for m in rmx])
        save()
    log("done: %d molecules in %.0f s" % (len(mols), time.time() - t0))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
