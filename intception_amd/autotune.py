"""Autotuner for the per-class launch shape.

For every class bin of a plan it times the candidate team sizes T (groups of
G lanes that split one quartet's primitive quartets; only T whose teams keep
>= 90% of a wave's lanes busy) and grid sizes (one quartet per team, or a
persistent grid of k x 256 CUs), with HIP events on the launch stream, and
keeps the fastest.  Results are cached by (class, output kind, Boys mode,
log2 of the mean primitive-quartet count) in a JSON file so later plans with
the same shape start from the tuned choice.
"""
import json
import math
import os

import numpy as np

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
N_CU = 256


def lane_use(G, T):
    tl = G * T
    return 0.0 if tl > 64 else (64 // tl) * tl / 64.0


def team_candidates(G):
    best = max(lane_use(G, T) for T in range(1, 64 // G + 1))
    return [T for T in range(1, 64 // G + 1) if lane_use(G, T) >= 0.9 * best]


def grid_for(nq, G, T, mode):
    tpw = 64 // (G * T)
    per_block = 4 * tpw
    full = max(1, -(-nq // per_block))
    if mode == "auto":
        return full
    return max(1, min(full, N_CU * int(mode)))


def _key(b, spherical, boys_mode, mean_p, prec=0):
    return "%s%d%d%d%d_%s_b%d_p%d" % ("d" if prec == 0 else "f", *b["cls"], "s" if spherical else "c", boys_mode,
                                     int(round(math.log2(max(mean_p, 1.0)))))


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def save_cache(cache, path=CACHE):
    with open(path, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)


def bin_mean_prims(plan):
    """Mean primitive-quartet count of each bin, host side."""
    from .codegen import canonical_np
    b = plan.batch
    P = b.shell_nprim[b.quartets].astype(np.int64).prod(axis=1)
    canon = canonical_np(b.classes())
    key = canon[:, 0] * 1000 + canon[:, 1] * 100 + canon[:, 2] * 10 + canon[:, 3]
    out = {}
    for r in plan.bins():
        c = r["cls"]
        sel = key == c[0] * 1000 + c[1] * 100 + c[2] * 10 + c[3]
        out[r["bin"]] = float(P[sel].mean()) if sel.any() else 1.0
    return out


def apply_cache(plan, boys_mode=0, cache=None, precs=(0, 1)):
    cache = load_cache() if cache is None else cache
    means = bin_mean_prims(plan)
    n = 0
    for prec in precs:
        for r in plan.bins(prec):
            k = _key(r, plan.batch.spherical, boys_mode, means[r["bin"]], prec)
            if k in cache:
                T, mode = cache[k]["T"], cache[k]["grid"]
                if G_ok(r["G"], T):
                    plan.set_tuning(r["bin"], T, grid_for(r["nq"], r["G"], T, mode), prec)
                    n += 1
    return n


def G_ok(G, T):
    return 1 <= T and G * T <= 64


def tune(plan, out, reps=3, boys_mode=0, grid_modes=("auto", 3, 4, 8, 16), save=True, log=None, save_path=None,
         precs=(0,)):
    """Time every candidate per bin (full bins) and set the fastest; returns the table."""
    import torch
    stream = torch.cuda.current_stream(plan.device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    means = bin_mean_prims(plan)
    cache = load_cache()
    table = []
    plan.build_pairs()   # the class kernels read the primitive-pair arrays
    for prec in precs:
        for r in plan.bins(prec):
            best = None
            for T in team_candidates(r["G"]):
                for mode in grid_modes:
                    grid = grid_for(r["nq"], r["G"], T, mode)
                    plan.set_tuning(r["bin"], T, grid, prec)
                    # two untimed launches: after a synchronisation a bin's
                    # launches start up to 15 % slower and settle over a few
                    plan.run_bin(r["bin"], out, prec)
                    plan.run_bin(r["bin"], out, prec)
                    ev0.record(stream)
                    for _ in range(reps):
                        plan.run_bin(r["bin"], out, prec)
                    ev1.record(stream)
                    torch.cuda.synchronize(plan.device)
                    ms = ev0.elapsed_time(ev1) / reps
                    if best is None or ms < best[0]:
                        best = (ms, T, mode, grid)
            ms, T, mode, grid = best
            plan.set_tuning(r["bin"], T, grid, prec)
            cache[_key(r, plan.batch.spherical, boys_mode, means[r["bin"]], prec)] = {"T": T, "grid": mode,
                                                                                      "ms": round(ms, 4)}
            table.append({"cls": r["cls"], "nq": r["nq"], "prec": prec, "T": T, "grid": mode, "ms": ms})
            if log:
                log("tune %s%s nq=%d -> T=%d grid=%s %.3f ms" % ("d" if prec == 0 else "f", r["cls"], r["nq"], T,
                                                                mode, ms))
    if save:
        try:
            save_cache(cache)
        except OSError:
            pass
    if save_path:
        save_cache(cache, save_path)
    return table
