"""Autotuner of the KRP sampler's launch shapes and FP32/FP64 level split.

For every (precision, rank bucket, tree-depth bucket, sample-count bucket) it
builds a sampler on a representative shape, times every candidate setting of
the draw-kernel options (``draw_kernel``, ``group_lanes``, ``block_threads``,
``grid_blocks``) and of the computeM block size (``jacobi_threads``) on the
GPU, and for FP32 samplers picks the largest ``fp32_levels`` (the most levels
stored FP32, the fewest bytes) whose index mismatch against the FP64 sampler
on identical uniforms stays at or below ``mismatch_bar`` on Gaussian and on
ill-conditioned factors (s geometric from 1 to 1e-6).  The fastest setting of
each bucket is written as one row of the tuning table the library reads
(``ftl_tuning.tsv`` beside ``libftl_hip.so``; format in csrc/tuning.hpp).

Buckets cover the shapes of the reference's benchmark sweeps
(benchmarks/runtime_benchmark.py:26-58: I up to 2^25, R 16-128, N 2-9) and of
the serving stream (R in {8, 16, 32, 64}, I in [2^10, 2^22], J in
[2^10, 2^16]).  Usage: ``python -m fast_tensor_leverage_amd.autotune [--out FILE]
[--ranks 32,64]``.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "ftl_tuning.tsv")

# (R_lo, R_hi, representative R)
RANK_BUCKETS = [(1, 8, 8), (9, 16, 16), (17, 32, 32), (33, 64, 64), (65, 128, 128)]
# (depth_lo, depth_hi, representative depth): tree levels below the root
DEPTH_BUCKETS = [(0, 9, 8), (10, 14, 13), (15, 40, 18)]
# (J_lo, J_hi, representative J)
J_BUCKETS = [(1, 8191, 4096), (8192, 2 ** 40, 50_000)]
N_MODES = 3
MAX_FACTOR_ROWS = 2 ** 23


def tree_depth(I, R):
    """Levels below the root of the Gram tree over I rows with leaf size
    min(R, I) (partition_tree.hpp:83-105): the tuning table's depth key."""
    F = min(R, I)
    leaves = -(-I // F)
    lfill = leaves.bit_length() - 1
    return lfill + (0 if leaves == 1 << lfill else 1)


def representative_rows(R, depth):
    """Rows per factor whose tree (leaf size R) has `depth` levels below the
    root, capped so that one factor stays at most 2^23 rows."""
    return min(R * (1 << depth), MAX_FACTOR_ROWS)


def candidates(R, precision):
    """Draw-kernel settings to time for rank R (each a dict of options)."""
    out = []
    if R <= 8:
        out += [{"draw_kernel": 0, "group_lanes": 0}, {"draw_kernel": 1, "group_lanes": 0}]
    elif R <= 16:
        out += [{"draw_kernel": 0, "group_lanes": 0}, {"draw_kernel": 1, "group_lanes": 0}]
    elif R <= 32:
        out += [{"draw_kernel": 0, "group_lanes": g} for g in (16, 32, 64)]
        out += [{"draw_kernel": 1, "group_lanes": 0}]
    elif R <= 64:
        out += [{"draw_kernel": 0, "group_lanes": g} for g in (32, 64)]
        out += [{"draw_kernel": 2, "block_threads": b} for b in (128, 256)]
    else:
        out += [{"draw_kernel": 2, "block_threads": b} for b in (128, 256, 512)]
    return out


def _gaussian(torch, N, I, R, dev, seed):
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    return [torch.randn((I, R), generator=gen, device=dev, dtype=torch.float64) for _ in range(N)]


def _time_draws(smp, torch, d_s, d_w, draws, steps):
    """Median wall time per call over `steps` calls after two warm-ups, and
    the mean computeM kernel time (ms)."""
    for _ in range(2):
        smp.draw_device(N_MODES + 1, d_s, None, d_w, draws)
    torch.cuda.synchronize()
    smp.enable_timing(True)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        smp.draw_device(N_MODES + 1, d_s, None, d_w, draws)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    tm = smp.last_timing()
    smp.enable_timing(False)
    return float(np.median(times)) * 1e3, tm["gram_pinv_ms"] + tm["eigen_modes_ms"]


def _mismatch(ftl, torch, U, R, J, levels_list, dev, seed, rounds):
    """Index mismatch rate of FP32 samplers at each fp32_levels against the
    FP64 sampler, `rounds` draws of J samples on shared uniforms."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    draws = [torch.rand(2 * N_MODES * J, generator=gen, device=dev, dtype=torch.float64) for _ in range(rounds)]
    ref = ftl.Sampler.from_device(U, J, R, precision="fp64")
    d_s = torch.zeros((N_MODES, J), dtype=torch.int64, device=dev)
    s64 = []
    for d in draws:
        ref.draw_device(N_MODES + 1, d_s, None, None, d)
        s64.append(d_s.clone())
    ref.close()
    smp = ftl.Sampler.from_device(U, J, R, precision="fp32")
    rates = {}
    for lv in levels_list:
        smp.set_option("fp32_levels", lv)
        bad = 0
        for d, ref_s in zip(draws, s64):
            smp.draw_device(N_MODES + 1, d_s, None, None, d)
            bad += int(torch.any(d_s != ref_s, dim=0).sum())
        rates[lv] = bad / (rounds * J)
    smp.close()
    return rates


_LEVEL_CACHE = {}


def tune_bucket(ftl, torch, dev, precision, rb, db, jb, steps, mismatch_bar, rounds, log):
    R_lo, R_hi, R = rb
    d_lo, d_hi, depth = db
    j_lo, j_hi, J = jb
    I = representative_rows(R, depth)
    U = _gaussian(torch, N_MODES, I, R, dev, seed=7 + R + depth)
    opts = {}
    if precision == "fp32":
        # the most FP32 levels under the bar, on Gaussian and ill-conditioned
        # factors (mismatch rates at this I stand for the bucket's depths)
        from .synth import ill_conditioned_factors
        key = (R, I)
        if key not in _LEVEL_CACHE:
            levels = [lv for lv in (7, 6, 5, 4, 3, 2, 1) if lv <= depth + 1]
            r_g = _mismatch(ftl, torch, U, R, 50_000, levels, dev, 11, rounds)
            Ui, _ = ill_conditioned_factors(torch, N_MODES, I, R, dev, seed=31 + R)
            r_i = _mismatch(ftl, torch, Ui, R, 50_000, levels, dev, 12, rounds)
            del Ui
            ok = [lv for lv in levels if max(r_g[lv], r_i[lv]) <= mismatch_bar]
            # none under the bar: the fewest FP32 levels an FP32 sampler has
            _LEVEL_CACHE[key] = (max(ok) if ok else 1, r_g, r_i)
        lv, r_g, r_i = _LEVEL_CACHE[key]
        opts["fp32_levels"] = lv
        log(f"  fp32_levels: gaussian {r_g} ill-conditioned {r_i} -> {lv}")
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    draws = torch.rand(2 * N_MODES * J, generator=gen, device=dev, dtype=torch.float64)
    smp = ftl.Sampler.from_device(U, J, R, precision=precision)
    if "fp32_levels" in opts:
        smp.set_option("fp32_levels", opts["fp32_levels"])
    d_s = torch.zeros((N_MODES, J), dtype=torch.int64, device=dev)
    d_w = torch.empty((J,), dtype=torch.float64, device=dev)
    results = []
    for cand in candidates(R, precision):
        base = {"draw_kernel": 0, "group_lanes": 0, "block_threads": 0, "grid_blocks": 0}
        base.update(cand)
        for k, v in base.items():
            smp.set_option(k, v)
        ms, _ = _time_draws(smp, torch, d_s, d_w, draws, steps)
        results.append((ms, base))
        log(f"  {base}: {ms:.3f} ms")
    results.sort(key=lambda t: t[0])
    best_ms, best = results[0]
    for k, v in best.items():
        smp.set_option(k, v)
    jac = []
    for jt in (256, 512, 1024):
        smp.set_option("jacobi_threads", jt)
        _, cm = _time_draws(smp, torch, d_s, d_w, draws, max(3, steps // 2))
        jac.append((cm, jt))
    jac.sort()
    smp.close()
    opts.update({k: v for k, v in best.items() if v})
    opts["jacobi_threads"] = jac[0][1]
    row = {"precision": precision, "R": [R_lo, R_hi], "depth": [d_lo, d_hi], "J": [j_lo, j_hi],
           "opts": opts, "ms": best_ms, "shape": {"I": I, "R": R, "J": J, "N": N_MODES},
           "candidates": [{"ms": ms, **b} for ms, b in results], "jacobi_computeM_ms": dict((j, c) for c, j in jac)}
    del U
    torch.cuda.empty_cache()
    return row


def write_table(rows, path, note=""):
    lines = ["# KRP sampler tuning table, written by fast_tensor_leverage_amd/autotune.py",
             "# <precision> <R_lo> <R_hi> <depth_lo> <depth_hi> <J_lo> <J_hi> <options>"]
    if note:
        lines.append(f"# {note}")
    for r in rows:
        opt = ",".join(f"{k}={v}" for k, v in sorted(r["opts"].items()))
        lines.append(f"{r['precision']} {r['R'][0]} {r['R'][1]} {r['depth'][0]} {r['depth'][1]} "
                     f"{r['J'][0]} {r['J'][1]} {opt}  # {r['shape']} {r['ms']:.3f} ms")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--json", default=None, help="also write every bucket's measurements here")
    ap.add_argument("--steps", type=int, default=7)
    ap.add_argument("--rounds", type=int, default=4, help="50,000-sample draws per mismatch measurement")
    ap.add_argument("--mismatch-bar", type=float, default=5e-5)
    ap.add_argument("--precisions", default="fp32,fp64")
    ap.add_argument("--ranks", default="", help="comma-separated representative ranks to tune (default all)")
    args = ap.parse_args(argv)
    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import fast_tensor_leverage_amd as ftl
    dev = torch.device("cuda", 0)
    # time without the table in effect, so the candidates start from the
    # built-in defaults
    os.environ.setdefault("FTL_TUNING_TABLE", "")

    def log(msg):
        print(f"[autotune] {msg}", flush=True)

    ranks = {int(r) for r in args.ranks.split(",") if r}
    rows = []
    t0 = time.perf_counter()
    for precision in args.precisions.split(","):
        for rb in RANK_BUCKETS:
            if ranks and rb[2] not in ranks:
                continue
            for db in DEPTH_BUCKETS:
                for jb in J_BUCKETS:
                    log(f"{precision} R={rb[2]} depth={db[2]} J={jb[2]}")
                    rows.append(tune_bucket(ftl, torch, dev, precision, rb, db, jb, args.steps,
                                            args.mismatch_bar, args.rounds, log))
    write_table(rows, args.out, note=f"mismatch bar {args.mismatch_bar:g}; tuned in {time.perf_counter() - t0:.0f} s")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f, indent=1)
    log(f"wrote {len(rows)} rows to {args.out}")


if __name__ == "__main__":
    main()
