#!/usr/bin/env python3
"""Headline benchmark: exact KRP leverage-score draws on one MI355X.

Workload (one "step" = one KRPDrawSamples over the whole batch): N=3 factor
matrices of I=2^23 x R=32, FP64 N(0,1) entries (rng 1000+k), J=50,000 samples
of the KRP of all modes (j = N+1, ~5.9e20 rows), uniforms rng(7) resident in
HBM, outputs (samples, h, weights) written to HBM.  Each step runs computeM
(two kernels) and the fused draw kernel.  Both the FP32 sampler (FP64 top
tree levels, FP32 bottom levels; the headline value) and the plain FP64
sampler are timed; the FP32/FP64 ratio and the index mismatch rate between
them are reported.  The CPU baseline is the FP64 oracle (OpenMP C++
restatement of the reference) timed on this host's cores on a bounded sample
of the same workload, and its indices are compared with the GPU's.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N_MODES, I_ROWS, RANK, J_SAMPLES = 3, 2 ** 23, 32, 50_000
HBM_PEAK = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(I, R, J, N, s_G, s_U):
    """Algorithmic HBM bytes of one draw (all N modes sampled): per sample and
    mode, the packed upper triangle (T = R(R+1)/2 entries of s_G bytes) of the
    left child read at each of the L levels below the root, the F x R leaf
    block and the R-entry factor row of the h update (s_U bytes per entry):
    B = sum_k J [L_k T s_G + F R s_U + R s_U].  The root Gram and the eigen
    trees (read by every sample) are excluded.  s_G = s_U = 4: 6.336 GB at
    N=3, I=2^23, R=32, J=50,000; s_G = s_U = 8: 12.672 GB."""
    F = min(R, I)
    leaves = -(-I // F)
    depth = leaves.bit_length() - 1 + (0 if leaves & (leaves - 1) == 0 else 1)
    T = R * (R + 1) // 2
    per = depth * T * s_G + F * R * s_U + R * s_U
    return per * J * N


def lscpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(U, draws, J, j, gpu_samples, sample_j, trials=7):
    """Oracle (FP64 OpenMP restatement) on J' = sample_j samples of the same
    workload, trees prebuilt: 1 warm-up and `trials` timed draws; the value is
    J' over the median time (the minimum is reported beside it).  Returns
    the baseline record and the index comparison with the GPU FP64 sampler."""
    sys.path.insert(0, HERE)
    from oracle import oracle as orc
    orc.build()
    N = len(U)
    t0 = time.perf_counter()
    ref = orc.OracleSampler(U, sample_j, U[0].shape[1])
    t_build = time.perf_counter() - t0
    sub = np.concatenate([draws[s * J: s * J + sample_j] for s in range(2 * N)])
    ref.draw(j, sub, want_margins=False)  # warm-up
    times = []
    for _ in range(trials):
        t0 = time.perf_counter()
        ref.draw(j, sub, want_margins=False)
        times.append(time.perf_counter() - t0)
    out = ref.draw(j, sub, want_margins=True)
    t_med, t_min = float(np.median(times)), float(min(times))
    diff = np.zeros(sample_j, dtype=bool)
    for k in range(N):
        diff |= gpu_samples[k, :sample_j] != out["samples"][k]
    sure = out["margins"] > 1e-10
    rec = {"value": sample_j / t_med, "unit": "samples/s", "cores": orc.num_threads(), "kind": "port",
           "value_best": sample_j / t_min, "trial_s": times,
           "sample": f"KRPDrawSamples over J'={sample_j} samples of the same N=3, I=2^23, R=32 workload "
                     f"(trees prebuilt in {t_build:.1f} s, not timed); median of {trials} timed draws after "
                     f"1 warm-up; host {lscpu_model()}"}
    cmp = {"oracle_vs_gpu_fp64_mismatches": int(diff.sum()),
           "oracle_vs_gpu_fp64_mismatches_outside_margin": int((diff & sure).sum()),
           "oracle_compared_samples": sample_j}
    del ref
    return rec, cmp


# ---------------------------------------------------------------------------
# sampled CP-ALS workload (--workload als)
# ---------------------------------------------------------------------------
ALS_SIZES = {
    # dims, nnz, R, J, sweeps
    "c5": ((2 ** 22, 2 ** 21, 2 ** 17), 200_000_000, 25, 2 ** 16, 10),
    "small": ((2 ** 16, 2 ** 15, 2 ** 12), 2_000_000, 25, 2 ** 16, 10),
    # the shapes and nonzero counts of the FROSTT tensors of the reference's
    # STS-CP timings, filled with synthetic Zipf data (the real tensors are
    # not available here)
    "uber": ((183, 24, 1140, 1717), 3_309_490, 25, 2 ** 16, 10),
    "enron": ((6066, 5699, 244268, 1176), 54_202_099, 25, 2 ** 16, 10),
    "nell2": ((12092, 9184, 28818), 76_879_419, 25, 2 ** 16, 10),
}
# the reference's published STS-CP seconds per sweep on the real tensors (R=25,
# J=2^16, 128 threads of a Perlmutter CPU node; plotting/generate_images.ipynb:1230-1234)
ALS_REFERENCE_S_PER_SWEEP = {"uber": 0.224, "enron": 0.462, "nell2": 0.570}


def als_main(args):
    import ctypes
    import torch
    sys.path.insert(0, HERE)
    import fast_tensor_leverage_amd as ftl
    from fast_tensor_leverage_amd import _lib
    dims, nnz, R, J, sweeps = ALS_SIZES[args.als_size]
    if args.sweeps:
        sweeps = args.sweeps
    dev = torch.device("cuda", 0)
    N = len(dims)
    log(f"ALS workload {args.als_size}: dims={dims} nnz={nnz} R={R} J={J} sweeps={sweeps}")
    t0 = time.perf_counter()
    from fast_tensor_leverage_amd.synth import zipf_sparse
    idx, vals = zipf_sparse(torch, dims, nnz, R, dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    log(f"generated {nnz} nonzeros in {t_gen:.1f} s")
    L = _lib.load()
    h = ctypes.c_void_p()
    dims_arr = np.array(dims, dtype=np.uint64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.ftl_sparse_create_device(N, nnz, idx.data_ptr(), vals.data_ptr(), dims_arr.ctypes.data,
                                          ctypes.byref(h)))
    t_build = time.perf_counter() - t0
    log(f"sparse tensor built in {t_build:.2f} s")
    t_hash = None
    if args.als_lookup == "hash":
        t0 = time.perf_counter()
        _lib.check(L.ftl_sparse_set_lookup(h, 1), "set_lookup")
        t_hash = time.perf_counter() - t0
        log(f"coordinate hash of every mode built in {t_hash:.3f} s")
    X = ftl.SparseTensor.__new__(ftl.SparseTensor)
    X._h, X.N, X.nnz, X.dims, X.last_found = h, N, nnz, list(dims), 0
    X._destroy = L.ftl_sparse_destroy
    host_idx = host_vals = None
    if args.als_oracle:
        host_idx = idx.cpu().numpy().view(np.uint32).copy()
        host_vals = vals.cpu().numpy().copy()
    del idx, vals
    torch.cuda.empty_cache()

    lhs = ftl.PyLowRank(dims, R, seed=5)
    U_init = [u.copy() for u in lhs.U]
    als = ftl.ALS(lhs.ten, X, host_sync=False)
    lhs.ten.renormalize_columns(-1)
    sigma0 = lhs.ten.sigma.copy()
    U_norm = [u.copy() for u in lhs.U]
    als.initialize_ds_als(J, "efficient", precision=args.als_precision)
    als.enable_timing(True)
    rng = np.random.default_rng(99)
    draws_all = [[rng.random(2 * N * J) for j in range(N)] for _ in range(sweeps)]
    fits, fit_s, upd_ms, phases, stats = [], [], [], [], []

    def fit():
        t = time.perf_counter()
        f = 1.0 - np.sqrt(als.residual_normsq()) / np.sqrt(X.get_normsq())
        fit_s.append(time.perf_counter() - t)
        return f

    fits.append(fit())
    log(f"initial fit {fits[-1]:.6f} ({fit_s[-1] * 1e3:.1f} ms)")
    gpu_samples = []  # per update, for the oracle comparison (copied outside the timed call)
    t_all = time.perf_counter()
    for sw in range(sweeps):
        for j in range(N):
            t = time.perf_counter()
            als.execute_ds_als_update(j, True, True, random_draws=draws_all[sw][j])
            upd_ms.append((time.perf_counter() - t) * 1e3)
            st = als.last_stats()
            stats.append(st)
            if args.als_oracle:
                gpu_samples.append(als.last_draw()[0])
        if (sw + 1) % 5 == 0 or sw + 1 == sweeps:
            fits.append(fit())
        log(f"sweep {sw + 1}: last update {upd_ms[-1]:.2f} ms, found {stats[-1]['found']}, fit {fits[-1]:.6f}")
    t_total = time.perf_counter() - t_all
    upd = np.array(upd_ms)
    ph = {k: float(np.mean([s[k] for s in stats])) for k in
          ("draw_ms", "dedup_pinv_probe_ms", "mttkrp_solve_ms", "renormalize_rebuild_ms")}
    line = {
        "metric": "sampled CP-ALS (STS-CP) seconds per sweep, sparse Zipf(1.1) tensor",
        "value": float(upd.sum() / 1e3 / sweeps),
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": sweeps,
        "warmup": 0,
        "ms_per_step": float(upd.sum() / sweeps),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.als_precision == "fp64" else "f64 (FP32 sampler trees)",
        "data": "synthetic: Zipf(1.1) coordinates, rank-25 CP values + 1% noise, generated on the device",
        "config": {"workload": f"ALS {args.als_size}: dims={list(dims)} nnz={nnz} R={R} J={J}",
                   "sweeps": sweeps, "precision": args.als_precision, "lookup": args.als_lookup},
        "update_ms": {"mean": float(upd.mean()), "median": float(np.median(upd)), "max": float(upd.max())},
        "phase_ms_mean": ph,
        "unique_samples_mean": float(np.mean([s["unique_samples"] for s in stats])),
        "matched_nonzeros_mean": float(np.mean([s["found"] for s in stats])),
        "touched_rows_mean": float(np.mean([s["touched_rows"] for s in stats])),
        "fits": fits,
        "fit_eval_s": fit_s,
        "wall_s_updates_and_fits": t_total,
        "sparse_construction_s": t_build,
        "hash_construction_s": t_hash,
        "eigensolves_capped": ftl.eigensolver_capped(),
        # published on the real tensor of this shape (different data, different machine)
        "reference_real_tensor_s_per_sweep": ALS_REFERENCE_S_PER_SWEEP.get(args.als_size),
        "generation_s": t_gen,
    }
    if args.exact_sweeps:
        # execute_exact_als_update sweeps continuing from the sampled result;
        # the first update of each mode also sorts the nonzeros by that mode
        ex_ms = []
        for sw in range(args.exact_sweeps):
            for j in range(N):
                t = time.perf_counter()
                als.execute_exact_als_update(j, True, True)
                ex_ms.append((time.perf_counter() - t) * 1e3)
            log(f"exact sweep {sw + 1}: updates {', '.join(f'{x:.1f}' for x in ex_ms[-N:])} ms")
        steady = ex_ms[N:] if len(ex_ms) > N else ex_ms
        line["exact"] = {"sweeps": args.exact_sweeps, "first_update_per_mode_ms": ex_ms[:N],
                         "update_ms_mean": float(np.mean(steady)), "s_per_sweep": float(np.mean(steady)) * N / 1e3,
                         "fit_after": float(fit())}
    if args.als_oracle:
        line.update(als_oracle_compare(host_idx, host_vals, dims, R, J, U_norm, sigma0, draws_all, fits, sweeps,
                                       gpu_samples))
    print(json.dumps(line), flush=True)


def als_oracle_compare(idx, vals, dims, R, J, U_norm, sigma0, draws_all, gpu_fits, sweeps, gpu_samples):
    """The same sweeps on the CPU FP64 oracle from the same initial factors.
    At every update the oracle first draws with its own factors and the same
    uniforms, and its samples are compared with the GPU's (mismatches and the
    oracle's CDF margin at each); the oracle then updates with the GPU's
    samples, so that a sample flipped by eigensolver rounding near a CDF
    boundary does not send the two runs down different random trajectories.
    Returns the oracle's fits, their relative difference to the GPU's, the
    sample comparison and the time of one oracle sweep (the CPU baseline)."""
    sys.path.insert(0, HERE)
    from oracle import oracle as orc
    orc.build()
    N = len(dims)
    log(f"oracle: sorted lookups of {idx.shape[0]} nonzeros on {orc.num_threads()} threads")
    t0 = time.perf_counter()
    S = orc.OracleSparse(idx, vals)
    t_build = time.perf_counter() - t0
    log(f"oracle: lookups built in {t_build:.1f} s; sampler trees and the initial fit")
    ref = orc.OracleALS(S, [u.copy() for u in U_norm], J, sigma=sigma0)
    fits = [ref.fit()]
    log(f"oracle: initial fit {fits[0]:.6f}")
    times, mism, margins = [], [], []
    u = 0
    for sw in range(sweeps):
        for j in range(N):
            t = time.perf_counter()
            d = ref.sampler.draw(j, draws_all[sw][j])
            g = gpu_samples[u]
            ref.update_with_samples(j, g)
            times.append(time.perf_counter() - t)
            rows = [k for k in range(N) if k != j]
            bad = np.any(d["samples"][rows] != g[rows], axis=0)
            mism.append(int(bad.sum()))
            margins.extend(float(x) for x in d["margins"][bad])
            u += 1
        if (sw + 1) % 5 == 0 or sw + 1 == sweeps:
            fits.append(ref.fit())
        log(f"oracle sweep {sw + 1}: fit {fits[-1]:.6f}; sample mismatches {mism[-N:]}")
    rel = [abs(a - b) / abs(b) for a, b in zip(gpu_fits, fits)]
    return {"oracle_fits": fits, "fit_rel_diff_vs_oracle": rel,
            "oracle_sample_comparison": {
                "mismatches_per_update": mism, "samples_per_update": J,
                "mismatch_rate": float(np.sum(mism) / (len(mism) * J)),
                "max_oracle_margin_of_a_mismatch": max(margins) if margins else None,
                "note": "oracle draws with its own factors and the GPU's uniforms; the oracle then updates "
                        "with the GPU's samples"},
            "cpu_baseline": {"value": float(np.sum(times) / sweeps), "unit": "s/sweep", "cores": orc.num_threads(),
                             "kind": "port",
                             "sample": f"the same {sweeps} sweeps of updates on the oracle (sparse lookups built in "
                                       f"{t_build:.1f} s; each update timed with its draw); host {lscpu_model()}"}}


# ---------------------------------------------------------------------------
# serving stream (--workload serve, and one rate in the default run)
# ---------------------------------------------------------------------------
def serve_run(rate, n_requests, budget_gb, slots, precision="fp32", seed=0):
    sys.path.insert(0, HERE)
    from fast_tensor_leverage_amd.serve import Server, random_requests, serve_stream
    reqs = random_requests(n_requests, seed=seed)
    srv = Server(int(budget_gb * (1 << 30)), slots)
    # warm-up: one small request (kernel loading)
    srv.wait(srv.submit([1024, 1024, 1024], 8, 1024, None, 1, precision))
    out = serve_stream(srv, reqs, rate, seed=seed + 1, precision=precision)
    lat = out["latency_s"] * 1e3
    total = sum(q["J"] for q in reqs)
    st = srv.stats()
    return {"rate_per_s": rate, "requests": n_requests, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "samples_per_s": total / out["makespan_s"], "makespan_s": out["makespan_s"],
            "peak_arena_GB": st["peak_bytes"] / 2 ** 30, "budget_GB": budget_gb, "slots": slots,
            "precision": precision}


def serve_resident_run(rates, n_requests, n_sets, budget_gb, slots, precision="fp32", seed=0):
    """Draw requests on registered factor sets whose trees stay resident: n_sets
    sets with the serving stream's shapes (N in {3,4,5}, I log-uniform in
    [2^10, 2^22], R in {8,16,32,64}; FP64 N(0,1) factors generated on the
    device) are registered once; each request draws J (log-uniform in
    [2^10, 2^16]) samples from a set picked uniformly.  No tree is built while
    requests are served, so every draw goes through the packed launches."""
    import torch
    sys.path.insert(0, HERE)
    from fast_tensor_leverage_amd.serve import Server, random_requests, serve_stream
    dev = torch.device("cuda", 0)
    srv = Server(int(budget_gb * (1 << 30)), slots)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 5)
    sets = []
    t0 = time.perf_counter()
    for q in random_requests(n_sets, seed=seed + 7):
        U = [torch.randn((I, q["R"]), generator=gen, device=dev, dtype=torch.float64) for I in q["dims"]]
        sets.append(srv.add_factors(U, 2 ** 16, precision))
        del U
    for fs in sets:  # trees built, kernels loaded
        srv.wait(srv.submit_on(fs, 1024, None, 1))
    t_reg = time.perf_counter() - t0
    rng = np.random.default_rng(seed + 11)
    reqs = [{"fs": sets[int(rng.integers(n_sets))], "J": int(2 ** rng.uniform(10, 16)), "seed": 200_000 + i}
            for i in range(n_requests)]
    total = sum(q["J"] for q in reqs)
    rows = []
    for rate in rates:
        out = serve_stream(srv, reqs, rate, seed=seed + 1, precision=precision)
        lat = out["latency_s"] * 1e3
        rows.append({"rate_per_s": rate, "requests": n_requests, "factor_sets": n_sets,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
                     "samples_per_s": total / out["makespan_s"], "makespan_s": out["makespan_s"],
                     "peak_arena_GB": srv.stats()["peak_bytes"] / 2 ** 30, "budget_GB": budget_gb,
                     "registration_s": t_reg, "precision": precision})
        log(f"resident sets, rate {rate}/s: p50 {rows[-1]['p50_ms']:.2f} ms p99 {rows[-1]['p99_ms']:.2f} ms "
            f"{rows[-1]['samples_per_s']:,.0f} samples/s")
    for fs in sets:
        srv.remove_factors(fs)
    return rows


def serve_main(args):
    rates = [float(r) for r in args.serve_rates.split(",")]
    rows = []
    if args.serve_sets:
        rrates = [float(r) for r in args.serve_resident_rates.split(",")]
        res = serve_resident_run(rrates, args.serve_requests, args.serve_sets, args.serve_budget_gb,
                                 args.serve_slots)
        best = max(res, key=lambda x: x["samples_per_s"])
        line = {"metric": "KRP draw-request serving on resident factor sets: aggregate samples/s; "
                          "p50/p99 request latency",
                "value": best["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.serve_requests,
                "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32 (FP64 top tree levels, computeM and h in f64)",
                "data": "synthetic: FP64 N(0,1) factor sets generated on the device, uniforms from the request seed",
                "config": {"workload": f"{args.serve_sets} registered factor sets (N in {{3,4,5}}, I log-uniform "
                                       "[2^10,2^22], R in {8,16,32,64}); Poisson requests with J log-uniform "
                                       "[2^10,2^16]", "requests": args.serve_requests},
                "p99_latency_ms": best["p99_ms"], "rates": res}
        print(json.dumps(line), flush=True)
        return
    for r in rates:
        rows.append(serve_run(r, args.serve_requests, args.serve_budget_gb, args.serve_slots))
        log(f"rate {r}/s: p50 {rows[-1]['p50_ms']:.2f} ms p99 {rows[-1]['p99_ms']:.2f} ms "
            f"{rows[-1]['samples_per_s']:,.0f} samples/s")
    best = max(rows, key=lambda x: x["samples_per_s"])
    line = {"metric": "KRP draw-request serving: aggregate samples/s; p50/p99 request latency",
            "value": best["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.serve_requests,
            "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (FP64 top tree levels, computeM and h in f64)",
            "data": "synthetic: per-request N(0,1) factors (FP32-rounded) and uniforms from the request seed",
            "config": {"workload": "serving stream: Poisson arrivals, N in {3,4,5}, I log-uniform [2^10,2^22], "
                                   "R in {8,16,32,64}, J log-uniform [2^10,2^16]",
                       "requests": args.serve_requests},
            "p99_latency_ms": best["p99_ms"], "rates": rows}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# the reference's runtime sweeps (--workload sweep)
# ---------------------------------------------------------------------------
# Published CPU means (seconds; 128 OpenMP threads of a Perlmutter node) of
# benchmarks/runtime_benchmark.py, from submission_data/oct_30/runtime_bench.json
# (I_trace: R=32, N=3; R_trace and N_trace: I=2^22): (I, R, N) -> (sampling, construction)
REF_SWEEP = {
    (2 ** 6, 32, 3): (0.0412, 0.0010), (2 ** 7, 32, 3): (0.0336, 0.0008), (2 ** 8, 32, 3): (0.0345, 0.0010),
    (2 ** 9, 32, 3): (0.0418, 0.0014), (2 ** 10, 32, 3): (0.0437, 0.0014), (2 ** 11, 32, 3): (0.0451, 0.0016),
    (2 ** 12, 32, 3): (0.0491, 0.0023), (2 ** 13, 32, 3): (0.0749, 0.0053), (2 ** 14, 32, 3): (0.0810, 0.0058),
    (2 ** 15, 32, 3): (0.0686, 0.0074), (2 ** 16, 32, 3): (0.0705, 0.0097), (2 ** 17, 32, 3): (0.0742, 0.0141),
    (2 ** 18, 32, 3): (0.1408, 0.0296), (2 ** 19, 32, 3): (0.1563, 0.0428), (2 ** 20, 32, 3): (0.1762, 0.0639),
    (2 ** 21, 32, 3): (0.1771, 0.1244), (2 ** 22, 32, 3): (0.1885, 0.2438), (2 ** 23, 32, 3): (0.2189, 0.4806),
    (2 ** 24, 32, 3): (0.2635, 0.9738), (2 ** 25, 32, 3): (0.3109, 2.0254),
    (2 ** 22, 16, 3): (0.0835, 0.1505), (2 ** 22, 64, 3): (0.6744, 0.4978), (2 ** 22, 128, 3): (2.6808, 1.0201),
    (2 ** 22, 32, 2): (0.1285, 0.1770), (2 ** 22, 32, 4): (0.2421, 0.3313), (2 ** 22, 32, 5): (0.3044, 0.4160),
    (2 ** 22, 32, 6): (0.3804, 0.4933), (2 ** 22, 32, 7): (0.4460, 0.5732), (2 ** 22, 32, 8): (0.5099, 0.6523),
    (2 ** 22, 32, 9): (0.5724, 0.7361),
}


def sweep_main(args):
    """The configurations of the reference's runtime benchmark
    (benchmarks/runtime_benchmark.py:33-118: I from 2^6 to 2^25 at R=32, N=3;
    R from 16 to 128 and N from 2 to 9 at I=2^22; J=50,000; every mode sampled,
    j = N+1), each with `trials` constructions and draws as there.  Timed like
    the reference: construction = the sampler built from factors already in
    memory (here HBM, N(0,1) from torch's generator), sampling = one
    KRPDrawSamples with its built-in uniforms into a host buffer, synchronised.
    Reported per configuration: the median over the trials for the FP64 and FP32
    samplers, the reference's published CPU mean and the ratio."""
    import torch
    sys.path.insert(0, HERE)
    import fast_tensor_leverage_amd as ftl
    dev = torch.device("cuda", 0)
    J = J_SAMPLES
    rows = []
    t_all = time.perf_counter()
    for (I, R, N), (ref_s, ref_c) in REF_SWEEP.items():
        gen = torch.Generator(device=dev)
        gen.manual_seed(I + 7 * R + 131 * N)
        U = [torch.randn((I, R), generator=gen, device=dev, dtype=torch.float64) for _ in range(N)]
        row = {"I": I, "R": R, "N": N, "ref_sampling_s": ref_s, "ref_construction_s": ref_c}
        for prec in ("fp64", "fp32"):
            buf = np.zeros((N, J), dtype=np.uint64)
            build, draw = [], []
            for t in range(args.trials + 1):  # the first one is a warm-up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                smp = ftl.Sampler.from_device(U, J, R, precision=prec)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                smp.KRPDrawSamples(N + 1, buf)
                t2 = time.perf_counter()
                smp.close()
                if t:
                    build.append(t1 - t0)
                    draw.append(t2 - t1)
            row[f"{prec}_sampling_s"] = float(np.median(draw))
            row[f"{prec}_construction_s"] = float(np.median(build))
            row[f"{prec}_samples_per_s"] = J / row[f"{prec}_sampling_s"]
            assert np.all(buf < I), "sample out of range"
        row["sampling_ratio_fp64"] = ref_s / row["fp64_sampling_s"]
        row["sampling_ratio_fp32"] = ref_s / row["fp32_sampling_s"]
        log(f"sweep I=2^{I.bit_length() - 1} R={R} N={N}: draw fp64 {1e3 * row['fp64_sampling_s']:.2f} ms, "
            f"fp32 {1e3 * row['fp32_sampling_s']:.2f} ms (reference CPU {1e3 * ref_s:.1f} ms: "
            f"x{row['sampling_ratio_fp64']:.0f} / x{row['sampling_ratio_fp32']:.0f}); construction fp64 "
            f"{1e3 * row['fp64_construction_s']:.1f} ms, fp32 {1e3 * row['fp32_construction_s']:.1f} ms "
            f"(reference {1e3 * ref_c:.1f} ms)")
        rows.append(row)
        del U
        torch.cuda.empty_cache()
    head = [r for r in rows if (r["I"], r["R"], r["N"]) == (2 ** 23, 32, 3)][0]
    line = {"metric": "KRPDrawSamples over the reference's runtime sweeps (I, R, N), j=N+1, J=50,000",
            "value": head["fp32_samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.trials,
            "warmup": 1, "ms_per_step": 1e3 * head["fp32_sampling_s"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 top tree levels) and f64",
            "data": "synthetic: N(0,1) factors in HBM, the sampler's built-in Philox uniforms",
            "config": {"workload": "benchmarks/runtime_benchmark.py sweeps: I 2^6..2^25 (R=32, N=3), "
                                   "R 16..128 and N 2..9 (I=2^22); median of "
                                   f"{args.trials} after 1 warm-up, host-timed with a host sample buffer"},
            "rows": rows, "wall_s": time.perf_counter() - t_all}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# mixed precision on ill-conditioned factors (--workload c3)
# ---------------------------------------------------------------------------
def c3_main(args):
    """N=3, I=2^23, J=50,000 at R in {16, 64, 128}: U_k = Q_U diag(s) Q_V^T with
    Q_U (I x R) and Q_V (R x R) orthonormal from a seeded QR and s geometric
    from 1 to 1e-6.  The FP64 and FP32 samplers draw on identical uniforms;
    reported: samples/s of each, the index mismatch rate, and the relative
    difference of the leverage probabilities of the drawn rows
    (p = hᵀ pinv h / R = 1 / (J w))."""
    import torch
    sys.path.insert(0, HERE)
    import fast_tensor_leverage_amd as ftl
    dev = torch.device("cuda", 0)
    N, I, J = N_MODES, args.I, J_SAMPLES
    rows = []
    for R in [int(r) for r in args.c3_ranks.split(",")]:
        log(f"C3 R={R}: building ill-conditioned factors")
        from fast_tensor_leverage_amd.synth import ill_conditioned_factors
        U, gen = ill_conditioned_factors(torch, N, I, R, dev, seed=3000 + R)
        draws = torch.rand(2 * N * J, generator=gen, device=dev, dtype=torch.float64)
        res = {}
        for prec in ("fp64", "fp32"):
            smp = ftl.Sampler.from_device(U, J, R, precision=prec)
            d_s = torch.zeros((N, J), dtype=torch.int64, device=dev)
            d_w = torch.empty((J,), dtype=torch.float64, device=dev)
            smp.draw_device(N + 1, d_s, None, d_w, draws)
            torch.cuda.synchronize()
            smp.enable_timing(True)
            t0 = time.perf_counter()
            for _ in range(args.c3_steps):
                smp.draw_device(N + 1, d_s, None, d_w, draws)
            tm = smp.last_timing()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            res[prec] = {"samples_per_s": args.c3_steps * J / el, "kernel_ms": tm,
                         "samples": d_s.cpu().numpy().copy(), "p": (1.0 / (J * d_w)).cpu().numpy(),
                         "fp32_levels": smp.get_option("fp32_levels") if prec == "fp32" else None}
            log(f"C3 R={R} {prec}: {res[prec]['samples_per_s']:,.0f} samples/s, draw {tm['draw_ms']:.3f} ms")
            smp.close()
            del smp, d_s, d_w
            torch.cuda.synchronize()
        diff = np.any(res["fp32"]["samples"] != res["fp64"]["samples"], axis=0)
        # the FP32 path's probability of each row it drew, 1 / (J w), against
        # the FP64 probability of the same row, h^T pinv(G) h / R (h the FP64
        # product of the drawn rows, pinv with the 1e-8 cut of the Hadamard
        # product of the FP64 Grams)
        G = torch.ones((R, R), dtype=torch.float64, device=dev)
        for M in U:
            G *= M.T @ M
        lam, V = torch.linalg.eigh(G)
        inv = torch.where(lam > 1e-8, 1.0 / torch.where(lam > 1e-8, lam, torch.ones_like(lam)),
                          torch.zeros_like(lam))
        pinv = (V * inv) @ V.T
        rows32 = torch.from_numpy(res["fp32"]["samples"]).to(dev)
        hh = torch.ones((J, R), dtype=torch.float64, device=dev)
        for k in range(N):
            hh *= U[k][rows32[k]]
        p64 = (torch.einsum("ir,rs,is->i", hh, pinv, hh) / R).cpu().numpy()
        p32 = res["fp32"]["p"]
        del hh, rows32
        # rows whose leverage under the cut pinv is zero (hᵀ pinv h = 0: the
        # row lies in the discarded eigenspace) have no relative error
        ok = (p64 > 0) & np.isfinite(p64) & np.isfinite(p32)
        rel = np.abs(p32[ok] - p64[ok]) / p64[ok]
        rows.append({"R": R, "fp64_samples_per_s": res["fp64"]["samples_per_s"],
                     "fp32_samples_per_s": res["fp32"]["samples_per_s"],
                     "fp32_over_fp64": res["fp32"]["samples_per_s"] / res["fp64"]["samples_per_s"],
                     "draw_ms": {"fp64": res["fp64"]["kernel_ms"]["draw_ms"],
                                 "fp32": res["fp32"]["kernel_ms"]["draw_ms"]},
                     "computeM_ms": res["fp64"]["kernel_ms"]["gram_pinv_ms"] + res["fp64"]["kernel_ms"]["eigen_modes_ms"],
                     "index_mismatch_rate": float(diff.mean()),
                     "fp32_levels": res["fp32"]["fp32_levels"],
                     "leverage_prob_rel_diff_max": float(rel.max()) if rel.size else None,
                     "leverage_prob_rel_diff_mean": float(rel.mean()) if rel.size else None,
                     "drawn_rows_with_zero_leverage_fp64": float(1.0 - ok.mean())})
        del U, draws
        torch.cuda.empty_cache()
    line = {"metric": "KRP samples/s, FP32 vs FP64 sampler, ill-conditioned factors (s from 1 to 1e-6)",
            "value": rows[0]["fp32_samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.c3_steps,
            "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (FP64 top tree levels) and f64",
            "data": "synthetic: U = Q_U diag(geomspace(1,1e-6,R)) Q_V^T from seeded QR, uniforms seeded",
            "config": {"workload": f"KRPDrawSamples N=3 I={I} J={J} all modes, R in {args.c3_ranks}"},
            "ranks": rows}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Larsen-Kolda sampler (--workload lk)
# ---------------------------------------------------------------------------
def lk_update_bytes(I, R):
    """Algorithmic HBM bytes of update_sampler(k) on an I x R FP64 factor: U
    read twice (its Gram, then the leverage scores, which need the pinv of
    the whole Gram), scores written and read once by the prefix pass, prefix
    sums written."""
    return 2 * I * R * 8 + 3 * I * 8


def lk_main(args):
    """Sampler(U, J, R, "larsen_kolda") at N=3, I=2^23, R=32, J=50,000, all
    modes sampled (j = N+1), with h and weights: draw samples/s (factors and
    uniforms resident in HBM) and update_sampler(k) on the resident factor,
    both timed with HIP events on the stream the kernels run on."""
    import torch
    sys.path.insert(0, HERE)
    import fast_tensor_leverage_amd as ftl
    N, I, R, J = N_MODES, args.I, RANK, J_SAMPLES
    j = N + 1
    log(f"LK: generating factors N={N} I={I} R={R}")
    U = [np.random.default_rng(1000 + k).standard_normal((I, R)) for k in range(N)]
    draws = np.random.default_rng(7).random(2 * N * J)
    dev = torch.device("cuda", 0)
    d_draws = torch.from_numpy(draws).to(dev)
    t0 = time.perf_counter()
    s = ftl.Sampler(U, J, R, "larsen_kolda")
    t_build = time.perf_counter() - t0
    d_s = torch.zeros((N, J), dtype=torch.int64, device=dev)
    d_h = torch.empty((J, R), dtype=torch.float64, device=dev)
    d_w = torch.empty((J,), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream()
    for _ in range(args.warmup):
        s.draw_device(j, d_s, d_h, d_w, d_draws, st)
        for k in range(N):
            s.update_sampler(k)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(args.steps):
        s.draw_device(j, d_s, d_h, d_w, d_draws, st)
    e1.record(st)
    torch.cuda.synchronize()
    draw_el = time.perf_counter() - t0
    draw_ms = e0.elapsed_time(e1) / args.steps
    # update_sampler(k) on the resident factor, on the sampler's own stream
    ext = torch.cuda.ExternalStream(s.stream(), device=dev)
    u0, u1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    u0.record(ext)
    for _ in range(args.steps):
        for k in range(N):
            s.update_sampler(k)
    u1.record(ext)
    torch.cuda.synchronize()
    upd_ms = u0.elapsed_time(u1) / (args.steps * N)
    ubytes = lk_update_bytes(I, R)
    achieved = ubytes / (upd_ms * 1e-3)
    gpu_samples = d_s.cpu().numpy().view(np.uint64).copy()
    gpu_w = d_w.cpu().numpy().copy()
    log(f"LK: draw {draw_ms * 1e3:.1f} us ({J / (draw_ms * 1e-3):,.0f} samples/s); update_sampler "
        f"{upd_ms:.3f} ms per mode ({achieved / 1e9:.0f} GB/s)")
    # the hybrid variant on the same factors (its draw synchronises: the
    # candidate count sizes the enumeration)
    del s
    hs = ftl.Sampler(U, J, R, "larsen_kolda_hybrid")
    hs.draw_device(j, d_s, d_h, d_w, d_draws, st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hs.draw_device(j, d_s, d_h, d_w, d_draws, st)
    torch.cuda.synchronize()
    hyb_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    hyb_info = hs.last_info()
    del hs
    log(f"LK hybrid: draw {hyb_ms:.3f} ms (host-clocked, synchronising); {hyb_info}")
    cpu = None
    cmp = {}
    if not args.skip_cpu:
        # oracle on the first 2^20 rows of each factor (its update scales
        # linearly in I): construction, one update_sampler, the draw of all J
        # samples; the index check runs the GPU sampler on the same rows
        from oracle import oracle as orc
        orc.build()
        Is = min(I, 2 ** 20)
        Us = [u[:Is].copy() for u in U]
        t0 = time.perf_counter()
        ref = orc.OracleLKSampler(Us, J, R)
        t_create = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref.update_sampler(0)
        t_upd = time.perf_counter() - t0
        ref.draw(j, draws)
        t0 = time.perf_counter()
        out = ref.draw(j, draws)
        t_draw = time.perf_counter() - t0
        small = ftl.Sampler(Us, J, R, "larsen_kolda")
        gs = np.zeros((N, J), dtype=np.uint64)
        gh = np.zeros((J, R))
        gw = np.zeros(J)
        small.KRPDrawSamples_materialize(j, gs, gh, gw, draws)
        sure = out["margins"] > 1e-9
        diff = np.any(gs != out["samples"], axis=0)
        cpu = {"value": J / t_draw, "unit": "samples/s", "cores": 1, "kind": "port",
               "sample": f"oracle LarsenKoldaSampler on the first {Is} rows of each factor: draw of J={J} "
                         f"(all modes, h and weights) {t_draw * 1e3:.1f} ms; update_sampler(0) {t_upd:.2f} s "
                         f"(x{I / Is:.0f} for I={I}: {t_upd * I / Is:.1f} s); construction {t_create:.1f} s; "
                         f"host {lscpu_model()}",
               "update_s_scaled_to_I": t_upd * I / Is}
        cmp = {"oracle_vs_gpu_mismatches": int(diff.sum()),
               "oracle_vs_gpu_mismatches_outside_margin": int((diff & sure).sum()),
               "oracle_compared_samples": J,
               "oracle_weights_max_rel_diff": float(np.max(np.abs(gw[~diff] - out["weights"][~diff])
                                                           / out["weights"][~diff]))}
        log(f"LK CPU oracle: draw {t_draw * 1e3:.1f} ms, update {t_upd:.2f} s at I={Is}; {cmp}")
        del small
    line = {"metric": "Larsen-Kolda KRP samples/s (N=3,I=2^23,R=32), update_sampler ms",
            "value": J / (draw_ms * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * draw_el / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic: N(0,1) factors rng(1000+k), uniforms rng(7), resident in HBM",
            "config": {"workload": "Larsen-Kolda KRPDrawSamples N=3 I=2^23 R=32 J=50000 j=N+1 (all modes)",
                       "N": N, "I": I, "R": R, "J": J, "j": j},
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK, "traffic": None,
                         "kernel": "update_sampler(k): k_gram_mfma + k_pinv_partials + k_lev_rows + tile scan",
                         "kernel_ms": upd_ms, "algorithmic_bytes_per_update": ubytes},
            "draw_ms": draw_ms, "update_ms_per_mode": upd_ms, "construction_s": t_build,
            "hybrid_draw_ms": hyb_ms, "hybrid_last_draw": hyb_info,
            "weights_finite": bool(np.all(np.isfinite(gpu_w))),
            "samples_in_range": bool(np.all(gpu_samples < I)),
            "cpu_baseline": cpu}
    line.update(cmp)
    print(json.dumps(line), flush=True)


def device_copy_bandwidth(dev, gib=4, reps=5):
    """Bytes read plus written per second by a device-to-device copy of a
    `gib` GiB tensor (torch's copy kernel, HIP events on its stream): the
    bandwidth a streaming kernel reaches on this device, a second denominator
    beside the 8 TB/s peak."""
    import torch
    n = gib * 2 ** 30 // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * n * 8 / sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--trials", type=int, default=7, help="single synchronised draws after the timed run")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=J_SAMPLES)
    ap.add_argument("--cpu-trials", type=int, default=7)
    ap.add_argument("--I", type=int, default=I_ROWS)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with a 1 GiB device-to-device copy (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    ap.add_argument("--c3-ranks", default="16,64,128")
    ap.add_argument("--c3-steps", type=int, default=5)
    ap.add_argument("--workload", choices=["krp", "als", "serve", "c3", "lk", "sweep"], default="krp",
                    help="krp: the KRP leverage-sampling headline (default); als: sampled CP-ALS sweeps; "
                         "serve: the draw-request stream at several arrival rates; c3: FP32 vs FP64 on "
                         "ill-conditioned factors; lk: the Larsen-Kolda sampler; sweep: the reference's "
                         "runtime-benchmark configurations (I, R, N) against its published CPU times")
    ap.add_argument("--als-size", choices=sorted(ALS_SIZES), default="c5")
    ap.add_argument("--als-precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--als-oracle", action="store_true", help="rerun the sweeps on the CPU oracle and compare fits")
    ap.add_argument("--als-lookup", choices=["sort", "hash"], default="sort",
                    help="probe the sorted keys by binary search, or the per-mode coordinate hash")
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--exact-sweeps", type=int, default=0,
                    help="ALS workload: exact (unsampled) update sweeps after the sampled ones")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=500.0, help="arrival rate of the default run's serving leg")
    ap.add_argument("--serve-rates", default="50,100,200,500,1000,2000")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-budget-gb", type=float, default=96.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-sets", type=int, default=0,
                    help="serve workload: draw on this many registered factor sets with resident trees "
                         "instead of per-request generated factors")
    ap.add_argument("--serve-resident-rates", default="500,1000,2000,4000,8000")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the sampler runs on a single MI355X; sharding the draws and "
              "tree builds over several GPUs is not implemented, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "als":
        als_main(args)
        return
    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "c3":
        c3_main(args)
        return
    if args.workload == "lk":
        lk_main(args)
        return
    if args.workload == "sweep":
        sweep_main(args)
        return

    import torch  # device buffers only; imported before libftl_hip.so so both share one HIP runtime
    sys.path.insert(0, HERE)
    import fast_tensor_leverage_amd as ftl

    N, I, R, J = N_MODES, args.I, RANK, J_SAMPLES
    j = N + 1
    log(f"generating factors N={N} I={I} R={R}")
    U = [np.random.default_rng(1000 + k).standard_normal((I, R)) for k in range(N)]
    draws = np.random.default_rng(7).random(2 * N * J)

    dev = torch.device("cuda", 0)
    d_draws = torch.from_numpy(draws).to(dev)
    results = {}
    samples_out = {}
    for prec in ("fp64", "fp32"):
        log(f"building {prec} sampler")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = ftl.Sampler(U, J, R, "efficient", precision=prec)
        t_build = time.perf_counter() - t0
        info = s.info()
        d_samples = torch.zeros((N, J), dtype=torch.int64, device=dev)
        d_h = torch.empty((J, R), dtype=torch.float64, device=dev)
        d_w = torch.empty((J,), dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            s.draw_device(j, d_samples, d_h, d_w, d_draws)
        torch.cuda.synchronize()
        torch.cuda.synchronize()
        s.enable_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.draw_device(j, d_samples, d_h, d_w, d_draws)
        timing = s.last_timing()  # waits for the timed draws (events on the sampler's stream)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        s.enable_timing(False)
        # single calls, each synchronised on both sides: median and best
        trial_s = []
        for _ in range(args.trials):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            s.draw_device(j, d_samples, d_h, d_w, d_draws)
            torch.cuda.synchronize()
            trial_s.append(time.perf_counter() - t1)
        results[prec] = {"elapsed": elapsed, "timing": timing, "build_s": t_build, "info": info,
                         "group_lanes": s.get_option("group_lanes") or 32, "trial_s": trial_s}
        samples_out[prec] = d_samples.cpu().numpy().view(np.uint64).copy()
        log(f"{prec}: {args.steps * J / elapsed:,.0f} samples/s; kernels {timing}; build {t_build:.2f} s")
        del s, d_samples, d_h, d_w
        torch.cuda.synchronize()

    f64, f32 = results["fp64"], results["fp32"]
    value = args.steps * J / f32["elapsed"]
    v64 = args.steps * J / f64["elapsed"]
    mism = float(np.mean(np.any(samples_out["fp32"] != samples_out["fp64"], axis=0)))
    fp64_levels = f32["info"]["fp64_levels"]
    # algorithmic bytes with every level at the stored width the algorithm
    # needs (FP32 nodes and rows; FP64 for the plain sampler)
    bytes32 = algorithmic_bytes(I, R, J, N, 4, 4)
    bytes64 = algorithmic_bytes(I, R, J, N, 8, 8)
    t_kernel = f32["timing"]["draw_ms"] * 1e-3
    achieved = bytes32 / t_kernel
    if achieved > HBM_PEAK:
        raise SystemExit(f"roofline check failed: {bytes32 / 1e9:.3f} GB in {t_kernel * 1e3:.3f} ms exceeds the "
                         f"HBM peak; the byte count or the kernel time is wrong")
    # HBM bytes of the FP32 draw (its per-mode launches together), from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated)
    traffic = None
    traffic64 = None
    tfile = os.path.join(HERE, "profiles", "k_krp_draw_N3_I2p23_R32_J50000_traffic.json")
    if os.path.exists(tfile) and I == I_ROWS:
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj["fp32"]["hbm_bytes_per_draw"]
        traffic64 = tj["fp64"]["hbm_bytes_per_draw"]
    t64 = f64["timing"]["draw_ms"] * 1e-3
    copy_bw = device_copy_bandwidth(dev)
    log(f"device-to-device copy: {copy_bw / 1e9:,.0f} GB/s (read + write)")

    serving = None
    if not args.skip_serve:
        log(f"serving stream at {args.serve_rate}/s")
        serving = serve_run(args.serve_rate, args.serve_requests, args.serve_budget_gb, args.serve_slots)
        log(f"serving: p50 {serving['p50_ms']:.2f} ms p99 {serving['p99_ms']:.2f} ms")

    cpu = None
    extra_cmp = {}
    if not args.skip_cpu:
        log("CPU baseline (oracle)")
        cpu, extra_cmp = cpu_baseline(U, draws, J, j, samples_out["fp64"], min(args.cpu_sample, J),
                                      args.cpu_trials)

    line = {
        "metric": "KRP samples/s (N=3,I=2^23,R=32) FP32-MFMA vs FP64 GPU; p99 req latency",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f32["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 top tree levels, computeM and h in f64)",
        "data": "synthetic: N(0,1) factors rng(1000+k), uniforms rng(7), resident in HBM",
        "config": {"workload": "KRPDrawSamples N=3 I=2^23 R=32 J=50000 j=N+1 (all modes)",
                   "N": N, "I": I, "R": R, "J": J, "j": j},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic / N if traffic else None,
                     "traffic_source": "profiles/k_krp_draw_N3_I2p23_R32_J50000_traffic.json (rocprofv3 --pmc "
                                       "FETCH_SIZE and WRITE_SIZE per draw launch, copy-calibrated), not this run",
                     "traffic_GBps": (traffic / t_kernel / 1e9) if traffic else None,
                     "kernel": f"k_krp_draw<32,{f32['group_lanes']},256,fp32>",
                     "kernel_ms_per_launch": f32["timing"]["draw_ms"] / N,
                     "algorithmic_bytes_per_launch": bytes32 / N,
                     "algorithmic_bytes_per_sample_mode": bytes32 / (N * J),
                     "launches_per_draw": N,
                     "stored_fp64_top_levels": fp64_levels,
                     "copy_GBps": copy_bw / 1e9,
                     "frac_of_copy": achieved / copy_bw},
        "trials": {p: {"calls": len(results[p]["trial_s"]),
                       "median_samples_per_s": J / float(np.median(results[p]["trial_s"])),
                       "best_samples_per_s": J / float(np.min(results[p]["trial_s"])),
                       "trial_ms": [1e3 * t for t in results[p]["trial_s"]]} for p in ("fp32", "fp64")}
        if args.trials > 0 else None,
        "cpu_baseline": cpu,
        "fp64_gpu_samples_per_s": v64,
        "fp32_over_fp64": value / v64,
        "fp32_vs_fp64_index_mismatch_rate": mism,
        # the plain sampler's algorithmic bytes exceed what HBM could deliver
        # in its kernel time (its top levels are served from L2 / MALL), so
        # its fraction is taken over the counter-measured HBM traffic
        "fp64_algorithmic_GBps": bytes64 / t64 / 1e9,
        "fp64_traffic_frac": (traffic64 / t64 / HBM_PEAK) if traffic64 else None,
        "kernel_ms": {"fp32": f32["timing"], "fp64": f64["timing"]},
        "construction_s": {"fp32": f32["build_s"], "fp64": f64["build_s"]},
        "p99_latency_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
    }
    line.update(extra_cmp)
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        src = torch.ones(2 ** 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB device-to-device copy done")


if __name__ == "__main__":
    main()
