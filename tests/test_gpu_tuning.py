"""The tuning table's pick against the built-in defaults on the benchmarked
shapes: the headline (N=3, I=2^23, R=32, J=50,000, both precisions), the
mixed-precision ranks (R = 16, 64, 128 at I=2^23) and serving-stream shapes
(N in {3,4,5}, I from 2^12 to 2^20, R in {8,16,64}, J from 2^11 to 2^13).  A
sampler created with the table applied is timed with the row's launch
options and with the built-in ones (same FP32 level split), alternating, and
the table's median wall time per draw may not exceed the defaults' by more
than measurement noise (5% + 20 us)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    # N, I, R, J, precision
    (3, 2 ** 23, 32, 50_000, "fp32"),
    (3, 2 ** 23, 32, 50_000, "fp64"),
    (3, 2 ** 23, 16, 50_000, "fp32"),
    (3, 2 ** 23, 64, 50_000, "fp32"),
    (3, 2 ** 23, 128, 50_000, "fp32"),
    (4, 2 ** 14, 8, 4096, "fp32"),
    (3, 2 ** 20, 64, 8192, "fp32"),
    (5, 2 ** 12, 16, 2048, "fp32"),
]
PERF_KEYS = ("group_lanes", "grid_blocks", "draw_kernel", "block_threads", "jacobi_threads")


def _median_ms(smp, torch, N, d_s, d_w, draws, steps=5):
    for _ in range(2):
        smp.draw_device(N + 1, d_s, None, d_w, draws)
    torch.cuda.synchronize()
    import time
    t = []
    for _ in range(steps):
        t0 = time.perf_counter()
        smp.draw_device(N + 1, d_s, None, d_w, draws)
        torch.cuda.synchronize()
        t.append(time.perf_counter() - t0)
    return float(np.median(t)) * 1e3


@pytest.mark.parametrize("N,I,R,J,precision", SHAPES, ids=[f"N{s[0]}_I{s[1]}_R{s[2]}_J{s[3]}_{s[4]}" for s in SHAPES])
def test_table_pick_not_slower_than_defaults(N, I, R, J, precision):
    import torch
    import fast_tensor_leverage_amd as ftl
    from fast_tensor_leverage_amd import _lib
    from fast_tensor_leverage_amd.autotune import tree_depth
    dev = torch.device("cuda", 0)
    path, row, opts = _lib.tuning_lookup(_lib.FTL_FP32 if precision == "fp32" else _lib.FTL_FP64, R,
                                         tree_depth(I, R), J)
    assert row >= 0, f"no tuning row for R={R} I={I} J={J} {precision} ({path})"
    gen = torch.Generator(device=dev)
    gen.manual_seed(R)
    U = [torch.randn((I, R), generator=gen, device=dev, dtype=torch.float64) for _ in range(N)]
    draws = torch.rand(2 * N * J, generator=gen, device=dev, dtype=torch.float64)
    smp = ftl.Sampler.from_device(U, J, R, precision=precision)
    assert smp.get_option("tuning_row") == row
    if precision == "fp32":
        assert smp.get_option("fp32_levels") == opts["fp32_levels"]
    d_s = torch.zeros((N, J), dtype=torch.int64, device=dev)
    d_w = torch.empty((J,), dtype=torch.float64, device=dev)
    picks, defaults = [], []
    for _ in range(3):
        for k in PERF_KEYS:
            smp.set_option(k, opts.get(k, 0))
        picks.append(_median_ms(smp, torch, N, d_s, d_w, draws))
        for k in PERF_KEYS:
            smp.set_option(k, 0)
        defaults.append(_median_ms(smp, torch, N, d_s, d_w, draws))
    tp, td = float(np.median(picks)), float(np.median(defaults))
    print(f"[gpu] tuning N={N} I={I} R={R} J={J} {precision}: table {opts} {tp:.3f} ms, defaults {td:.3f} ms",
          flush=True)
    assert tp <= 1.05 * td + 0.02
