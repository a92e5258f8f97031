"""The sampler's tuning table (host side, no GPU): the parser and the first-
match lookup behind ftl_tuning_lookup, and the committed table written by
the autotuner (fast_tensor_leverage_amd/autotune.py)."""
import os

import pytest

from fast_tensor_leverage_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "fast_tensor_leverage_amd", "ftl_tuning.tsv")

TEXT = """# comment line
fp32 17 32 15 40 8192 1099511627776 fp32_levels=5,group_lanes=32,jacobi_threads=1024  # trailing comment
fp32 17 32 0 40 1 1099511627776 fp32_levels=6
fp64 1 128 0 40 1 1099511627776 draw_kernel=2,block_threads=256
fp32 1 8 0 9 1 10 bad_value=x
fp16 1 8 0 9 1 10 group_lanes=8
"""


def test_lookup_first_matching_row():
    path, row, opts = _lib.tuning_lookup(_lib.FTL_FP32, 32, 18, 50_000, TEXT)
    assert row == 0 and opts == {"fp32_levels": 5, "group_lanes": 32, "jacobi_threads": 1024}
    _, row, opts = _lib.tuning_lookup(_lib.FTL_FP32, 20, 18, 4096, TEXT)  # J below row 0's range
    assert row == 1 and opts == {"fp32_levels": 6}
    _, row, opts = _lib.tuning_lookup(_lib.FTL_FP64, 128, 3, 1, TEXT)
    assert row == 2 and opts["draw_kernel"] == 2


def test_lookup_skips_malformed_rows():
    # the row with a non-integer value and the unknown precision are dropped
    _, row, opts = _lib.tuning_lookup(_lib.FTL_FP32, 8, 5, 5, TEXT)
    assert row == -1 and opts == {}


def test_committed_table_covers_the_benchmarked_shapes():
    if not os.path.exists(TABLE):
        pytest.skip("no tuning table committed")
    text = open(TABLE).read()
    rows = [ln for ln in text.splitlines() if ln.split("#")[0].strip()]
    keys = {"group_lanes", "grid_blocks", "draw_kernel", "block_threads", "jacobi_threads", "fp32_levels",
            "mode_split", "leaf_mfma"}
    for ln in rows:
        f = ln.split("#")[0].split()
        assert f[0] in ("fp32", "fp64") and len(f) == 8, ln
        for kv in f[7].split(","):
            k, v = kv.split("=")
            assert k in keys and int(v) >= 0, ln
    # every shape of the headline, mixed-precision and serving configurations
    # finds a row; FP32 rows fix the level split
    shapes = [(32, 18, 50_000), (16, 19, 50_000), (64, 17, 50_000), (128, 16, 50_000), (8, 4, 1024),
              (64, 14, 65536), (16, 10, 2048)]
    for prec in (_lib.FTL_FP32, _lib.FTL_FP64):
        for R, depth, J in shapes:
            _, row, opts = _lib.tuning_lookup(prec, R, depth, J, text)
            assert row >= 0, (prec, R, depth, J)
            if prec == _lib.FTL_FP32:
                assert 1 <= opts["fp32_levels"] <= 7
