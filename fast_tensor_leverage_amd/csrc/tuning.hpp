// tuning.hpp -- the sampler's tuning table: launch shapes and precision split
// per (precision, rank, tree depth, sample count) bucket, written by the
// package's autotuner (fast_tensor_leverage_amd/autotune.py) into
// ftl_tuning.tsv beside libftl_hip.so and read once per process.
//
// File format: one row per line, '#' starts a comment, fields separated by
// whitespace:
//   <precision fp64|fp32> <R_lo> <R_hi> <depth_lo> <depth_hi> <J_lo> <J_hi> key=value[,key=value...]
// The first row whose closed ranges hold (R, depth, J) applies; depth is the
// largest tree depth over the factors (levels below the root).  Keys are the
// sampler options of ftl_sampler_set_option ("group_lanes", "grid_blocks",
// "draw_kernel", "block_threads", "jacobi_threads", "fp32_levels",
// "mode_split", "leaf_mfma").  The environment variable FTL_TUNING_TABLE
// names another file; set to the empty string it disables the table.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace ftl {

struct TuningRow {
    int precision = 0;
    uint64_t r_lo = 0, r_hi = 0, d_lo = 0, d_hi = 0, j_lo = 0, j_hi = 0;
    std::vector<std::pair<std::string, int64_t>> opts;
};

struct TuningTable {
    std::string path;  // file read ("" when none)
    std::vector<TuningRow> rows;
    // index of the first row that applies, or -1
    int find(int precision, uint64_t R, uint64_t depth, uint64_t J) const {
        for (size_t i = 0; i < rows.size(); i++) {
            const TuningRow& t = rows[i];
            if (t.precision == precision && R >= t.r_lo && R <= t.r_hi && depth >= t.d_lo && depth <= t.d_hi &&
                J >= t.j_lo && J <= t.j_hi)
                return (int)i;
        }
        return -1;
    }
};

// The process-wide table (parsed on first use; malformed rows are skipped).
const TuningTable& tuning_table();

// Parses table text (exposed for the host-side tests through the C ABI).
TuningTable parse_tuning(const std::string& text);

}  // namespace ftl
