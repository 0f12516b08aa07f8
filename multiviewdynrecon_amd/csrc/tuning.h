// Launch shapes chosen per (node count, batch size): a process-wide table
// filled by the autotuner (mvdr_*_autotune) or a saved table, with rule
// defaults for shapes it has no entry for.  Every parameter here changes only
// how work is spread over workgroups, never a result: the fused CG's partial
// sums are per 32-row group, the term evaluation and the CG transform are
// grid-stride loops over independent items.
#pragma once
#include <cstdint>
#include <string>

namespace mvdr {

struct LaunchParams {
    uint32_t cg_workgroups;         // fused CG row kernels, per frame (0: rule)
    uint32_t eval_workgroups;       // term evaluation, per frame (0: rule)
    uint32_t transform_workgroups;  // CG transform, per frame (0: rule)
    uint32_t gram_rows;             // rows per Gram row wave: 1, 2 or 4 (0: rule)
};

// Shapes are keyed by the batch's node count in 1024-node buckets and its
// frame count.
uint32_t node_bucket(uint32_t nodes);

// The table's entry for the shape, zeros where it has none.
LaunchParams launch_params(uint32_t nodes, uint32_t frames);
void set_launch_params(uint32_t nodes, uint32_t frames, const LaunchParams& p);
void clear_launch_params();
// Bumped by every change: captured graphs recorded with older shapes are
// re-captured.
uint64_t launch_table_version();

// Text form: one line per entry, "bucket frames cg eval transform gram"
// (lines without the last field read it as 0).
std::string launch_table_text();
bool load_launch_table_text(const std::string& text, std::string& err);

}  // namespace mvdr
