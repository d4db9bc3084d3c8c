// autotune.hpp -- config selection per (form, m, n, k) size bucket.
#pragma once

namespace sgemm_amd {

// Size bucket of one dimension: ceil(log2(x)) capped to [0, 15]
// (1 -> 0, 2 -> 1, 3..4 -> 2, ..., 8193..16384 -> 14, larger -> 15).
int size_bucket(int x);

// Config id for a single-problem launch of this shape: the table entry of the
// shape's bucket if there is one, else the built-in heuristic.  `aligned`:
// A, B, lda and ldb give 16-byte aligned operand columns, so the FP32
// configs 5, 9 and 10 run their direct-to-LDS tile (four workgroups per CU
// instead of three), which moves the FP32 heuristic's thresholds.
int choose_config(int form, int m, int n, int k, bool aligned = true);

// Same for a problem inside a grouped launch (restricted to the configs the
// grouped kernel carries).
int choose_grouped_config(int form, int m, int n, int k);

}  // namespace sgemm_amd
