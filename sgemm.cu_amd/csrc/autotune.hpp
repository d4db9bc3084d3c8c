// autotune.hpp -- config selection per (form, m, n, k).
#pragma once

namespace sgemm_amd {

// Config id for a single problem of this shape: the autotune table's entry
// for the shape's size bucket, else the built-in heuristic.
int choose_config(int form, int m, int n, int k);

}  // namespace sgemm_amd
