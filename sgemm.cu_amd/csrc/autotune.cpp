// autotune.cpp -- config selection per (form, m, n, k).
#include "autotune.hpp"

#include "kernels.hpp"

namespace sgemm_amd {

// Built-in heuristic: the largest tile that still gives every CU at least two
// workgroups; small problems fall back to smaller tiles.
int choose_config(int form, int m, int n, int k) {
    (void)k;
    if (form != kFormF32) return 0;
    const long long t128 = (long long)div_up(m, 128) * div_up(n, 128);
    if (t128 >= 512) return 0;
    const long long t64 = (long long)div_up(m, 128) * div_up(n, 64);
    if (t64 >= 512) return 2;
    return 1;
}

}  // namespace sgemm_amd
