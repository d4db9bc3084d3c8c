// grouped.hpp -- the grouped-launch plan behind the sgemm_grouped_* C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "sgemm.h"
#include "kernels.hpp"

namespace sgemm_amd {

int validate_problem(const sgemm_problem& q);
Problem to_problem(const sgemm_problem& q);

// Device descriptor buffer + pinned staging copy + optional captured graph.
// set() may be called again (up to `capacity` problems) without invalidating a
// captured graph: the kernel reads count and tile totals from device memory.
class GroupedPlan {
  public:
    static int create(GroupedPlan** out, int capacity, int form);
    void release();
    // Validates, picks configs (autotune table, or `forced_cfg` per problem),
    // orders the tiles and uploads the descriptors on `stream`.
    int set(int count, const sgemm_problem* problems, hipStream_t stream, const int* forced_cfg);
    int run(hipStream_t stream) const;
    int capture(hipStream_t stream);
    int replay(hipStream_t stream) const;

    int count() const { return count_; }
    int total_tiles() const { return total_tiles_; }
    int grid() const { return grid_; }
    double flops() const { return flops_; }
    int form() const { return form_; }
    int capacity() const { return capacity_; }

  private:
    int launch(hipStream_t stream) const;

    int form_ = 0;
    int capacity_ = 0;
    int count_ = 0;
    int total_tiles_ = 0;
    int grid_ = 0;
    double flops_ = 0.0;
    size_t bytes_ = 0;
    void* dev_ = nullptr;
    void* staging_ = nullptr;
    hipEvent_t uploaded_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace sgemm_amd
