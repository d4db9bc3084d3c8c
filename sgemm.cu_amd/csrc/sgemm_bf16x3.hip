// sgemm_bf16x3.hip -- split-precision BF16x3 SGEMM on the BF16 matrix cores.
#include "kernels.hpp"

namespace sgemm_amd {

static const ConfigInfo kBf16x3Configs[] = {
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2"},
};

int bf16x3_config_count() { return 0; }
const ConfigInfo& bf16x3_config(int id) { return kBf16x3Configs[id]; }
hipError_t launch_bf16x3(int, const Problem&, hipStream_t) { return hipErrorNotSupported; }

}  // namespace sgemm_amd
