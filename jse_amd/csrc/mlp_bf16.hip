// BF16-MFMA variant of the per-species MLP (written next).
#include "common.h"
#include "kernels.h"

namespace jnn {
void launch_mlp_bf16(hipStream_t, const MlpArgs &, const uint16_t *, const uint16_t *) {
    throw Error(ERR_UNSUPPORTED, "the BF16 MLP is not available in this build");
}
}  // namespace jnn
