// txg_fp32.hip -- FP32-compute instantiations of the step kernels with FP32
// population storage (the shifted-FP16 storage ones are in txg_fp16.hip).
#include "txg_instantiate.inc"

namespace txg {
bool kernels_fp32_half(int dims, int ncomp, int iso, KernelSet* ks);

bool kernels_fp32(int dims, int ncomp, int iso, bool half_storage, KernelSet* ks) {
  if (half_storage) return kernels_fp32_half(dims, ncomp, iso, ks);
  return pick<float, float>(dims, ncomp, iso, ks);
}
}  // namespace txg
