// txg_fp32.hip -- FP32-compute instantiations of the step kernels, with
// FP32 or shifted-FP16 population storage.
#include "txg_instantiate.inc"

namespace txg {
bool kernels_fp32(int dims, int ncomp, int iso, bool half_storage, KernelSet* ks) {
  if (half_storage) return pick<__half, float>(dims, ncomp, iso, ks);
  return pick<float, float>(dims, ncomp, iso, ks);
}
}  // namespace txg
