// txg_fp64.hip -- FP64 (parity mode) instantiation of the step kernels.
// Compiled with -ffp-contract=off: no multiply-add fusion, so each
// expression rounds exactly as the reference's FP64 code does.
#include "txg_instantiate.inc"

namespace txg {
bool kernels_fp64(int dims, int ncomp, int iso, KernelSet* ks) { return pick<double, double>(dims, ncomp, iso, ks); }
}  // namespace txg
