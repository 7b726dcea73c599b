// FP64 instantiations of the fused block-tridiagonal LU kernels (one
// translation unit per type so the two compile in parallel).
#include "btri_lu_factor.h"

namespace slab {
SLAB_BTRI_LU_FACTOR_INST(double)
}  // namespace slab
