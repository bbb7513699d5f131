// The FP32-local random walk (precision 1 and 2): cell-local positions and
// sub-step arithmetic in FP32, times and the Newton exit solve in FP64.
//
// The kernels are those of rwpt_walk_kernels.h instantiated for float.  This
// translation unit is built with -ffp-contract=fast: the FP32 walk carries
// no bit-exactness contract (its error against the FP64 walk is measured),
// so products and sums contract into fused multiply-add.
#include "rwpt_walk_kernels.h"

namespace rwpt {

hipError_t walk_launch32(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                         const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                         int32_t *nv) {
  if (e->opt.precision == 1)
    return launch_walk_typed<float, 1, 0>(e, W, grid, block, work, cnt, budget, nk, nv);
  return launch_walk_typed<float, 2, 0>(e, W, grid, block, work, cnt, budget, nk, nv);
}

hipError_t walk_prepare32(rwpt_engine *e, const WalkParams &W) {
  return launch_prepare_typed<float>(e, W);
}

hipError_t walk_terminate32(rwpt_engine *e, const int32_t *work, int64_t cnt) {
  return launch_terminate_typed<float>(e, work, cnt);
}

}  // namespace rwpt
