// single-precision instantiation of the differentiable timestep
#include "vx3mi_diff_impl.h"

namespace vx3mi {
DiffBase *make_diff_fp32(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts, int *rc) {
    return make_diff<float>(models, n, opts, rc);
}
}  // namespace vx3mi
