// vx3mi_batch_fp64.hip -- FP64 instantiation of the device timestep.  Built
// with -ffp-contract=off so every operation rounds as in the reference's
// double arithmetic.
#include "vx3mi_batch_impl.h"

namespace vx3mi {
BatchBase *make_batch_fp64(vx3mi_model *const *models, int n, int *rc) { return make_batch<double>(models, n, rc); }
}  // namespace vx3mi
