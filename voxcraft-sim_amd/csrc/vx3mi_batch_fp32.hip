// vx3mi_batch_fp32.hip -- FP32 instantiation of the device timestep.
#include "vx3mi_batch_impl.h"

namespace vx3mi {
BatchBase *make_batch_fp32(vx3mi_model *const *models, int n, int *rc) { return make_batch<float>(models, n, rc); }
}  // namespace vx3mi
