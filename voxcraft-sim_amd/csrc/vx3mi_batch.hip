// vx3mi_batch.hip -- C ABI of device batches (include/vx3mi.h).
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "vx3mi_batch.h"

using namespace vx3mi;

// Allocation-free contact record (vx3mi_batch_record_contacts_static): one
// device queue per GPU in a file-scope array, here `scratch` -- a counter in its
// first 256 bytes, then at most kPairScratch pairs (8 MiB in all).  In host
// code `scratch` names a host-side stand-in; its device address is taken once
// per device with hipGetSymbolAddress and only that address is handed to
// kernels.  Every batch on a GPU that records this way shares that GPU's queue.
// (A static __device__ variable that host code names through HIP_SYMBOL is
// given an external device symbol by the compiler, which the runtime looks up;
// marking it `used` would keep that symbol local, and the lookup would fail.)
constexpr int kPairScratch = 1 << 20;
static __device__ __align__(256) int2 scratch[32 + kPairScratch];
namespace {
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices];
std::mutex g_scratchLock;
}  // namespace


extern "C" {

int vx3mi_batch_configure(vx3mi_batch *b, int mode, int tile, int threads) {
    if (!b || !b->b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    return b->b->configure(mode, tile, threads);
}

int vx3mi_batch_autotune(vx3mi_batch *b, int trial_steps, double *ms6, int *choice3) {
    if (!b || !b->b || !ms6 || !choice3) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return b->b->autotune(trial_steps, ms6, choice3);
}

int vx3mi_batch_create(vx3mi_model *const *models, int n, int precision, vx3mi_batch **out) {
    if (!out || !models || n <= 0) { set_error("vx3mi_batch_create: invalid arguments"); return VX3MI_E_ARG; }
    *out = nullptr;
    if (precision != VX3MI_FP32 && precision != VX3MI_FP64) { set_error("precision must be 32 or 64"); return VX3MI_E_ARG; }
    for (int i = 0; i < n; i++)
        if (!models[i] || models[i]->m.nv() == 0) {
            // CUDA_Simulation aborts empty models (VX3_SimulationManager.cu:15-19)
            set_error("model " + std::to_string(i) + " has no voxels");
            return VX3MI_E_ARG;
        }
    int devs = 0;
    if (hipGetDeviceCount(&devs) != hipSuccess || devs <= 0) {
        set_error("no HIP device available");
        return VX3MI_E_HIP;
    }
    vx3mi_batch *h = new (std::nothrow) vx3mi_batch();
    if (!h) { set_error("out of host memory"); return VX3MI_E_NOMEM; }
    int rc = VX3MI_OK;
    h->b = precision == VX3MI_FP32 ? make_batch_fp32(models, n, &rc) : make_batch_fp64(models, n, &rc);
    if (!h->b) {
        delete h;
        return rc;
    }
    *out = h;
    return VX3MI_OK;
}

void vx3mi_batch_destroy(vx3mi_batch *b) {
    if (!b) return;
    if (b->b && b->b->stream) (void)hipStreamSynchronize(b->b->stream);
    delete b->b;
    delete b;
}

int vx3mi_batch_counts(const vx3mi_batch *b, int *nrobot, int *nvox, int *nlink) {
    if (!b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    if (nrobot) *nrobot = b->b->R;
    if (nvox) *nvox = b->b->V;
    if (nlink) *nlink = b->b->L;
    return VX3MI_OK;
}

int vx3mi_batch_recommended_dt(vx3mi_batch *b, double *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return b->b->recommended(out);
}

int vx3mi_step(vx3mi_batch *b, int nsteps, float dt, int mode) {
    if (!b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    if (mode < 0 || mode > 4 || nsteps < 0) { set_error("invalid step mode or step count"); return VX3MI_E_ARG; }
    return b->b->step(nsteps, dt, mode);
}

int vx3mi_batch_sync(vx3mi_batch *b) {
    if (!b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    hipError_t e = hipStreamSynchronize(b->b->stream);
    if (e != hipSuccess) { set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e)); return VX3MI_E_HIP; }
    return VX3MI_OK;
}

static int check_robot(vx3mi_batch *b, int robot) {
    if (!b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    if (robot >= b->b->R) { set_error("robot index out of range"); return VX3MI_E_ARG; }
    return VX3MI_OK;
}

int vx3mi_get_state(vx3mi_batch *b, int robot, double *pos3, double *quat4, double *linmom3, double *angmom3) {
    int rc = check_robot(b, robot);
    return rc ? rc : b->b->get_state(robot, pos3, quat4, linmom3, angmom3);
}

int vx3mi_set_state(vx3mi_batch *b, int robot, const double *pos3, const double *quat4, const double *linmom3,
                    const double *angmom3) {
    int rc = check_robot(b, robot);
    return rc ? rc : b->b->set_state(robot, pos3, quat4, linmom3, angmom3);
}

int vx3mi_get_voxel_flags(vx3mi_batch *b, int robot, int *flags) {
    int rc = check_robot(b, robot);
    if (rc) return rc;
    if (!flags) { set_error("flags is NULL"); return VX3MI_E_ARG; }
    return b->b->get_flags(robot, flags);
}

int vx3mi_get_link_flags(vx3mi_batch *b, int robot, int *flags, float *strain) {
    int rc = check_robot(b, robot);
    if (rc) return rc;
    return b->b->get_link_flags(robot, flags, strain);
}

int vx3mi_batch_removed(vx3mi_batch *b, int robot, int *voxel_step, int *link_step) {
    int rc = check_robot(b, robot);
    if (rc) return rc;
    return b->b->get_removed(robot, voxel_step, link_step);
}

int vx3mi_batch_signals(vx3mi_batch *b, int robot, double *local) {
    int rc = check_robot(b, robot);
    if (rc) return rc;
    if (!local) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return b->b->get_signals(robot, local);
}

static int read_rt(vx3mi_batch *b, std::vector<RobotTime> &rt) {
    rt.resize(b->b->R);
    hipError_t e = hipStreamSynchronize(b->b->stream);
    if (e == hipSuccess) e = hipMemcpy(rt.data(), b->b->d_rt, sizeof(RobotTime) * b->b->R, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error(std::string("reading robot time: ") + hipGetErrorString(e)); return VX3MI_E_HIP; }
    return VX3MI_OK;
}

int vx3mi_batch_time(vx3mi_batch *b, double *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    std::vector<RobotTime> rt;
    int rc = read_rt(b, rt);
    if (rc) return rc;
    for (int r = 0; r < b->b->R; r++) out[r] = rt[r].time[rt[r].steps & 1];
    return VX3MI_OK;
}

int vx3mi_batch_diverged(vx3mi_batch *b, int *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    std::vector<RobotTime> rt;
    int rc = read_rt(b, rt);
    if (rc) return rc;
    int any = 0;
    for (int r = 0; r < b->b->R; r++) { out[r] = rt[r].diverged; any |= rt[r].diverged; }
    return any ? VX3MI_DIVERGED : VX3MI_OK;
}

int vx3mi_batch_hits(vx3mi_batch *b, int *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    std::vector<RobotTime> rt;
    int rc = read_rt(b, rt);
    if (rc) return rc;
    for (int r = 0; r < b->b->R; r++) out[r] = rt[r].collisionCount;
    return VX3MI_OK;
}

int vx3mi_batch_listings(vx3mi_batch *b, int *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    std::vector<RobotTime> rt;
    int rc = read_rt(b, rt);
    if (rc) return rc;
    for (int r = 0; r < b->b->R; r++) out[r] = rt[r].listings;
    return VX3MI_OK;
}

int vx3mi_batch_record_contacts(vx3mi_batch *b, int cap) {
    if (!b || !b->b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    return b->b->record_contacts(cap);
}

int vx3mi_batch_record_contacts_static(vx3mi_batch *b, int cap) {
    if (!b || !b->b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    if (cap < 0 || cap > kPairScratch) {
        set_error("contact record capacity must be 0.." + std::to_string(kPairScratch) + " pairs");
        return cap < 0 ? VX3MI_E_ARG : VX3MI_E_LIMIT;
    }
    if (cap == 0) return b->b->record_contacts_at(nullptr, nullptr, 0);
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error(std::string("hipGetDevice: ") + (e != hipSuccess ? hipGetErrorString(e) : "device index out of range"));
        return VX3MI_E_HIP;
    }
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchLock);
        if (!g_scratch[dev]) {  // this device's copy of the array, with the device current
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
                return VX3MI_E_HIP;
            }
            g_scratch[dev] = p;
        }
        base = g_scratch[dev];
    }
    return b->b->record_contacts_at(reinterpret_cast<int2 *>(static_cast<char *>(base) + 256), static_cast<int *>(base),
                                    cap);
}

int vx3mi_batch_contact_pairs(vx3mi_batch *b, int robot, int *pairs2, int cap) {
    int rc = check_robot(b, robot);
    if (rc) return rc;
    if (!pairs2 && cap > 0) { set_error("pairs is NULL"); return VX3MI_E_ARG; }
    return b->b->contact_pairs(robot, pairs2, cap);
}

int vx3mi_batch_attach_state(vx3mi_batch *b, int robot, int *voxel_flags, int *link_neg, int *link_pos,
                             int *link_axis, int *new_left, int cap, int *overflow) {
    if (!b) { set_error("batch is NULL"); return VX3MI_E_ARG; }
    if (robot < 0 || robot >= b->b->R) { set_error("robot index out of range"); return VX3MI_E_ARG; }
    if (cap < 0) { set_error("negative capacity"); return VX3MI_E_ARG; }
    return b->b->attach_state(robot, voxel_flags, link_neg, link_pos, link_axis, new_left, cap, overflow);
}

int vx3mi_center_of_mass(vx3mi_batch *b, double *out) {
    if (!b || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return b->b->com(out);
}

int vx3mi_eval_link_forces(vx3mi_batch *b, double *out12) {
    if (!b || !out12) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return b->b->eval_links(out12);
}

int vx3mi_last_step_ms(vx3mi_batch *b, float *ms) {
    if (!b || !ms) { set_error("NULL argument"); return VX3MI_E_ARG; }
    hipError_t e = hipEventSynchronize(b->b->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->b->ev0, b->b->ev1);
    if (e != hipSuccess) { set_error(std::string("event timing: ") + hipGetErrorString(e)); return VX3MI_E_HIP; }
    return VX3MI_OK;
}

int vx3mi_last_step_launches(vx3mi_batch *b, long long *launches) {
    if (!b || !launches) { set_error("NULL argument"); return VX3MI_E_ARG; }
    *launches = b->b->launches;
    return VX3MI_OK;
}

}  // extern "C"
