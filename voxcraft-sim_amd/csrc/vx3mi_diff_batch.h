// vx3mi_diff_batch.h -- host-side interface of the differentiable timestep,
// shared by the FP32 / FP64 device translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "vx3mi_batch.h"
#include "vx3mi_diff.h"

namespace vx3mi {

struct DiffBase {
    int precision = 32;
    int R = 0, V = 0, L = 0;
    int K = 1;               // checkpoint interval of the last forward
    int steps = 0;           // steps of the last forward
    int chunk = 1000;        // steps per persistent launch
    bool tapeValid = false;  // the tape holds the whole last forward (no recompute needed)
    float dt = 0;
    float fwdMs = 0, bwdMs = 0;
    float recomputeMs = 0, reverseMs = 0;  // stream time of the last backward's recompute and reverse kernels
    std::vector<hipEvent_t> segEv;         // per segment: recompute start/end, reverse start/end
    double ckBytes = 0, tapeBytes = 0;
    DiffOpts o{};
    vx3mi_diff_opts optsIn{};
    hipStream_t stream = nullptr;   // forward, reverse sweep
    hipStream_t rstream = nullptr;  // recompute of the next segment, beside the reverse sweep
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void *> allocs;
    std::vector<RobotParams> hrp;
    std::vector<int> hlat;
    std::vector<double> hphase0, hE0, hphase, hE;
    virtual ~DiffBase() {
        for (hipEvent_t e : segEv) (void)hipEventDestroy(e);
        for (void *p : allocs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        if (rstream) (void)hipStreamDestroy(rstream);
    }
    template <typename P> int alloc(P **p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void **)p, n * sizeof(P));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return VX3MI_E_NOMEM;
        }
        allocs.push_back((void *)*p);
        return VX3MI_OK;
    }
    template <typename P> void free_one(P *&p) {
        auto it = std::find(allocs.begin(), allocs.end(), (void *)p);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree((void *)p);
        p = nullptr;
    }
    virtual int set_params(const double *phase, const double *E) = 0;
    virtual int forward(int nsteps) = 0;
    virtual int get_state(int robot, double *pos, double *quat, double *lm, double *am) = 0;
    virtual int backward(const double *gpos, const double *gquat, const double *glm, const double *gam,
                         double *gphase, double *gE) = 0;
    virtual size_t bytes_ck_slot() const = 0;
    virtual size_t bytes_tape_slot() const = 0;
};

DiffBase *make_diff_fp32(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts, int *rc);
DiffBase *make_diff_fp64(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts, int *rc);

}  // namespace vx3mi

struct vx3mi_diff {
    vx3mi::DiffBase *b;
};
