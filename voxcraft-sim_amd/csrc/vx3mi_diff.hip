// vx3mi_diff.hip -- C ABI of the differentiable timestep (include/vx3mi.h)
#include <new>

#include "vx3mi_diff_batch.h"

using namespace vx3mi;

extern "C" {

void vx3mi_diff_default_opts(vx3mi_diff_opts *o) {
    if (!o) return;
    o->floor_stiffness = 1e5;
    o->floor_width = 1e-4;
    o->friction_speed = 1e-3;
    o->dt = 0;
    o->e_max = 1e7;
    o->checkpoint_every = 0;
    o->hbm_budget = 16.0 * 1024 * 1024 * 1024;
    o->reverse_threads = 0;
}

int vx3mi_diff_create(vx3mi_model *const *models, int n, int precision, const vx3mi_diff_opts *opts, vx3mi_diff **out) {
    if (!out) { set_error("out is NULL"); return VX3MI_E_ARG; }
    *out = nullptr;
    if (!models || n <= 0) { set_error("no models"); return VX3MI_E_ARG; }
    for (int i = 0; i < n; i++)
        if (!models[i] || models[i]->m.nv() == 0) { set_error("model is NULL or has no voxels"); return VX3MI_E_ARG; }
    if (precision != VX3MI_FP32 && precision != VX3MI_FP64) { set_error("precision must be 32 or 64"); return VX3MI_E_ARG; }
    vx3mi_diff_opts o;
    vx3mi_diff_default_opts(&o);
    if (opts) o = *opts;
    const int rt = o.reverse_threads;
    if (rt != 0 && rt != -1 && !(precision == VX3MI_FP32 ? (rt == 512 || rt == 1024) : rt == 256)) {
        set_error("reverse_threads must be 0, -1 (flat), or 512 / 1024 (FP32), 256 (FP64)");
        return VX3MI_E_ARG;
    }
    if (!(o.floor_stiffness > 0) || !(o.floor_width > 0) || !(o.friction_speed > 0) || !(o.e_max > 0)) {
        set_error("floor stiffness, floor width, friction speed and e_max must be positive");
        return VX3MI_E_ARG;
    }
    int rc = VX3MI_OK;
    DiffBase *b = precision == VX3MI_FP32 ? make_diff_fp32(models, n, o, &rc) : make_diff_fp64(models, n, o, &rc);
    if (!b) return rc;
    vx3mi_diff *h = new (std::nothrow) vx3mi_diff{b};
    if (!h) { delete b; set_error("out of host memory"); return VX3MI_E_NOMEM; }
    *out = h;
    return VX3MI_OK;
}

void vx3mi_diff_destroy(vx3mi_diff *d) {
    if (!d) return;
    delete d->b;
    delete d;
}

int vx3mi_diff_counts(const vx3mi_diff *d, int *nrobot, int *nvox, int *nlink) {
    if (!d) { set_error("diff is NULL"); return VX3MI_E_ARG; }
    if (nrobot) *nrobot = d->b->R;
    if (nvox) *nvox = d->b->V;
    if (nlink) *nlink = d->b->L;
    return VX3MI_OK;
}

int vx3mi_diff_set_params(vx3mi_diff *d, const double *phase, const double *E) {
    if (!d || !phase || !E) { set_error("NULL argument"); return VX3MI_E_ARG; }
    return d->b->set_params(phase, E);
}

int vx3mi_diff_forward(vx3mi_diff *d, int nsteps) {
    if (!d) { set_error("diff is NULL"); return VX3MI_E_ARG; }
    return d->b->forward(nsteps);
}

int vx3mi_diff_get_state(vx3mi_diff *d, int robot, double *pos, double *quat, double *linmom, double *angmom) {
    if (!d) { set_error("diff is NULL"); return VX3MI_E_ARG; }
    return d->b->get_state(robot, pos, quat, linmom, angmom);
}

int vx3mi_diff_backward(vx3mi_diff *d, const double *grad_pos, const double *grad_quat, const double *grad_linmom,
                        const double *grad_angmom, double *grad_phase, double *grad_E) {
    if (!d) { set_error("diff is NULL"); return VX3MI_E_ARG; }
    return d->b->backward(grad_pos, grad_quat, grad_linmom, grad_angmom, grad_phase, grad_E);
}

int vx3mi_diff_info(const vx3mi_diff *d, double *o) {
    // (the last two entries: recompute and reverse-sweep parts of the backward)
    if (!d || !o) { set_error("NULL argument"); return VX3MI_E_ARG; }
    const DiffBase &b = *d->b;
    o[0] = b.K;
    o[1] = b.ckBytes;
    o[2] = b.tapeBytes;
    o[3] = b.dt;
    o[4] = b.fwdMs;
    o[5] = b.bwdMs;
    o[6] = b.steps;
    o[7] = b.precision;
    o[8] = b.recomputeMs;
    o[9] = b.reverseMs;
    return VX3MI_OK;
}

}  // extern "C"
