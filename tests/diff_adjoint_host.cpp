// Host run of the differentiable timestep's forward, checkpoint/recompute and
// reverse sweeps for the CPU tests: the work items of vx3mi_diff_steps.h (the
// ones the GPU kernels execute) driven by plain loops in the kernels' phase
// order, on one robot built by the product's model builder (vx3mi_model.cpp,
// compiled in by the test).  Not part of the product.
#include <cstring>
#include <vector>

#include "vx3mi_diff_steps.h"
#include "vx3mi_internal.h"

using namespace vx3mi;
typedef double D;

extern "C" int harness_adjoint(const vx3mi_lattice *spec, const double *phase, const double *E, int nsteps, double dt,
                               int K, const double *gpos, double *gphase, double *gE, double *finalPos) {
    vx3mi_model *mh = nullptr;
    if (vx3mi_model_from_lattice(spec, &mh) != VX3MI_OK) return -1;
    const Model &m = mh->m;
    const int V = m.nv(), L = m.nl();
    std::vector<VoxelMat> vm((size_t)V);
    std::vector<int> iz((size_t)V), axis((size_t)std::max(L, 1));
    std::vector<int2v> ends((size_t)std::max(L, 1));
    std::vector<double> nom((size_t)std::max(L, 1));
    for (int v = 0; v < V; v++) { vm[v] = m.vmats[m.vmat[v]]; iz[v] = m.iz[v]; }
    for (int l = 0; l < L; l++) {
        ends[l] = int2v{m.lneg[l], m.lpos[l]};
        axis[l] = m.laxis[l];
        nom[l] = m.lmats[m.lmat[l]].nomSize;
    }
    std::vector<DVox<D>> vc((size_t)V);
    std::vector<DLink<D>> lc((size_t)std::max(L, 1));
    diff_constants<D>(vm, ends.data(), axis.data(), nom.data(), iz.data(), V, L, E, vc.data(), lc.data());
    std::vector<D> Ev(E, E + V);
    std::vector<double> times((size_t)nsteps + 1, 0.0);
    const float dtf = (float)dt;
    for (int n = 0; n < nsteps; n++) times[n + 1] = times[n] + (double)dtf;
    const size_t NV = (size_t)V, NL = (size_t)std::max(L, 1);
    std::vector<T4<D>> st(NV * 4), ls(NL * 3), lf(NL * 3), adj(NV * 6), al(NL * 3), cc(NL * 4);
    std::vector<D> gph(NV, 0.0);
    DD<D> d;
    d.vlinks = m.vlinks.data();
    d.lends = ends.data();
    d.vc = vc.data();
    d.lc = lc.data();
    d.phase = phase;
    d.E = Ev.data();
    d.times = times.data();
    d.o.floorK = 1e5; d.o.floorW = 1e-4; d.o.fricV = 1e-3;
    d.o.floorEnabled = m.floorEnabled;
    d.o.enableExpansion = m.enableExpansion;
    d.o.tempAmp = m.tempAmplitude;
    d.o.tempPeriod = m.tempPeriod;
    d.o.varyTemp = m.varyTemp;
    d.dt = (D)dtf;
    d.V = V;
    d.L = L;
    d.su = &st[0]; d.sq = &st[NV]; d.sp = &st[2 * NV]; d.sh = &st[3 * NV];
    d.sl0 = &ls[0]; d.sl1 = &ls[NL]; d.sl2 = &ls[2 * NL];
    d.lf0 = &lf[0]; d.lf1 = &lf[NL]; d.lf2 = &lf[2 * NL];
    d.au = &adj[0]; d.aq = &adj[NV]; d.ap = &adj[2 * NV]; d.ah = &adj[3 * NV]; d.lamF0 = &adj[4 * NV]; d.lamF1 = &adj[5 * NV];
    d.al0 = &al[0]; d.al1 = &al[NL]; d.al2 = &al[2 * NL];
    d.c0 = &cc[0]; d.c1 = &cc[NL]; d.c2 = &cc[2 * NL]; d.c3 = &cc[3 * NL];
    d.gPhase = gph.data();

    const int nseg = (nsteps + K - 1) / K;
    std::vector<T4<D>> ckV((size_t)nseg * NV * 4), ckL((size_t)nseg * NL * 3), tapeV((size_t)K * NV * 5),
        tapeL((size_t)K * NL * 3);
    auto run = [&](int n0, int ns, bool ck, bool tape, int tapeStart) {
        for (int s = 0; s < ns; s++) {
            const int n = n0 + s;
            const size_t ts = (size_t)(n - tapeStart);
            if (ck && n % K == 0) {
                for (int v = 0; v < V; v++) diff_save_voxel(d, v, &ckV[((size_t)(n / K) * NV + v) * 4]);
                for (int l = 0; l < L; l++) diff_save_link(d, l, &ckL[((size_t)(n / K) * NL + l) * 3]);
            }
            if (tape)
                for (int l = 0; l < L; l++) diff_save_link(d, l, &tapeL[(ts * NL + l) * 3]);
            const D prevDt = n > 0 ? d.dt : 0.0;
            for (int l = 0; l < L; l++) diff_link_fwd(d, l, prevDt);
            for (int v = 0; v < V; v++) diff_voxel_fwd(d, v, times[n + 1], tape ? &tapeV[(ts * NV + v) * 5] : nullptr);
        }
    };
    for (int v = 0; v < V; v++) diff_init_voxel(d, v);
    for (int l = 0; l < L; l++) diff_init_link(d, l);
    run(0, nsteps, true, false, 0);
    for (int v = 0; v < V; v++)
        for (int k = 0; k < 3; k++) {
            const int lat = k == 0 ? m.ix[v] : (k == 1 ? m.iy[v] : m.iz[v]);
            finalPos[3 * v + k] = lat * m.vmats[m.vmat[v]].nomSize + (&d.su[v].x)[k];
        }
    for (int v = 0; v < V; v++) {
        d.au[v] = T4<D>{gpos[3 * v], gpos[3 * v + 1], gpos[3 * v + 2], 0};
        d.aq[v] = d.ap[v] = d.ah[v] = T4<D>{0, 0, 0, 0};
    }
    for (int l = 0; l < L; l++) d.al0[l] = d.al1[l] = d.al2[l] = T4<D>{0, 0, 0, 0};
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int s0 = seg * K, len = std::min(K, nsteps - s0);
        for (int v = 0; v < V; v++) diff_restore_voxel(d, v, &ckV[((size_t)seg * NV + v) * 4]);
        for (int l = 0; l < L; l++) diff_restore_link(d, l, &ckL[((size_t)seg * NL + l) * 3]);
        run(s0, len, false, true, s0);
        for (int n = s0 + len - 1; n >= s0; n--) {
            const size_t ts = (size_t)(n - s0);
            const bool gather = n < s0 + len - 1 || seg != nseg - 1;
            for (int v = 0; v < V; v++) {
                if (gather) diff_gather_adjoint(d, v, times[n + 1]);
                diff_voxel_rev(d, v, &tapeV[(ts * NV + v) * 5]);
            }
            const D prevDt = n > 0 ? d.dt : 0.0;
            for (int l = 0; l < L; l++) diff_link_rev(d, l, prevDt, &tapeV[ts * NV * 5], &tapeL[(ts * NL + l) * 3]);
        }
    }
    for (int v = 0; v < V; v++) gE[v] = diff_finish_voxel(d, v);
    for (int v = 0; v < V; v++) gphase[v] = gph[v];
    vx3mi_model_free(mh);
    return 0;
}
