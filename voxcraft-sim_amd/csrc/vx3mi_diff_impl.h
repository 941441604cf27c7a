// vx3mi_diff_impl.h -- device side and host driver of the differentiable
// timestep (instantiated for float and double by vx3mi_diff_fp32.hip /
// vx3mi_diff_fp64.hip).
//
// Forward: one workgroup per robot runs many steps (link phase, barrier, voxel
// phase, barrier), writing a checkpoint of the full state every K steps.
// Backward, per segment of K steps from the last: the segment is recomputed
// from its checkpoint into a tape (state entering each step and the link force
// sums of each voxel), then swept in reverse.  A reverse step is two
// barrier-separated phases:
//   voxel phase  gather the link adjoints of step n+1 (deterministic, <=6 per
//                voxel, direction order), then the voxel adjoint of step n;
//   link phase   recompute the link forward of step n from the tape and apply
//                its adjoint; each link writes its end contributions to its
//                own slots (no atomics).
// Layout (structure of arrays, CSR by robot):
//   state  su = (u, T)  sq = q  sp = (p, -)  sh = (h, -)
//   link   sl0 = (P, A1v.x)  sl1 = (A1v.yz, A2v.xy)  sl2 = (A2v.z, flags, -, -)
//          lf0 = (FN, MN.x)  lf1 = (MN.yz, FP.xy)  lf2 = (FP.z, MP)
//   tape   per voxel 5: (u, F.x) q (p, F.y) (h, F.z) (M, -); per link 3 as sl
//   ckpt   per voxel 4: su sq sp sh; per link 3 as sl
//   adjoint au = (ub, Tbar of the step) aq ap = (pb, Esq) ah; lamF0/1 = Fbar/Mbar;
//          al0..2 = (Pb, A1b, A2b, Elin in al2.w); contributions c0 = (-dub, tNb)
//          c1 = qnb  c2 = (dub, tPb)  c3 = qpb
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vx3mi_batch.h"
#include "vx3mi_diff.h"
#include "vx3mi_diff_batch.h"

namespace vx3mi {

template <typename T> struct DD {
    const RobotParams *rp;
    const int *vlinks;
    const int2 *lends;
    const DVox<T> *vc;
    const DLink<T> *lc;
    const double *phase;
    const T *E;
    const double *times;
    DiffOpts o;
    T dt;
    int V, L;
    T4<T> *su, *sq, *sp, *sh;
    T4<T> *sl0, *sl1, *sl2;
    T4<T> *lf0, *lf1, *lf2;
    // adjoint
    T4<T> *au, *aq, *ap, *ah, *lamF0, *lamF1;
    T4<T> *al0, *al1, *al2;
    T4<T> *c0, *c1, *c2, *c3;
    T *gPhase;
};

template <typename T> __device__ __forceinline__ V3<T> xyz(const T4<T> &a) { return V3<T>{a.x, a.y, a.z}; }
template <typename T> __device__ __forceinline__ Q4<T> wxyz(const T4<T> &a) { return Q4<T>{a.x, a.y, a.z, a.w}; }
template <typename T> __device__ __forceinline__ T4<T> t4(V3<T> v, T w) { return T4<T>{v.x, v.y, v.z, w}; }
template <typename T> __device__ __forceinline__ T4<T> t4(Q4<T> q) { return T4<T>{q.w, q.x, q.y, q.z}; }

// link state entering a step, from (sl0, sl1, sl2)
template <typename T>
__device__ __forceinline__ void unpack_ls(const T4<T> &a, const T4<T> &b, const T4<T> &c, V3<T> &P, V3<T> &A1, V3<T> &A2,
                                          int &flags) {
    P = V3<T>{a.x, a.y, a.z};
    A1 = V3<T>{a.w, b.x, b.y};
    A2 = V3<T>{b.z, b.w, c.x};
    flags = (int)c.y;
}

template <typename T> __device__ __forceinline__ T temp_at(const DD<T> &d, int v, double t, T *dTdPhase) {
    if (d.vc[v].fixed) { *dTdPhase = 0; return 0; }
    return d_temperature<T>(d.o, t, d.phase[v], dTdPhase);
}

// ---- forward -----------------------------------------------------------------------
template <typename T> __global__ __launch_bounds__(256) void k_diff_init(DD<T> d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.V) {
        T dd;
        d.su[i] = T4<T>{0, 0, 0, temp_at(d, i, d.times[0], &dd)};
        d.sq[i] = T4<T>{1, 0, 0, 0};
        d.sp[i] = T4<T>{0, 0, 0, 0};
        d.sh[i] = T4<T>{0, 0, 0, 0};
    }
    if (i < d.L) {
        d.sl0[i] = T4<T>{0, 0, 0, 0};
        d.sl1[i] = T4<T>{0, 0, 0, 0};
        d.sl2[i] = T4<T>{0, (T)LF_SMALL_ANGLE, 0, 0};
        d.lf0[i] = d.lf1[i] = d.lf2[i] = T4<T>{0, 0, 0, 0};
    }
}

template <typename T>
__device__ __forceinline__ void diff_link_fwd(const DD<T> &d, int l, T prevDt) {
    const DLink<T> &c = d.lc[l];
    if (c.fixedBoth) {
        d.lf0[l] = d.lf1[l] = d.lf2[l] = T4<T>{0, 0, 0, 0};
        return;
    }
    const int2 e = d.lends[l];
    const T4<T> un = d.su[e.x], up = d.su[e.y];
    V3<T> P0, A10, A20;
    int flags;
    unpack_ls(d.sl0[l], d.sl1[l], d.sl2[l], P0, A10, A20, flags);
    DLinkEval<T> ev;
    d_link_forward(c, V3<T>{up.x - un.x, up.y - un.y, up.z - un.z}, wxyz(d.sq[e.x]), wxyz(d.sq[e.y]), un.w, up.w, P0, A10,
                   A20, flags, prevDt, ev);
    d.sl0[l] = t4(ev.P, ev.A1v.x);
    d.sl1[l] = T4<T>{ev.A1v.y, ev.A1v.z, ev.A2v.x, ev.A2v.y};
    d.sl2[l] = T4<T>{ev.A2v.z, (T)ev.flagsOut, 0, 0};
    d.lf0[l] = t4(ev.FN, ev.MN.x);
    d.lf1[l] = T4<T>{ev.MN.y, ev.MN.z, ev.FP.x, ev.FP.y};
    d.lf2[l] = T4<T>{ev.FP.z, ev.MP.x, ev.MP.y, ev.MP.z};
}

template <typename T> __device__ __forceinline__ void gather_forces(const DD<T> &d, int v, V3<T> &F, V3<T> &M) {
    F = V3<T>{0, 0, 0};
    M = V3<T>{0, 0, 0};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int l = d.vlinks[6 * v + k];
        if (l < 0) continue;
        if (k & 1) {  // this voxel is the link's pVPos end
            const T4<T> f1 = d.lf1[l], f2 = d.lf2[l];
            F = F + V3<T>{f1.z, f1.w, f2.x};
            M = M + V3<T>{f2.y, f2.z, f2.w};
        } else {
            const T4<T> f0 = d.lf0[l], f1 = d.lf1[l];
            F = F + V3<T>{f0.x, f0.y, f0.z};
            M = M + V3<T>{f0.w, f1.x, f1.y};
        }
    }
}

// nsteps steps from step n0; checkpoints (every ckEvery) and/or a tape of the
// steps [tapeStart, tapeStart + nsteps)
template <typename T>
__global__ __launch_bounds__(256) void k_diff_forward(DD<T> d, int n0, int nsteps, int ckEvery, T4<T> *ckV, T4<T> *ckL,
                                                      T4<T> *tapeV, T4<T> *tapeL, int tapeStart) {
    const RobotParams rp = d.rp[blockIdx.x];
    const int vb = rp.voff, ve = rp.voff + rp.vcount, lb = rp.loff, le = rp.loff + rp.lcount;
    for (int s = 0; s < nsteps; s++) {
        const int n = n0 + s;
        if (ckV && n % ckEvery == 0) {
            const size_t slot = (size_t)(n / ckEvery);
            for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) {
                T4<T> *o = ckV + (slot * d.V + v) * 4;
                o[0] = d.su[v]; o[1] = d.sq[v]; o[2] = d.sp[v]; o[3] = d.sh[v];
            }
            for (int l = lb + threadIdx.x; l < le; l += blockDim.x) {
                T4<T> *o = ckL + (slot * d.L + l) * 3;
                o[0] = d.sl0[l]; o[1] = d.sl1[l]; o[2] = d.sl2[l];
            }
        }
        const size_t ts = (size_t)(n - tapeStart);
        if (tapeV) {
            for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) {
                T4<T> *o = tapeV + (ts * d.V + v) * 5;
                o[0] = d.su[v]; o[1] = d.sq[v]; o[2] = d.sp[v]; o[3] = d.sh[v];
            }
            for (int l = lb + threadIdx.x; l < le; l += blockDim.x) {
                T4<T> *o = tapeL + (ts * d.L + l) * 3;
                o[0] = d.sl0[l]; o[1] = d.sl1[l]; o[2] = d.sl2[l];
            }
        }
        const T prevDt = n > 0 ? d.dt : (T)0;
        for (int l = lb + threadIdx.x; l < le; l += blockDim.x) diff_link_fwd(d, l, prevDt);
        __syncthreads();
        const double tNext = d.times[n + 1];
        for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) {
            V3<T> F, M;
            gather_forces(d, v, F, M);
            const T4<T> u = d.su[v];
            if (tapeV) {  // the tape's state entries are this thread's own writes above
                T4<T> *o = tapeV + (ts * d.V + v) * 5;
                o[0].w = F.x;
                o[2].w = F.y;
                o[3].w = F.z;
                o[4] = t4(M, (T)0);
            }
            DVoxState<T> s, out;
            s.u = xyz(u);
            s.q = wxyz(d.sq[v]);
            s.p = xyz(d.sp[v]);
            s.h = xyz(d.sh[v]);
            d_voxel_forward(d.vc[v], d.o, d.dt, u.w, F, M, s, out);
            T dd;
            d.su[v] = t4(out.u, temp_at(d, v, tNext, &dd));
            d.sq[v] = t4(out.q);
            d.sp[v] = t4(out.p, (T)0);
            d.sh[v] = t4(out.h, (T)0);
        }
        __syncthreads();
    }
}

// restore the state of checkpoint slot
template <typename T> __global__ __launch_bounds__(256) void k_diff_restore(DD<T> d, const T4<T> *ckV, const T4<T> *ckL, int slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.V) {
        const T4<T> *o = ckV + ((size_t)slot * d.V + i) * 4;
        d.su[i] = o[0]; d.sq[i] = o[1]; d.sp[i] = o[2]; d.sh[i] = o[3];
    }
    if (i < d.L) {
        const T4<T> *o = ckL + ((size_t)slot * d.L + i) * 3;
        d.sl0[i] = o[0]; d.sl1[i] = o[1]; d.sl2[i] = o[2];
    }
}

// ---- reverse -----------------------------------------------------------------------
// add the link contributions of step n+1 to the state adjoint of voxel v and turn
// the step's temperature adjoint into the phase gradient
template <typename T> __device__ __forceinline__ void gather_adjoint(const DD<T> &d, int v, double tStep) {
    T4<T> au = d.au[v], aq = d.aq[v];
    T tb = au.w;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int l = d.vlinks[6 * v + k];
        if (l < 0) continue;
        const T4<T> a = (k & 1) ? d.c2[l] : d.c0[l], b = (k & 1) ? d.c3[l] : d.c1[l];
        au.x += a.x; au.y += a.y; au.z += a.z;
        tb += a.w;
        aq.x += b.x; aq.y += b.y; aq.z += b.z; aq.w += b.w;
    }
    T dTdPhase;
    (void)temp_at(d, v, tStep, &dTdPhase);
    d.gPhase[v] += tb * dTdPhase;
    au.w = 0;
    d.au[v] = au;
    d.aq[v] = aq;
}

template <typename T>
__global__ __launch_bounds__(256) void k_diff_reverse(DD<T> d, int segStart, int segLen, int gatherFirst,
                                                      const T4<T> *tapeV, const T4<T> *tapeL) {
    const RobotParams rp = d.rp[blockIdx.x];
    const int vb = rp.voff, ve = rp.voff + rp.vcount, lb = rp.loff, le = rp.loff + rp.lcount;
    for (int n = segStart + segLen - 1; n >= segStart; n--) {
        const size_t ts = (size_t)(n - segStart);
        const double t = d.times[n];
        // ---- voxel phase
        for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) {
            if (n < segStart + segLen - 1 || gatherFirst) gather_adjoint(d, v, d.times[n + 1]);
            const T4<T> *tp = tapeV + (ts * d.V + v) * 5;
            const T4<T> t0 = tp[0], t2 = tp[2], t3 = tp[3], t4v = tp[4];
            DVoxState<T> s, ob;
            s.u = xyz(t0);
            s.q = wxyz(tp[1]);
            s.p = xyz(t2);
            s.h = xyz(t3);
            const V3<T> F = {t0.w, t2.w, t3.w}, M = xyz(t4v);
            const T4<T> au = d.au[v], ap = d.ap[v];
            ob.u = xyz(au);
            ob.q = wxyz(d.aq[v]);
            ob.p = xyz(ap);
            ob.h = xyz(d.ah[v]);
            DVoxBar<T> r;
            d_voxel_vjp(d.vc[v], d.o, d.dt, t0.w, F, M, s, ob, r);
            d.au[v] = t4(r.s.u, r.temp);
            d.aq[v] = t4(r.s.q);
            d.ap[v] = t4(r.s.p, ap.w + r.Esq);
            d.ah[v] = t4(r.s.h, (T)0);
            d.lamF0[v] = t4(r.F, (T)0);
            d.lamF1[v] = t4(r.M, (T)0);
        }
        __syncthreads();
        // ---- link phase
        const T prevDt = n > 0 ? d.dt : (T)0;
        for (int l = lb + threadIdx.x; l < le; l += blockDim.x) {
            const DLink<T> &c = d.lc[l];
            if (c.fixedBoth) {
                d.c0[l] = d.c1[l] = d.c2[l] = d.c3[l] = T4<T>{0, 0, 0, 0};
                continue;
            }
            const int2 e = d.lends[l];
            const T4<T> *tn = tapeV + (ts * d.V + e.x) * 5, *tpp = tapeV + (ts * d.V + e.y) * 5;
            const T4<T> un = tn[0], up = tpp[0];
            const Q4<T> qn = wxyz(tn[1]), qp = wxyz(tpp[1]);
            const T4<T> *tl = tapeL + (ts * d.L + l) * 3;
            V3<T> P0, A10, A20;
            int flags;
            unpack_ls(tl[0], tl[1], tl[2], P0, A10, A20, flags);
            DLinkEval<T> ev;
            d_link_forward(c, V3<T>{up.x - un.x, up.y - un.y, up.z - un.z}, qn, qp, un.w, up.w, P0, A10, A20, flags,
                           prevDt, ev);
            const T4<T> a0 = d.al0[l], a1 = d.al1[l], a2 = d.al2[l];
            DLinkBar<T> b;
            d_link_vjp(c, ev, qn, qp, P0, A10, A20, prevDt, xyz(d.lamF0[e.x]), xyz(d.lamF0[e.y]), xyz(d.lamF1[e.x]),
                       xyz(d.lamF1[e.y]), V3<T>{a0.x, a0.y, a0.z}, V3<T>{a0.w, a1.x, a1.y}, V3<T>{a1.z, a1.w, a2.x}, b);
            d.c0[l] = t4(neg(b.du), b.tN);
            d.c1[l] = t4(b.qn);
            d.c2[l] = t4(b.du, b.tP);
            d.c3[l] = t4(b.qp);
            d.al0[l] = t4(b.P0, b.A10.x);
            d.al1[l] = T4<T>{b.A10.y, b.A10.z, b.A20.x, b.A20.y};
            d.al2[l] = T4<T>{b.A20.z, 0, 0, a2.w + b.Elin};
        }
        __syncthreads();
    }
}

// step-0 gather and the modulus chain rule: dL/dE_v = Esq_v / (2 E_v) +
// sum over links of (E_link dL/dE_link) / E_link * dE_link/dE_v
template <typename T> __global__ __launch_bounds__(256) void k_diff_finish(DD<T> d, T *gE) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.V) return;
    gather_adjoint(d, v, d.times[0]);
    const T Ev = d.E[v];
    T g = (T)0.5 * d.ap[v].w / Ev;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int l = d.vlinks[6 * v + k];
        if (l < 0) continue;
        const int2 e = d.lends[l];
        const T Eo = d.E[(k & 1) ? e.x : e.y];
        const T El = d.lc[l].E, s = Ev + Eo;
        g += d.al2[l].w / El * ((T)2 * Eo * Eo / (s * s));
    }
    gE[v] = g;
}

// ------------------------------------------------------------------- host
template <typename T> struct DiffBatch : DiffBase {
    DD<T> d{};
    RobotParams *d_rp = nullptr;
    DVox<T> *d_vc = nullptr;
    DLink<T> *d_lc = nullptr;
    double *d_phase = nullptr, *d_times = nullptr;
    T *d_E = nullptr, *d_gE = nullptr;
    T4<T> *ckV = nullptr, *ckL = nullptr, *tapeV = nullptr, *tapeL = nullptr;
    int ckSlots = 0, tapeSlots = 0, timesCap = 0;
    std::vector<VoxelMat> hvm;           // material of each voxel
    T4<T> *fin = nullptr;                // final state of the last forward (su, sq, sp, sh)
    std::vector<int> hiz, hvl;           // lattice z; 6 link ids per voxel
    std::vector<int2> hends;
    std::vector<int> haxis;
    std::vector<double> hnom;            // link nominal length

    int create(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts) {
        R = n;
        optsIn = opts;
        std::vector<int> vl;
        std::vector<int2> ends;
        hrp.resize(n);
        double maxF2 = 0;
        double dtFrac = 0.9;
        const Model &m0 = models[0]->m;
        for (int r = 0; r < n; r++) {
            const Model &m = models[r]->m;
            if (m.varyTemp != m0.varyTemp || m.tempAmplitude != m0.tempAmplitude || m.tempPeriod != m0.tempPeriod ||
                m.enableExpansion != m0.enableExpansion || m.floorEnabled != m0.floorEnabled) {
                set_error("differentiable batches need one actuation and floor setting for every robot");
                return VX3MI_E_ARG;
            }
            RobotParams &p = hrp[r];
            std::memset(&p, 0, sizeof(p));
            p.voff = (int)hvm.size();
            p.vcount = m.nv();
            p.loff = (int)hends.size();
            p.lcount = m.nl();
            for (int v = 0; v < m.nv(); v++) {
                const VoxelMat &vm = m.vmats[m.vmat[v]];
                if (vm.thermalOnAfter > 0.0) {
                    set_error("differentiable mode: TurnOnThermalExpansionAfterThisManySeconds is not supported");
                    return VX3MI_E_ARG;
                }
                hvm.push_back(vm);
                hiz.push_back(m.iz[v]);
                hlat.push_back(m.ix[v] | (m.iy[v] << 10) | (m.iz[v] << 20));
                hphase0.push_back(m.phase[v]);
                hE0.push_back(m.vmatE[m.vmat[v]]);
                for (int k = 0; k < 6; k++) {
                    const int l = m.vlinks[(size_t)v * 6 + k];
                    hvl.push_back(l < 0 ? -1 : p.loff + l);
                }
            }
            for (int l = 0; l < m.nl(); l++) {
                if (m.lmats[m.lmat[l]].nu != 0.0f) {
                    set_error("differentiable mode: volume effects (Poisson's ratio) are not supported");
                    return VX3MI_E_ARG;
                }
                hends.push_back(int2{p.voff + m.lneg[l], p.voff + m.lpos[l]});
                haxis.push_back(m.laxis[l]);
                hnom.push_back(m.lmats[m.lmat[l]].nomSize);
                const double m1 = m.vmats[m.vmat[m.lneg[l]]].mass, m2 = m.vmats[m.vmat[m.lpos[l]]].mass;
                const double k = (double)(float)(opts.e_max * m.lmats[m.lmat[l]].nomSize);
                maxF2 = std::max(maxF2, k / std::min(m1, m2));
            }
            if (m.nl() == 0)
                for (int v = 0; v < m.nv(); v++) {
                    const VoxelMat &vm = m.vmats[m.vmat[v]];
                    maxF2 = std::max(maxF2, opts.e_max * vm.nomSize / vm.mass);
                }
            dtFrac = m.dtFrac;
        }
        V = (int)hvm.size();
        L = (int)hends.size();
        o.floorK = opts.floor_stiffness;
        o.floorW = opts.floor_width;
        o.fricV = opts.friction_speed;
        o.floorEnabled = m0.floorEnabled;
        o.enableExpansion = m0.enableExpansion;
        o.tempAmp = m0.tempAmplitude;
        o.tempPeriod = m0.tempPeriod;
        o.varyTemp = m0.varyTemp;
        if (opts.dt > 0) dt = (float)opts.dt;
        else dt = (float)(dtFrac * (maxF2 > 0 ? 1.0f / (6.283185f * std::sqrt(maxF2)) : 1e-10));

        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        int rc;
        int *vld;
        int2 *endsd;
        if ((rc = alloc(&d_rp, R)) || (rc = alloc(&vld, (size_t)6 * V)) || (rc = alloc(&endsd, L)) ||
            (rc = alloc(&d_vc, V)) || (rc = alloc(&d_lc, L)) || (rc = alloc(&d_phase, V)) || (rc = alloc(&d_E, V)) ||
            (rc = alloc(&d_gE, V)))
            return rc;
        T4<T> *arr[27];
        const size_t cnt[27] = {(size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)L, (size_t)L, (size_t)L,
                                (size_t)L, (size_t)L, (size_t)L, (size_t)V, (size_t)V, (size_t)V, (size_t)V,
                                (size_t)V, (size_t)V, (size_t)L, (size_t)L, (size_t)L, (size_t)L, (size_t)L,
                                (size_t)L, (size_t)L, 1, 1, 1, 1};
        for (int k = 0; k < 23; k++)
            if ((rc = alloc(&arr[k], cnt[k]))) return rc;
        T *gph;
        if ((rc = alloc(&gph, V)) || (rc = alloc(&fin, (size_t)4 * V))) return rc;
        HIP_TRY(hipMemcpy(d_rp, hrp.data(), sizeof(RobotParams) * R, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(vld, hvl.data(), sizeof(int) * 6 * V, hipMemcpyHostToDevice));
        if (L) HIP_TRY(hipMemcpy(endsd, hends.data(), sizeof(int2) * L, hipMemcpyHostToDevice));
        d.rp = d_rp; d.vlinks = vld; d.lends = endsd; d.vc = d_vc; d.lc = d_lc; d.phase = d_phase; d.E = d_E;
        d.o = o; d.dt = (T)dt; d.V = V; d.L = L;
        d.su = arr[0]; d.sq = arr[1]; d.sp = arr[2]; d.sh = arr[3];
        d.sl0 = arr[4]; d.sl1 = arr[5]; d.sl2 = arr[6];
        d.lf0 = arr[7]; d.lf1 = arr[8]; d.lf2 = arr[9];
        d.au = arr[10]; d.aq = arr[11]; d.ap = arr[12]; d.ah = arr[13]; d.lamF0 = arr[14]; d.lamF1 = arr[15];
        d.al0 = arr[16]; d.al1 = arr[17]; d.al2 = arr[18];
        d.c0 = arr[19]; d.c1 = arr[20]; d.c2 = arr[21]; d.c3 = arr[22];
        d.gPhase = gph;
        return set_params(hphase0.data(), hE0.data());
    }

    int set_params(const double *phase, const double *E) override {
        std::vector<DVox<T>> vc((size_t)V);
        std::vector<DLink<T>> lc((size_t)L);
        std::vector<T> Ev((size_t)V);
        hphase.assign(phase, phase + V);
        hE.assign(E, E + V);
        for (int v = 0; v < V; v++) {
            if (!(E[v] > 0)) {
                set_error("Young's modulus must be positive");
                return VX3MI_E_ARG;
            }
            const VoxelMat &vm = hvm[v];
            const double m = vm.mass, Lv = vm.nomSize, I = m * Lv * Lv / 6.0;
            DVox<T> &c = vc[v];
            c.mInv = (T)(1.0 / m);
            c.IInv = (T)(1.0 / I);
            c.gDampT = (T)(vm.zetaG * 2.0 * std::sqrt(m * E[v] * Lv));
            c.gDampR = (T)(vm.zetaG * 2.0 * std::sqrt(I * E[v] * Lv * Lv * Lv));
            c.cDampT = (T)(vm.zetaC * 2.0 * std::sqrt(m * E[v] * Lv));
            c.grav = (T)vm.gravF;
            c.L = (T)Lv;
            c.alpha = (T)vm.alphaCTE;
            c.muK = (T)vm.muK;
            c.dmNum = (T)(2.0 * std::sqrt(m) * vm.zetaI);
            c.iz = hiz[v];
            c.fixed = vm.fixed;
            Ev[v] = (T)E[v];
        }
        for (int l = 0; l < L; l++) {
            const int a = hends[l].x, b = hends[l].y;
            const double Ea = E[a], Eb = E[b], El = 2.0 * Ea * Eb / (Ea + Eb), Ll = hnom[l];
            DLink<T> &c = lc[l];
            const double a2 = El * Ll * Ll * Ll / 12.0, b1 = El * Ll, b2 = El * Ll * Ll / 2.0, b3 = El * Ll * Ll * Ll / 6.0;
            c.E = (T)El;
            c.a2 = (T)a2; c.b1 = (T)b1; c.b2 = (T)b2; c.b3 = (T)b3;
            c.sqA1 = (T)std::sqrt(El * Ll);
            c.sqA2xIp = (T)std::sqrt(a2 * Ll * Ll / 6.0);
            c.sqB1 = (T)std::sqrt(b1);
            c.sqB2xFMp = (T)std::sqrt(b2 * Ll / 2.0);
            c.sqB3xIp = (T)std::sqrt(b3 * Ll * Ll / 6.0);
            c.tArea = (T)(Ll * Ll);
            c.L = (T)Ll;
            c.nomN = (T)hvm[a].nomSize;
            c.nomP = (T)hvm[b].nomSize;
            c.alphaN = (T)hvm[a].alphaCTE;
            c.alphaP = (T)hvm[b].alphaCTE;
            c.dmNumN = vc[a].dmNum;
            c.dmNumP = vc[b].dmNum;
            c.axis = haxis[l];
            c.fixedBoth = hvm[a].fixed && hvm[b].fixed;
        }
        HIP_TRY(hipMemcpy(d_vc, vc.data(), sizeof(DVox<T>) * V, hipMemcpyHostToDevice));
        if (L) HIP_TRY(hipMemcpy(d_lc, lc.data(), sizeof(DLink<T>) * L, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_phase, phase, sizeof(double) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_E, Ev.data(), sizeof(T) * V, hipMemcpyHostToDevice));
        steps = 0;
        return VX3MI_OK;
    }

    size_t bytes_ck_slot() const override { return sizeof(T4<T>) * ((size_t)4 * V + (size_t)3 * L); }
    size_t bytes_tape_slot() const override { return sizeof(T4<T>) * ((size_t)5 * V + (size_t)3 * L); }

    int ensure_buffers(int nsteps) {
        // checkpoint interval: as given, else the whole run if its tape fits the
        // budget (no recompute), else the interval minimising checkpoint + tape bytes
        int kk = optsIn.checkpoint_every;
        if (kk <= 0) {
            const double budget = optsIn.hbm_budget;
            const double full = (double)bytes_tape_slot() * nsteps + (double)bytes_ck_slot();
            if (full <= budget) kk = nsteps;
            else {
                kk = (int)std::ceil(std::sqrt((double)nsteps * bytes_ck_slot() / (double)bytes_tape_slot()));
                kk = std::max(1, std::min(kk, nsteps));
                const double need = (double)bytes_ck_slot() * ((nsteps + kk - 1) / kk) + (double)bytes_tape_slot() * kk;
                if (need > budget) {
                    set_error("checkpoints and tape of this run exceed the memory budget");
                    return VX3MI_E_LIMIT;
                }
            }
        }
        K = std::max(1, std::min(kk, nsteps));
        const int needCk = (nsteps + K - 1) / K, needTape = K;
        int rc;
        if (needCk > ckSlots) {
            if (ckV) { free_one(ckV); free_one(ckL); }
            if ((rc = alloc(&ckV, (size_t)needCk * 4 * V)) || (rc = alloc(&ckL, (size_t)needCk * 3 * std::max(L, 1))))
                return rc;
            ckSlots = needCk;
        }
        if (needTape > tapeSlots) {
            if (tapeV) { free_one(tapeV); free_one(tapeL); }
            if ((rc = alloc(&tapeV, (size_t)needTape * 5 * V)) ||
                (rc = alloc(&tapeL, (size_t)needTape * 3 * std::max(L, 1))))
                return rc;
            tapeSlots = needTape;
        }
        if (nsteps + 1 > timesCap) {
            if (d_times) free_one(d_times);
            if ((rc = alloc(&d_times, (size_t)nsteps + 1))) return rc;
            timesCap = nsteps + 1;
        }
        ckBytes = (double)bytes_ck_slot() * needCk;
        tapeBytes = (double)bytes_tape_slot() * needTape;
        return VX3MI_OK;
    }

    int forward(int nsteps) override {
        if (nsteps <= 0) { set_error("nsteps must be positive"); return VX3MI_E_ARG; }
        int rc = ensure_buffers(nsteps);
        if (rc) return rc;
        std::vector<double> times((size_t)nsteps + 1);
        times[0] = 0.0;
        for (int n = 0; n < nsteps; n++) times[n + 1] = times[n] + (double)dt;  // currentTime += dt
        HIP_TRY(hipMemcpyAsync(d_times, times.data(), sizeof(double) * (nsteps + 1), hipMemcpyHostToDevice, stream));
        d.times = d_times;
        HIP_TRY(hipEventRecord(ev0, stream));
        const int nb = (std::max(V, L) + 255) / 256;
        k_diff_init<T><<<std::max(nb, 1), 256, 0, stream>>>(d);
        // the forward keeps checkpoints; with one segment it also writes the tape
        const bool oneSeg = K >= nsteps;
        for (int s0 = 0; s0 < nsteps; s0 += chunk) {
            const int k = std::min(chunk, nsteps - s0);
            k_diff_forward<T><<<R, 256, 0, stream>>>(d, s0, k, K, ckV, ckL, oneSeg ? tapeV : nullptr,
                                                    oneSeg ? tapeL : nullptr, 0);
        }
        HIP_TRY(hipEventRecord(ev1, stream));
        T4<T> *src[4] = {d.su, d.sq, d.sp, d.sh};
        for (int k = 0; k < 4; k++)
            HIP_TRY(hipMemcpyAsync(fin + (size_t)k * V, src[k], sizeof(T4<T>) * V, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventElapsedTime(&fwdMs, ev0, ev1));
        steps = nsteps;
        tapeValid = oneSeg;
        return VX3MI_OK;
    }

    int get_state(int robot, double *pos, double *quat, double *lm, double *am) override {
        int v0 = 0, nv = V;
        if (robot >= 0) {
            if (robot >= R) { set_error("robot index out of range"); return VX3MI_E_ARG; }
            v0 = hrp[robot].voff;
            nv = hrp[robot].vcount;
        }
        std::vector<T4<T>> a((size_t)nv), b((size_t)nv), c((size_t)nv), e((size_t)nv);
        if (steps <= 0) { set_error("no forward run yet"); return VX3MI_E_ARG; }
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(a.data(), fin + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b.data(), fin + (size_t)V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(c.data(), fin + (size_t)2 * V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(e.data(), fin + (size_t)3 * V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        for (int i = 0; i < nv; i++) {
            const int v = v0 + i, pk = hlat[v];
            const double Lv = hvm[v].nomSize;
            if (pos) {
                pos[3 * i] = (pk & 1023) * Lv + (double)a[i].x;
                pos[3 * i + 1] = ((pk >> 10) & 1023) * Lv + (double)a[i].y;
                pos[3 * i + 2] = (pk >> 20) * Lv + (double)a[i].z;
            }
            if (quat) { quat[4 * i] = b[i].x; quat[4 * i + 1] = b[i].y; quat[4 * i + 2] = b[i].z; quat[4 * i + 3] = b[i].w; }
            if (lm) { lm[3 * i] = c[i].x; lm[3 * i + 1] = c[i].y; lm[3 * i + 2] = c[i].z; }
            if (am) { am[3 * i] = e[i].x; am[3 * i + 1] = e[i].y; am[3 * i + 2] = e[i].z; }
        }
        return VX3MI_OK;
    }

    int backward(const double *gpos, const double *gquat, const double *glm, const double *gam, double *gphase,
                 double *gE) override {
        if (steps <= 0) { set_error("backward needs a forward run first"); return VX3MI_E_ARG; }
        // adjoint of the final state
        std::vector<T4<T>> au((size_t)V, T4<T>{0, 0, 0, 0}), aq = au, ap = au, ah = au;
        for (int v = 0; v < V; v++) {
            if (gpos) au[v] = T4<T>{(T)gpos[3 * v], (T)gpos[3 * v + 1], (T)gpos[3 * v + 2], 0};
            if (gquat) aq[v] = T4<T>{(T)gquat[4 * v], (T)gquat[4 * v + 1], (T)gquat[4 * v + 2], (T)gquat[4 * v + 3]};
            if (glm) ap[v] = T4<T>{(T)glm[3 * v], (T)glm[3 * v + 1], (T)glm[3 * v + 2], 0};
            if (gam) ah[v] = T4<T>{(T)gam[3 * v], (T)gam[3 * v + 1], (T)gam[3 * v + 2], 0};
        }
        HIP_TRY(hipEventRecord(ev0, stream));
        HIP_TRY(hipMemcpyAsync(d.au, au.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.aq, aq.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.ap, ap.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.ah, ah.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        if (L) {
            HIP_TRY(hipMemsetAsync(d.al0, 0, sizeof(T4<T>) * L, stream));
            HIP_TRY(hipMemsetAsync(d.al1, 0, sizeof(T4<T>) * L, stream));
            HIP_TRY(hipMemsetAsync(d.al2, 0, sizeof(T4<T>) * L, stream));
        }
        HIP_TRY(hipMemsetAsync(d.gPhase, 0, sizeof(T) * V, stream));
        const int nseg = (steps + K - 1) / K;
        const int nb = (std::max(V, L) + 255) / 256;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int s0 = seg * K, len = std::min(K, steps - s0);
            if (!(tapeValid && nseg == 1)) {
                k_diff_restore<T><<<std::max(nb, 1), 256, 0, stream>>>(d, ckV, ckL, seg);
                for (int c = 0; c < len; c += chunk) {
                    const int k = std::min(chunk, len - c);
                    k_diff_forward<T><<<R, 256, 0, stream>>>(d, s0 + c, k, K, nullptr, nullptr, tapeV, tapeL, s0);
                }
            }
            const int gatherFirst = seg != nseg - 1;
            k_diff_reverse<T><<<R, 256, 0, stream>>>(d, s0, len, gatherFirst, tapeV, tapeL);
        }
        k_diff_finish<T><<<(V + 255) / 256, 256, 0, stream>>>(d, d_gE);
        HIP_TRY(hipEventRecord(ev1, stream));
        std::vector<T> gp((size_t)V), ge((size_t)V);
        HIP_TRY(hipMemcpyAsync(gp.data(), d.gPhase, sizeof(T) * V, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(ge.data(), d_gE, sizeof(T) * V, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventElapsedTime(&bwdMs, ev0, ev1));
        for (int v = 0; v < V; v++) {
            if (gphase) gphase[v] = (double)gp[v];
            if (gE) gE[v] = (double)ge[v];
        }
        tapeValid = tapeValid && nseg == 1;  // a recompute leaves the tape of segment 0
        return VX3MI_OK;
    }
};

template <typename T> DiffBase *make_diff(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts, int *rc) {
    DiffBatch<T> *b = new (std::nothrow) DiffBatch<T>();
    if (!b) { *rc = VX3MI_E_NOMEM; set_error("out of host memory"); return nullptr; }
    b->precision = sizeof(T) == 4 ? 32 : 64;
    *rc = b->create(models, n, opts);
    if (*rc != VX3MI_OK) { delete b; return nullptr; }
    return b;
}

}  // namespace vx3mi
