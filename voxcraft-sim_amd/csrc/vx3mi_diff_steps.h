// vx3mi_diff_steps.h -- the per-voxel and per-link work items of the
// differentiable timestep's forward and reverse sweeps, as host/device
// functions over a structure-of-arrays view (DD).  The GPU kernels
// (vx3mi_diff_impl.h) loop over them with barriers between phases; a host
// build runs the same items in plain loops.
//
// Layout (structure of arrays, CSR by robot):
//   state  su = (u, T)  sq = q  sp = (p, -)  sh = (h, -)   (T: temperature of the step about to run)
//   link   sl0 = (P, A1v.x)  sl1 = (A1v.yz, A2v.xy)  sl2 = (A2v.z, flags, -, -)
//          lf0 = (FN, MN.x)  lf1 = (MN.yz, FP.xy)  lf2 = (FP.z, MP)
//   tape   per voxel 5 T4: su, sq, (p, F.x), (h, F.y), (M, F.z) -- the state
//          entering the step and the link force / moment sums of the step;
//          per link 3 T4: sl0..2 entering the step
//   ckpt   per voxel 4 T4: su sq sp sh; per link 3 T4: sl0..2
//   adjoint au = (ub, Tbar of the step) aq  ap = (pb, Esq)  ah; lamF0/1 = (Fbar, -), (Mbar, -);
//          al0..2 = (Pb, A1b, A2b; al2.w accumulates E_link dL/dE_link);
//          link contributions to its ends: c0 = (-dub, tNb) c1 = qnb c2 = (dub, tPb) c3 = qpb
#pragma once

#include "vx3mi_diff.h"

namespace vx3mi {

template <typename T> struct DD {
    const int *vlinks;    // 6 per voxel, X_POS..Z_NEG, -1 none
    const int2v *lends;   // (pVNeg, pVPos) per link
    const DVox<T> *vc;
    const DLink<T> *lc;
    const double *phase;  // cycles
    const T *E;
    const double *times;  // time before step n
    DiffOpts o;
    T dt;
    int V, L;
    T4<T> *su, *sq, *sp, *sh;
    T4<T> *sl0, *sl1, *sl2;
    T4<T> *lf0, *lf1, *lf2;
    T4<T> *au, *aq, *ap, *ah, *lamF0, *lamF1;
    T4<T> *al0, *al1, *al2;
    T4<T> *c0, *c1, *c2, *c3;
    T *gPhase;
};

template <typename T> VX_HD V3<T> xyz(const T4<T> &a) { return V3<T>{a.x, a.y, a.z}; }
template <typename T> VX_HD Q4<T> wxyz(const T4<T> &a) { return Q4<T>{a.x, a.y, a.z, a.w}; }
template <typename T> VX_HD T4<T> t4(V3<T> v, T w) { return T4<T>{v.x, v.y, v.z, w}; }
template <typename T> VX_HD T4<T> t4(Q4<T> q) { return T4<T>{q.w, q.x, q.y, q.z}; }

template <typename T>
VX_HD void unpack_ls(const T4<T> &a, const T4<T> &b, const T4<T> &c, V3<T> &P, V3<T> &A1, V3<T> &A2, int &flags) {
    P = V3<T>{a.x, a.y, a.z};
    A1 = V3<T>{a.w, b.x, b.y};
    A2 = V3<T>{b.z, b.w, c.x};
    flags = (int)c.y;
}

// temperature of voxel v at time t (fixed voxels stay at 0) and dT/dphase
template <typename T> VX_HD T temp_at(const DD<T> &d, int v, double t, T *dTdPhase) {
    if (d.vc[v].fixed) { *dTdPhase = 0; return 0; }
    return d_temperature<T>(d.o, t, d.phase[v], dTdPhase);
}

// ---- forward items ---------------------------------------------------------------
template <typename T> VX_HD void diff_init_voxel(const DD<T> &d, int v) {
    T dd;
    d.su[v] = T4<T>{0, 0, 0, temp_at(d, v, d.times[0], &dd)};
    d.sq[v] = T4<T>{1, 0, 0, 0};
    d.sp[v] = T4<T>{0, 0, 0, 0};
    d.sh[v] = T4<T>{0, 0, 0, 0};
}
template <typename T> VX_HD void diff_init_link(const DD<T> &d, int l) {
    d.sl0[l] = T4<T>{0, 0, 0, 0};
    d.sl1[l] = T4<T>{0, 0, 0, 0};
    d.sl2[l] = T4<T>{0, (T)LF_SMALL_ANGLE, 0, 0};
    d.lf0[l] = d.lf1[l] = d.lf2[l] = T4<T>{0, 0, 0, 0};
}

// snapshot of a voxel (4 T4, checkpoint) or the first 4 T4 of a tape slot
template <typename T> VX_HD void diff_save_voxel(const DD<T> &d, int v, T4<T> *o) {
    o[0] = d.su[v]; o[1] = d.sq[v]; o[2] = d.sp[v]; o[3] = d.sh[v];
}
template <typename T> VX_HD void diff_save_link(const DD<T> &d, int l, T4<T> *o) {
    o[0] = d.sl0[l]; o[1] = d.sl1[l]; o[2] = d.sl2[l];
}
template <typename T> VX_HD void diff_restore_voxel(const DD<T> &d, int v, const T4<T> *o) {
    d.su[v] = o[0]; d.sq[v] = o[1]; d.sp[v] = o[2]; d.sh[v] = o[3];
}
template <typename T> VX_HD void diff_restore_link(const DD<T> &d, int l, const T4<T> *o) {
    d.sl0[l] = o[0]; d.sl1[l] = o[1]; d.sl2[l] = o[2];
}

template <typename T> VX_HD void diff_link_fwd(const DD<T> &d, int l, T prevDt) {
    const DLink<T> &c = d.lc[l];
    if (c.fixedBoth) {
        d.lf0[l] = d.lf1[l] = d.lf2[l] = T4<T>{0, 0, 0, 0};
        return;
    }
    const int2v e = d.lends[l];
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

template <typename T> VX_HD void gather_forces(const DD<T> &d, int v, V3<T> &F, V3<T> &M) {
    F = V3<T>{0, 0, 0};
    M = V3<T>{0, 0, 0};
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

// voxel phase of step n; tape (5 T4 of this voxel's slot, or null) receives
// the state entering the step and the force sums
template <typename T> VX_HD void diff_voxel_fwd(const DD<T> &d, int v, double tNext, T4<T> *tape) {
    V3<T> F, M;
    gather_forces(d, v, F, M);
    const T4<T> u = d.su[v];
    DVoxState<T> s, out;
    s.u = xyz(u);
    s.q = wxyz(d.sq[v]);
    s.p = xyz(d.sp[v]);
    s.h = xyz(d.sh[v]);
    if (tape) {
        tape[0] = u;
        tape[1] = d.sq[v];
        tape[2] = t4(s.p, F.x);
        tape[3] = t4(s.h, F.y);
        tape[4] = t4(M, F.z);
    }
    d_voxel_forward(d.vc[v], d.o, d.dt, u.w, F, M, s, out);
    T dd;
    d.su[v] = t4(out.u, temp_at(d, v, tNext, &dd));
    d.sq[v] = t4(out.q);
    d.sp[v] = t4(out.p, (T)0);
    d.sh[v] = t4(out.h, (T)0);
}

// ---- reverse items ------------------------------------------------------------------
// add the link contributions of the step at time tStep to voxel v's state
// adjoint and turn that step's temperature adjoint into the phase gradient
template <typename T> VX_HD void diff_gather_adjoint(const DD<T> &d, int v, double tStep) {
    T4<T> au = d.au[v], aq = d.aq[v];
    T tb = au.w;
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

// voxel adjoint of step n from its tape slot (5 T4)
template <typename T> VX_HD void diff_voxel_rev(const DD<T> &d, int v, const T4<T> *tape) {
    const T4<T> t0 = tape[0], t2 = tape[2], t3 = tape[3], t4v = tape[4];
    DVoxState<T> s, ob;
    s.u = xyz(t0);
    s.q = wxyz(tape[1]);
    s.p = xyz(t2);
    s.h = xyz(t3);
    const V3<T> F = {t2.w, t3.w, t4v.w}, M = xyz(t4v);
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

// link adjoint of step n: the voxel tape of the step (tapeV, 5 T4 per voxel
// at the slot) and the link's tape entry (3 T4)
template <typename T> VX_HD void diff_link_rev(const DD<T> &d, int l, T prevDt, const T4<T> *tapeV, const T4<T> *tl) {
    const DLink<T> &c = d.lc[l];
    if (c.fixedBoth) {
        d.c0[l] = d.c1[l] = d.c2[l] = d.c3[l] = T4<T>{0, 0, 0, 0};
        return;
    }
    const int2v e = d.lends[l];
    const T4<T> *tn = tapeV + (size_t)e.x * 5, *tp = tapeV + (size_t)e.y * 5;
    const T4<T> un = tn[0], up = tp[0];
    const Q4<T> qn = wxyz(tn[1]), qp = wxyz(tp[1]);
    V3<T> P0, A10, A20;
    int flags;
    unpack_ls(tl[0], tl[1], tl[2], P0, A10, A20, flags);
    DLinkEval<T> ev;
    d_link_forward(c, V3<T>{up.x - un.x, up.y - un.y, up.z - un.z}, qn, qp, un.w, up.w, P0, A10, A20, flags, prevDt, ev);
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

// step-0 gather and the modulus chain rule: dL/dE_v = Esq_v / (2 E_v) + sum over
// the voxel's links of (E_link dL/dE_link) / E_link * dE_link/dE_v
template <typename T> VX_HD T diff_finish_voxel(const DD<T> &d, int v) {
    diff_gather_adjoint(d, v, d.times[0]);
    const T Ev = d.E[v];
    T g = (T)0.5 * d.ap[v].w / Ev;
    for (int k = 0; k < 6; k++) {
        const int l = d.vlinks[6 * v + k];
        if (l < 0) continue;
        const int2v e = d.lends[l];
        const T Eo = d.E[(k & 1) ? e.x : e.y];
        const T El = d.lc[l].E, s = Ev + Eo;
        g += d.al2[l].w / El * ((T)2 * Eo * Eo / (s * s));
    }
    return g;
}

// ---- host: per-voxel / per-link constants from the per-voxel moduli -------------
// vm: material of each voxel; ends/axis/nom: link ends, axis, nominal length;
// iz: lattice z of each voxel.  Constants in double, then rounded to T.
template <typename T, typename VM>
inline void diff_constants(const VM &vm, const int2v *ends, const int *axis, const double *nom, const int *iz, int V,
                           int L, const double *E, DVox<T> *vc, DLink<T> *lc) {
    for (int v = 0; v < V; v++) {
        const VoxelMat &mt = vm[v];
        const double m = mt.mass, Lv = mt.nomSize, I = m * Lv * Lv / 6.0;
        DVox<T> &c = vc[v];
        c.mInv = (T)(1.0 / m);
        c.IInv = (T)(1.0 / I);
        c.gDampT = (T)(mt.zetaG * 2.0 * sqrt(m * E[v] * Lv));
        c.gDampR = (T)(mt.zetaG * 2.0 * sqrt(I * E[v] * Lv * Lv * Lv));
        c.cDampT = (T)(mt.zetaC * 2.0 * sqrt(m * E[v] * Lv));
        c.grav = (T)mt.gravF;
        c.L = (T)Lv;
        c.alpha = (T)mt.alphaCTE;
        c.muK = (T)mt.muK;
        c.dmNum = (T)(2.0 * sqrt(m) * mt.zetaI);
        c.iz = iz[v];
        c.fixed = mt.fixed;
    }
    for (int l = 0; l < L; l++) {
        const int a = ends[l].x, b = ends[l].y;
        const double Ea = E[a], Eb = E[b], El = 2.0 * Ea * Eb / (Ea + Eb), Ll = nom[l];
        DLink<T> &c = lc[l];
        const double a2 = El * Ll * Ll * Ll / 12.0, b1 = El * Ll, b2 = El * Ll * Ll / 2.0, b3 = El * Ll * Ll * Ll / 6.0;
        c.E = (T)El;
        c.a2 = (T)a2; c.b1 = (T)b1; c.b2 = (T)b2; c.b3 = (T)b3;
        c.sqA1 = (T)sqrt(El * Ll);
        c.sqA2xIp = (T)sqrt(a2 * Ll * Ll / 6.0);
        c.sqB1 = (T)sqrt(b1);
        c.sqB2xFMp = (T)sqrt(b2 * Ll / 2.0);
        c.sqB3xIp = (T)sqrt(b3 * Ll * Ll / 6.0);
        c.tArea = (T)(Ll * Ll);
        c.L = (T)Ll;
        c.nomN = (T)vm[a].nomSize;
        c.nomP = (T)vm[b].nomSize;
        c.alphaN = (T)vm[a].alphaCTE;
        c.alphaP = (T)vm[b].alphaCTE;
        c.dmNumN = vc[a].dmNum;
        c.dmNumP = vc[b].dmNum;
        c.axis = axis[l];
        c.fixedBoth = vm[a].fixed && vm[b].fixed;
    }
}

}  // namespace vx3mi
