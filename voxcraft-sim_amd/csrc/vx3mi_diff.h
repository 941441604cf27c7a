// vx3mi_diff.h -- the differentiable timestep: forward equations and their
// hand-written adjoints (vector-Jacobian products), templated on T.
//
// The differentiable mode keeps the link equations of VX3_Link::updateForces
// (VX3_Link.cu:90-218) and the integrator of VX3_Voxel::timeStep
// (VX3_Voxel.cu:162-277) and changes only what is not differentiable:
//   * the floor is a softplus penalty: N = k w softplus(pen / w), floor damping
//     -cDampT vz sigmoid(pen / w), with pen the floor penetration of
//     VX3_Voxel.h:185-188 (no pen >= 0 switch);
//   * friction is kinetic only with a regularised direction,
//     F = -muK N v_xy / sqrt(|v_xy|^2 + vEps^2) (no static-friction state and no
//     angular-momentum reset);
//   * material constants are per voxel, from a per-voxel Young's modulus, and
//     held in T (link E is the harmonic mean of its end moduli,
//     VX3_MaterialLink.cu:78; Poisson's ratio 0);
//   * temperatures are held in T; the orientation is renormalised every step.
// Discrete decisions (small-angle hysteresis, local-velocity validity) are
// taken by the forward pass and treated as constants by the adjoint.
//
// Both precisions use the cancellation-free forms of the float timestep
// (positions as displacements from the lattice point, rotation vectors through
// atan2, |pos2| - restLen as a quotient).
#pragma once

#include "vx3mi_physics.h"

namespace vx3mi {

template <typename T> VX_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> VX_HD V3<T> cross(V3<T> a, V3<T> b) {
    return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> VX_HD Q4<T> qadd(Q4<T> a, Q4<T> b) { Q4<T> r = {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
template <typename T> VX_HD V3<T> neg(V3<T> a) { return v3<T>(-a.x, -a.y, -a.z); }
// quotients of T: exact in double, through the hardware reciprocal in float
// (about 1 ulp; the FP32 differentiable mode is held to the FP64 one by
// tolerance, its gradients to finite differences at 1e-4)
template <typename T> VX_HD T ddiv(T a, T b) { return qdiv(a, b); }
VX_HD float matan(float a) { return atanf(a); }
VX_HD double matan(double a) { return atan(a); }
VX_HD float mexp(float a) { return expf(a); }
VX_HD double mexp(double a) { return exp(a); }
VX_HD float mlog1p(float a) { return log1pf(a); }
VX_HD double mlog1p(double a) { return log1p(a); }

// link end pair (pVNeg, pVPos), the layout of HIP's int2
struct int2v {
    int x, y;
};

// ---- per-voxel / per-link constants of the differentiable mode -------------
template <typename T> struct DVox {
    T mInv, IInv;      // 1/mass, 1/inertia
    T gDampT, gDampR;  // zetaG 2 sqrt(m E L), zetaG 2 sqrt(I E L^3)  (prop. to sqrt E)
    T cDampT;          // zetaC 2 sqrt(m E L)                          (prop. to sqrt E)
    T grav;            // gravity force (z)
    T L, alpha, muK;   // nominal size, CTE, kinetic friction
    T dmNum;           // 2 sqrt(m) zetaI (damping multiplier numerator)
    int iz, fixed;
};
template <typename T> struct DLink {
    T E;                                   // harmonic mean of the end moduli
    T a2, b1, b2, b3;                      // prop. to E
    T sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;  // prop. to sqrt E
    T tArea, L, nomN, nomP, alphaN, alphaP, dmNumN, dmNumP;
    int axis, fixedBoth;
};

struct DiffOpts {
    double floorK, floorW, fricV;  // softplus floor stiffness, width; friction regularisation speed
    int floorEnabled, enableExpansion;
    double tempAmp, tempPeriod;    // actuation (VaryTemp on)
    int varyTemp;
};

// ---- temperature ---------------------------------------------------------------
// T = amp sin(2 (float)pi (t/P + phase)), clamped to <= 0 unless expansion is on;
// *dTdPhase = dT/dphase (phase in cycles).
template <typename T>
VX_HD T d_temperature(const DiffOpts &o, double t, double phase, T *dTdPhase) {
    *dTdPhase = 0;
    if (!o.varyTemp) return 0;
    const double twoPi = (double)(2 * 3.1415926f);
    T val, der;
    if constexpr (std::is_same<T, double>::value) {
        const double a = twoPi * (t / o.tempPeriod + phase);
        val = o.tempAmp * vxm_sin(a);
        der = o.tempAmp * twoPi * vxm_cos(a);
    } else {
        double c = t / o.tempPeriod + phase;
        c -= floor(c);
        // sin and cos of 2 * 3.1415926f * c (cos as the sine a quarter cycle on)
        val = (float)o.tempAmp * sin_cycles_f32((float)c);
        der = (float)o.tempAmp * (2 * 3.1415926f) * sin_cycles_f32((float)c + 0.25f);
    }
    if (!o.enableExpansion && val > 0) return 0;
    *dTdPhase = der;
    return val;
}

// ---- adjoints of the quaternion primitives ------------------------------------
// y = qrot(q, f) = (w^2 - v.v) f + 2 (v.f) v + 2 w (v x f)
template <typename T> VX_HD void qrot_vjp(Q4<T> q, V3<T> f, V3<T> yb, Q4<T> &qb, V3<T> &fb) {
    const V3<T> v = v3<T>(q.x, q.y, q.z);
    fb = fb + qrotinv(q, yb);
    qb.w += dot(yb, f * (T)(2 * q.w) + cross(v, f) * (T)2);
    const V3<T> g = v * (T)(-2 * dot(f, yb)) + (f * dot(v, yb) + yb * dot(v, f)) * (T)2 + cross(f, yb) * (T)(2 * q.w);
    qb.x += g.x; qb.y += g.y; qb.z += g.z;
}
// y = qrotinv(q, f) = (w^2 - v.v) f + 2 (v.f) v - 2 w (v x f)
template <typename T> VX_HD void qrotinv_vjp(Q4<T> q, V3<T> f, V3<T> yb, Q4<T> &qb, V3<T> &fb) {
    const V3<T> v = v3<T>(q.x, q.y, q.z);
    fb = fb + qrot(q, yb);
    qb.w += dot(yb, f * (T)(2 * q.w) - cross(v, f) * (T)2);
    const V3<T> g = v * (T)(-2 * dot(f, yb)) + (f * dot(v, yb) + yb * dot(v, f)) * (T)2 - cross(f, yb) * (T)(2 * q.w);
    qb.x += g.x; qb.y += g.y; qb.z += g.z;
}
// r = a (x) b:  ab += rb (x) conj(b),  bb += conj(a) (x) rb
template <typename T> VX_HD void qmul_vjp(Q4<T> a, Q4<T> b, Q4<T> rb, Q4<T> &ab, Q4<T> &bb) {
    ab = qadd(ab, qmul(rb, qconj(b)));
    bb = qadd(bb, qmul(qconj(a), rb));
}
template <typename T> VX_HD V3<T> to_axis_x_vjp(V3<T> yb, int axis) { return to_axis_original(yb, axis); }
template <typename T> VX_HD V3<T> to_axis_original_vjp(V3<T> yb, int axis) { return to_axis_x(yb, axis); }
template <typename T> VX_HD Q4<T> to_axis_x_q_vjp(Q4<T> rb, int axis) {
    Q4<T> qb = rb;
    if (axis == 1) { qb.x = -rb.y; qb.y = rb.x; }
    if (axis == 2) { qb.x = -rb.z; qb.z = rb.x; }
    return qb;
}

// rotation vector of a unit quaternion: v k(w, |v|) with k = 2 atan2(|v|, w)/|v|
// (k = 2 sqrt(2/(1+w)) for |v|^2 below 2.4e-3), the float form of ToRotationVector
template <typename T> VX_HD V3<T> d_rotvec(Q4<T> q) {
    const T s2 = q.x * q.x + q.y * q.y + q.z * q.z;
    if (q.w <= (T)-1) return v3<T>(0, 0, 0);
    T k;
    if (s2 < (T)kSLTHRESH_ACOS2SQRT) k = (T)2 * msqrt(ddiv((T)2, (T)1 + q.w));
    else {
        const T s = msqrt(s2);
        k = ddiv((T)2 * matan2_nonneg(s, q.w), s);
    }
    return v3<T>(q.x * k, q.y * k, q.z * k);
}
template <typename T> VX_HD void d_rotvec_vjp(Q4<T> q, V3<T> ob, Q4<T> &qb) {
    if (q.w <= (T)-1) return;
    const T s2 = q.x * q.x + q.y * q.y + q.z * q.z;
    const V3<T> v = v3<T>(q.x, q.y, q.z);
    T k, dkdw, dkds_over_s;  // dk/ds / s
    if (s2 < (T)kSLTHRESH_ACOS2SQRT) {
        k = (T)2 * msqrt(ddiv((T)2, (T)1 + q.w));
        dkdw = ddiv(-k, (T)2 * ((T)1 + q.w));
        dkds_over_s = 0;
    } else {
        const T s = msqrt(s2), at = matan2_nonneg(s, q.w), r2 = s2 + q.w * q.w;
        k = ddiv((T)2 * at, s);
        dkdw = ddiv((T)-2, r2);
        dkds_over_s = ddiv((T)2 * (ddiv(q.w, r2 * s) - ddiv(at, s2)), s);
    }
    const T kb = dot(ob, v);
    qb.x += k * ob.x + kb * dkds_over_s * q.x;
    qb.y += k * ob.y + kb * dkds_over_s * q.y;
    qb.z += k * ob.z + kb * dkds_over_s * q.z;
    qb.w += kb * dkdw;
}

// FromRotationVector (VX3_Quat3D.h:363-378) and its adjoint
template <typename T> VX_HD void from_rotvec_vjp(V3<T> in, Q4<T> ob, V3<T> &inb) {
    const V3<T> th = v3<T>(in.x / 2, in.y / 2, in.z / 2);
    const T m2 = dot(th, th);
    T s, m2b;
    const T sb = ob.x * th.x + ob.y * th.y + ob.z * th.z;
    if (m2 * m2 < (T)kDBL_EPSILONx24) {
        s = (T)1 - m2 / (T)6;
        m2b = -ob.w * (T)0.5 - sb / (T)6;
    } else {
        const T m = msqrt(m2), sn = msin(m), cs = mcos(m);
        s = sn / m;
        const T mb = -ob.w * sn + sb * (m * cs - sn) / m2;
        m2b = mb / ((T)2 * m);
    }
    const V3<T> thb = v3<T>(ob.x * s + (T)2 * m2b * th.x, ob.y * s + (T)2 * m2b * th.y, ob.z * s + (T)2 * m2b * th.z);
    inb = inb + thb * (T)0.5;
}

// rotation taking the direction of rf onto +x (FromAngleToPosX, VX3_Quat3D.h:384-432)
template <typename T> VX_HD Q4<T> d_angle_to_posx(V3<T> rf) {
    Q4<T> q = {1, 0, 0, 0};
    if (rf.x == 0 && rf.y == 0 && rf.z == 0) return q;
    const T yx = ddiv(rf.y, rf.x), zx = ddiv(rf.z, rf.x);
    const T sa = (T)kSMALL_ANGLE_RAD;
    if (yx < sa && yx > -sa && zx < sa && zx > -sa) {
        q.y = (T)0.5 * zx;
        q.z = (T)-0.5 * yx;
        q.w = (T)1 + (T)0.5 * (-q.y * q.y - q.z * q.z);
        return q;
    }
    const T rho = msqrt(rf.y * rf.y + rf.z * rf.z);
    const T theta = matan2_nonneg(rho, rf.x);
    if (theta > (T)(kPI - kDISCARD_ANGLE_RAD)) { q.w = 0; q.y = 1; return q; }
    const T h = (T)0.5 * theta, sn = msin(h);
    q.w = mcos(h);
    q.y = ddiv(rf.z, rho) * sn;
    q.z = -ddiv(rf.y, rho) * sn;
    return q;
}
template <typename T> VX_HD void d_angle_to_posx_vjp(V3<T> rf, Q4<T> qb, V3<T> &rb) {
    if (rf.x == 0 && rf.y == 0 && rf.z == 0) return;
    const T yx = ddiv(rf.y, rf.x), zx = ddiv(rf.z, rf.x);
    const T sa = (T)kSMALL_ANGLE_RAD;
    if (yx < sa && yx > -sa && zx < sa && zx > -sa) {
        const T qy = (T)0.5 * zx, qz = (T)-0.5 * yx;
        const T qyb = qb.y - qb.w * qy, qzb = qb.z - qb.w * qz;
        const T zxb = (T)0.5 * qyb, yxb = (T)-0.5 * qzb;
        rb.x += ddiv(-(yxb * yx + zxb * zx), rf.x);
        rb.y += ddiv(yxb, rf.x);
        rb.z += ddiv(zxb, rf.x);
        return;
    }
    const T rho2 = rf.y * rf.y + rf.z * rf.z, rho = msqrt(rho2);
    const T theta = matan2_nonneg(rho, rf.x);
    if (theta > (T)(kPI - kDISCARD_ANGLE_RAD)) return;
    const T h = (T)0.5 * theta, sn = msin(h), cs = mcos(h);
    const T ey = ddiv(rf.y, rho), ez = ddiv(rf.z, rho);
    const T thb = qb.w * (-sn * (T)0.5) + (qb.y * ez - qb.z * ey) * (cs * (T)0.5);
    const T ezb = qb.y * sn, eyb = -qb.z * sn;
    const T r2 = rho2 + rf.x * rf.x;
    const T rhob = ddiv(thb * rf.x, r2);
    rb.x += ddiv(-thb * rho, r2);
    const T rho3 = rho2 * rho;
    rb.y += rhob * ey + ddiv(eyb * rf.z * rf.z - ezb * rf.y * rf.z, rho3);
    rb.z += rhob * ez + ddiv(ezb * rf.y * rf.y - eyb * rf.y * rf.z, rho3);
}

// ---- link -------------------------------------------------------------------------
// All intermediates of one link evaluation (forward), kept for the adjoint.
template <typename T> struct DLinkEval {
    V3<T> d;           // displacement difference, link frame
    Q4<T> A2, R, B, G, Bf;
    T restLen, lmr;    // rest length, L - rest length
    V3<T> rd, pos2;
    T ext;
    int small, lvv, flagsOut;
    V3<T> P, A1v, A2v;  // new link state
    T strain;
    V3<T> fN1, mN1, fP1, mP1;  // link-frame forces before the final rotations
    V3<T> FN, FP, MN, MP;      // outputs (voxel frames)
};

// Forward of one link.  du = u_pos - u_neg (displacements), tN/tP end
// temperatures, (P0, A10, A20, flags) the link state of the previous step.
template <typename T>
VX_HD void d_link_forward(const DLink<T> &c, V3<T> du, Q4<T> qn, Q4<T> qp, T tN, T tP, V3<T> P0, V3<T> A10, V3<T> A20,
                          int flags, T prevDt, DLinkEval<T> &e) {
    const int axis = c.axis;
    e.d = to_axis_x(du, axis);
    const Q4<T> A1 = to_axis_x(qn, axis);
    e.A2 = to_axis_x(qp, axis);
    e.R = qconj(A1);
    e.B = qmul(e.R, e.A2);
    const T L = c.L;
    e.restLen = (T)0.5 * (c.nomN * ((T)1 + tN * c.alphaN) + c.nomP * ((T)1 + tP * c.alphaP));
    e.lmr = (T)-0.5 * (c.nomN * (tN * c.alphaN) + c.nomP * (tP * c.alphaP));
    const Q4<T> R = e.R;
    e.rd = qrot(R, e.d);
    const V3<T> rel = v3<T>(e.rd.x - (T)2 * L * (R.y * R.y + R.z * R.z), e.rd.y + (T)2 * L * (R.x * R.y + R.w * R.z),
                            e.rd.z + (T)2 * L * (R.x * R.z - R.w * R.y));
    e.pos2 = v3<T>(L + rel.x, rel.y, rel.z);
    e.ext = rel.x + e.lmr;

    const float smallTurn = (float)ddiv(mabs(e.pos2.z) + mabs(e.pos2.y), e.pos2.x);
    const float extendPerc = (float)mabs(ddiv(e.ext, e.restLen));
    bool small = (flags & LF_SMALL_ANGLE) != 0;
    if (!small && smallTurn < kSA_BOND_BEND_RAD && extendPerc < kSA_BOND_EXT_PERC) {
        small = true;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    } else if (small && (smallTurn > kHYSTERESIS_FACTOR * kSA_BOND_BEND_RAD ||
                         extendPerc > kHYSTERESIS_FACTOR * kSA_BOND_EXT_PERC)) {
        small = false;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    }
    e.small = small;
    if (small) {
        e.G = Q4<T>{1, 0, 0, 0};
        e.Bf = e.B;
        e.P = v3<T>(e.ext, e.pos2.y, e.pos2.z);
        e.A1v = v3<T>(0, 0, 0);
    } else {
        e.G = d_angle_to_posx(e.pos2);
        e.Bf = qmul(e.G, e.B);
        const T l = len(e.pos2);
        const T num = e.ext * (e.pos2.x + e.restLen) + e.pos2.y * e.pos2.y + e.pos2.z * e.pos2.z;
        e.P = v3<T>(ddiv(num, l + e.restLen), 0, 0);
        e.A1v = d_rotvec(e.G);
    }
    e.A2v = d_rotvec(e.Bf);
    e.strain = ddiv(e.P.x, e.restLen);
    const T stress = c.E * e.strain;
    const V3<T> P = e.P, a1 = e.A1v, a2 = e.A2v;
    V3<T> fN = v3<T>(stress * c.tArea, c.b1 * P.y - c.b2 * (a1.z + a2.z), c.b1 * P.z + c.b2 * (a1.y + a2.y));
    V3<T> fP = neg(fN);
    V3<T> mN = v3<T>(c.a2 * (a2.x - a1.x), -c.b2 * P.z - c.b3 * ((T)2 * a1.y + a2.y), c.b2 * P.y - c.b3 * ((T)2 * a1.z + a2.z));
    V3<T> mP = v3<T>(c.a2 * (a1.x - a2.x), -c.b2 * P.z - c.b3 * (a1.y + (T)2 * a2.y), c.b2 * P.y - c.b3 * (a1.z + (T)2 * a2.z));
    e.lvv = (flags & LF_LOCAL_VELOCITY_VALID) != 0;
    if (e.lvv) {
        const V3<T> dP = (P - P0) * (T)0.5, dA1 = (a1 - A10) * (T)0.5, dA2 = (a2 - A20) * (T)0.5;
        const V3<T> pc = v3<T>(c.sqA1 * dP.x, c.sqB1 * dP.y - c.sqB2xFMp * (dA1.z + dA2.z),
                               c.sqB1 * dP.z + c.sqB2xFMp * (dA1.y + dA2.y));
        const T dmN = ddiv(c.dmNumN, prevDt), dmP = ddiv(c.dmNumP, prevDt);
        fN = fN + pc * dmN;
        fP = fP - pc * dmP;
        const V3<T> tNv = v3<T>(-c.sqA2xIp * (dA2.x - dA1.x), c.sqB2xFMp * dP.z + c.sqB3xIp * ((T)2 * dA1.y + dA2.y),
                                -c.sqB2xFMp * dP.y + c.sqB3xIp * ((T)2 * dA1.z + dA2.z));
        const V3<T> tPv = v3<T>(c.sqA2xIp * (dA2.x - dA1.x), c.sqB2xFMp * dP.z + c.sqB3xIp * (dA1.y + (T)2 * dA2.y),
                                -c.sqB2xFMp * dP.y + c.sqB3xIp * (dA1.z + (T)2 * dA2.z));
        mN = mN - tNv * ((T)0.5 * dmN);
        mP = mP - tPv * ((T)0.5 * dmP);
    } else {
        flags |= LF_LOCAL_VELOCITY_VALID;
    }
    e.fN1 = fN; e.mN1 = mN; e.fP1 = fP; e.mP1 = mP;
    if (!small) {
        fN = qrotinv(e.G, fN);
        mN = qrotinv(e.G, mN);
    }
    fP = qrotinv(e.Bf, fP);
    mP = qrotinv(e.Bf, mP);
    e.FN = to_axis_original(fN, axis);
    e.FP = to_axis_original(fP, axis);
    e.MN = to_axis_original(mN, axis);
    e.MP = to_axis_original(mP, axis);
    e.flagsOut = small ? (flags | LF_SMALL_ANGLE) : (flags & ~LF_SMALL_ANGLE);
}

// Adjoints produced by one link: inputs of the forward above.
template <typename T> struct DLinkBar {
    V3<T> du;       // d/d(u_pos - u_neg)
    Q4<T> qn, qp;
    T tN, tP;
    V3<T> P0, A10, A20;  // previous link state
    T Elin;         // sum over E-proportional constants c of cbar * c (sqrt ones halved): E * dL/dE
};

// Adjoint of d_link_forward given output adjoints (FNb..MPb for the four end
// forces/moments, Pb/A1b/A2b for the new link state).  e holds the forward.
template <typename T>
VX_HD void d_link_vjp(const DLink<T> &c, const DLinkEval<T> &e, Q4<T> qn, Q4<T> qp, V3<T> P0, V3<T> A10, V3<T> A20,
                      T prevDt, V3<T> FNb, V3<T> FPb, V3<T> MNb, V3<T> MPb, V3<T> Pb, V3<T> A1b, V3<T> A2b,
                      DLinkBar<T> &o) {
    (void)qn; (void)qp;
    const int axis = c.axis;
    o.du = v3<T>(0, 0, 0);
    o.qn = o.qp = Q4<T>{0, 0, 0, 0};
    o.tN = o.tP = 0;
    o.P0 = o.A10 = o.A20 = v3<T>(0, 0, 0);
    T Elin = 0, Esq = 0;  // sum cbar * c over E-proportional / sqrt(E)-proportional constants

    V3<T> fNb = to_axis_original_vjp(FNb, axis), fPb = to_axis_original_vjp(FPb, axis);
    V3<T> mNb = to_axis_original_vjp(MNb, axis), mPb = to_axis_original_vjp(MPb, axis);
    Q4<T> Bfb = {0, 0, 0, 0}, Gb = {0, 0, 0, 0};
    // final rotations
    V3<T> fP1b = v3<T>(0, 0, 0), mP1b = v3<T>(0, 0, 0), fN1b = v3<T>(0, 0, 0), mN1b = v3<T>(0, 0, 0);
    qrotinv_vjp(e.Bf, e.fP1, fPb, Bfb, fP1b);
    qrotinv_vjp(e.Bf, e.mP1, mPb, Bfb, mP1b);
    if (!e.small) {
        qrotinv_vjp(e.G, e.fN1, fNb, Gb, fN1b);
        qrotinv_vjp(e.G, e.mN1, mNb, Gb, mN1b);
    } else {
        fN1b = fNb;
        mN1b = mNb;
    }
    // local damping
    V3<T> Pbar = Pb, a1b = A1b, a2b = A2b;
    if (e.lvv) {
        const V3<T> dP = (e.P - P0) * (T)0.5, dA1 = (e.A1v - A10) * (T)0.5, dA2 = (e.A2v - A20) * (T)0.5;
        const T dmN = ddiv(c.dmNumN, prevDt), dmP = ddiv(c.dmNumP, prevDt);
        const V3<T> tNb = mN1b * (-(T)0.5 * dmN), tPb = mP1b * (-(T)0.5 * dmP);
        const V3<T> pcb = fN1b * dmN - fP1b * dmP;
        V3<T> dPb = v3<T>(0, 0, 0), dA1b = v3<T>(0, 0, 0), dA2b = v3<T>(0, 0, 0);
        // pc
        dPb.x += c.sqA1 * pcb.x;
        dPb.y += c.sqB1 * pcb.y;
        dA1b.z -= c.sqB2xFMp * pcb.y; dA2b.z -= c.sqB2xFMp * pcb.y;
        dPb.z += c.sqB1 * pcb.z;
        dA1b.y += c.sqB2xFMp * pcb.z; dA2b.y += c.sqB2xFMp * pcb.z;
        Esq += pcb.x * c.sqA1 * dP.x + pcb.y * (c.sqB1 * dP.y - c.sqB2xFMp * (dA1.z + dA2.z)) +
               pcb.z * (c.sqB1 * dP.z + c.sqB2xFMp * (dA1.y + dA2.y));
        // tN, tP
        dA2b.x += -c.sqA2xIp * tNb.x + c.sqA2xIp * tPb.x;
        dA1b.x += c.sqA2xIp * tNb.x - c.sqA2xIp * tPb.x;
        dPb.z += c.sqB2xFMp * (tNb.y + tPb.y);
        dA1b.y += c.sqB3xIp * ((T)2 * tNb.y + tPb.y);
        dA2b.y += c.sqB3xIp * (tNb.y + (T)2 * tPb.y);
        dPb.y += -c.sqB2xFMp * (tNb.z + tPb.z);
        dA1b.z += c.sqB3xIp * ((T)2 * tNb.z + tPb.z);
        dA2b.z += c.sqB3xIp * (tNb.z + (T)2 * tPb.z);
        const V3<T> tNv = v3<T>(-c.sqA2xIp * (dA2.x - dA1.x), c.sqB2xFMp * dP.z + c.sqB3xIp * ((T)2 * dA1.y + dA2.y),
                                -c.sqB2xFMp * dP.y + c.sqB3xIp * ((T)2 * dA1.z + dA2.z));
        const V3<T> tPv = v3<T>(c.sqA2xIp * (dA2.x - dA1.x), c.sqB2xFMp * dP.z + c.sqB3xIp * (dA1.y + (T)2 * dA2.y),
                                -c.sqB2xFMp * dP.y + c.sqB3xIp * (dA1.z + (T)2 * dA2.z));
        Esq += dot(tNb, tNv) + dot(tPb, tPv);
        Pbar = Pbar + dPb * (T)0.5;
        a1b = a1b + dA1b * (T)0.5;
        a2b = a2b + dA2b * (T)0.5;
        o.P0 = dPb * (T)-0.5;
        o.A10 = dA1b * (T)-0.5;
        o.A20 = dA2b * (T)-0.5;
    }
    // elastic forces: fP = -fN
    const V3<T> fNb0 = fN1b - fP1b;
    const V3<T> mNb0 = mN1b, mPb0 = mP1b;
    const V3<T> P = e.P, a1 = e.A1v, a2 = e.A2v;
    const T stress = c.E * e.strain;
    T stressb = fNb0.x * c.tArea;
    Pbar.y += c.b1 * fNb0.y;
    a1b.z -= c.b2 * fNb0.y; a2b.z -= c.b2 * fNb0.y;
    Pbar.z += c.b1 * fNb0.z;
    a1b.y += c.b2 * fNb0.z; a2b.y += c.b2 * fNb0.z;
    Elin += fNb0.x * stress * c.tArea + fNb0.y * (c.b1 * P.y - c.b2 * (a1.z + a2.z)) +
            fNb0.z * (c.b1 * P.z + c.b2 * (a1.y + a2.y));
    // mN
    a2b.x += c.a2 * mNb0.x; a1b.x -= c.a2 * mNb0.x;
    Pbar.z -= c.b2 * mNb0.y; a1b.y -= (T)2 * c.b3 * mNb0.y; a2b.y -= c.b3 * mNb0.y;
    Pbar.y += c.b2 * mNb0.z; a1b.z -= (T)2 * c.b3 * mNb0.z; a2b.z -= c.b3 * mNb0.z;
    // mP
    a1b.x += c.a2 * mPb0.x; a2b.x -= c.a2 * mPb0.x;
    Pbar.z -= c.b2 * mPb0.y; a1b.y -= c.b3 * mPb0.y; a2b.y -= (T)2 * c.b3 * mPb0.y;
    Pbar.y += c.b2 * mPb0.z; a1b.z -= c.b3 * mPb0.z; a2b.z -= (T)2 * c.b3 * mPb0.z;
    Elin += mNb0.x * (c.a2 * (a2.x - a1.x)) + mNb0.y * (-c.b2 * P.z - c.b3 * ((T)2 * a1.y + a2.y)) +
            mNb0.z * (c.b2 * P.y - c.b3 * ((T)2 * a1.z + a2.z)) + mPb0.x * (c.a2 * (a1.x - a2.x)) +
            mPb0.y * (-c.b2 * P.z - c.b3 * (a1.y + (T)2 * a2.y)) + mPb0.z * (c.b2 * P.y - c.b3 * (a1.z + (T)2 * a2.z));
    // stress = E strain, strain = P.x / restLen
    const T strainb = stressb * c.E;
    Pbar.x += ddiv(strainb, e.restLen);
    T restLenb = ddiv(-strainb * P.x, e.restLen * e.restLen);
    // rotation vectors
    Q4<T> Bb = {0, 0, 0, 0};
    d_rotvec_vjp(e.Bf, a2b, Bfb);
    V3<T> pos2b = v3<T>(0, 0, 0);
    T extb = 0;
    if (e.small) {
        Bb = qadd(Bb, Bfb);
        extb += Pbar.x;
        pos2b.y += Pbar.y;
        pos2b.z += Pbar.z;
    } else {
        d_rotvec_vjp(e.G, a1b, Gb);
        qmul_vjp(e.G, e.B, Bfb, Gb, Bb);
        const T l = len(e.pos2), den = l + e.restLen;
        const T num = e.ext * (e.pos2.x + e.restLen) + e.pos2.y * e.pos2.y + e.pos2.z * e.pos2.z;
        const T numb = ddiv(Pbar.x, den);
        const T denb = ddiv(-Pbar.x * num, den * den);
        const T lb = denb;
        restLenb += denb + numb * e.ext;
        extb += numb * (e.pos2.x + e.restLen);
        pos2b.x += numb * e.ext;
        pos2b.y += (T)2 * numb * e.pos2.y;
        pos2b.z += (T)2 * numb * e.pos2.z;
        if (l > 0) pos2b = pos2b + e.pos2 * ddiv(lb, l);
        d_angle_to_posx_vjp(e.pos2, Gb, pos2b);
    }
    // ext = rel.x + lmr; pos2 = (L + rel.x, rel.y, rel.z)
    const V3<T> relb = v3<T>(pos2b.x + extb, pos2b.y, pos2b.z);
    const T lmrb = extb;
    const Q4<T> R = e.R;
    const T L = c.L;
    Q4<T> Rb = {0, 0, 0, 0};
    Rb.y += (T)-4 * L * R.y * relb.x + (T)2 * L * R.x * relb.y - (T)2 * L * R.w * relb.z;
    Rb.z += (T)-4 * L * R.z * relb.x + (T)2 * L * R.w * relb.y + (T)2 * L * R.x * relb.z;
    Rb.x += (T)2 * L * (R.y * relb.y + R.z * relb.z);
    Rb.w += (T)2 * L * (R.z * relb.y - R.y * relb.z);
    V3<T> db = v3<T>(0, 0, 0);
    qrot_vjp(R, e.d, relb, Rb, db);
    // restLen and L - restLen from the end temperatures
    o.tN = (T)0.5 * c.nomN * c.alphaN * (restLenb - lmrb);
    o.tP = (T)0.5 * c.nomP * c.alphaP * (restLenb - lmrb);
    // B = R (x) A2, R = conj(A1)
    Q4<T> qA2b = {0, 0, 0, 0};
    qmul_vjp(R, e.A2, Bb, Rb, qA2b);
    o.qn = to_axis_x_q_vjp(qconj(Rb), axis);
    o.qp = to_axis_x_q_vjp(qA2b, axis);
    o.du = to_axis_x_vjp(db, axis);
    o.Elin = Elin + (T)0.5 * Esq;
}

// ---- voxel --------------------------------------------------------------------------
template <typename T> struct DVoxState {
    V3<T> u;  // displacement from the lattice point
    Q4<T> q;
    V3<T> p, h;
};

template <typename T> VX_HD T softplus(T x) { return x > 0 ? x + mlog1p(mexp(-x)) : mlog1p(mexp(x)); }
template <typename T> VX_HD T sigmoid(T x) {
    if (x >= 0) return ddiv((T)1, (T)1 + mexp(-x));
    const T e = mexp(x);
    return ddiv(e, (T)1 + e);
}

// Forward of one voxel (temp = temperature of this step, F/M the link sums in
// the voxel frame).  Fixed voxels do not move.
template <typename T>
VX_HD void d_voxel_forward(const DVox<T> &c, const DiffOpts &o, T dt, T temp, V3<T> F, V3<T> M, const DVoxState<T> &s,
                           DVoxState<T> &out) {
    if (c.fixed) { out = s; return; }
    V3<T> tf = qrot(s.q, F);
    const V3<T> vel = s.p * c.mInv;
    tf = tf - vel * c.gDampT;
    tf.z += c.grav;
    if (o.floorEnabled) {
        const T pen = (c.L * (temp * c.alpha) * (T)0.5 - (T)c.iz * c.L) - s.u.z;
        const T x = ddiv(pen, (T)o.floorW);
        const T N = (T)(o.floorK * o.floorW) * softplus(x);
        tf.z += N - c.cDampT * vel.z * sigmoid(x);
        const T r = msqrt(vel.x * vel.x + vel.y * vel.y + (T)(o.fricV * o.fricV));
        const T g = ddiv(c.muK * N, r);
        tf.x -= g * vel.x;
        tf.y -= g * vel.y;
    }
    out.p = s.p + tf * dt;
    out.u = s.u + out.p * (dt * c.mInv);
    V3<T> tm = qrot(s.q, M) - s.h * (c.IInv * c.gDampR);
    out.h = s.h + tm * dt;
    const Q4<T> q1 = qmul(from_rotvec(out.h * (dt * c.IInv)), s.q);
    const T n = ddiv((T)1, msqrt(q1.w * q1.w + q1.x * q1.x + q1.y * q1.y + q1.z * q1.z));
    out.q = Q4<T>{q1.w * n, q1.x * n, q1.y * n, q1.z * n};
}

template <typename T> struct DVoxBar {
    DVoxState<T> s;   // adjoint of the state entering the step
    V3<T> F, M;       // adjoint of the link force / moment sums
    T temp;           // adjoint of this step's temperature
    T Esq;            // sum over sqrt(E)-proportional constants c of cbar * c
};

template <typename T>
VX_HD void d_voxel_vjp(const DVox<T> &c, const DiffOpts &o, T dt, T temp, V3<T> F, V3<T> M, const DVoxState<T> &s,
                       const DVoxState<T> &ob, DVoxBar<T> &r) {
    if (c.fixed) {
        r.s = ob;
        r.F = r.M = v3<T>(0, 0, 0);
        r.temp = 0;
        r.Esq = 0;
        return;
    }
    // forward intermediates
    const V3<T> tf0 = qrot(s.q, F);
    const V3<T> vel = s.p * c.mInv;
    V3<T> tf = tf0 - vel * c.gDampT;
    tf.z += c.grav;
    T pen = 0, x = 0, N = 0, sg = 0, r0 = 1, g = 0;
    if (o.floorEnabled) {
        pen = (c.L * (temp * c.alpha) * (T)0.5 - (T)c.iz * c.L) - s.u.z;
        x = ddiv(pen, (T)o.floorW);
        N = (T)(o.floorK * o.floorW) * softplus(x);
        sg = sigmoid(x);
        tf.z += N - c.cDampT * vel.z * sg;
        r0 = msqrt(vel.x * vel.x + vel.y * vel.y + (T)(o.fricV * o.fricV));
        g = ddiv(c.muK * N, r0);
        tf.x -= g * vel.x;
        tf.y -= g * vel.y;
    }
    const V3<T> tmM = qrot(s.q, M);
    const V3<T> tm = tmM - s.h * (c.IInv * c.gDampR);
    const V3<T> h1 = s.h + tm * dt;
    const V3<T> rv = h1 * (dt * c.IInv);
    const Q4<T> dq = from_rotvec(rv);
    const Q4<T> q1 = qmul(dq, s.q);
    const T nn = msqrt(q1.w * q1.w + q1.x * q1.x + q1.y * q1.y + q1.z * q1.z), ni = ddiv((T)1, nn);
    const Q4<T> qh = {q1.w * ni, q1.x * ni, q1.y * ni, q1.z * ni};

    // reverse: normalisation
    const T proj = qh.w * ob.q.w + qh.x * ob.q.x + qh.y * ob.q.y + qh.z * ob.q.z;
    const Q4<T> q1b = {(ob.q.w - proj * qh.w) * ni, (ob.q.x - proj * qh.x) * ni, (ob.q.y - proj * qh.y) * ni,
                       (ob.q.z - proj * qh.z) * ni};
    Q4<T> dqb = {0, 0, 0, 0}, qb = {0, 0, 0, 0};
    qmul_vjp(dq, s.q, q1b, dqb, qb);
    V3<T> rvb = v3<T>(0, 0, 0);
    from_rotvec_vjp(rv, dqb, rvb);
    const V3<T> h1b = ob.h + rvb * (dt * c.IInv);
    const V3<T> tmb = h1b * dt;
    V3<T> hb = h1b - tmb * (c.IInv * c.gDampR);
    const T gDampRb = -c.IInv * dot(s.h, tmb);
    V3<T> Mb = v3<T>(0, 0, 0);
    qrot_vjp(s.q, M, tmb, qb, Mb);
    // u1 = u + p1 dt mInv; p1 = p + tf dt
    V3<T> ub = ob.u;
    const V3<T> p1b = ob.p + ob.u * (dt * c.mInv);
    V3<T> pb = p1b;
    V3<T> tfb = p1b * dt;
    V3<T> velb = v3<T>(0, 0, 0);
    T cDampTb = 0, tempb = 0;
    if (o.floorEnabled) {
        // friction
        velb.x -= g * tfb.x;
        velb.y -= g * tfb.y;
        const T gb = -(vel.x * tfb.x + vel.y * tfb.y);
        T Nb = ddiv(gb * c.muK, r0);
        const T rb = ddiv(-gb * g, r0);
        velb.x += ddiv(rb * vel.x, r0);
        velb.y += ddiv(rb * vel.y, r0);
        // normal force and floor damping
        Nb += tfb.z;
        velb.z -= c.cDampT * sg * tfb.z;
        const T sgb = -c.cDampT * vel.z * tfb.z;
        cDampTb -= vel.z * sg * tfb.z;
        const T xb = Nb * (T)(o.floorK * o.floorW) * sg + sgb * sg * ((T)1 - sg);
        const T penb = ddiv(xb, (T)o.floorW);
        ub.z -= penb;
        tempb += penb * c.L * c.alpha * (T)0.5;
    }
    velb = velb - tfb * c.gDampT;
    const T gDampTb = -dot(vel, tfb);
    pb = pb + velb * c.mInv;
    V3<T> Fb = v3<T>(0, 0, 0);
    qrot_vjp(s.q, F, tfb, qb, Fb);

    r.s.u = ub;
    r.s.q = qb;
    r.s.p = pb;
    r.s.h = hb;
    r.F = Fb;
    r.M = Mb;
    r.temp = tempb;
    r.Esq = gDampTb * c.gDampT + gDampRb * c.gDampR + cDampTb * c.cDampT;
}

}  // namespace vx3mi
