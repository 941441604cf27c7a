// vx3mi_physics.h -- per-link and per-voxel update equations of the VX3 timestep,
// templated on the state precision T (float or double).
//
// The equations are those of the reference (file:line given per function).  The
// double instantiation keeps the reference's operand order and float/double
// promotions, so it reproduces an FP64 run of the reference to rounding.  The
// float instantiation evaluates the same mathematical functions in forms that
// do not cancel catastrophically in single precision:
//   * positions are stored as displacements u from the nominal lattice point
//     (ix*L, iy*L, iz*L), so link vectors are L*e_axis + (u_pos - u_neg)
//     (the double instantiation stores absolute positions, as the reference
//     does, because the reference's own rounding of ix*L - (ix-1)*L is part
//     of what an FP64 run of it computes);
//   * the quaternion -> rotation vector map uses |v|^2 and atan2 instead of
//     1-w^2 and acos(w); the reference's small-angle approximation
//     2*sqrt((2-2w)/(1-w^2)) is evaluated as the identical 2*sqrt(2/(1+w));
//   * the align-to-+x rotation takes its angle from atan2, not acos.
#pragma once

#include <math.h>
#include <stdint.h>
#include <type_traits>

#include "vx3mi_math.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define VX_HD __host__ __device__ __forceinline__
#else
#define VX_HD inline
#endif

namespace vx3mi {

// the position field of the state holds absolute positions (double) or
// displacements from the nominal lattice point (float)
template <typename T> constexpr bool kAbsolutePositions = std::is_same<T, double>::value;

template <typename T> struct V3 { T x, y, z; };
// four-component device storage unit (one 16 / 32 byte load)
template <typename T> struct alignas(4 * sizeof(T)) T4 { T x, y, z, w; };
template <typename T> struct Q4 { T w, x, y, z; };

VX_HD float msqrt(float a) { return sqrtf(a); }
VX_HD double msqrt(double a) { return sqrt(a); }
// double transcendentals round identically on host and device (vx3mi_math.h)
VX_HD float msin(float a) { return sinf(a); }
VX_HD double msin(double a) { return vxm_sin(a); }
VX_HD float mcos(float a) { return cosf(a); }
VX_HD double mcos(double a) { return vxm_cos(a); }
VX_HD float macos(float a) { return acosf(a); }
VX_HD double macos(double a) { return vxm_acos(a); }
VX_HD float matan2(float a, float b) { return atan2f(a, b); }
VX_HD double matan2(double a, double b) { return atan2(a, b); }
VX_HD float mabs(float a) { return fabsf(a); }
VX_HD double mabs(double a) { return fabs(a); }

template <typename T> VX_HD V3<T> v3(T x, T y, T z) { V3<T> r = {x, y, z}; return r; }
template <typename T> VX_HD V3<T> operator+(V3<T> a, V3<T> b) { return v3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> VX_HD V3<T> operator-(V3<T> a, V3<T> b) { return v3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> VX_HD V3<T> operator*(V3<T> a, T s) { return v3<T>(a.x * s, a.y * s, a.z * s); }
template <typename T> VX_HD T len(V3<T> a) { return msqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// ---- material constants (float, as the reference stores them) ------------
// Voxel material: VX3_MaterialVoxel.h:29-44, VX3_Voxel.h:185-208.
struct VoxelMat {
    double nomSize;        // nominalSize()
    double thermalOnAfter; // TurnOnThermalExpansionAfterThisManySeconds
    float alphaCTE;
    float mass, massInv, inertiaInv;
    float dmNum;    // 2*_sqrtMass*zetaInternal   (dampingMultiplier numerator)
    float gDampT;   // globalDampingTranslateC
    float gDampR;   // globalDampingRotateC
    float cDampT;   // collisionDampingTranslateC
    float gravF;    // gravityForce()
    float penK;     // penetrationStiffness()
    float penKinv;  // 1 / penK, rounded as the contact force's float division rounds it
    float muS, muK;
    int fixed, isMeasured;
    float zetaI, zetaG, zetaC;  // damping ratios (the differentiable mode derives its own constants)
    int ext;        // index of the voxel external (VoxelExt table), -1 none
};

// External of a voxel (VX3_External.h:27-31, src/old/VX_External.h): fixed
// degrees of freedom, the prescribed translation / rotation of the fixed ones,
// an applied force and moment.  Voxels with an external get their own
// material table entry pointing at it, so voxels without one cost nothing.
enum { DOF_X = 1, DOF_Y = 2, DOF_Z = 4, DOF_TX = 8, DOF_TY = 16, DOF_TZ = 32, DOF_ALL = 63 };
struct VoxelExt {
    double t[3];       // extTranslation (m)
    double q[4];       // rotationQuat() of extRotation: w, x, y, z
    double r[3];       // extRotation (rotation vector, rad)
    float f[3], m[3];  // extForce (N), extMoment (N m)
    int dof;           // dofFixed
    int pad;
};

// Cilia of a voxel (VX3_Voxel.h:298-300, read from the VXA's BaseCiliaForce /
// ShiftCiliaForce layers): the force is coef * R(q) (base + localSignal *
// shift) on surface voxels once the time reaches onAfter
// (gpu_update_cilia_force VX3_VoxelyzeKernel.cu:846-858); coef is the
// material's Cilia, 0 for voxels of robots without EnableCilia.
struct VoxelCilia {
    double base[3], shift[3];
    double coef, onAfter;
};

// Signal state of a voxel (VX3_Voxel.h:304-309) and the signal it sent in the
// current step (value 0: none), with its material's signal constants.
struct VoxelSignal {
    double local, localDt;          // localSignal, localSignaldt
    double value, activeTime;       // d_signal
    double inactiveUntil, nextPulse;
    double outValue, outTime;
};
struct SignalMat {
    double period, decay, delay, inactive;  // PaceMakerPeriod, signalValueDecay, signalTimeDelay, inactivePeriod
    int pacemaker, pad;
};

// VX3_Voxel::receiveSignal (VX3_Voxel.cu:301-322); sm: the receiver's material
VX_HD void signal_receive(VoxelSignal &s, const SignalMat &sm, double value, double activeTime, bool force) {
    if (!force && s.inactiveUntil > activeTime) return;
    if (value < 0.1) return;
    s.inactiveUntil = activeTime + sm.inactive;
    s.local = value;
    s.value = value * sm.decay;
    if (s.value < 0.1) s.value = 0;
    s.activeTime = activeTime;
}
// a voxel's own signal section at the end of VX3_Voxel::timeStep
// (VX3_Voxel.cu:271-276): propagateSignal (:323-348; the send is kept in
// outValue / outTime for the receive phase), packMaker (:290-299),
// localSignalDecay (:279-288); t is the time at the start of the step
VX_HD void signal_own(VoxelSignal &s, const SignalMat &sm, double t) {
    s.outValue = 0;
    if (!(s.activeTime > t) && !(s.value < 0.1)) {
        s.outValue = s.value;
        s.outTime = t + sm.delay;
        s.value = 0;
        s.activeTime = 0;
        s.inactiveUntil = t + 2 * sm.delay + sm.inactive;
    }
    if (sm.pacemaker && !(s.nextPulse > t)) {
        signal_receive(s, sm, 100, t, true);
        s.nextPulse = t + sm.period;
    }
    if (!(s.localDt > t)) {
        if (s.local < 0.1) {
            s.local = 0;
        } else {
            s.local = s.local * 0.9;
            s.localDt = t + 0.01;
        }
    }
}

// Link material (VX3_MaterialLink.h:27-36) plus per-end constants the link
// kernel needs so it never touches the voxel material table.
struct LinkMat {
    double nomSize;
    double nomN, nomP;      // endpoint voxel nominal sizes (rest length)
    float E, nu, eHat, epsFail;
    float a1, a2, b1, b2, b3;
    float sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;
    float tArea, tStrainSum; // frozen transverse info (VX3_Link.cu:147-150)
    float alphaN, alphaP;    // endpoint CTEs
    float dmNumN, dmNumP;    // endpoint damping multiplier numerators
    int fixedBoth;
    int curve;               // stress-strain curve: first point << 8 | points (0: linear material)
};

// One point of a non-linear stress-strain curve (CVX_Material strainData /
// stressData); a curve starts at (0, 0).
struct CurvePt {
    float strain, stress;
};
constexpr int kMaxCurvePoints = 72;  // host capacity of a material or link curve (VX3MI_MAX_CURVE_POINTS)
VX_HD int curve_points(const LinkMat &lm) { return lm.curve & 255; }
VX_HD int curve_first(const LinkMat &lm) { return lm.curve >> 8; }

// ---- constants --------------------------------------------------------------
// src/Utils/VX3_Quat3D.h:30-56, src/old/types.h:57-59, VX3_Collision.h:6
constexpr double kPI = 3.14159265358979;
constexpr double kDBL_EPSILONx24 = 5.328e-15;
constexpr double kDISCARD_ANGLE_RAD = 1e-7;
constexpr double kSMALL_ANGLE_RAD = 1.732e-2;
constexpr double kSLTHRESH_ACOS2SQRT = 2.4e-3;
constexpr float kHYSTERESIS_FACTOR = 1.2f;
constexpr float kSA_BOND_BEND_RAD = 0.05f;
constexpr float kSA_BOND_EXT_PERC = 0.50f;
constexpr double kCOLLISION_ENVELOPE_RADIUS = 0.625;

// voxel flag bits (src/old/types.h:34-38); link flag bits (types.h:6-8 + smallAngle)
constexpr int VF_FLOOR_ENABLED = 1 << 2;
constexpr int VF_FLOOR_STATIC_FRICTION = 1 << 3;
constexpr int LF_LOCAL_VELOCITY_VALID = 1;
constexpr int LF_SMALL_ANGLE = 2;
// branches the last update took, kept in the flag word for comparisons (no
// equation reads them): ToRotationVector of angle1 / angle2 by acos (above the
// 2.4e-3 threshold), FromAngleToPosX by its exact form (above 1.732e-2)
constexpr int LF_ROTVEC_ACOS1 = 4;
constexpr int LF_ROTVEC_ACOS2 = 8;
constexpr int LF_POSX_EXACT = 16;
constexpr int LF_DECISIONS = LF_ROTVEC_ACOS1 | LF_ROTVEC_ACOS2 | LF_POSX_EXACT;
// a link made by attach (its ends are not lattice neighbours), and the updates
// left whose forces are scaled by 0.01 (isNewLink, VX3_Link.cu:206-213) in the
// bits from LF_NEW_SHIFT; the flags travel as a float, exact below 2^24
constexpr int LF_ATTACHED = 32;
constexpr int LF_NEW_SHIFT = 8;
constexpr int kMaxSafetyGuard = 65535;

// ---- discrete decisions of the piecewise equations ------------------------------
// Each branch point of the reference's equations where a continuous quantity is
// compared with a threshold.  A host test build may define
// VX3MI_TEST_DECISION(kind, b) before including this header to observe or
// replace the decision b (a bool lvalue) -- the attribution of FP32-vs-FP64
// differences to branches; otherwise it expands to nothing.
enum DecisionKind {
    DEC_SMALL_ANGLE = 0,  // link small-angle formulation with hysteresis (VX3_Link.cu:117-131)
    DEC_ROTVEC_SERIES,    // ToRotationVector square-root form below 2.4e-3 (VX3_Quat3D.h:347-359)
    DEC_POSX_LINEAR,      // FromAngleToPosX first-order form below 1.732e-2 (VX3_Quat3D.h:384-400)
    DEC_FLOOR,            // floor contact, floorPenetration() >= 0 (VX3_Voxel.cu:188, 399-426)
    DEC_STATIC_BREAK,     // static friction broken, |F_xy|^2 > (mu_s N)^2 (VX3_Voxel.cu:409-415)
    DEC_STATIC_MAKE,      // static friction entered, hKE + work <= 0 (VX3_Voxel.cu:202-209)
    DEC_FLOOR_AFTER,      // floor contact after the move, the angMom reset (VX3_Voxel.cu:259-264)
    DEC_CONTACT,          // collision pair in contact, relDist > 0 (VX3_Collision.cu:17-31)
    DEC_KINDS
};
#ifndef VX3MI_TEST_DECISION
#define VX3MI_TEST_DECISION(kind, b) ((void)0)
#endif

// ---- quaternion algebra (VX3_Quat3D.h:217-222, 439-447, 465-474) -----------
template <typename T> VX_HD Q4<T> qmul(Q4<T> a, Q4<T> f) {
    Q4<T> r;
    r.w = a.w * f.w - a.x * f.x - a.y * f.y - a.z * f.z;
    r.x = a.w * f.x + a.x * f.w + a.y * f.z - a.z * f.y;
    r.y = a.w * f.y - a.x * f.z + a.y * f.w + a.z * f.x;
    r.z = a.w * f.z + a.x * f.y - a.y * f.x + a.z * f.w;
    return r;
}
template <typename T> VX_HD Q4<T> qconj(Q4<T> a) { Q4<T> r = {a.w, -a.x, -a.y, -a.z}; return r; }
template <typename T> VX_HD V3<T> qrot(Q4<T> q, V3<T> f) {
    T tw = f.x * q.x + f.y * q.y + f.z * q.z;
    T tx = f.x * q.w - f.y * q.z + f.z * q.y;
    T ty = f.x * q.z + f.y * q.w - f.z * q.x;
    T tz = -f.x * q.y + f.y * q.x + f.z * q.w;
    return v3<T>(q.w * tx + q.x * tw + q.y * tz - q.z * ty, q.w * ty - q.x * tz + q.y * tw + q.z * tx,
                 q.w * tz + q.x * ty - q.y * tx + q.z * tw);
}
template <typename T> VX_HD V3<T> qrotinv(Q4<T> q, V3<T> f) {
    T tw = q.x * f.x + q.y * f.y + q.z * f.z;
    T tx = q.w * f.x - q.y * f.z + q.z * f.y;
    T ty = q.w * f.y + q.x * f.z - q.z * f.x;
    T tz = q.w * f.z - q.x * f.y + q.y * f.x;
    return v3<T>(tw * q.x + tx * q.w + ty * q.z - tz * q.y, tw * q.y - tx * q.z + ty * q.w + tz * q.x,
                 tw * q.z + tx * q.y - ty * q.x + tz * q.w);
}

// ---- MathTree formulas (src/Utils/VX3_MathTree.h) ---------------------------------
enum MtOp {
    MT_END, MT_CONST, MT_E, MT_PI, MT_VAR, MT_ADD, MT_SUB, MT_MUL, MT_DIV, MT_POW, MT_SQRT, MT_SIN, MT_COS,
    MT_TAN, MT_ATAN, MT_LOG, MT_INT, MT_ABS, MT_NOT, MT_GT, MT_LT, MT_AND, MT_OR, MT_NORMALCDF
};
// variables in the order of VX3_MathTree::eval's arguments
enum MtVar { MV_X, MV_Y, MV_Z, MV_HIT, MV_T, MV_ANGLE, MV_CLOSENESS, MV_PAIRS, MV_NUMVOX, MV_COUNT };
struct MtToken {
    int op;
    double value;
};
constexpr int kMaxFormulaStack = 16;

// A formula compiled to postfix (operands before their operator, left before
// right) evaluated with a bounded stack instead of VX3_MathTree::eval's
// 1024-entry value buffer; every operation gets the operands the reference
// gives it.  sin and cos are the shared vxm_ functions.
VX_HD double formula_eval(const MtToken *p, int n, const double *vars) {
    double st[kMaxFormulaStack];
    int sp = 0;
    for (int i = 0; i < n; i++) {
        const int op = p[i].op;
        if (op == MT_CONST) st[sp++] = p[i].value;
        else if (op == MT_E) st[sp++] = 2.71828182845904523536;
        else if (op == MT_PI) st[sp++] = 3.14159265358979323846;
        else if (op == MT_VAR) st[sp++] = vars[(int)p[i].value];
        else if ((op >= MT_SQRT && op <= MT_NOT) || op == MT_NORMALCDF) {
            const double a = st[sp - 1];
            double r;
            switch (op) {
            case MT_SQRT: r = sqrt(a); break;
            case MT_SIN: r = vxm_sin(a); break;
            case MT_COS: r = vxm_cos(a); break;
            case MT_TAN: r = tan(a); break;
            case MT_ATAN: r = atan(a); break;
            case MT_LOG: r = log(a); break;
            case MT_INT: r = rint(a); break;
            case MT_ABS: r = fabs(a); break;
            case MT_NOT: r = (a == 0.0) ? 1.0 : 0.0; break;
            default: r = 0.5 * erfc(-a / sqrt(2.0)); break;  // standard normal CDF
            }
            st[sp - 1] = r;
        } else {
            const double b = st[--sp], a = st[sp - 1];
            double r;
            switch (op) {
            case MT_ADD: r = a + b; break;
            case MT_SUB: r = a - b; break;
            case MT_MUL: r = a * b; break;
            case MT_DIV: r = a / b; break;
            case MT_POW: r = pow(a, b); break;
            case MT_GT: r = a > b; break;
            case MT_LT: r = a < b; break;
            case MT_AND: r = (a != 0.0) && (b != 0.0); break;
            default: r = (a != 0.0) || (b != 0.0); break;
            }
            st[sp - 1] = r;
        }
    }
    return sp > 0 ? st[0] : 0.0;
}

// ---- FP32 device arithmetic ------------------------------------------------------
// The FP32 path is held to the FP64 result by tolerance, not by bits: on the
// device its quotients and square roots use the hardware reciprocal / square
// root / reciprocal square root (v_rcp_f32, v_sqrt_f32, v_rsq_f32, about 1 ulp)
// instead of the IEEE-exact instruction sequences.  FP64 code and host builds
// keep exact division.
VX_HD float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
VX_HD float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
VX_HD float frsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
// atan2(y, x) in float for y >= 0 and (x, y) != (0, 0), in [0, pi]: the ratio
// of the smaller to the larger of y and |x| (in [0, 1]) is reduced below
// tan(pi/8) by atan(t) = pi/4 + atan((t - 1)/(t + 1)), formed from the two
// magnitudes with one reciprocal; atan(u) = u + u^3 P(u^2) with P of degree 4
// fitted by least squares (relative error weight) on Chebyshev nodes of
// [0, tan^2(pi/8)], within 6.3e-8 relative in float; then the octant and the
// sign of x, with pi/4, pi/2 and pi each split into a float and the float of
// its remainder (rounded constants alone bias the angle by +0.3 ulp on
// average).  Within 1.9e-7 relative (about 2.5 ulp): the rotation-vector angle
// of the FP32 differentiable mode, whose gradients are held to finite
// differences.  The product timestep keeps atan2f there: with this function its
// FP32 run of config 1 shows orientation excursions above 1e-5 that no branch
// of the reference's piecewise formulas explains (a correctly rounded atan2
// shows none), so the product path gives up the 3.5 % it saved.
VX_HD float atan2_nonneg_f32(float y, float x) {
    const float ax = mabs(x);
    const bool swap = y > ax;
    const float num = swap ? ax : y, den = swap ? y : ax;
    const bool hi = num > 0.41421356f * den;
    const float u = (hi ? num - den : num) * frcp(hi ? num + den : den);
    const float z = u * u;
    const float p = -0.33333313f + z * (0.19998442f + z * (-0.14243065f + z * (0.10590773f + z * -0.06071304f)));
    float r = u + (u * z) * p;
    if (hi) r = 0.785398126f + (r + 3.77489471e-08f);   // pi/4 +
    if (swap) r = 1.57079625f - (r - 7.54978942e-08f);  // pi/2 -
    if (x < 0) r = 3.14159250f - (r - 1.50995788e-07f); // pi -
    return r;
}

// atan2(a, b) for a >= 0: the function above in float, libm's in double
VX_HD float matan2_nonneg(float a, float b) { return atan2_nonneg_f32(a, b); }
VX_HD double matan2_nonneg(double a, double b) { return atan2(a, b); }

// a / b: exact in double, through the reciprocal in float
template <typename T> VX_HD T qdiv(T a, T b) {
    if constexpr (std::is_same<T, float>::value) return a * frcp(b);
    else return a / b;
}

// ToRotationVector, VX3_Quat3D.h:347-359; *acos_branch (when given): 1 when the
// angle is above the threshold of the square-root form
template <typename T> VX_HD V3<T> to_rotvec(Q4<T> q, int *acos_branch = nullptr) {
    if (acos_branch) *acos_branch = 0;
    if constexpr (std::is_same<T, double>::value) {
        if (q.w >= 1.0 || q.w <= -1.0) return v3<T>(0, 0, 0);
        double sl = 1.0 - q.w * q.w;
        bool series = sl < kSLTHRESH_ACOS2SQRT;
        VX3MI_TEST_DECISION(DEC_ROTVEC_SERIES, series);
        if (series) {
            double s = sqrt((2 - 2 * q.w) / sl);
            return v3<T>(q.x * 2.0 * s, q.y * 2.0 * s, q.z * 2.0 * s);
        }
        if (acos_branch) *acos_branch = 1;
        double a = vxm_acos(q.w), d = sqrt(sl);
        return v3<T>(q.x * 2.0 * a / d, q.y * 2.0 * a / d, q.z * 2.0 * a / d);
    } else {
        float s2 = q.x * q.x + q.y * q.y + q.z * q.z;  // == 1 - w^2 for a unit quaternion
        if (s2 == 0.0f || q.w <= -1.0f) return v3<T>(0, 0, 0);
        float k;
        bool series = s2 < (float)kSLTHRESH_ACOS2SQRT;
        VX3MI_TEST_DECISION(DEC_ROTVEC_SERIES, series);
        // 2 sqrt((2 - 2w) / (1 - w^2)) = 2 sqrt(2) / sqrt(1 + w)
        if (series) k = 2.8284271247461903f * frsqrt(1.0f + q.w);
        else {
            if (acos_branch) *acos_branch = 1;
            float s = fsqrt(s2);
            k = 2.0f * atan2f(s, q.w) * frcp(s);
        }
        return v3<T>(q.x * k, q.y * k, q.z * k);
    }
}

// FromRotationVector, VX3_Quat3D.h:363-378
template <typename T> VX_HD Q4<T> from_rotvec(V3<T> in) {
    V3<T> th = v3<T>(in.x / 2, in.y / 2, in.z / 2);
    T s, m2 = th.x * th.x + th.y * th.y + th.z * th.z;
    Q4<T> r;
    if constexpr (std::is_same<T, float>::value) {
        // half angles below 0.25 rad (every per-step rotation): cos m and
        // sin(m)/m by their Maclaurin series to m^6 (truncation < 4e-10)
        if (m2 < 0.0625f) {
            r.w = 1.0f + m2 * (-0.5f + m2 * (1.0f / 24.0f + m2 * (-1.0f / 720.0f)));
            s = 1.0f + m2 * (-1.0f / 6.0f + m2 * (1.0f / 120.0f + m2 * (-1.0f / 5040.0f)));
            r.x = th.x * s;
            r.y = th.y * s;
            r.z = th.z * s;
            return r;
        }
    }
    if (m2 * m2 < (T)kDBL_EPSILONx24) {
        r.w = (T)1.0 - (T)0.5 * m2;
        s = (T)1.0 - m2 / (T)6.0;
    } else {
        T m = msqrt(m2);
        r.w = mcos(m);
        s = msin(m) / m;
    }
    r.x = th.x * s;
    r.y = th.y * s;
    r.z = th.z * s;
    return r;
}

// FromAngleToPosX, VX3_Quat3D.h:384-432 (the target starts as the identity,
// VX3_Link.cu:103); *exact (when given): 1 unless the first-order form is taken
template <typename T> VX_HD Q4<T> from_angle_to_posx(V3<T> rf, int *exact = nullptr) {
    Q4<T> q = {1, 0, 0, 0};
    if (exact) *exact = 0;
    if (rf.x == 0 && rf.y == 0 && rf.z == 0) return q;
    T yx, zx;
    if constexpr (std::is_same<T, float>::value) {
        const float ix = frcp(rf.x);
        yx = rf.y * ix;
        zx = rf.z * ix;
    } else {
        yx = rf.y / rf.x;
        zx = rf.z / rf.x;
    }
    const T sa = (T)kSMALL_ANGLE_RAD;
    bool linear = yx < sa && yx > -sa && zx < sa && zx > -sa;
    VX3MI_TEST_DECISION(DEC_POSX_LINEAR, linear);
    if (linear) {
        q.x = 0;
        q.y = (T)0.5 * zx;
        q.z = (T)-0.5 * yx;
        q.w = (T)1 + (T)0.5 * (-q.y * q.y - q.z * q.z);
        return q;
    }
    if (exact) *exact = 1;
    if constexpr (std::is_same<T, float>::value) {
        // the same rotation without trigonometry: for the unit vector n = rf/l
        // at angle theta from +x, cos(theta/2) = sqrt((1 + n.x)/2) and the axis
        // (0, n.z, -n.y) scaled by sin(theta/2)/sin(theta) = 1/(2 cos(theta/2)):
        // q = ((l + x), 0, z, -y) / sqrt(2 l (l + x)).  Well conditioned while
        // l + x is not small against l (theta below ~150 degrees); beyond, the
        // angle path below.
        const float l = fsqrt(rf.x * rf.x + rf.y * rf.y + rf.z * rf.z);
        const float lpx = l + rf.x;
        if (lpx > 0.125f * l) {
            const float r = frsqrt(2.0f * l * lpx);
            q.w = lpx * r;
            q.x = 0;
            q.y = rf.z * r;
            q.z = -rf.y * r;
            return q;
        }
    }
    V3<T> n = rf;
    T l = msqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    if (l > 0) {
        T li = (T)1.0 / l;
        n.x *= li;
        n.y *= li;
        n.z *= li;
    }
    T theta;
    if constexpr (std::is_same<T, double>::value) theta = vxm_acos(n.x);
    else theta = atan2f(sqrtf(n.y * n.y + n.z * n.z), n.x);
    if (theta > (T)(kPI - kDISCARD_ANGLE_RAD)) {
        q.w = 0; q.x = 0; q.y = 1; q.z = 0;
        return q;
    }
    T ami = (T)1.0 / msqrt(n.z * n.z + n.y * n.y);
    T a = (T)0.5 * theta, s = msin(a);
    q.w = mcos(a);
    q.x = 0;
    q.y = n.z * ami * s;
    q.z = -n.y * ami * s;
    return q;
}

// toAxisX / toAxisOriginal, VX3_Link.h:134-191
template <typename T> VX_HD V3<T> to_axis_x(V3<T> v, int axis) {
    if (axis == 1) return v3<T>(v.y, -v.x, v.z);
    if (axis == 2) return v3<T>(v.z, v.y, -v.x);
    return v;
}
template <typename T> VX_HD Q4<T> to_axis_x(Q4<T> q, int axis) {
    Q4<T> r = q;
    if (axis == 1) { r.x = q.y; r.y = -q.x; }
    if (axis == 2) { r.x = q.z; r.z = -q.x; }
    return r;
}
template <typename T> VX_HD V3<T> to_axis_original(V3<T> v, int axis) {
    if (axis == 1) return v3<T>(-v.y, v.x, v.z);
    if (axis == 2) return v3<T>(-v.z, v.y, v.x);
    return v;
}

// ---- temperature (gpu_update_temperature VX3_VoxelyzeKernel.cu:625-650) -----
// sin(K c) in float, K = 2 * 3.1415926f (the reference's constant), for a phase
// c in [0, 1.25) counted in cycles: the argument x = K c is formed in float, as
// the float path always formed it, and brought into [-pi/2, pi/2] by pi - x or
// x - 2 pi with pi split into a float and the float of its remainder (Cody and
// Waite's reduction: the subtraction of the float part is exact here), then sin
// is its odd Taylor series to the x^13 term (remainder below (pi/2)^15 / 15! <
// 7e-10).  Within about 1 ulp of the sine of the same float argument, with no
// bias; about a third of the instructions of sinf.
VX_HD float sin_cycles_f32(float c) {
    constexpr float kPiHi = 3.14159274101257324f, kPiLo = -8.74227766e-08f;  // pi = kPiHi + kPiLo
    float x = (2 * 3.1415926f) * c;
    if (x > 4.71238898f) x = (x - 2 * kPiHi) - 2 * kPiLo;  // sin(x - 2 pi)
    else if (x > 1.57079633f) x = (kPiHi - x) + kPiLo;     // sin(pi - x)
    const float x2 = x * x;
    return x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f +
               x2 * (1.0f / 362880.0f + x2 * (-1.0f / 39916800.0f + x2 * (1.0f / 6227020800.0f)))))));
}

// Returns the new temperature of a voxel at time t (the caller skips fixed
// voxels and voxels whose TurnOnThermalExpansion time has not come).
// FP32 path: the cycle count t / period + phase is formed in double as
// t * invPeriod + phase and reduced to [0, 1) before the float sine
VX_HD float actuation_temperature_f32(double amp, double invPeriod, double t, double phase, int enableExpansion) {
    double c = t * invPeriod + phase;
    c -= floor(c);
    float Tc = (float)amp * sin_cycles_f32((float)c);
    if (!enableExpansion && Tc > 0) Tc = 0;
    return Tc;
}

template <typename T> VX_HD float actuation_temperature(double amp, double period, double t, double phase, int enableExpansion) {
    if constexpr (std::is_same<T, double>::value) {
        double Tc = amp * vxm_sin((double)(2 * 3.1415926f) * (t / period + phase));
        if (!enableExpansion && Tc > 0) Tc = 0;
        return (float)Tc;
    } else {
        return actuation_temperature_f32(amp, 1.0 / period, t, phase, enableExpansion);
    }
}

// baseSize(axis) thermal factor, VX3_Voxel.h:96-100: (1 + tempe*alphaCTE) is float
VX_HD float thermal_factor(float temp, float alpha) { return 1 + temp * alpha; }

// rest length, VX3_Link.cu:80-83
template <typename T> VX_HD T rest_length(const LinkMat &lm, float tN, float tP) {
    float fN = thermal_factor(tN, lm.alphaN), fP = thermal_factor(tP, lm.alphaP);
    if constexpr (std::is_same<T, double>::value) return 0.5 * (lm.nomN * (double)fN + lm.nomP * (double)fP);
    else return 0.5f * ((float)lm.nomN * fN + (float)lm.nomP * fP);
}

// ---- link state -----------------------------------------------------------------
template <typename T> struct LinkState {
    V3<T> pos2, a1v, a2v;
    float maxStrain;
    int flags;
};

template <typename T> struct LinkOut {
    V3<T> fN, fP, mN, mP;  // forceNeg, forcePos, momentNeg, momentPos (voxel frames)
    float strain;
};

// VX3_Material::stress (VX3_Material.cu:90-124) of a link material with a
// stress-strain curve c of curve_points(lm) points (at least three: non-linear).
VX_HD float curve_stress(const LinkMat &lm, const CurvePt *c, float strain, float tss, bool forceLinear) {
    if (lm.epsFail != -1.0f && strain > lm.epsFail) return 0.0f;
    if (strain <= c[1].strain || forceLinear) {
        if (lm.nu == 0.0f) return lm.E * strain;
        return lm.eHat * ((1 - lm.nu) * strain + lm.nu * tss);
    }
    const int n = curve_points(lm);
    for (int i = 2; i < n; i++) {
        if (strain <= c[i].strain || i == n - 1) {  // the segment ending at point i (the last one extrapolates)
            const float perc = (strain - c[i - 1].strain) / (c[i].strain - c[i - 1].strain);
            const float basic = c[i - 1].stress + perc * (c[i].stress - c[i - 1].stress);
            if (lm.nu == 0.0f) return basic;
            const float modulus = (c[i].stress - c[i - 1].stress) / (c[i].strain - c[i - 1].strain);
            const float modulusHat = modulus / ((1 - 2 * lm.nu) * (1 + lm.nu));
            const float effStrain = basic / modulus;
            const float effTss = tss * (effStrain / strain);
            return modulusHat * ((1 - lm.nu) * effStrain + lm.nu * effTss);
        }
    }
    return 0.0f;
}

// VX3_Link::updateForces VX3_Link.cu:135-218 with orientLink :90-133,
// updateStrain :220-260, VX3_Material::stress VX3_Material.cu:90-124.
// du = position field of pVPos minus that of pVNeg: the global link vector in
// double, the displacement difference (nominal offset nomSize along the link
// axis) in float.  tN/tP are the end temperatures (rest length,
// VX3_Link.cu:80-83).  curves: the stress-strain curve table (read only for
// non-linear materials).  The strain offset of a non-linear link that backed
// off its curve, maxStrain - stress(maxStrain) / E (:239-242), is a function
// of maxStrain alone and is recomputed instead of stored.  LINK_FRAME: the
// four end vectors are left in the link's frame (x along the link; the caller
// applies toAxisOriginal, an exact permutation with sign changes, where the
// axis is known at compile time).
template <typename T, bool LINK_FRAME = false>
VX_HD void link_forces(const LinkMat &lm, const CurvePt *curves, int axis, V3<T> du, Q4<T> qn, Q4<T> qp, float tN,
                       float tP, float prevDt, LinkState<T> &st, LinkOut<T> &out) {
    V3<T> oldPos2 = st.pos2, oldA1v = st.a1v, oldA2v = st.a2v;
    int flags = st.flags;

    Q4<T> angle1 = to_axis_x(qn, axis);
    Q4<T> angle2 = to_axis_x(qp, axis);
    Q4<T> totalRot = qconj(angle1);
    angle2 = qmul(totalRot, angle2);
    angle1.w = 1; angle1.x = angle1.y = angle1.z = 0;

    // pos2 (the pVPos position in the pVNeg frame, VX3_Link.cu:92-101) and
    // ext = pos2.x - restLength
    V3<T> pos2;
    T restLen, ext;
    if constexpr (std::is_same<T, double>::value) {
        restLen = rest_length<T>(lm, tN, tP);
        pos2 = qrot(totalRot, to_axis_x(du, axis));
        ext = pos2.x - restLen;
    } else {
        // R(L e_x + d) = R d + L (R e_x - e_x) + L e_x with R e_x - e_x written
        // for a unit quaternion, so the small extension never cancels
        const float L = (float)lm.nomSize;
        const float fN = thermal_factor(tN, lm.alphaN), fP = thermal_factor(tP, lm.alphaP);
        restLen = 0.5f * ((float)lm.nomN * fN + (float)lm.nomP * fP);
        const float lMinusRest = -0.5f * ((float)lm.nomN * (fN - 1.0f) + (float)lm.nomP * (fP - 1.0f));
        const Q4<T> R = totalRot;
        V3<T> rd = qrot(R, to_axis_x(du, axis));
        V3<T> rel = v3<T>(rd.x - 2.0f * L * (R.y * R.y + R.z * R.z), rd.y + 2.0f * L * (R.x * R.y + R.w * R.z),
                          rd.z + 2.0f * L * (R.x * R.z - R.w * R.y));
        pos2 = v3<T>(L + rel.x, rel.y, rel.z);
        ext = rel.x + lMinusRest;
    }

    float smallTurn = (float)qdiv(mabs(pos2.z) + mabs(pos2.y), pos2.x);
    float extendPerc;
    if constexpr (std::is_same<T, double>::value) extendPerc = (float)(mabs((T)1 - pos2.x / restLen));
    else extendPerc = mabs(ext * frcp(restLen));
    const bool wasSmall = (flags & LF_SMALL_ANGLE) != 0;
    bool smallAngle = wasSmall;
    if (!smallAngle && smallTurn < kSA_BOND_BEND_RAD && extendPerc < kSA_BOND_EXT_PERC) smallAngle = true;
    else if (smallAngle && (smallTurn > kHYSTERESIS_FACTOR * kSA_BOND_BEND_RAD ||
                            extendPerc > kHYSTERESIS_FACTOR * kSA_BOND_EXT_PERC))
        smallAngle = false;
    VX3MI_TEST_DECISION(DEC_SMALL_ANGLE, smallAngle);
    // a change of formulation invalidates the local velocity (:117-131)
    if (smallAngle != wasSmall) flags &= ~LF_LOCAL_VELOCITY_VALID;
    int exact = 0, acos1, acos2;
    if (smallAngle) {
        if constexpr (std::is_same<T, double>::value) pos2.x -= restLen;
        else pos2.x = ext;
    } else {
        angle1 = from_angle_to_posx(pos2, &exact);
        angle2 = qmul(angle1, angle2);
        if constexpr (std::is_same<T, double>::value) {
            pos2 = v3<T>(len(pos2) - restLen, 0, 0);
        } else {
            // |pos2| - restLen = (|pos2|^2 - restLen^2) / (|pos2| + restLen)
            T l = fsqrt(pos2.x * pos2.x + pos2.y * pos2.y + pos2.z * pos2.z);
            T num = ext * (pos2.x + restLen) + pos2.y * pos2.y + pos2.z * pos2.z;
            pos2 = v3<T>(num * frcp(l + restLen), 0, 0);
        }
    }
    V3<T> a1v = to_rotvec(angle1, &acos1);
    V3<T> a2v = to_rotvec(angle2, &acos2);
    flags = (flags & ~(LF_SMALL_ANGLE | LF_DECISIONS)) | (smallAngle ? LF_SMALL_ANGLE : 0) |
            (acos1 ? LF_ROTVEC_ACOS1 : 0) | (acos2 ? LF_ROTVEC_ACOS2 : 0) | (exact ? LF_POSX_EXACT : 0);
    st.pos2 = pos2;
    st.a1v = a1v;
    st.a2v = a2v;

    V3<T> dPos2 = (pos2 - oldPos2) * (T)0.5;
    V3<T> dA1 = (a1v - oldA1v) * (T)0.5;
    V3<T> dA2 = (a2v - oldA2v) * (T)0.5;

    float axialStrain = (float)qdiv(pos2.x, restLen);
    out.strain = axialStrain;
    float stress;
    if (lm.curve == 0) {  // linear material
        if (axialStrain > st.maxStrain) st.maxStrain = axialStrain;
        if (lm.epsFail != -1.0f && axialStrain > lm.epsFail) stress = 0.0f;
        else if (lm.nu == 0.0f) stress = lm.E * axialStrain;
        else stress = lm.eHat * ((1 - lm.nu) * axialStrain + lm.nu * lm.tStrainSum);
    } else {
        const CurvePt *c = curves + curve_first(lm);
        if (axialStrain > st.maxStrain) {  // new territory on the curve
            st.maxStrain = axialStrain;
            stress = curve_stress(lm, c, axialStrain, lm.tStrainSum, false);
        } else {  // backed off: linear with the offset of the largest plastic strain
            const float offset = lm.nu != 0.0f
                                     ? st.maxStrain - curve_stress(lm, c, st.maxStrain, 0.0f, false) / (lm.eHat * (1 - lm.nu))
                                     : st.maxStrain - curve_stress(lm, c, st.maxStrain, lm.tStrainSum, false) / lm.E;
            const float rel = axialStrain - offset;
            stress = lm.nu != 0.0f ? curve_stress(lm, c, rel, lm.tStrainSum, true) : lm.E * rel;
        }
    }

    if (lm.epsFail != -1.0f && st.maxStrain > lm.epsFail) {
        out.fN = out.fP = out.mN = out.mP = v3<T>(0, 0, 0);
        st.flags = flags;
        return;
    }
    const float b1 = lm.b1, b2 = lm.b2, b3 = lm.b3, a2 = lm.a2;
    V3<T> fN = v3<T>((T)(stress * lm.tArea), b1 * pos2.y - b2 * (a1v.z + a2v.z), b1 * pos2.z + b2 * (a1v.y + a2v.y));
    V3<T> fP = v3<T>(-fN.x, -fN.y, -fN.z);
    V3<T> mN = v3<T>(a2 * (a2v.x - a1v.x), -b2 * pos2.z - b3 * (2 * a1v.y + a2v.y), b2 * pos2.y - b3 * (2 * a1v.z + a2v.z));
    V3<T> mP = v3<T>(a2 * (a1v.x - a2v.x), -b2 * pos2.z - b3 * (a1v.y + 2 * a2v.y), b2 * pos2.y - b3 * (a1v.z + 2 * a2v.z));

    if (flags & LF_LOCAL_VELOCITY_VALID) {
        const float sqA1 = lm.sqA1, sqA2xIp = lm.sqA2xIp, sqB1 = lm.sqB1, sqB2xFMp = lm.sqB2xFMp, sqB3xIp = lm.sqB3xIp;
        V3<T> posCalc = v3<T>(sqA1 * dPos2.x, sqB1 * dPos2.y - sqB2xFMp * (dA1.z + dA2.z),
                              sqB1 * dPos2.z + sqB2xFMp * (dA1.y + dA2.y));
        float dmN, dmP;  // dampingMultiplier(), VX3_Voxel.h:206-208
        if constexpr (std::is_same<T, float>::value) {
            const float ipd = frcp(prevDt);
            dmN = lm.dmNumN * ipd;
            dmP = lm.dmNumP * ipd;
        } else {
            dmN = lm.dmNumN / prevDt;
            dmP = lm.dmNumP / prevDt;
        }
        fN = fN + posCalc * (T)dmN;
        fP = fP - posCalc * (T)dmP;
        T hN = (T)0.5 * dmN, hP = (T)0.5 * dmP;
        V3<T> tN = v3<T>(-sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * (2 * dA1.y + dA2.y),
                         -sqB2xFMp * dPos2.y + sqB3xIp * (2 * dA1.z + dA2.z));
        V3<T> tP = v3<T>(sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * (dA1.y + 2 * dA2.y),
                         -sqB2xFMp * dPos2.y + sqB3xIp * (dA1.z + 2 * dA2.z));
        mN = mN - tN * hN;
        mP = mP - tP * hP;
    } else {
        flags |= LF_LOCAL_VELOCITY_VALID;
    }
    if (!smallAngle) {
        fN = qrotinv(angle1, fN);
        mN = qrotinv(angle1, mN);
    }
    fP = qrotinv(angle2, fP);
    mP = qrotinv(angle2, mP);
    if constexpr (LINK_FRAME) {
        out.fN = fN;
        out.fP = fP;
        out.mN = mN;
        out.mP = mP;
    } else {
        out.fN = to_axis_original(fN, axis);
        out.fP = to_axis_original(fP, axis);
        out.mN = to_axis_original(mN, axis);
        out.mP = to_axis_original(mP, axis);
    }
    st.flags = flags;
}

// ---- voxel integrator -----------------------------------------------------------
// Floor penetration, VX3_Voxel.h:185-188: (float)(baseSizeAverage()/2 -
// nominalSize()/2 - pos.z).  uz is the z position field (absolute in double).
template <typename T> VX_HD float floor_penetration(const VoxelMat &vm, float temp, int iz, T uz) {
    float f = thermal_factor(temp, vm.alphaCTE);
    if constexpr (std::is_same<T, double>::value) {
        (void)iz;
        double b = vm.nomSize * (double)f;
        double bsa = (b + b + b) / 3.0f;
        return (float)(bsa / 2 - vm.nomSize / 2 - uz);
    } else {
        float L = (float)vm.nomSize;
        return (L * (f - 1.0f) * 0.5f - (float)iz * L) - uz;
    }
}

template <typename T> struct VoxelState {
    V3<T> u;     // position field (see kAbsolutePositions)
    Q4<T> q;     // orientation
    V3<T> p;     // linear momentum
    V3<T> h;     // angular momentum
    int flags;
};

// position field of a voxel fixed at lattice coordinate c plus translation t
// (pos = ix * nominalSize() + t, VX3_Voxel.cu:169, 225-233)
template <typename T> VX_HD T fixed_coordinate(int c, double nomSize, double t) {
    if constexpr (kAbsolutePositions<T>) return (T)(c * nomSize + t);
    else { (void)c; (void)nomSize; return (T)t; }
}

// VX3_Voxel::timeStep VX3_Voxel.cu:162-277 with force() :350-380, moment()
// :382-397, floorForce :399-426.  fLink / mLink are the link force / moment
// sums in the voxel frame (summed in link direction order X_POS..Z_NEG),
// contact is the collision force to subtract (VX3_Voxel.cu:372-373), field the
// force-field force (force_field).  lat is the packed lattice coordinate
// (ix | iy << 10 | iz << 20); exts the external table (read only when
// vm.ext >= 0).

// The force field of a voxel (VX3_Voxel.cu:177-183): formulas f (postfix,
// off / len per axis, len 0 = none) at the position entering the step and the
// time before it; hit is the robot's collisionCount; angle, targetCloseness
// and numClosePairs are sampled inside doTimeStep and formulas using them are
// rejected when read.
template <typename T>
VX_HD V3<T> force_field(const MtToken *f, const int *off, const int *len, double x, double y, double z, double t,
                        int numVoxels, int hit, double angle = 0.0, double closeness = 0.0, int pairs = 0) {
    const double vars[MV_COUNT] = {x, y, z, (double)hit, t, angle, closeness, (double)pairs, (double)numVoxels};
    V3<T> r = {0, 0, 0};
    if (len[0]) r.x = (T)formula_eval(f + off[0], len[0], vars);
    if (len[1]) r.y = (T)formula_eval(f + off[1], len[1], vars);
    if (len[2]) r.z = (T)formula_eval(f + off[2], len[2], vars);
    return r;
}
template <typename T>
VX_HD void voxel_step(const VoxelMat &vm, const VoxelExt *exts, float dtf, int lat, float temp, V3<T> fLink,
                      V3<T> mLink, V3<T> contact, V3<T> field, VoxelState<T> &s, const V3<T> *cilia = nullptr) {
    if (dtf == 0.0f) return;
    const T dt = (T)dtf;
    const int iz = lat >> 20;
    int flags = s.flags;
    const VoxelExt *ext = vm.ext >= 0 ? exts + vm.ext : nullptr;
    if (ext && ext->dof == DOF_ALL) {
        // all six fixed: pos = originalPosition() + translation, orient =
        // rotationQuat(), haltMotion() (VX3_Voxel.cu:167-172)
        s.u = V3<T>{fixed_coordinate<T>(lat & 1023, vm.nomSize, ext->t[0]),
                    fixed_coordinate<T>((lat >> 10) & 1023, vm.nomSize, ext->t[1]),
                    fixed_coordinate<T>(iz, vm.nomSize, ext->t[2])};
        s.q = Q4<T>{(T)ext->q[0], (T)ext->q[1], (T)ext->q[2], (T)ext->q[3]};
        s.p = s.h = V3<T>{0, 0, 0};
        return;
    }

    V3<T> tf = qrot(s.q, fLink);
    if (ext) tf = tf + V3<T>{(T)ext->f[0], (T)ext->f[1], (T)ext->f[2]};  // external forces (:363-364)
    V3<T> vel = s.p * (T)vm.massInv;
    tf = tf - vel * (T)vm.gDampT;
    tf.z += vm.gravF;
    tf = tf - contact;
    if (cilia) tf = tf + *cilia;  // CiliaForce * mat->Cilia (VX3_Voxel.cu:376-377)
    // force field (VX3_Voxel.cu:177-183); zeros without one, as the reference's
    // constant-zero default trees (VX3_ForceField.h:11-20) add +0.0
    tf.x += field.x; tf.y += field.y; tf.z += field.z;

    V3<T> cur = tf;
    // floorForce (:399-426) and the friction test of timeStep (:198-209) both
    // act while the voxel touches the floor at its position entering the step
    const float pen = floor_penetration<T>(vm, temp, iz, s.u.z);
    bool onFloor = (flags & VF_FLOOR_ENABLED) && pen >= 0;
    VX3MI_TEST_DECISION(DEC_FLOOR, onFloor);
    if (onFloor) {
        float normalForce = vm.penK * pen;
        cur.z += normalForce - vm.cDampT * vel.z;
        if (flags & VF_FLOOR_STATIC_FRICTION) {
            float sfs = (float)(cur.x * cur.x + cur.y * cur.y);
            float ffs = (vm.muS * normalForce) * (vm.muS * normalForce);
            bool slip = sfs > ffs;
            VX3MI_TEST_DECISION(DEC_STATIC_BREAK, slip);
            if (slip) flags &= ~VF_FLOOR_STATIC_FRICTION;
        } else {
            V3<T> hv = v3<T>(vel.x, vel.y, 0);
            T l = len(hv);
            V3<T> hn = hv;
            if (l > 0) {
                if constexpr (std::is_same<T, float>::value) hn = hv * frcp(l);
                else hn = v3<T>(hv.x / l, hv.y / l, hv.z / l);
            }
            cur = cur - hn * (T)(vm.muK * normalForce);
        }
    } else {
        flags &= ~VF_FLOOR_STATIC_FRICTION;
    }
    V3<T> fric = cur - tf;
    V3<T> p = s.p + cur * dt;
    V3<T> tr = p * (dt * (T)vm.massInv);
    if (onFloor) {
        T work = fric.x * tr.x + fric.y * tr.y;
        T hKe = (T)0.5 * (T)vm.massInv * (p.x * p.x + p.y * p.y);
        bool stick = hKe + work <= 0;
        VX3MI_TEST_DECISION(DEC_STATIC_MAKE, stick);
        if (stick) flags |= VF_FLOOR_STATIC_FRICTION;
        if (flags & VF_FLOOR_STATIC_FRICTION) {
            p.x = p.y = 0;
            tr.x = tr.y = 0;
        }
    }
    s.u = s.u + tr;
    s.p = p;

    V3<T> tm = qrot(s.q, mLink);
    if (ext) tm = tm + V3<T>{(T)ext->m[0], (T)ext->m[1], (T)ext->m[2]};  // external moments (:393-394)
    tm = tm - (s.h * (T)vm.inertiaInv) * (T)vm.gDampR;
    V3<T> h = s.h + tm * dt;
    s.q = qmul(from_rotvec(h * (dt * (T)vm.inertiaInv)), s.q);
    if (ext && ext->dof) {
        // fixed degrees of freedom (VX3_Voxel.cu:222-256)
        const int dof = ext->dof;
        if (dof & DOF_X) { s.u.x = fixed_coordinate<T>(lat & 1023, vm.nomSize, ext->t[0]); s.p.x = 0; }
        if (dof & DOF_Y) { s.u.y = fixed_coordinate<T>((lat >> 10) & 1023, vm.nomSize, ext->t[1]); s.p.y = 0; }
        if (dof & DOF_Z) { s.u.z = fixed_coordinate<T>(iz, vm.nomSize, ext->t[2]); s.p.z = 0; }
        if (dof & (DOF_TX | DOF_TY | DOF_TZ)) {
            if ((dof & (DOF_TX | DOF_TY | DOF_TZ)) == (DOF_TX | DOF_TY | DOF_TZ)) {
                s.q = Q4<T>{(T)ext->q[0], (T)ext->q[1], (T)ext->q[2], (T)ext->q[3]};
                h = V3<T>{0, 0, 0};
            } else {  // partial rotation fixes: through the rotation vector
                V3<T> rv = to_rotvec(s.q);
                if (dof & DOF_TX) { rv.x = 0; h.x = 0; }
                if (dof & DOF_TY) { rv.y = 0; h.y = 0; }
                if (dof & DOF_TZ) { rv.z = 0; h.z = 0; }
                s.q = from_rotvec(rv);
            }
        }
    }
    if constexpr (std::is_same<T, float>::value) {
        // single precision only: keep the orientation a unit quaternion (the
        // FP64 reference never renormalises, VX3_Voxel.cu:221, and drifts by
        // ~1e-16 per step; float would drift by ~1e-7 per step)
        float n = frsqrt(s.q.w * s.q.w + s.q.x * s.q.x + s.q.y * s.q.y + s.q.z * s.q.z);
        s.q.w *= n; s.q.x *= n; s.q.y *= n; s.q.z *= n;
    }
    // VX3-only: no rotation while held by static floor friction (VX3_Voxel.cu:259-264)
    bool onFloorAfter = (flags & VF_FLOOR_ENABLED) && floor_penetration<T>(vm, temp, iz, s.u.z) >= 0;
    VX3MI_TEST_DECISION(DEC_FLOOR_AFTER, onFloorAfter);
    if (onFloorAfter && (flags & VF_FLOOR_STATIC_FRICTION)) h = v3<T>(0, 0, 0);
    s.h = h;
    s.flags = flags;
}


// ---- collisions ---------------------------------------------------------------
// baseSizeAverage, VX3_Voxel.h:102-105
template <typename T> VX_HD T base_size_average(double nomSize, float temp, float alpha) {
    const float f = thermal_factor(temp, alpha);
    if constexpr (std::is_same<T, double>::value) {
        const double b = nomSize * (double)f;
        return (b + b + b) / 3.0f;
    } else {
        return (float)nomSize * f;
    }
}

// Contact force on voxel i from surface voxel j: handle_collision_attachment
// (VX3_VoxelyzeKernel.cu:682-729) with voxel1 = i, voxel2 = j and
// VX3_Collision::updateContactForce (VX3_Collision.cu:17-31); returns false
// when the pair is outside the envelope.  diff = pos_i - pos_j.  The caller
// has already excluded directly linked pairs and pairs of two fixed voxels.
// The force computed with i as voxel1 is exactly the negation of the one the
// reference adds to voxel2 when i is voxel2, so each voxel can sum its own
// contacts.
template <typename T>
VX_HD bool in_contact_envelope(V3<T> diff, T bsaI, T bsaJ) {
    const T watch = (bsaI + bsaJ) * (T)kCOLLISION_ENVELOPE_RADIUS;
    if (diff.x > watch || diff.x < -watch) return false;
    if (diff.y > watch || diff.y < -watch) return false;
    if (diff.z > watch || diff.z < -watch) return false;
    return len(diff) < watch;  // RelDist > 0 (equality contributes a zero force)
}

template <typename T>
VX_HD V3<T> contact_force(V3<T> diff, T bsaI, T bsaJ, V3<T> velI, V3<T> velJ, float ipkI, float ipkJ, float cdI,
                          float cdJ) {
    const T penStiff = (T)(2.0f / (ipkI + ipkJ));  // 2 / (1 / pkI + 1 / pkJ), the reciprocals from the materials
    const T dampC = (T)(0.5f * (cdI + cdJ));
    const V3<T> offset = v3<T>(-diff.x, -diff.y, -diff.z);
    const T nomDist = (bsaI + bsaJ) * (T)kCOLLISION_ENVELOPE_RADIUS;
    const T l = len(offset);
    const T relDist = nomDist - l;
    bool touch = relDist > 0;
    VX3MI_TEST_DECISION(DEC_CONTACT, touch);
    if (!touch) return v3<T>(0, 0, 0);
    const V3<T> unit = l > 0 ? v3<T>(offset.x / l, offset.y / l, offset.z / l) : offset;
    const T rv = (velI.x * unit.x + velI.y * unit.y + velI.z * unit.z) - (velJ.x * unit.x + velJ.y * unit.y + velJ.z * unit.z);
    return unit * (penStiff * relDist + dampC * rv);
}

}  // namespace vx3mi
