/*
 * vx3_oracle.c -- FP64 CPU restatement of VX3_VoxelyzeKernel::doTimeStep.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function cites the reference lines it restates.  Types follow the
 * reference exactly: state is double, material constants, temperature, stress
 * and dt are float, and every mixed float/double expression is written with the
 * same operand order and promotions as the reference so that the rounding is
 * the reference's.  Build with -ffp-contract=off (no FMA contraction).
 */
#include "vx3_oracle.h"
#include "vx3mi_math.h" /* sin/cos/acos rounding identically on host and device */

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } v3;
typedef struct { double w, x, y, z; } q4;

/* src/Utils/VX3_Quat3D.h:30-56 */
#define Q_PI 3.14159265358979
#define Q_DBL_EPSILONx24 5.328e-15
#define Q_DISCARD_ANGLE_RAD 1e-7
#define Q_SMALL_ANGLE_RAD 1.732e-2
#define Q_SLTHRESH_ACOS2SQRT 2.4e-3
/* src/old/types.h:57-59 */
static const float HYSTERESIS_FACTOR = 1.2f;
static const float SA_BOND_BEND_RAD = 0.05f;
static const float SA_BOND_EXT_PERC = 0.50f;
/* src/VX3/VX3_Collision.h:6 */
#define COLLISION_ENVELOPE_RADIUS 0.625

/* voxel flag bits, src/old/types.h:34-38 */
#define VF_FLOOR_ENABLED (1 << 2)
#define VF_FLOOR_STATIC_FRICTION (1 << 3)
#define OM_MAX_POINTS VXO_MAX_CURVE_POINTS  /* stress-strain points of a material or link */

/* link flag bits: src/old/types.h:6-8 (LOCAL_VELOCITY_VALID) + smallAngle */
#define LF_LOCAL_VELOCITY_VALID 1
#define LF_SMALL_ANGLE 2
/* branches the last update took (recorded for comparisons, read by nothing):
 * ToRotationVector of angle1 / angle2 by acos (above the 2.4e-3 threshold),
 * FromAngleToPosX by its exact form (above 1.732e-2) */
#define LF_ROTVEC_ACOS1 4
#define LF_ROTVEC_ACOS2 8
#define LF_POSX_EXACT 16
#define LF_DECISIONS (LF_ROTVEC_ACOS1 | LF_ROTVEC_ACOS2 | LF_POSX_EXACT)

typedef struct {
    double nomSize;
    float E, nu, rho, alphaCTE, muS, muK, zetaI, zetaG, zetaC, eHat, gravMult;
    float mass, massInv, sqrtMass, inertia, inertiaInv, c2xSqMxExS, c2xSqIxExSxSxS;
    float sigmaFail, epsFail;
    int fixed, isMeasured;
    double thermalOnAfter;
    double cilia, ciliaOnAfter, removeAfter;
    int isPaceMaker, isTarget, sticky;
    double paceMakerPeriod, signalDecay, signalDelay, inactivePeriod;
    /* stress-strain model: CVX_Material linear / strainData / stressData */
    int linear, npts;
    float sd[OM_MAX_POINTS], td[OM_MAX_POINTS];
} vmat;

typedef struct {
    double nomSize;
    float E, nu, eHat, epsFail;
    float a1, a2, b1, b2, b3, sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;
    int linear, npts;
    float sd[OM_MAX_POINTS], td[OM_MAX_POINTS];
} lmat;

struct vxo_model {
    int nv, nl, ns;
    double voxSize;
    /* voxels */
    int *ix, *iy, *iz;
    int *vmat_of;          /* index into vm */
    int (*vlinks)[6];      /* link index per direction X_POS..Z_NEG, -1 none */
    v3 *pos, *linMom, *angMom, *contact;
    q4 *orient;
    float *temp;
    float *prevDt;
    int *vflags;
    double *phase;
    /* externals (VX3_External.h:27-31): fixed degrees of freedom (bits X, Y, Z
     * translation, X, Y, Z rotation; -1 = the voxel has no external), prescribed
     * translation and rotation (and its quaternion), force and moment */
    int ff_n[3];           /* force field tokens per axis (0: none) */
    int *ff_op[3];
    double *ff_val[3];
    int *ext_dof;
    v3 *ext_t, *ext_r;
    q4 *ext_q;
    v3 *ext_f, *ext_m;
    int *surface;          /* ns voxel indices, voxel order */
    int *is_surface;
    /* cilia (BaseCiliaForce / ShiftCiliaForce per voxel) and removal: the step
     * (1-based) at whose end a voxel or link was removed, 0 while present */
    v3 *ciliaBase, *ciliaShift;
    int *vremoved, *lremoved;
    long steps;
    /* signals (VX3_Voxel.h:304-309): localSignal, localSignaldt, d_signal,
     * inactiveUntil, packmakerNextPulse; the signal a voxel sent this step */
    double *sLocal, *sLocalDt, *sValue, *sActive, *sInactive, *sNextPulse;
    double *outValue, *outTime;
    long collisionCount;
    /* the pairs in contact in the last step's collision phase (voxel1, voxel2) */
    int *pairs, npairs, pairCap;
    /* links */
    int *lneg, *lpos, *laxis;
    int lcap;              /* capacity of the per-link arrays (links attach at run time) */
    int *newLink;          /* isNewLink: updates left whose forces are scaled by 0.01 */
    lmat *lm;
    v3 *pos2, *a1v, *a2v;
    float *strain, *maxStrain, *strainOffset, *tArea, *tStrainSum;
    double *restLen;
    int *lflags;
    v3 *fNeg, *fPos, *mNeg, *mPos;
    /* materials */
    int nvm;
    vmat *vm;
    /* settings */
    double dtFrac, optimalDt, time;
    int enableCollision, enableExpansion, varyTemp;
    int enableCilia, secondary, enableSignals, enableDetach;
    double tempAmplitude, tempPeriod;
    /* attach: EnableAttach, SafetyGuard, per-voxel enableAttach (VX3_Voxel.h:302,
     * true until the first voxel update), AttachCondition Condition_0..4 */
    int enableAttach, safetyGuard;
    int *vattach;
    /* CycleStep samples (VX3_VoxelyzeKernel.cu:314-334): centre of mass now and
     * at the last two samples, recentAngle, targetCloseness, numClosePairs */
    double maxDistPairs;
    v3 comCur, comHist[2];
    double recentAngle, closeness;
    int closePairs, angleSamples;
    int ac_n[5];
    int *ac_op[5];
    double *ac_val[5];
};

/* ------------------------------------------------------------------ math */
static inline v3 v3_(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return v3_(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return v3_(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vscale(v3 a, double s) { return v3_(a.x * s, a.y * s, a.z * s); }
static inline double vlen(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* VX3_Quat3D.h:217-222 */
static inline q4 qmul(q4 a, q4 f) {
    q4 r;
    r.w = a.w * f.w - a.x * f.x - a.y * f.y - a.z * f.z;
    r.x = a.w * f.x + a.x * f.w + a.y * f.z - a.z * f.y;
    r.y = a.w * f.y - a.x * f.z + a.y * f.w + a.z * f.x;
    r.z = a.w * f.z + a.x * f.y - a.y * f.x + a.z * f.w;
    return r;
}
static inline q4 qconj(q4 a) { q4 r = {a.w, -a.x, -a.y, -a.z}; return r; }
/* VX3_Quat3D.h:439-447 */
static inline v3 qrot(q4 q, v3 f) {
    double fx = f.x, fy = f.y, fz = f.z;
    double tw = fx * q.x + fy * q.y + fz * q.z;
    double tx = fx * q.w - fy * q.z + fz * q.y;
    double ty = fx * q.z + fy * q.w - fz * q.x;
    double tz = -fx * q.y + fy * q.x + fz * q.w;
    return v3_(q.w * tx + q.x * tw + q.y * tz - q.z * ty, q.w * ty - q.x * tz + q.y * tw + q.z * tx,
               q.w * tz + q.x * ty - q.y * tx + q.z * tw);
}
/* VX3_Quat3D.h:465-474 */
static inline v3 qrotinv(q4 q, v3 f) {
    double fx = f.x, fy = f.y, fz = f.z;
    double tw = q.x * fx + q.y * fy + q.z * fz;
    double tx = q.w * fx - q.y * fz + q.z * fy;
    double ty = q.w * fy + q.x * fz - q.z * fx;
    double tz = q.w * fz - q.x * fy + q.y * fx;
    return v3_(tw * q.x + tx * q.w + ty * q.z - tz * q.y, tw * q.y - tx * q.z + ty * q.w + tz * q.x,
               tw * q.z + tx * q.y - ty * q.x + tz * q.w);
}
/* VX3_Quat3D.h:347-359 */
static inline v3 q_to_rotvec(q4 q, int *acos_branch) {
    if (acos_branch) *acos_branch = 0;
    if (q.w >= 1.0 || q.w <= -1.0) return v3_(0, 0, 0);
    double sl = 1.0 - q.w * q.w;
    if (sl < Q_SLTHRESH_ACOS2SQRT) {
        double s = sqrt((2 - 2 * q.w) / sl);
        return v3_(q.x * 2.0 * s, q.y * 2.0 * s, q.z * 2.0 * s);
    }
    if (acos_branch) *acos_branch = 1;
    double a = vxm_acos(q.w), d = sqrt(sl);
    return v3_(q.x * 2.0 * a / d, q.y * 2.0 * a / d, q.z * 2.0 * a / d);
}
/* VX3_Quat3D.h:363-378 */
static inline q4 q_from_rotvec(v3 in) {
    v3 th = v3_(in.x / 2, in.y / 2, in.z / 2);
    double s, m2 = th.x * th.x + th.y * th.y + th.z * th.z;
    q4 r;
    if (m2 * m2 < Q_DBL_EPSILONx24) {
        r.w = 1.0 - 0.5 * m2;
        s = 1.0 - m2 / 6.0;
    } else {
        double m = sqrt(m2);
        r.w = vxm_cos(m);
        s = vxm_sin(m) / m;
    }
    r.x = th.x * s;
    r.y = th.y * s;
    r.z = th.z * s;
    return r;
}
/* VX3_Quat3D.h:384-432; q is the identity on entry (VX3_Link.cu:103,120) */
static inline q4 q_from_angle_to_posx(v3 rf, int *exact) {
    q4 q = {1, 0, 0, 0};
    if (exact) *exact = 0;
    if (rf.x == 0 && rf.y == 0 && rf.z == 0) return q;
    double yx = rf.y / rf.x, zx = rf.z / rf.x;
    if (yx < Q_SMALL_ANGLE_RAD && yx > -Q_SMALL_ANGLE_RAD && zx < Q_SMALL_ANGLE_RAD && zx > -Q_SMALL_ANGLE_RAD) {
        q.x = 0;
        q.y = 0.5 * zx;
        q.z = -0.5 * yx;
        q.w = 1 + 0.5 * (-q.y * q.y - q.z * q.z);
        return q;
    }
    if (exact) *exact = 1;
    v3 n = rf;
    double l = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    if (l > 0) {
        double li = 1.0 / l;
        n.x *= li;
        n.y *= li;
        n.z *= li;
    }
    double theta = vxm_acos(n.x);
    if (theta > Q_PI - Q_DISCARD_ANGLE_RAD) {
        q.w = 0; q.x = 0; q.y = 1; q.z = 0;
        return q;
    }
    double ami = 1.0 / sqrt(n.z * n.z + n.y * n.y);
    double a = 0.5 * theta, s = vxm_sin(a);
    q.w = vxm_cos(a);
    q.x = 0;
    q.y = n.z * ami * s;
    q.z = -n.y * ami * s;
    return q;
}
/* VX3_Link.h:134-153 (by value) */
static inline v3 to_axis_x(v3 v, int axis) {
    if (axis == 1) return v3_(v.y, -v.x, v.z);
    if (axis == 2) return v3_(v.z, v.y, -v.x);
    return v;
}
static inline q4 to_axis_x_q(q4 q, int axis) {
    q4 r = q;
    if (axis == 1) { r.x = q.y; r.y = -q.x; }
    if (axis == 2) { r.x = q.z; r.z = -q.x; }
    return r;
}
/* VX3_Link.h:156-173 (in place) */
static inline v3 to_axis_original(v3 v, int axis) {
    if (axis == 1) return v3_(-v.y, v.x, v.z);
    if (axis == 2) return v3_(-v.z, v.y, v.x);
    return v;
}

/* ------------------------------------------------------------- materials */
/* CVX_Material::updateDerived src/old/VX_Material.cpp:542-548 and
 * CVX_MaterialVoxel::updateDerived src/old/VX_MaterialVoxel.cpp:57-79 */
static void vmat_derive(vmat *m) {
    m->eHat = m->E / ((1 - 2 * m->nu) * (1 + m->nu));
    double ns = m->nomSize;
    double volume = ns * ns * ns;
    m->mass = (float)(volume * m->rho);
    m->inertia = (float)(m->mass * ns * ns / 6.0f);
    if (volume == 0 || m->mass == 0 || m->inertia == 0) {
        m->massInv = m->sqrtMass = m->inertiaInv = m->c2xSqMxExS = m->c2xSqIxExSxSxS = 0.0f;
        return;
    }
    m->massInv = 1.0f / m->mass;
    m->sqrtMass = (float)sqrt((double)m->mass);
    m->inertiaInv = 1.0f / m->inertia;
    m->c2xSqMxExS = (float)(2.0f * sqrt((double)(m->mass * m->E) * ns));
    m->c2xSqIxExSxSxS = (float)(2.0f * sqrt((double)(m->inertia * m->E) * ns * ns * ns));
}

/* CVX_Material::setPoissonsRatio src/old/VX_Material.cpp:488-494 */
static float clamp_nu(float nu) {
    if (nu < 0) nu = 0;
    if (nu >= 0.5) nu = (float)(0.5 - FLT_EPSILON * 2);
    return nu;
}

/* CVX_MaterialLink::updateAll/updateDerived src/old/VX_MaterialLink.cpp:45-141 */
/* CVX_Material::setModelLinear (src/old/VX_Material.cpp:372-401), arguments valid */
static void om_set_linear(vmat *m, float E, float fail) {
    float tf = fail == -1 ? 1000000 : fail;
    float tfs = tf / E;
    m->npts = 2;
    m->sd[0] = 0; m->sd[1] = tfs;
    m->td[0] = 0; m->td[1] = tf;
    m->linear = 1;
    m->E = E;
    m->sigmaFail = fail;
    m->epsFail = fail == -1 ? -1 : tfs;
}
/* CVX_Material::setModelBilinear (src/old/VX_Material.cpp:406-451); 0 when it
 * rejects the arguments (the model is then left as it was) */
static int om_set_bilinear(vmat *m, float E, float plastic, float yield, float fail) {
    if (E <= 0 || plastic <= 0 || plastic >= E || yield <= 0) return 0;
    if (fail != -1.0f && fail <= yield) return 0;
    float yieldStrain = yield / E;
    float tf = fail == -1 ? 3 * yield : fail;
    float tM = plastic;
    float tB = yield - tM * yieldStrain;
    float tfs = (tf - tB) / tM;
    m->npts = 3;
    m->sd[0] = 0; m->sd[1] = yieldStrain; m->sd[2] = tfs;
    m->td[0] = 0; m->td[1] = yield; m->td[2] = tf;
    m->linear = 0;
    m->E = E;
    m->sigmaFail = fail;
    m->epsFail = fail == -1.0f ? -1.0f : tfs;
    return 1;
}
/* CVX_Material::setModel (src/old/VX_Material.cpp:300-367) on float data: a
 * leading (0, 0) is dropped; it rejects no points, a first point not above
 * zero and strains that do not increase (0: the model is then left as it
 * was).  Decreasing stresses only leave a message, and its slope test compares
 * with 0/0 and never fails.  E is the first slope, the last point is the
 * failure point, one point is a linear material. */
static int om_set_data(vmat *m, int n, const float *strain, const float *stress) {
    if (n > 0 && strain[0] == 0 && stress[0] == 0) { strain++; stress++; n--; }
    if (n <= 0 || n + 1 > OM_MAX_POINTS) return 0;
    if (strain[0] <= 0 || stress[0] <= 0) return 0;
    float sweep = 0.0f;
    for (int i = 0; i < n; i++) {
        if (strain[i] <= sweep) return 0;
        sweep = strain[i];
    }
    m->npts = n + 1;
    m->sd[0] = 0; m->td[0] = 0;
    for (int i = 0; i < n; i++) { m->sd[i + 1] = strain[i]; m->td[i + 1] = stress[i]; }
    m->E = m->td[1] / m->sd[1];
    m->sigmaFail = m->td[n];
    m->epsFail = m->sd[n];
    m->linear = n == 1;
    return 1;
}
/* CVX_Material::modulus (src/old/VX_Material.cpp:231-243) */
static float om_modulus(const vmat *m, float strain) {
    if (m->epsFail != -1.0f && strain > m->epsFail) return 0.0f;
    if (strain <= m->sd[1] || m->linear) return m->E;
    for (int i = 2; i < m->npts; i++)
        if (strain <= m->sd[i] || i == m->npts - 1) return (m->td[i] - m->td[i - 1]) / (m->sd[i] - m->sd[i - 1]);
    return 0.0f;
}
/* CVX_Material::strain(stress) (src/old/VX_Material.cpp:214-228) of a curve */
static float om_strain_of(int linear, int n, const float *sd, const float *td, float E, float stress) {
    if (stress <= td[1] || linear) return stress / E;
    for (int i = 2; i < n; i++)
        if (stress <= td[i] || i == n - 1) {
            float perc = (stress - td[i - 1]) / (td[i] - td[i - 1]);
            return sd[i - 1] + perc * (sd[i] - sd[i - 1]);
        }
    return 0.0f;
}
/* VX3_Material::stress (src/VX3/VX3_Material.cu:90-124) of a link material */
static float om_stress(const lmat *m, float strain, float tss, int forceLinear) {
    if (m->epsFail != -1.0f && strain > m->epsFail) return 0.0f;
    if (strain <= m->sd[1] || m->linear || forceLinear) {
        if (m->nu == 0.0f) return m->E * strain;
        return m->eHat * ((1 - m->nu) * strain + m->nu * tss);
    }
    for (int i = 2; i < m->npts; i++) {
        if (strain <= m->sd[i] || i == m->npts - 1) {
            float perc = (strain - m->sd[i - 1]) / (m->sd[i] - m->sd[i - 1]);
            float basic = m->td[i - 1] + perc * (m->td[i] - m->td[i - 1]);
            if (m->nu == 0.0f) return basic;
            float modulus = (m->td[i] - m->td[i - 1]) / (m->sd[i] - m->sd[i - 1]);
            float modulusHat = modulus / ((1 - 2 * m->nu) * (1 + m->nu));
            float effStrain = basic / modulus;
            float effTss = tss * (effStrain / strain);
            return modulusHat * ((1 - m->nu) * effStrain + m->nu * effTss);
        }
    }
    return 0.0f;
}

/* CVX_MaterialLink::updateAll (src/old/VX_MaterialLink.cpp:45-125) */
static void lmat_combine(lmat *l, const vmat *a, const vmat *b) {
    l->nomSize = 0.5 * (a->nomSize + b->nomSize);
    float f1 = a->sigmaFail, f2 = b->sigmaFail, sf;
    if (f1 == -1.0f) sf = f2;
    else if (f2 == -1.0f) sf = f1;
    else sf = f1 < f2 ? f1 : f2;
    if (a->linear && b->linear) {
        l->E = 2.0f * a->E * b->E / (a->E + b->E);
        l->epsFail = (sf == -1.0f) ? -1.0f : sf / l->E;  /* setModelLinear VX_Material.cpp */
        l->linear = 1;
        l->npts = 2;
        l->sd[0] = 0; l->td[0] = 0;
        l->sd[1] = (sf == -1.0f ? 1000000 : sf) / l->E;
        l->td[1] = sf == -1.0f ? 1000000 : sf;
    } else {
        /* series springs over the union of both materials' strain points */
        float st[2 * OM_MAX_POINTS], ss[2 * OM_MAX_POINTS];
        int n = 1, i1 = 1, i2 = 1;
        st[0] = ss[0] = 0;
        while (i1 < a->npts && i2 < b->npts) {
            float e = FLT_MAX;
            if (i1 < a->npts) e = a->sd[i1];
            if (i2 < b->npts && b->sd[i2] < e) e = b->sd[i2];
            if (e == a->sd[i1]) i1++;
            if (e == b->sd[i2]) i2++;
            float m1 = om_modulus(a, e - FLT_EPSILON), m2 = om_modulus(b, e - FLT_EPSILON);
            float mm = 2.0f * m1 * m2 / (m1 + m2);
            ss[n] = ss[n - 1] + mm * (e - st[n - 1]);
            st[n] = e;
            n++;
        }
        /* setModel (VX_Material.cpp:300-370): E from the first segment */
        l->npts = n < OM_MAX_POINTS ? n : OM_MAX_POINTS;
        for (int k = 0; k < l->npts; k++) { l->sd[k] = st[k]; l->td[k] = ss[k]; }
        l->linear = (n - 1 == 1);
        l->E = l->td[1] / l->sd[1];
        l->epsFail = sf == -1.0f ? -1.0f : om_strain_of(l->linear, l->npts, l->sd, l->td, l->E, sf);
    }
    if (a->nu == 0 && b->nu == 0) l->nu = 0;
    else {
        float tmpEHat = 2 * a->eHat * b->eHat / (a->eHat + b->eHat);
        float tmpE = l->E;
        float c2 = (tmpEHat - tmpE) / (2 * tmpEHat) + 0.0625;
        l->nu = (float)(sqrtf(c2) - 0.25);
    }
    l->eHat = l->E / ((1 - 2 * l->nu) * (1 + l->nu));
    float L = (float)l->nomSize;
    l->a1 = l->E * L;
    l->a2 = l->E * L * L * L / (12.0f * (1 + l->nu));
    l->b1 = l->E * L;
    l->b2 = l->E * L * L / 2.0f;
    l->b3 = l->E * L * L * L / 6.0f;
    l->sqA1 = (float)sqrt((double)l->a1);
    l->sqA2xIp = (float)sqrt((double)(l->a2 * L * L / 6.0f));
    l->sqB1 = (float)sqrt((double)l->b1);
    l->sqB2xFMp = (float)sqrt((double)(l->b2 * L / 2.0f));
    l->sqB3xIp = (float)sqrt((double)(l->b3 * L * L / 6.0f));
}

/* ----------------------------------------------------------------- build */
static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) abort();
    return p;
}

/* baseSize(axis): VX3_Voxel.h:96-100.  (1 + tempe*alphaCTE) is float math. */
static inline double base_size(const vxo_model *m, int v) {
    const vmat *mt = &m->vm[m->vmat_of[v]];
    float f = 1 + m->temp[v] * mt->alphaCTE;
    return mt->nomSize * (double)f;
}
/* baseSizeAverage: VX3_Voxel.h:102-105 */
static inline double base_size_average(const vxo_model *m, int v) {
    double b = base_size(m, v);
    return (b + b + b) / 3.0f;
}
/* floorPenetration: VX3_Voxel.h:185-188 */
static inline float floor_penetration(const vxo_model *m, int v) {
    const vmat *mt = &m->vm[m->vmat_of[v]];
    return (float)(base_size_average(m, v) / 2 - mt->nomSize / 2 - m->pos[v].z);
}
/* updateRestLength: VX3_Link.cu:80-83 */
static inline void update_rest_length(vxo_model *m, int l) {
    m->restLen[l] = 0.5 * (base_size(m, m->lneg[l]) + base_size(m, m->lpos[l]));
}

/* recommendedTimeStep: VX3_VoxelyzeKernel.cu:184-217 (axialStiffness with
 * nu == 0 returns _a1, VX3_Link.cu:269-278; with nu != 0 the frozen
 * transverse area and the current rest length are used) */
double vxo_recommended_dt(const vxo_model *m) {
    double maxFreq2 = 0.0f;
    for (int i = 0; i < m->nl; i++) {
        const lmat *lm = &m->lm[i];
        double m1 = m->vm[m->vmat_of[m->lneg[i]]].mass, m2 = m->vm[m->vmat_of[m->lpos[i]]].mass;
        float k;
        if (lm->nu == 0.0f) k = lm->a1;
        else k = (float)(lm->eHat * m->tArea[i] / ((m->strain[i] + 1) * m->restLen[i]));
        double f2 = k / (m1 < m2 ? m1 : m2);
        if (f2 > maxFreq2) maxFreq2 = f2;
    }
    if (maxFreq2 <= 0.0f) {
        for (int i = 0; i < m->nv; i++) {
            const vmat *mt = &m->vm[m->vmat_of[i]];
            double f2 = mt->E * mt->nomSize / mt->mass;
            if (f2 > maxFreq2) maxFreq2 = f2;
        }
    }
    if (maxFreq2 <= 0.0f) return 0.0f;
    return 1.0f / (6.283185f * sqrt(maxFreq2));
}

float vxo_step_dt(vxo_model *m) {
    /* VX3_VoxelyzeKernel.cu:244-253 */
    if (!m->optimalDt) m->optimalDt = vxo_recommended_dt(m);
    if (m->optimalDt < 1e-10) m->optimalDt = 1e-10;
    return (float)(m->dtFrac * m->optimalDt);
}

/*
 * Model construction follows CVX_Sim::Import (src/VXA/VX_Sim.cpp:59-153):
 * palette materials -> CopyMat, damping, voxel size, voxels in lattice order
 * (x fastest; CVoxelyze::addVoxel src/old/Voxelyze.cpp:439-461), links created
 * when the second voxel of a pair is added in the X_POS..Z_NEG loop
 * (addLink :508-539; CVX_Link ctor src/old/VX_Link.cpp:21-48 puts the lower
 * lattice coordinate at pVNeg), initial temperatures (UpdateMatTemps
 * VX_Sim.cpp:467-483 with CVX_Environment::UpdateCurTemp
 * src/old/VX_Environment.cpp:327-346), volume effects (VX_Sim.cpp:445-458).
 */
/* surface voxels: fewer than 6 links (VX3_Voxel::updateSurface
 * VX3_Voxel.cu:515-524), removed ones left out of the list
 * (regenerateSurfaceVoxels VX3_VoxelyzeKernel.cu:495-513) */
static void regenerate_surface(vxo_model *m) {
    int ns = 0;
    for (int v = 0; v < m->nv; v++) {
        int c = 0;
        for (int d = 0; d < 6; d++) c += m->vlinks[v][d] >= 0;
        m->is_surface[v] = c < 6;
        if (c < 6 && !m->vremoved[v]) m->surface[ns++] = v;
    }
    m->ns = ns;
}

vxo_model *vxo_build(const vxo_spec *sp) {
    vxo_model *m = (vxo_model *)xcalloc(1, sizeof(vxo_model));
    int nx = sp->nx, ny = sp->ny, nz = sp->nz, ncell = nx * ny * nz;
    m->voxSize = sp->lattice_dim;
    m->dtFrac = sp->dt_frac;
    m->enableCollision = sp->enable_collision;
    m->enableExpansion = sp->enable_expansion;
    m->varyTemp = sp->vary_temp;
    m->tempAmplitude = sp->temp_amplitude;
    m->tempPeriod = sp->temp_period;
    m->enableCilia = sp->enable_cilia;
    m->secondary = sp->secondary_experiment;
    m->enableSignals = sp->enable_signals;
    m->enableDetach = sp->enable_detach;
    m->enableAttach = sp->enable_attach;
    m->safetyGuard = sp->safety_guard;
    m->maxDistPairs = sp->max_dist_pairs;

    int nv = 0;
    for (int i = 0; i < ncell; i++)
        if (sp->cells[i] > 0 && sp->cells[i] <= sp->nmat) nv++;
    m->nv = nv;

    /* gravity multiplier: CVX_Sim::SetGravityAccel VX_Sim.cpp:485 */
    float gravMult = (float)(-(float)sp->grav_acc / 9.80665);
    float zI = sp->bond_damp_z <= 0 ? 0 : sp->bond_damp_z;
    float zG = sp->slow_damp_z <= 0 ? 0 : sp->slow_damp_z;
    float zC = sp->col_damp_z <= 0 ? 0 : sp->col_damp_z;

    /* one voxel material per palette entry, plus one per voxel when E is overridden */
    int npal = sp->nmat;
    m->nvm = npal + (sp->voxel_E ? nv : 0);
    m->vm = (vmat *)xcalloc(m->nvm, sizeof(vmat));
    for (int k = 0; k < npal; k++) {
        const vxo_material *in = &sp->mats[k];
        vmat *vm = &m->vm[k];
        vm->nomSize = sp->lattice_dim;
        /* CVX_Sim::CopyMat (src/VXA/VX_Sim.cpp:390-405): a setter that rejects its
         * arguments leaves the addMaterial default, linear at 1e6 Pa */
        om_set_linear(vm, 1e6f, -1.0f);
        if (in->model == 0) {
            if (in->E > 0) om_set_linear(vm, in->E, -1.0f);
        } else if (in->model == 1) {
            if (in->E > 0 && (in->fail_stress == -1.0f || in->fail_stress > 0)) om_set_linear(vm, in->E, in->fail_stress);
        } else if (in->model == 2) {
            om_set_bilinear(vm, in->E, in->plastic_mod, in->yield_stress, in->fail_stress);
        } else {
            om_set_data(vm, in->ss_points, in->ss_strain, in->ss_stress);
        }
        vm->nu = sp->volume_effects ? clamp_nu(in->nu) : 0.0f;
        vm->rho = in->rho <= 0 ? FLT_MIN : in->rho;
        vm->alphaCTE = in->cte;
        vm->muS = in->mu_s <= 0 ? 0 : in->mu_s;
        vm->muK = in->mu_k <= 0 ? 0 : in->mu_k;
        vm->zetaI = zI;
        vm->zetaG = zG;
        vm->zetaC = zC;
        vm->gravMult = gravMult;
        vm->fixed = in->fixed;
        vm->isMeasured = in->is_measured;
        vm->thermalOnAfter = in->thermal_on_after;
        vm->cilia = in->cilia;
        vm->ciliaOnAfter = in->cilia_on_after;
        vm->removeAfter = in->remove_after;
        vm->isPaceMaker = in->is_pacemaker;
        vm->paceMakerPeriod = in->pacemaker_period;
        vm->signalDecay = in->signal_decay;
        vm->signalDelay = in->signal_delay;
        vm->inactivePeriod = in->inactive_period;
        vm->isTarget = in->is_target;
        vm->sticky = in->sticky;
        vmat_derive(vm);
    }

    m->ix = (int *)xcalloc(nv, sizeof(int));
    m->iy = (int *)xcalloc(nv, sizeof(int));
    m->iz = (int *)xcalloc(nv, sizeof(int));
    m->vmat_of = (int *)xcalloc(nv, sizeof(int));
    m->vlinks = (int(*)[6])xcalloc(nv, sizeof(int[6]));
    m->pos = (v3 *)xcalloc(nv, sizeof(v3));
    m->linMom = (v3 *)xcalloc(nv, sizeof(v3));
    m->angMom = (v3 *)xcalloc(nv, sizeof(v3));
    m->contact = (v3 *)xcalloc(nv, sizeof(v3));
    m->orient = (q4 *)xcalloc(nv, sizeof(q4));
    m->temp = (float *)xcalloc(nv, sizeof(float));
    m->prevDt = (float *)xcalloc(nv, sizeof(float));
    m->vflags = (int *)xcalloc(nv, sizeof(int));
    m->phase = (double *)xcalloc(nv, sizeof(double));
    m->is_surface = (int *)xcalloc(nv, sizeof(int));
    m->ext_dof = (int *)xcalloc(nv, sizeof(int));
    m->ext_t = (v3 *)xcalloc(nv, sizeof(v3));
    m->ext_r = (v3 *)xcalloc(nv, sizeof(v3));
    m->ext_q = (q4 *)xcalloc(nv, sizeof(q4));
    m->ext_f = (v3 *)xcalloc(nv, sizeof(v3));
    m->ext_m = (v3 *)xcalloc(nv, sizeof(v3));
    for (int i = 0; i < nv; i++) m->ext_dof[i] = -1;
    m->ciliaBase = (v3 *)xcalloc(nv, sizeof(v3));
    m->ciliaShift = (v3 *)xcalloc(nv, sizeof(v3));
    m->vremoved = (int *)xcalloc(nv, sizeof(int));
    m->vattach = (int *)xcalloc(nv, sizeof(int));
    for (int i = 0; i < nv; i++) m->vattach[i] = 1;
    double **sig[8] = {&m->sLocal, &m->sLocalDt, &m->sValue, &m->sActive, &m->sInactive, &m->sNextPulse,
                       &m->outValue, &m->outTime};
    for (int k = 0; k < 8; k++) *sig[k] = (double *)xcalloc(nv, sizeof(double));
    for (int i = 0; i < nv; i++) {
        if (sp->cilia_base) m->ciliaBase[i] = v3_(sp->cilia_base[3 * i], sp->cilia_base[3 * i + 1], sp->cilia_base[3 * i + 2]);
        if (sp->cilia_shift)
            m->ciliaShift[i] = v3_(sp->cilia_shift[3 * i], sp->cilia_shift[3 * i + 1], sp->cilia_shift[3 * i + 2]);
    }

    int *cellvox = (int *)xcalloc(ncell, sizeof(int));
    int v = 0;
    for (int i = 0; i < ncell; i++) {
        cellvox[i] = -1;
        if (!(sp->cells[i] > 0 && sp->cells[i] <= sp->nmat)) continue;
        int z = i / (nx * ny), y = (i - z * nx * ny) / nx, x = i - z * nx * ny - y * nx;
        cellvox[i] = v;
        m->ix[v] = x; m->iy[v] = y; m->iz[v] = z;
        int pal = sp->cells[i] - 1;
        if (sp->voxel_E) {
            int k = npal + v;
            m->vm[k] = m->vm[pal];
            m->vm[k].E = sp->voxel_E[v];
            om_set_linear(&m->vm[k], sp->voxel_E[v], m->vm[k].sigmaFail);
            vmat_derive(&m->vm[k]);
            m->vmat_of[v] = k;
        } else {
            m->vmat_of[v] = pal;
        }
        m->pos[v] = v3_(x * m->voxSize, y * m->voxSize, z * m->voxSize);
        m->orient[v].w = 1;
        m->vflags[v] = VF_FLOOR_STATIC_FRICTION | (sp->floor_enabled ? VF_FLOOR_ENABLED : 0);
        m->phase[v] = sp->phase ? sp->phase[v] : 0.0;
        for (int d = 0; d < 6; d++) m->vlinks[v][d] = -1;
        v++;
    }

    /* links: for each new voxel, directions X_NEG, Y_NEG, Z_NEG find an existing neighbour */
    int nlmax = 3 * nv;
    m->lneg = (int *)xcalloc(nlmax, sizeof(int));
    m->lpos = (int *)xcalloc(nlmax, sizeof(int));
    m->laxis = (int *)xcalloc(nlmax, sizeof(int));
    int nl = 0;
    for (v = 0; v < nv; v++) {
        int x = m->ix[v], y = m->iy[v], z = m->iz[v];
        int nb[3] = {x > 0 ? cellvox[(x - 1) + nx * (y + ny * z)] : -1,
                     y > 0 ? cellvox[x + nx * ((y - 1) + ny * z)] : -1,
                     z > 0 ? cellvox[x + nx * (y + ny * (z - 1))] : -1};
        for (int a = 0; a < 3; a++) {
            if (nb[a] < 0) continue;
            m->lneg[nl] = nb[a];
            m->lpos[nl] = v;
            m->laxis[nl] = a;
            m->vlinks[v][2 * a + 1] = nl;     /* new voxel: link in the NEG direction */
            m->vlinks[nb[a]][2 * a] = nl;     /* neighbour: link in the POS direction */
            nl++;
        }
    }
    free(cellvox);
    m->nl = nl;
    m->lcap = nlmax;
    m->newLink = (int *)xcalloc(nlmax, sizeof(int));
    nl = nlmax;  /* the arrays below hold lcap links */
    m->lm = (lmat *)xcalloc(nl, sizeof(lmat));
    m->pos2 = (v3 *)xcalloc(nl, sizeof(v3));
    m->a1v = (v3 *)xcalloc(nl, sizeof(v3));
    m->a2v = (v3 *)xcalloc(nl, sizeof(v3));
    m->strain = (float *)xcalloc(nl, sizeof(float));
    m->maxStrain = (float *)xcalloc(nl, sizeof(float));
    m->strainOffset = (float *)xcalloc(nl, sizeof(float));
    m->tArea = (float *)xcalloc(nl, sizeof(float));
    m->tStrainSum = (float *)xcalloc(nl, sizeof(float));
    m->restLen = (double *)xcalloc(nl, sizeof(double));
    m->lflags = (int *)xcalloc(nl, sizeof(int));
    m->fNeg = (v3 *)xcalloc(nl, sizeof(v3));
    m->fPos = (v3 *)xcalloc(nl, sizeof(v3));
    m->mNeg = (v3 *)xcalloc(nl, sizeof(v3));
    m->mPos = (v3 *)xcalloc(nl, sizeof(v3));
    m->lremoved = (int *)xcalloc(nl > 0 ? nl : 1, sizeof(int));
    nl = m->nl;
    for (int l = 0; l < nl; l++) {
        /* combinedMaterial(voxel1 = new voxel, voxel2 = neighbour): symmetric */
        lmat_combine(&m->lm[l], &m->vm[m->vmat_of[m->lpos[l]]], &m->vm[m->vmat_of[m->lneg[l]]]);
        m->lflags[l] = LF_SMALL_ANGLE;  /* CVX_Link::reset VX_Link.cpp:50-63 */
        /* transverse info at creation (strains zero): VX3_Voxel.cu:496-513 */
        float s = (float)m->vm[m->vmat_of[m->lneg[l]]].nomSize;
        float s2 = (float)m->vm[m->vmat_of[m->lpos[l]]].nomSize;
        m->tArea[l] = 0.5f * (s * s + s2 * s2);
        m->tStrainSum[l] = 0.0f;
    }

    m->surface = (int *)xcalloc(nv, sizeof(int));
    regenerate_surface(m);

    /* initial temperature (UpdateMatTemps) */
    float t0 = 0.0f;
    if (sp->temp_enabled) {
        float cur;
        if (sp->vary_temp) cur = (sp->temp_period == 0) ? 0.0f : (float)(sp->temp_base + sp->temp_amplitude * vxm_sin(0.0));
        else cur = (float)(sp->temp_base + sp->temp_amplitude);
        t0 = (float)(cur - sp->temp_base);
    }
    for (v = 0; v < nv; v++) m->temp[v] = t0;
    for (int l = 0; l < nl; l++) update_rest_length(m, l);
    double c[3];
    vxo_center_of_mass(m, c);  /* InitializeCenterOfMass */
    m->comCur = v3_(c[0], c[1], c[2]);
    return m;
}

void vxo_free(vxo_model *m) {
    if (!m) return;
    void *ptrs[] = {m->ix, m->iy, m->iz, m->vmat_of, m->vlinks, m->pos, m->linMom, m->angMom, m->contact,
                    m->orient, m->temp, m->prevDt, m->vflags, m->phase, m->surface, m->is_surface, m->lneg,
                    m->lpos, m->laxis, m->lm, m->pos2, m->a1v, m->a2v, m->strain, m->maxStrain, m->tArea,
                    m->tStrainSum, m->restLen, m->lflags, m->fNeg, m->fPos, m->mNeg, m->mPos, m->vm, m->strainOffset,
                    m->ext_dof, m->ext_t, m->ext_r, m->ext_q, m->ext_f, m->ext_m,
                    m->ff_op[0], m->ff_op[1], m->ff_op[2], m->ff_val[0], m->ff_val[1], m->ff_val[2],
                    m->ciliaBase, m->ciliaShift, m->vremoved, m->lremoved, m->sLocal, m->sLocalDt, m->sValue,
                    m->sActive, m->sInactive, m->sNextPulse, m->outValue, m->outTime, m->pairs, m->newLink,
                    m->vattach, m->ac_op[0], m->ac_op[1], m->ac_op[2], m->ac_op[3], m->ac_op[4],
                    m->ac_val[0], m->ac_val[1], m->ac_val[2], m->ac_val[3], m->ac_val[4]};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(m);
}

/* VX3_MathTree::eval (src/Utils/VX3_MathTree.h:50-190): values appended in
 * token order; an operator consumes the oldest unconsumed values (a binary one
 * computes values[p + 1] op values[p]); END returns the oldest unconsumed one.
 * sin and cos are the shared vxm_ functions. */
static double mt_eval_fifo(const int *op, const double *val, int n, const double *vars) {
    double v[1024];
    int pc = 0, vc = 0;
    for (int i = 0; i < n && vc < 1024; i++) {
        double r;
        switch (op[i]) {
        case VXO_MT_END: return pc < vc ? v[pc] : 0.0;
        case VXO_MT_CONST: r = val[i]; break;
        case VXO_MT_E: r = 2.71828182845904523536; break;
        case VXO_MT_PI: r = 3.14159265358979323846; break;
        case VXO_MT_VAR: r = vars[(int)val[i]]; break;
        case VXO_MT_SIN: r = vxm_sin(v[pc++]); break;
        case VXO_MT_COS: r = vxm_cos(v[pc++]); break;
        case VXO_MT_TAN: r = tan(v[pc++]); break;
        case VXO_MT_ATAN: r = atan(v[pc++]); break;
        case VXO_MT_LOG: r = log(v[pc++]); break;
        case VXO_MT_INT: r = rint(v[pc++]); break;
        case VXO_MT_NORMALCDF: r = 0.5 * erfc(-v[pc] / sqrt(2.0)); pc++; break;
        case VXO_MT_SQRT: r = sqrt(v[pc++]); break;
        case VXO_MT_ABS: r = fabs(v[pc++]); break;
        case VXO_MT_NOT: r = !v[pc++]; break;
        case VXO_MT_ADD: r = v[pc + 1] + v[pc]; pc += 2; break;
        case VXO_MT_SUB: r = v[pc + 1] - v[pc]; pc += 2; break;
        case VXO_MT_MUL: r = v[pc + 1] * v[pc]; pc += 2; break;
        case VXO_MT_DIV: r = v[pc + 1] / v[pc]; pc += 2; break;
        case VXO_MT_POW: r = pow(v[pc + 1], v[pc]); pc += 2; break;
        case VXO_MT_GT: r = v[pc + 1] > v[pc]; pc += 2; break;
        case VXO_MT_LT: r = v[pc + 1] < v[pc]; pc += 2; break;
        case VXO_MT_AND: r = v[pc + 1] && v[pc]; pc += 2; break;
        case VXO_MT_OR: r = v[pc + 1] || v[pc]; pc += 2; break;
        default: return 0.0;
        }
        v[vc++] = r;
    }
    return 0.0;
}

void vxo_set_attach_condition(vxo_model *m, int k, int n, const int *ops, const double *values) {
    if (k < 0 || k > 4) return;
    free(m->ac_op[k]);
    free(m->ac_val[k]);
    m->ac_n[k] = n > 0 ? n : 0;
    m->ac_op[k] = (int *)xcalloc(m->ac_n[k], sizeof(int));
    m->ac_val[k] = (double *)xcalloc(m->ac_n[k], sizeof(double));
    for (int i = 0; i < m->ac_n[k]; i++) { m->ac_op[k][i] = ops[i]; m->ac_val[k][i] = values[i]; }
}

void vxo_get_samples(const vxo_model *m, double *out7) {
    out7[0] = m->recentAngle;
    out7[1] = m->closeness;
    out7[2] = m->closePairs;
    out7[3] = m->angleSamples;
    out7[4] = m->comCur.x; out7[5] = m->comCur.y; out7[6] = m->comCur.z;
}

void vxo_get_attach(const vxo_model *m, int *voxel_flags, int *link_new) {
    if (voxel_flags) memcpy(voxel_flags, m->vattach, sizeof(int) * m->nv);
    if (link_new) memcpy(link_new, m->newLink, sizeof(int) * m->nl);
}

void vxo_set_force_field(vxo_model *m, int axis, int n, const int *ops, const double *values) {
    if (axis < 0 || axis > 2) return;
    free(m->ff_op[axis]);
    free(m->ff_val[axis]);
    m->ff_n[axis] = n > 0 ? n : 0;
    m->ff_op[axis] = (int *)xcalloc(m->ff_n[axis], sizeof(int));
    m->ff_val[axis] = (double *)xcalloc(m->ff_n[axis], sizeof(double));
    for (int i = 0; i < m->ff_n[axis]; i++) { m->ff_op[axis][i] = ops[i]; m->ff_val[axis][i] = values[i]; }
}

/* CVX_Voxel::external() + CVX_External::setDisplacement / setForce / setMoment
 * (src/old/VX_External.cpp): the external of a voxel; the rotation quaternion
 * is cached as CVX_External::rotationChanged does (identity for a zero
 * rotation vector, else FromRotationVector). */
void vxo_set_external(vxo_model *m, int v, int dof, const double *t3, const double *r3, const float *f3,
                      const float *m3) {
    if (v < 0 || v >= m->nv) return;
    m->ext_dof[v] = dof < 0 ? -1 : (dof & 63);
    m->ext_t[v] = t3 ? v3_(t3[0], t3[1], t3[2]) : v3_(0, 0, 0);
    m->ext_r[v] = r3 ? v3_(r3[0], r3[1], r3[2]) : v3_(0, 0, 0);
    m->ext_f[v] = f3 ? v3_(f3[0], f3[1], f3[2]) : v3_(0, 0, 0);
    m->ext_m[v] = m3 ? v3_(m3[0], m3[1], m3[2]) : v3_(0, 0, 0);
    v3 r = m->ext_r[v];
    if (r.x != 0 || r.y != 0 || r.z != 0) m->ext_q[v] = q_from_rotvec(r);
    else { q4 id = {1, 0, 0, 0}; m->ext_q[v] = id; }
}

int vxo_get_link_curve(const vxo_model *m, int link, float *strain, float *stress) {
    const lmat *l = &m->lm[link];
    if (l->linear) return 0;
    for (int k = 0; k < l->npts; k++) { strain[k] = l->sd[k]; stress[k] = l->td[k]; }
    return l->npts;
}

int vxo_num_voxels(const vxo_model *m) { return m->nv; }
int vxo_num_links(const vxo_model *m) { return m->nl; }
int vxo_num_surface(const vxo_model *m) { return m->ns; }
double vxo_time(const vxo_model *m) { return m->time; }

/* ------------------------------------------------------------ temperature */
/* gpu_update_temperature: VX3_VoxelyzeKernel.cu:625-650 */
static void update_temperature(vxo_model *m) {
    if (!(m->varyTemp && m->tempPeriod > 0)) return;
    const double twoPiF = (double)(2 * 3.1415926f);
    for (int v = 0; v < m->nv; v++) {
        const vmat *mt = &m->vm[m->vmat_of[v]];
        if (m->vremoved[v]) continue;
        if (mt->thermalOnAfter > m->time) continue;
        if (mt->fixed) continue;
        double T = m->tempAmplitude * vxm_sin(twoPiF * (m->time / m->tempPeriod + m->phase[v]));
        if (!m->enableExpansion && T > 0) T = 0;
        m->temp[v] = (float)T;
    }
    for (int l = 0; l < m->nl; l++) update_rest_length(m, l);
}

/* ------------------------------------------------------------------ links */
/* VX3_Voxel::dampingMultiplier VX3_Voxel.h:206-208 */
static inline float damping_multiplier(const vxo_model *m, int v) {
    const vmat *mt = &m->vm[m->vmat_of[v]];
    return 2 * mt->sqrtMass * mt->zetaI / m->prevDt[v];
}

/* VX3_Link::updateForces VX3_Link.cu:135-218 with orientLink :90-133 and
 * updateStrain :220-260 (linear material model, VX3_Material.cu:90-124) */
static void link_update(vxo_model *m, int l) {
    const lmat *mat = &m->lm[l];
    int n = m->lneg[l], p = m->lpos[l], axis = m->laxis[l];
    v3 oldPos2 = m->pos2[l], oldA1v = m->a1v[l], oldA2v = m->a2v[l];
    double restLen = m->restLen[l];
    int flags = m->lflags[l];

    /* orientLink */
    v3 pos2 = to_axis_x(vsub(m->pos[p], m->pos[n]), axis);
    q4 angle1 = to_axis_x_q(m->orient[n], axis);
    q4 angle2 = to_axis_x_q(m->orient[p], axis);
    q4 totalRot = qconj(angle1);
    pos2 = qrot(totalRot, pos2);
    angle2 = qmul(totalRot, angle2);
    angle1.w = 1; angle1.x = angle1.y = angle1.z = 0;

    float smallTurn = (float)((fabs(pos2.z) + fabs(pos2.y)) / pos2.x);
    float extendPerc = (float)(fabs(1 - pos2.x / restLen));
    int smallAngle = (flags & LF_SMALL_ANGLE) != 0;
    if (!smallAngle && smallTurn < SA_BOND_BEND_RAD && extendPerc < SA_BOND_EXT_PERC) {
        smallAngle = 1;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    } else if (smallAngle && (smallTurn > HYSTERESIS_FACTOR * SA_BOND_BEND_RAD ||
                              extendPerc > HYSTERESIS_FACTOR * SA_BOND_EXT_PERC)) {
        smallAngle = 0;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    }
    int exact = 0, acos1, acos2;
    if (smallAngle) {
        pos2.x -= restLen;
    } else {
        angle1 = q_from_angle_to_posx(pos2, &exact);
        totalRot = qmul(angle1, totalRot);
        angle2 = qmul(angle1, angle2);
        pos2 = v3_(vlen(pos2) - restLen, 0, 0);
    }
    v3 a1v = q_to_rotvec(angle1, &acos1);
    v3 a2v = q_to_rotvec(angle2, &acos2);
    flags = (flags & ~(LF_SMALL_ANGLE | LF_DECISIONS)) | (smallAngle ? LF_SMALL_ANGLE : 0) |
            (acos1 ? LF_ROTVEC_ACOS1 : 0) | (acos2 ? LF_ROTVEC_ACOS2 : 0) | (exact ? LF_POSX_EXACT : 0);
    m->pos2[l] = pos2;
    m->a1v[l] = a1v;
    m->a2v[l] = a2v;

    v3 dPos2 = vscale(vsub(pos2, oldPos2), 0.5);
    v3 dA1 = vscale(vsub(a1v, oldA1v), 0.5);
    v3 dA2 = vscale(vsub(a2v, oldA2v), 0.5);
    /* dPos2 = 0.5*(pos2-oldPos2): the reference multiplies 0.5*vec, identical */

    /* updateStrain: VX3_Link.cu:220-260, VX3_Material.cu:90-124 */
    float axialStrain = (float)(pos2.x / restLen);
    m->strain[l] = axialStrain;
    float stress;
    if (mat->linear) {
        if (axialStrain > m->maxStrain[l]) m->maxStrain[l] = axialStrain;
        stress = om_stress(mat, axialStrain, m->tStrainSum[l], 0);
    } else if (axialStrain > m->maxStrain[l]) {  /* new territory on the stress-strain curve */
        m->maxStrain[l] = axialStrain;
        stress = om_stress(mat, axialStrain, m->tStrainSum[l], 0);
        if (mat->nu != 0.0f)
            m->strainOffset[l] = m->maxStrain[l] - om_stress(mat, axialStrain, 0.0f, 0) / (mat->eHat * (1 - mat->nu));
        else
            m->strainOffset[l] = m->maxStrain[l] - stress / mat->E;
    } else {  /* backed off: linear with the plastic strain offset */
        float rel = axialStrain - m->strainOffset[l];
        stress = mat->nu != 0.0f ? om_stress(mat, rel, m->tStrainSum[l], 1) : mat->E * rel;
    }

    if (mat->epsFail != -1.0f && m->maxStrain[l] > mat->epsFail) {
        m->fNeg[l] = m->fPos[l] = m->mNeg[l] = m->mPos[l] = v3_(0, 0, 0);
        m->lflags[l] = flags;
        return;
    }
    float b1 = mat->b1, b2 = mat->b2, b3 = mat->b3, a2 = mat->a2;
    v3 fNeg = v3_((double)(stress * m->tArea[l]), b1 * pos2.y - b2 * (a1v.z + a2v.z), b1 * pos2.z + b2 * (a1v.y + a2v.y));
    v3 fPos = v3_(-fNeg.x, -fNeg.y, -fNeg.z);
    v3 mNeg = v3_(a2 * (a2v.x - a1v.x), -b2 * pos2.z - b3 * (2 * a1v.y + a2v.y), b2 * pos2.y - b3 * (2 * a1v.z + a2v.z));
    v3 mPos = v3_(a2 * (a1v.x - a2v.x), -b2 * pos2.z - b3 * (a1v.y + 2 * a2v.y), b2 * pos2.y - b3 * (a1v.z + 2 * a2v.z));

    if (flags & LF_LOCAL_VELOCITY_VALID) {
        float sqA1 = mat->sqA1, sqA2xIp = mat->sqA2xIp, sqB1 = mat->sqB1, sqB2xFMp = mat->sqB2xFMp, sqB3xIp = mat->sqB3xIp;
        v3 posCalc = v3_(sqA1 * dPos2.x, sqB1 * dPos2.y - sqB2xFMp * (dA1.z + dA2.z),
                         sqB1 * dPos2.z + sqB2xFMp * (dA1.y + dA2.y));
        float dmN = damping_multiplier(m, n), dmP = damping_multiplier(m, p);
        fNeg = vadd(fNeg, vscale(posCalc, dmN));
        fPos = vsub(fPos, vscale(posCalc, dmP));
        double hN = 0.5 * dmN, hP = 0.5 * dmP;
        v3 tN = v3_(-sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * (2 * dA1.y + dA2.y),
                    -sqB2xFMp * dPos2.y + sqB3xIp * (2 * dA1.z + dA2.z));
        v3 tP = v3_(sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * (dA1.y + 2 * dA2.y),
                    -sqB2xFMp * dPos2.y + sqB3xIp * (dA1.z + 2 * dA2.z));
        mNeg = vsub(mNeg, vscale(tN, hN));
        mPos = vsub(mPos, vscale(tP, hP));
    } else {
        flags |= LF_LOCAL_VELOCITY_VALID;
    }
    if (!smallAngle) {
        fNeg = qrotinv(angle1, fNeg);
        mNeg = qrotinv(angle1, mNeg);
    }
    fPos = qrotinv(angle2, fPos);
    mPos = qrotinv(angle2, mPos);
    fNeg = to_axis_original(fNeg, axis);
    fPos = to_axis_original(fPos, axis);
    mNeg = to_axis_original(mNeg, axis);
    mPos = to_axis_original(mPos, axis);
    /* a link made by attach: its first SafetyGuard updates are scaled by 0.01
     * (VX3_Link.cu:206-213) */
    if (m->newLink[l]) {
        fNeg = vscale(fNeg, 0.01);
        fPos = vscale(fPos, 0.01);
        mNeg = vscale(mNeg, 0.01);
        mPos = vscale(mPos, 0.01);
        m->newLink[l] -= 1;
    }
    m->fNeg[l] = fNeg;
    m->fPos[l] = fPos;
    m->mNeg[l] = mNeg;
    m->mPos[l] = mPos;
    m->lflags[l] = flags;
}

static void update_links(vxo_model *m) {
    int nl = m->nl;
#pragma omp parallel for schedule(static)
    for (int l = 0; l < nl; l++) {
        /* gpu_update_links VX3_VoxelyzeKernel.cu:566-581: skip removed links and links between two fixed voxels */
        if (m->lremoved[l]) continue;
        if (m->vm[m->vmat_of[m->lpos[l]]].fixed && m->vm[m->vmat_of[m->lneg[l]]].fixed) continue;
        link_update(m, l);
    }
}

/* room for one more link in every per-link array (attach appends links) */
static void *grow(void *p, size_t old, size_t cap, size_t sz) {
    char *q = (char *)realloc(p, cap * sz);
    if (!q) abort();
    memset(q + old * sz, 0, (cap - old) * sz);
    return q;
}
static void link_reserve(vxo_model *m) {
    if (m->nl < m->lcap) return;
    size_t o = (size_t)m->lcap, c = o ? 2 * o : 64;
    m->lneg = (int *)grow(m->lneg, o, c, sizeof(int));
    m->lpos = (int *)grow(m->lpos, o, c, sizeof(int));
    m->laxis = (int *)grow(m->laxis, o, c, sizeof(int));
    m->newLink = (int *)grow(m->newLink, o, c, sizeof(int));
    m->lm = (lmat *)grow(m->lm, o, c, sizeof(lmat));
    m->pos2 = (v3 *)grow(m->pos2, o, c, sizeof(v3));
    m->a1v = (v3 *)grow(m->a1v, o, c, sizeof(v3));
    m->a2v = (v3 *)grow(m->a2v, o, c, sizeof(v3));
    m->strain = (float *)grow(m->strain, o, c, sizeof(float));
    m->maxStrain = (float *)grow(m->maxStrain, o, c, sizeof(float));
    m->strainOffset = (float *)grow(m->strainOffset, o, c, sizeof(float));
    m->tArea = (float *)grow(m->tArea, o, c, sizeof(float));
    m->tStrainSum = (float *)grow(m->tStrainSum, o, c, sizeof(float));
    m->restLen = (double *)grow(m->restLen, o, c, sizeof(double));
    m->lflags = (int *)grow(m->lflags, o, c, sizeof(int));
    m->fNeg = (v3 *)grow(m->fNeg, o, c, sizeof(v3));
    m->fPos = (v3 *)grow(m->fPos, o, c, sizeof(v3));
    m->mNeg = (v3 *)grow(m->mNeg, o, c, sizeof(v3));
    m->mPos = (v3 *)grow(m->mPos, o, c, sizeof(v3));
    m->lremoved = (int *)grow(m->lremoved, o ? o : 1, c, sizeof(int));
    m->lcap = (int)c;
}

/* ------------------------------------------------------------- collisions */
typedef struct { long long key; int s; } cellent;
static int cellent_cmp(const void *a, const void *b) {
    const cellent *x = (const cellent *)a, *y = (const cellent *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->s - y->s;
}
static inline long long cell_key(long long cx, long long cy, long long cz) {
    return ((cx + (1LL << 20)) << 42) | ((cy + (1LL << 20)) << 21) | (cz + (1LL << 20));
}

/* handle_collision_attachment (collision part) VX3_VoxelyzeKernel.cu:682-729
 * with VX3_Collision VX3_Collision.cu:3-31, for voxel1 = surface[first],
 * voxel2 = surface[second], first > second (gpu_update_attach :833-843). */
static void signal_receive(vxo_model *m, int v, const vmat *mt, double value, double activeTime, int force);

/* The contact force of the pair (voxel1 = v1, voxel2 = v2) if it is in contact:
 * the tests of handle_collision_attachment (VX3_VoxelyzeKernel.cu:682-729) in
 * its order, then VX3_Collision::updateContactForce (VX3_Collision.cu:17-31).
 * Returns 0 (no effect at all) for pairs the reference skips. */
/* is_neighbor (VX3_VoxelyzeKernel.cu:651-680): b is reached from a over at
 * most depth links without going straight back over the link just taken */
static int linked_within(const vxo_model *m, int a, int b, int incoming, int depth) {
    if (a == b) return 1;
    if (depth <= 0) return 0;
    for (int d = 0; d < 6; d++) {
        int l = m->vlinks[a][d];
        if (l < 0 || l == incoming) continue;
        int o = m->lneg[l] == a ? m->lpos[l] : m->lneg[l];
        if (linked_within(m, o, b, l, depth - 1)) return 1;
    }
    return 0;
}

/* The tests every effect of a pair passes first (:684-703): not both fixed,
 * inside the watch distance, not directly linked. */
static int pair_near(const vxo_model *m, int v1, int v2) {
    const vmat *m1 = &m->vm[m->vmat_of[v1]], *m2 = &m->vm[m->vmat_of[v2]];
    if (m1->fixed && m2->fixed) return 0;
    v3 diff = vsub(m->pos[v1], m->pos[v2]);
    double bsa1 = base_size_average(m, v1), bsa2 = base_size_average(m, v2);
    double watch = (bsa1 + bsa2) * COLLISION_ENVELOPE_RADIUS;
    if (diff.x > watch || diff.x < -watch) return 0;
    if (diff.y > watch || diff.y < -watch) return 0;
    if (diff.z > watch || diff.z < -watch) return 0;
    if (vlen(diff) > watch) return 0;
    return !linked_within(m, v1, v2, -1, 1);
}

/* A near pair may attach (:729-746): attach on, both voxels' enableAttach,
 * neither fixed, one material, sticky, and no path of at most 5 links between
 * them. */
static int attach_candidate(const vxo_model *m, int v1, int v2) {
    if (!m->enableAttach || !m->vattach[v1] || !m->vattach[v2]) return 0;
    const vmat *m1 = &m->vm[m->vmat_of[v1]], *m2 = &m->vm[m->vmat_of[v2]];
    if (m1->fixed || m2->fixed) return 0;
    if (m->vmat_of[v1] != m->vmat_of[v2]) return 0;
    if (!m1->sticky) return 0;
    return !linked_within(m, v1, v2, -1, 5);
}

/* VX3_Collision::updateContactForce (VX3_Collision.cu:17-31) of a near pair */
static v3 contact_force_of(const vxo_model *m, int v1, int v2) {
    const vmat *m1 = &m->vm[m->vmat_of[v1]], *m2 = &m->vm[m->vmat_of[v2]];
    double bsa1 = base_size_average(m, v1), bsa2 = base_size_average(m, v2);
    float pk1 = (float)(2 * m1->E * m1->nomSize), pk2 = (float)(2 * m2->E * m2->nomSize);
    double penStiff = 2.0f / (1.0f / pk1 + 1.0f / pk2);
    double dampC = 0.5f * (m1->zetaC * m1->c2xSqMxExS + m2->zetaC * m2->c2xSqMxExS);
    v3 offset = vsub(m->pos[v2], m->pos[v1]);
    double nomDist = (bsa1 + bsa2) * COLLISION_ENVELOPE_RADIUS;
    double relDist = nomDist - vlen(offset);
    v3 force = v3_(0, 0, 0);
    if (relDist > 0) {
        double l = vlen(offset);
        v3 unit = l > 0 ? v3_(offset.x / l, offset.y / l, offset.z / l) : offset;
        v3 vel1 = vscale(m->linMom[v1], m1->massInv), vel2 = vscale(m->linMom[v2], m2->massInv);
        double rv = (vel1.x * unit.x + vel1.y * unit.y + vel1.z * unit.z) - (vel2.x * unit.x + vel2.y * unit.y + vel2.z * unit.z);
        force = vscale(unit, penStiff * relDist + dampC * rv);
    }
    return force;
}

/* The effects of a pair in contact, in the reference's order: the contact
 * forces (voxel1 gains it, voxel2 loses it), then for a target touching a
 * non-target collisionCount and, with signals, a forced 100 received by the
 * non-target voxel (:711-726). */
static void apply_pair(vxo_model *m, int v1, int v2, v3 force) {
    if (m->npairs == m->pairCap) {
        m->pairCap = m->pairCap ? 2 * m->pairCap : 1024;
        m->pairs = (int *)realloc(m->pairs, sizeof(int) * 2 * (size_t)m->pairCap);
        if (!m->pairs) abort();
    }
    m->pairs[2 * m->npairs] = v1;
    m->pairs[2 * m->npairs + 1] = v2;
    m->npairs++;
    m->contact[v1] = vadd(m->contact[v1], force);
    m->contact[v2] = vsub(m->contact[v2], force);
    const vmat *m1 = &m->vm[m->vmat_of[v1]], *m2 = &m->vm[m->vmat_of[v2]];
    if (m1->isTarget != m2->isTarget) {
        m->collisionCount++;
        if (m->enableSignals) {
            int o = m1->isTarget ? v2 : v1;
            signal_receive(m, o, &m->vm[m->vmat_of[o]], 100, m->time, 1);
        }
    }
}

/* what a surface pair does this step: its contact force (contact) and whether
 * it is an attach candidate */
typedef struct { int v1, v2, contact, attach; v3 f; } pairhit;
static int pair_eval(const vxo_model *m, int v1, int v2, pairhit *h) {
    h->v1 = v1;
    h->v2 = v2;
    if (!pair_near(m, v1, v2)) return 0;
    h->contact = m->enableCollision;  /* with EnableCollision a near pair is in contact */
    if (h->contact) h->f = contact_force_of(m, v1, v2);
    h->attach = attach_candidate(m, v1, v2);
    return h->contact || h->attach;
}

/* The link of an attach candidate (handle_collision_attachment :748-830 with the
 * VX3_Link constructor and reset, VX3_Link.cu:31-76): the axis is the dominant
 * component of voxel2's offset in voxel1's frame; voxel1 takes the slot facing
 * voxel2 and voxel2 the opposite one, and the voxel holding a NEG-direction
 * slot is the link's positive end.  Made only when both slots are free; the
 * pair's contact force (f, when it had one) is given back.  Returns 1 when a
 * link was made. */
static int attach_pair(vxo_model *m, int v1, int v2, const v3 *f) {
    v3 e = vsub(m->pos[v1], m->pos[v2]);
    v3 ea = qrotinv(m->orient[v1], v3_(-e.x, -e.y, -e.z));
    double ax = fabs(ea.x), ay = fabs(ea.y), az = fabs(ea.z);
    int axis, neg1; /* neg1: voxel1 takes its NEG-direction slot (reverseOrder) */
    if (ax >= ay && ax >= az) { axis = 0; neg1 = ea.x < 0; }
    else if (ay >= ax && ay >= az) { axis = 1; neg1 = ea.y < 0; }
    else { axis = 2; neg1 = ea.z < 0; }
    int d1 = 2 * axis + (neg1 ? 1 : 0), d2 = 2 * axis + (neg1 ? 0 : 1);
    if (m->vlinks[v1][d1] >= 0 || m->vlinks[v2][d2] >= 0) return 0;
    link_reserve(m);
    int l = m->nl++;
    int neg = neg1 ? v2 : v1, pos = neg1 ? v1 : v2;
    m->lneg[l] = neg;
    m->lpos[l] = pos;
    m->laxis[l] = axis;
    /* combinedMaterial(positive end, negative end): one material here */
    lmat_combine(&m->lm[l], &m->vm[m->vmat_of[pos]], &m->vm[m->vmat_of[neg]]);
    m->lflags[l] = LF_SMALL_ANGLE;
    /* transverse info at creation with zero Poisson's ratio (VX3_Voxel.cu:478-513) */
    float sN = (float)m->vm[m->vmat_of[neg]].nomSize, sP = (float)m->vm[m->vmat_of[pos]].nomSize;
    m->tArea[l] = 0.5f * (sN * sN + sP * sP);
    m->tStrainSum[l] = 0.0f;
    update_rest_length(m, l);
    m->newLink[l] = m->safetyGuard;
    m->vlinks[v1][d1] = l;
    m->vlinks[v2][d2] = l;
    if (f) {
        m->contact[v1] = vsub(m->contact[v1], *f);
        m->contact[v2] = vsub(m->contact[v2], v3_(-f->x, -f->y, -f->z));
    }
    return 1;
}

/* The effects of the step's pairs, given in the all-pairs order: every contact
 * first, then the attach candidates one at a time.  Returns 1 when a link was
 * made (the surface changes). */
static int apply_hits(vxo_model *m, pairhit *h, long n) {
    int made = 0;
    for (long k = 0; k < n; k++)
        if (h[k].contact) apply_pair(m, h[k].v1, h[k].v2, h[k].f);
    for (long k = 0; k < n; k++)
        if (h[k].attach) made |= attach_pair(m, h[k].v1, h[k].v2, h[k].contact ? &h[k].f : NULL);
    return made;
}

/* updateAttach all-pairs branch VX3_VoxelyzeKernel.cu:448-458, literally: every
 * surface pair (first > second) in nested-loop order (gpu_update_attach
 * :833-843 serialised).  O(S^2): the check of the pruned search below. */
static int update_collisions_all_pairs(vxo_model *m) {
    m->npairs = 0;
    long n = 0, cap = 64;
    pairhit *h = (pairhit *)xcalloc(cap, sizeof(pairhit));
    for (int first = 1; first < m->ns; first++)
        for (int second = 0; second < first; second++) {
            if (n == cap) { cap *= 2; h = (pairhit *)realloc(h, cap * sizeof(pairhit)); if (!h) abort(); }
            if (pair_eval(m, m->surface[first], m->surface[second], &h[n])) n++;
        }
    int made = apply_hits(m, h, n);
    free(h);
    return made;
}

/* The same pairs in the same (first, second) order, pruned by a uniform grid
 * whose cell edge is the largest watch distance (pairs in non-adjacent cells
 * are farther apart than any watch distance, so the reference skips them
 * without an effect).  The candidate search and the pair forces run in
 * parallel over `first`; the effects are applied serially in the all-pairs
 * order, so the result does not depend on the thread count. */
static int update_collisions(vxo_model *m) {
    int ns = m->ns;
    m->npairs = 0;
    /* a pair has no effect without EnableCollision or EnableAttach */
    if (ns < 2 || !(m->enableCollision || m->enableAttach)) return 0;
    double h = 0;
    for (int i = 0; i < ns; i++) {
        double b = base_size_average(m, m->surface[i]);
        if (b > h) h = b;
    }
    h *= 2 * COLLISION_ENVELOPE_RADIUS * 1.0001;
    if (!(h > 0)) h = m->voxSize;
    cellent *ent = (cellent *)xcalloc(ns, sizeof(cellent));
    long long(*cc)[3] = (long long(*)[3])xcalloc(ns, sizeof(long long[3]));
    for (int i = 0; i < ns; i++) {
        v3 p = m->pos[m->surface[i]];
        cc[i][0] = (long long)floor(p.x / h);
        cc[i][1] = (long long)floor(p.y / h);
        cc[i][2] = (long long)floor(p.z / h);
        ent[i].key = cell_key(cc[i][0], cc[i][1], cc[i][2]);
        ent[i].s = i;
    }
    qsort(ent, ns, sizeof(cellent), cellent_cmp);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    pairhit **hits = (pairhit **)xcalloc(nt, sizeof(pairhit *));
    int *nhit = (int *)xcalloc(nt, sizeof(int));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        /* a contiguous block of `first` per thread, so that the blocks in thread
         * order are the all-pairs order */
        long lo_f = 1 + (long)(ns - 1) * t / nt, hi_f = 1 + (long)(ns - 1) * (t + 1) / nt;
        int cap = 64, hcap = 64, cnt, nh = 0;
        int *cand = (int *)xcalloc(cap, sizeof(int));
        pairhit *hb = (pairhit *)xcalloc(hcap, sizeof(pairhit));
        for (long first = lo_f; first < hi_f; first++) {
            cnt = 0;
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        long long key = cell_key(cc[first][0] + dx, cc[first][1] + dy, cc[first][2] + dz);
                        int lo = 0, hi = ns;
                        while (lo < hi) {
                            int mid = (lo + hi) / 2;
                            if (ent[mid].key < key) lo = mid + 1; else hi = mid;
                        }
                        for (int k = lo; k < ns && ent[k].key == key; k++) {
                            if (ent[k].s >= first) continue;
                            if (cnt == cap) { cap *= 2; cand = (int *)realloc(cand, cap * sizeof(int)); if (!cand) abort(); }
                            cand[cnt++] = ent[k].s;
                        }
                    }
            /* ascending second, as in the all-pairs order */
            for (int a = 1; a < cnt; a++) {
                int x = cand[a], b = a - 1;
                while (b >= 0 && cand[b] > x) { cand[b + 1] = cand[b]; b--; }
                cand[b + 1] = x;
            }
            for (int a = 0; a < cnt; a++) {
                if (nh == hcap) { hcap *= 2; hb = (pairhit *)realloc(hb, hcap * sizeof(pairhit)); if (!hb) abort(); }
                if (pair_eval(m, m->surface[first], m->surface[cand[a]], &hb[nh])) nh++;
            }
        }
        free(cand);
        hits[t] = hb;
        nhit[t] = nh;
    }
    long total = 0;
    for (int t = 0; t < nt; t++) total += nhit[t];
    pairhit *all = (pairhit *)xcalloc(total, sizeof(pairhit));
    long at = 0;
    for (int t = 0; t < nt; t++) {
        memcpy(all + at, hits[t], sizeof(pairhit) * nhit[t]);
        at += nhit[t];
        free(hits[t]);
    }
    int made = apply_hits(m, all, total);
    free(all);
    free(hits);
    free(nhit);
    free(cc);
    free(ent);
    return made;
}

/* ---------------------------------------------------------------- signals */
/* VX3_Voxel::receiveSignal VX3_Voxel.cu:301-322 (mat: the receiver's) */
static void signal_receive(vxo_model *m, int v, const vmat *mt, double value, double activeTime, int force) {
    if (!force && m->sInactive[v] > activeTime) return;
    if (value < 0.1) return;
    m->sInactive[v] = activeTime + mt->inactivePeriod;
    m->sLocal[v] = value;
    m->sValue[v] = value * mt->signalDecay;
    if (m->sValue[v] < 0.1) m->sValue[v] = 0;
    m->sActive[v] = activeTime;
}
/* the voxel's own signal section at the end of VX3_Voxel::timeStep
 * (VX3_Voxel.cu:271-276): propagateSignal (:323-348; the send is kept in
 * outValue / outTime for the receive phase), packMaker (:290-299),
 * localSignalDecay (:279-288) */
static void signal_own(vxo_model *m, int v, const vmat *mt) {
    const double t = m->time;
    m->outValue[v] = 0;
    if (!(m->sActive[v] > t) && !(m->sValue[v] < 0.1)) {
        m->outValue[v] = m->sValue[v];
        m->outTime[v] = t + mt->signalDelay;
        m->sValue[v] = 0;
        m->sActive[v] = 0;
        m->sInactive[v] = t + 2 * mt->signalDelay + mt->inactivePeriod;
    }
    if (mt->isPaceMaker && !(m->sNextPulse[v] > t)) {
        signal_receive(m, v, mt, 100, t, 1);
        m->sNextPulse[v] = t + mt->paceMakerPeriod;
    }
    if (!(m->sLocalDt[v] > t)) {
        if (m->sLocal[v] < 0.1) {
            m->sLocal[v] = 0;
        } else {
            m->sLocal[v] = m->sLocal[v] * 0.9;
            m->sLocalDt[v] = t + 0.01;
        }
    }
}
/* the receive phase: each present voxel takes the signals its linked
 * neighbours sent this step, in its link direction order */
static void signal_deliver(vxo_model *m) {
    for (int v = 0; v < m->nv; v++) {
        if (m->vremoved[v]) continue;
        const vmat *mt = &m->vm[m->vmat_of[v]];
        for (int d = 0; d < 6; d++) {
            int l = m->vlinks[v][d];
            if (l < 0) continue;
            int o = (d & 1) ? m->lneg[l] : m->lpos[l];
            if (m->outValue[o] != 0) signal_receive(m, v, mt, m->outValue[o], m->outTime[o], 0);
        }
    }
}

/* ----------------------------------------------------------------- voxels */
/* VX3_Voxel::timeStep VX3_Voxel.cu:162-277 with force() :350-380, moment()
 * :382-397, floorForce :399-426 (constant-zero force field); externals as at
 * :167-172 (all six fixed), :363-364 / :393-394 (force, moment) and :222-256
 * (fixed degrees of freedom) */
static void voxel_update(vxo_model *m, int v, float dtf, int cpu_semantics) {
    const vmat *mt = &m->vm[m->vmat_of[v]];
    double dt = dtf;
    m->prevDt[v] = dtf;
    if (dt == 0.0f) return;
    const int ed = m->ext_dof[v];
    if (ed == 63) {
        /* pos = originalPosition() + translation; orient = rotationQuat(); haltMotion() */
        double s = mt->nomSize;
        v3 t = m->ext_t[v];
        m->pos[v] = v3_(m->ix[v] * s + t.x, m->iy[v] * s + t.y, m->iz[v] * s + t.z);
        m->orient[v] = m->ext_q[v];
        m->linMom[v] = m->angMom[v] = v3_(0, 0, 0);
        return;
    }
    int flags = m->vflags[v];
    q4 orient = m->orient[v];
    v3 pos = m->pos[v], linMom = m->linMom[v], angMom = m->angMom[v];

    /* force() */
    v3 tf = v3_(0, 0, 0);
    for (int d = 0; d < 6; d++) {
        int l = m->vlinks[v][d];
        if (l < 0) continue;
        tf = vadd(tf, (d & 1) ? m->fPos[l] : m->fNeg[l]);
    }
    tf = qrot(orient, tf);
    if (ed >= 0) tf = vadd(tf, m->ext_f[v]);
    v3 vel = vscale(linMom, mt->massInv);
    float gdt = mt->zetaG * mt->c2xSqMxExS;
    tf = vsub(tf, vscale(vel, gdt));
    tf.z += -mt->mass * 9.80665f * mt->gravMult;
    tf = vsub(tf, m->contact[v]);
    m->contact[v] = v3_(0, 0, 0);
    /* cilia: CiliaForce = orient.RotateVec3D(baseCiliaForce + localSignal *
     * shiftCiliaForce) for surface voxels of a cilia material once
     * TurnOnCiliaAfterThisManySeconds has passed (gpu_update_cilia_force
     * VX3_VoxelyzeKernel.cu:846-858, before the voxel update); force() adds
     * CiliaForce * mat->Cilia (VX3_Voxel.cu:376-377). */
    if (m->enableCilia) {
        v3 cf = v3_(0, 0, 0);
        if (m->is_surface[v] && mt->cilia != 0 && !(mt->ciliaOnAfter > m->time))
            cf = qrot(orient, vadd(m->ciliaBase[v], vscale(m->ciliaShift[v], m->sLocal[v])));
        tf = vadd(tf, vscale(cf, mt->cilia));
    }
    /* force field (VX3_Voxel.cu:177-183) at the position entering the step and
     * currentTime; hit is collisionCount; angle, closeness and pairs are 0; absent
     * trees add +0.0 as the reference's constant-zero defaults do */
    const double vars[9] = {pos.x, pos.y, pos.z, (double)m->collisionCount, m->time, m->recentAngle, m->closeness,
                            (double)m->closePairs, (double)m->nv};
    tf.x += m->ff_n[0] ? mt_eval_fifo(m->ff_op[0], m->ff_val[0], m->ff_n[0], vars) : 0.0;
    tf.y += m->ff_n[1] ? mt_eval_fifo(m->ff_op[1], m->ff_val[1], m->ff_n[1], vars) : 0.0;
    tf.z += m->ff_n[2] ? mt_eval_fifo(m->ff_op[2], m->ff_val[2], m->ff_n[2], vars) : 0.0;

    v3 curForce = tf;
    if (flags & VF_FLOOR_ENABLED) {
        float pen = floor_penetration(m, v);
        if (pen >= 0) {
            v3 hv = v3_(vel.x, vel.y, 0);
            float normalForce = (float)(2 * mt->E * mt->nomSize) * pen;
            curForce.z += normalForce - (mt->zetaC * mt->c2xSqMxExS) * vel.z;
            if (flags & VF_FLOOR_STATIC_FRICTION) {
                float sfs = (float)(curForce.x * curForce.x + curForce.y * curForce.y);
                float ffs = (mt->muS * normalForce) * (mt->muS * normalForce);
                if (sfs > ffs) flags &= ~VF_FLOOR_STATIC_FRICTION;
            } else {
                double l = sqrt(hv.x * hv.x + hv.y * hv.y + hv.z * hv.z);
                v3 hn = l > 0 ? v3_(hv.x / l, hv.y / l, hv.z / l) : hv;
                double kf = mt->muK * normalForce;
                curForce = vsub(curForce, vscale(hn, kf));
            }
        } else {
            flags &= ~VF_FLOOR_STATIC_FRICTION;
        }
    }
    v3 fricForce = vsub(curForce, tf);
    linMom = vadd(linMom, vscale(curForce, dt));
    /* CVX_Voxel::timeStep takes a float dt, so its dt*_massInverse product is
     * float (src/old/VX_Voxel.cpp:182); VX3 passes a double (VX3_Voxel.cu:196) */
    double dtMinv = cpu_semantics ? (double)(dtf * mt->massInv) : dt * mt->massInv;
    double dtIinv = cpu_semantics ? (double)(dtf * mt->inertiaInv) : dt * mt->inertiaInv;
    v3 translate = vscale(linMom, dtMinv);
    if ((flags & VF_FLOOR_ENABLED) && floor_penetration(m, v) >= 0) {
        double work = fricForce.x * translate.x + fricForce.y * translate.y;
        double hKe = 0.5 * mt->massInv * (linMom.x * linMom.x + linMom.y * linMom.y);
        if (hKe + work <= 0) flags |= VF_FLOOR_STATIC_FRICTION;
        if (flags & VF_FLOOR_STATIC_FRICTION) {
            linMom.x = linMom.y = 0;
            translate.x = translate.y = 0;
        }
    } else {
        flags &= ~VF_FLOOR_STATIC_FRICTION;
    }
    pos = vadd(pos, translate);
    m->pos[v] = pos;

    /* moment() */
    v3 tm = v3_(0, 0, 0);
    for (int d = 0; d < 6; d++) {
        int l = m->vlinks[v][d];
        if (l < 0) continue;
        tm = vadd(tm, (d & 1) ? m->mPos[l] : m->mNeg[l]);
    }
    tm = qrot(orient, tm);
    if (ed >= 0) tm = vadd(tm, m->ext_m[v]);
    float gdr = mt->zetaG * mt->c2xSqIxExSxSxS;
    tm = vsub(tm, vscale(vscale(angMom, mt->inertiaInv), gdr));
    angMom = vadd(angMom, vscale(tm, dt));
    orient = qmul(q_from_rotvec(vscale(angMom, dtIinv)), orient);
    if (ed > 0) {
        double s = mt->nomSize;
        v3 t = m->ext_t[v];
        if (ed & 1) { pos.x = m->ix[v] * s + t.x; linMom.x = 0; }
        if (ed & 2) { pos.y = m->iy[v] * s + t.y; linMom.y = 0; }
        if (ed & 4) { pos.z = m->iz[v] * s + t.z; linMom.z = 0; }
        if (ed & 56) {  /* any rotation fixed */
            if ((ed & 56) == 56) {
                orient = m->ext_q[v];
                angMom = v3_(0, 0, 0);
            } else {
                v3 rv = q_to_rotvec(orient, NULL);
                if (ed & 8) { rv.x = 0; angMom.x = 0; }
                if (ed & 16) { rv.y = 0; angMom.y = 0; }
                if (ed & 32) { rv.z = 0; angMom.z = 0; }
                orient = q_from_rotvec(rv);
            }
        }
        m->pos[v] = pos;
    }

    /* VX3-only: no rotation while held by static floor friction (VX3_Voxel.cu:259-264) */
    if (!cpu_semantics && (flags & VF_FLOOR_ENABLED) && floor_penetration(m, v) >= 0) {
        if (flags & VF_FLOOR_STATIC_FRICTION) angMom = v3_(0, 0, 0);
    }
    m->orient[v] = orient;
    m->linMom[v] = linMom;
    m->angMom[v] = angMom;
    m->vflags[v] = flags;
    if (m->enableSignals) signal_own(m, v, mt);
}

static void update_voxels(vxo_model *m, float dt, int cpu_semantics) {
    int nv = m->nv;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < nv; v++) {
        if (m->vremoved[v]) continue;              /* gpu_update_voxels :582-589 */
        if (m->vm[m->vmat_of[v]].fixed) continue;
        voxel_update(m, v, dt, cpu_semantics);
    }
}

/* gpu_update_voxels VX3_VoxelyzeKernel.cu:607-621: after its update every
 * present, non-fixed voxel re-evaluates enableAttach, true when each of the
 * five AttachCondition formulas is > 0, at its new position with the step's
 * start time, the collision count so far and the last CycleStep samples */
static void update_attach_flags(vxo_model *m) {
    if (!m->enableAttach) return;
    for (int v = 0; v < m->nv; v++) {
        if (m->vremoved[v] || m->vm[m->vmat_of[v]].fixed) continue;
        const double vars[9] = {m->pos[v].x, m->pos[v].y, m->pos[v].z, (double)m->collisionCount, m->time,
                                m->recentAngle, m->closeness, (double)m->closePairs, (double)m->nv};
        int ok = 1;
        for (int k = 0; k < 5; k++)
            if (m->ac_n[k] && !(mt_eval_fifo(m->ac_op[k], m->ac_val[k], m->ac_n[k], vars) > 0)) ok = 0;
        m->vattach[v] = ok;
    }
}

/* The CycleStep block of doTimeStep (VX3_VoxelyzeKernel.cu:314-334) with
 * updateCurrentCenterOfMass (:477-493) and computeTargetCloseness (:545-562):
 * every CycleStep = int(TempPeriod / dt) steps (at least 1; the reference divides
 * by zero below that) the centre of mass is sampled, recentAngle is the angle
 * between the last two displacements of the sampled centre of mass (0 for the
 * first two samples or a repeated point), and with MaxDistInVoxelLengthsToCountAsPair
 * set, targetCloseness sums 1 / distance over the target voxel pairs and
 * numClosePairs counts those closer than that many voxel lengths. */
static void cycle_sample(vxo_model *m, float dt) {
    int cycle = (int)(m->tempPeriod / (double)dt);
    if (cycle < 1) cycle = 1;
    if ((m->steps + 1) % cycle != 0) return;  /* CurStepCount counts this step */
    m->angleSamples++;
    m->comHist[0] = m->comHist[1];
    m->comHist[1] = m->comCur;
    double c[3];
    vxo_center_of_mass(m, c);
    m->comCur = v3_(c[0], c[1], c[2]);
    v3 A = m->comHist[0], B = m->comHist[1], C = m->comCur;
    int bc = B.x == C.x && B.y == C.y && B.z == C.z, ab = A.x == B.x && A.y == B.y && A.z == B.z;
    if (bc || ab || m->angleSamples < 3) {
        m->recentAngle = 0;
    } else {
        v3 u = vsub(B, A), w = vsub(C, B);
        m->recentAngle = vxm_acos((u.x * w.x + u.y * w.y + u.z * w.z) / (vlen(u) * vlen(w)));
    }
    if (m->maxDistPairs == 0) return;
    double R = m->maxDistPairs * m->voxSize, ret = 0;
    m->closePairs = 0;
    for (int i = 0; i < m->nv; i++) {
        if (!m->vm[m->vmat_of[i]].isTarget) continue;
        for (int j = i + 1; j < m->nv; j++) {
            if (!m->vm[m->vmat_of[j]].isTarget) continue;
            double d = vlen(vsub(m->pos[i], m->pos[j]));
            if (d < R) m->closePairs++;
            ret += 1 / d;
        }
    }
    m->closeness = ret;
}

/* VX3_VoxelyzeKernel::removeVoxels VX3_VoxelyzeKernel.cu:365-399, called at
 * the end of doTimeStep before currentTime advances when SecondaryExperiment
 * is set (:336-349): every voxel of a material whose
 * RemoveFromSimulationAfterThisManySeconds is positive and below the current
 * time is removed with its links (both ends forget the link); the surface
 * list is regenerated before the next step's collisions and cilia (:282-287)
 * when this or a detach (changed) altered the topology. */
static void remove_voxels(vxo_model *m, int changed) {
    for (int v = 0; m->secondary && v < m->nv; v++) {
        const vmat *mt = &m->vm[m->vmat_of[v]];
        if (m->vremoved[v] || !(mt->removeAfter > 0 && mt->removeAfter < m->time)) continue;
        m->vremoved[v] = (int)(m->steps + 1);
        for (int d = 0; d < 6; d++) {
            int l = m->vlinks[v][d];
            if (l < 0) continue;
            m->lremoved[l] = (int)(m->steps + 1);
            int o = (d & 1) ? m->lneg[l] : m->lpos[l];
            m->vlinks[o][d ^ 1] = -1;
            m->vlinks[v][d] = -1;
        }
        changed = 1;
    }
    if (changed) regenerate_surface(m);
}

/* VX3_VoxelyzeKernel::updateDetach / gpu_update_detach
 * (VX3_VoxelyzeKernel.cu:461-475, 946-968): a link whose maxStrain has passed its
 * failure strain (VX3_Link::isFailed, VX3_Link.cu:78) leaves both voxels' link
 * lists.  The reference does it after the collisions of the step; the surface
 * list, cilia and collisions see it from the next step (regenerateSurfaceVoxels
 * runs then), and the voxel update of the step sums the failed link's forces,
 * which are zero (VX3_Link.cu:152-155), so it is applied here at the end of the
 * step. */
static int detach_links(vxo_model *m) {
    int changed = 0;
    for (int l = 0; l < m->nl; l++) {
        if (m->lremoved[l]) continue;
        const lmat *lm = &m->lm[l];
        if (!(lm->epsFail != -1.0f && m->maxStrain[l] > lm->epsFail)) continue;
        m->lremoved[l] = (int)(m->steps + 1);
        for (int d = 0; d < 6; d++) {
            if (m->vlinks[m->lneg[l]][d] == l) m->vlinks[m->lneg[l]][d] = -1;
            if (m->vlinks[m->lpos[l]][d] == l) m->vlinks[m->lpos[l]][d] = -1;
        }
        changed = 1;
    }
    return changed;
}

/* ------------------------------------------------------------------- step */
/* VX3_VoxelyzeKernel::doTimeStep VX3_VoxelyzeKernel.cu:237-359 */
int vxo_step(vxo_model *m, float dt, int nsteps, int flags) {
    for (int s = 0; s < nsteps; s++) {
        if (!(flags & VXO_NO_TEMPERATURE)) update_temperature(m);
        float d = dt;
        if (d < 0) d = vxo_step_dt(m);
        if (d == 0) {  /* CurStepCount still counts the step (:239-241) */
            m->steps++;
            continue;
        }
        update_links(m);
        int diverged = 0;
        for (int l = 0; l < m->nl; l++)
            if (m->strain[l] > 100) { diverged = 1; break; }
        if (diverged) return s;
        int attached = (flags & VXO_ALL_PAIRS) ? update_collisions_all_pairs(m) : update_collisions(m);
        update_voxels(m, d, flags & VXO_CPU_SEMANTICS);
        update_attach_flags(m);
        cycle_sample(m, d);
        /* the end of the step: detached links no longer carry signals (the
         * reference detaches before the voxel update), removal comes last; a
         * new link, a detach or a removal regenerates the surface list */
        int detached = m->enableDetach ? detach_links(m) : 0;
        if (m->enableSignals) signal_deliver(m);
        if (m->secondary || detached || attached) remove_voxels(m, detached || attached);
        m->time += d;
        m->steps++;
    }
    return nsteps;
}

void vxo_update_links(vxo_model *m, double *out) {
    update_temperature(m);  /* the link phase of the next doTimeStep (:240, :259-270) */
    update_links(m);
    if (!out) return;
    for (int l = 0; l < m->nl; l++) {
        v3 f[4] = {m->fNeg[l], m->fPos[l], m->mNeg[l], m->mPos[l]};
        for (int k = 0; k < 4; k++) {
            out[12 * l + 3 * k + 0] = f[k].x;
            out[12 * l + 3 * k + 1] = f[k].y;
            out[12 * l + 3 * k + 2] = f[k].z;
        }
    }
}

/* ----------------------------------------------------------------- access */
void vxo_get_voxels(const vxo_model *m, double *pos3, double *quat4, double *linmom3, double *angmom3) {
    for (int v = 0; v < m->nv; v++) {
        if (pos3) { pos3[3 * v] = m->pos[v].x; pos3[3 * v + 1] = m->pos[v].y; pos3[3 * v + 2] = m->pos[v].z; }
        if (quat4) {
            quat4[4 * v] = m->orient[v].w; quat4[4 * v + 1] = m->orient[v].x;
            quat4[4 * v + 2] = m->orient[v].y; quat4[4 * v + 3] = m->orient[v].z;
        }
        if (linmom3) { linmom3[3 * v] = m->linMom[v].x; linmom3[3 * v + 1] = m->linMom[v].y; linmom3[3 * v + 2] = m->linMom[v].z; }
        if (angmom3) { angmom3[3 * v] = m->angMom[v].x; angmom3[3 * v + 1] = m->angMom[v].y; angmom3[3 * v + 2] = m->angMom[v].z; }
    }
}

void vxo_set_voxels(vxo_model *m, const double *pos3, const double *quat4, const double *linmom3, const double *angmom3) {
    for (int v = 0; v < m->nv; v++) {
        if (pos3) m->pos[v] = v3_(pos3[3 * v], pos3[3 * v + 1], pos3[3 * v + 2]);
        if (quat4) { q4 q = {quat4[4 * v], quat4[4 * v + 1], quat4[4 * v + 2], quat4[4 * v + 3]}; m->orient[v] = q; }
        if (linmom3) m->linMom[v] = v3_(linmom3[3 * v], linmom3[3 * v + 1], linmom3[3 * v + 2]);
        if (angmom3) m->angMom[v] = v3_(angmom3[3 * v], angmom3[3 * v + 1], angmom3[3 * v + 2]);
    }
}

void vxo_get_flags(const vxo_model *m, int *voxel_flags, int *link_flags) {
    if (voxel_flags) memcpy(voxel_flags, m->vflags, sizeof(int) * m->nv);
    if (link_flags) memcpy(link_flags, m->lflags, sizeof(int) * m->nl);
}

void vxo_get_removed(const vxo_model *m, int *voxel_step, int *link_step) {
    if (voxel_step) memcpy(voxel_step, m->vremoved, sizeof(int) * m->nv);
    if (link_step) memcpy(link_step, m->lremoved, sizeof(int) * m->nl);
}

long vxo_collision_count(const vxo_model *m) { return m->collisionCount; }

int vxo_contact_pairs(const vxo_model *m, int *v1v2, int cap) {
    if (v1v2)
        for (int i = 0; i < m->npairs && i < cap; i++) {
            v1v2[2 * i] = m->pairs[2 * i];
            v1v2[2 * i + 1] = m->pairs[2 * i + 1];
        }
    return m->npairs;
}

void vxo_get_signals(const vxo_model *m, double *local) {
    memcpy(local, m->sLocal, sizeof(double) * m->nv);
}

void vxo_get_link_state(const vxo_model *m, double *state9, float *strain) {
    for (int l = 0; l < m->nl; l++) {
        if (state9) {
            double *s = state9 + 9 * l;
            s[0] = m->pos2[l].x; s[1] = m->pos2[l].y; s[2] = m->pos2[l].z;
            s[3] = m->a1v[l].x; s[4] = m->a1v[l].y; s[5] = m->a1v[l].z;
            s[6] = m->a2v[l].x; s[7] = m->a2v[l].y; s[8] = m->a2v[l].z;
        }
        if (strain) strain[l] = m->strain[l];
    }
}

void vxo_get_temperature(const vxo_model *m, float *temp, double *rest_length) {
    if (temp) memcpy(temp, m->temp, sizeof(float) * m->nv);
    if (rest_length) memcpy(rest_length, m->restLen, sizeof(double) * m->nl);
}

void vxo_set_temperature(vxo_model *m, const float *temp) {
    memcpy(m->temp, temp, sizeof(float) * m->nv);
    for (int l = 0; l < m->nl; l++) update_rest_length(m, l);
}

void vxo_get_topology(const vxo_model *m, int *lattice3, int *link_neg, int *link_pos, int *link_axis, int *surface) {
    for (int v = 0; v < m->nv; v++)
        if (lattice3) { lattice3[3 * v] = m->ix[v]; lattice3[3 * v + 1] = m->iy[v]; lattice3[3 * v + 2] = m->iz[v]; }
    if (link_neg) memcpy(link_neg, m->lneg, sizeof(int) * m->nl);
    if (link_pos) memcpy(link_pos, m->lpos, sizeof(int) * m->nl);
    if (link_axis) memcpy(link_axis, m->laxis, sizeof(int) * m->nl);
    if (surface) memcpy(surface, m->surface, sizeof(int) * m->ns);
}

/* updateCurrentCenterOfMass VX3_VoxelyzeKernel.cu:477-493 */
void vxo_center_of_mass(const vxo_model *m, double *com3) {
    double total = 0;
    v3 sum = v3_(0, 0, 0);
    for (int v = 0; v < m->nv; v++) {
        const vmat *mt = &m->vm[m->vmat_of[v]];
        if (!mt->isMeasured) continue;
        double w = mt->mass;
        sum = vadd(sum, vscale(m->pos[v], w));
        total += w;
    }
    if (total == 0) { com3[0] = com3[1] = com3[2] = 0; return; }
    com3[0] = sum.x / total; com3[1] = sum.y / total; com3[2] = sum.z / total;
}

void vxo_get_contact(const vxo_model *m, double *contact3) {
    for (int v = 0; v < m->nv; v++) {
        contact3[3 * v] = m->contact[v].x; contact3[3 * v + 1] = m->contact[v].y; contact3[3 * v + 2] = m->contact[v].z;
    }
}

/* voxel: E nu rho cte muS muK mass massInv sqrtMass inertiaInv 2sqrt(mEL) 2sqrt(IEL^3) */
void vxo_get_voxel_material(const vxo_model *m, int voxel, float *o) {
    const vmat *mt = &m->vm[m->vmat_of[voxel]];
    o[0] = mt->E; o[1] = mt->nu; o[2] = mt->rho; o[3] = mt->alphaCTE; o[4] = mt->muS; o[5] = mt->muK;
    o[6] = mt->mass; o[7] = mt->massInv; o[8] = mt->sqrtMass; o[9] = mt->inertiaInv;
    o[10] = mt->c2xSqMxExS; o[11] = mt->c2xSqIxExSxSxS;
}
/* link: E a1 a2 b1 b2 b3 sqA1 sqA2xIp sqB1 sqB2xFMp sqB3xIp */
void vxo_get_link_material(const vxo_model *m, int link, float *o) {
    const lmat *l = &m->lm[link];
    o[0] = l->E; o[1] = l->a1; o[2] = l->a2; o[3] = l->b1; o[4] = l->b2; o[5] = l->b3;
    o[6] = l->sqA1; o[7] = l->sqA2xIp; o[8] = l->sqB1; o[9] = l->sqB2xFMp; o[10] = l->sqB3xIp;
}

/* the elementary functions the oracle uses (for their accuracy test) */
double vxo_sin(double x) { return vxm_sin(x); }
double vxo_cos(double x) { return vxm_cos(x); }
double vxo_acos(double x) { return vxm_acos(x); }

/* ======================================================================
 * Differentiable mode (no counterpart in the reference; the variant the
 * adjoint differentiates).  Same link equations as link_update above and the
 * same integrator as voxel_update, in double throughout, with:
 *   - per-voxel Young's modulus E_i and phase phi_i (cycles); link modulus
 *     2 E_n E_p / (E_n + E_p); Poisson's ratio 0; constants in double from
 *     the float mass and damping ratios of the voxel's material;
 *   - floor: normal force k w softplus(pen / w), damping -cDampT vz
 *     sigmoid(pen / w), kinetic friction -muK N v_xy / sqrt(|v_xy|^2 + vEps^2);
 *     no static friction;
 *   - temperature held in double; orientation renormalised after each step.
 * ====================================================================== */
typedef struct {
    double mInv, IInv, gDampT, gDampR, cDampT, grav, L, alpha, muK, dmNum;
    int fixed;
} dvox;
typedef struct {
    double E, a2, b1, b2, b3, sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp, tArea, L;
    int fixedBoth;
} dlnk;

struct vxo_diff {
    const vxo_model *m;
    vxo_diff_opts o;
    dvox *vc;
    dlnk *lc;
    double *phase, *temp;
    v3 *pos, *linMom, *angMom;
    q4 *orient;
    v3 *pos2, *a1v, *a2v;
    int *lflags;
    v3 *fN, *fP, *mN, *mP;
    double time, prevDt;
};

static double softplus_d(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }
static double sigmoid_d(double x) {
    if (x >= 0) return 1.0 / (1.0 + exp(-x));
    double e = exp(x);
    return e / (1.0 + e);
}

vxo_diff *vxo_diff_new(const vxo_model *m, const double *phase, const double *E, const vxo_diff_opts *o) {
    vxo_diff *d = (vxo_diff *)xcalloc(1, sizeof(vxo_diff));
    int nv = m->nv, nl = m->nl;
    d->m = m;
    d->o = *o;
    d->vc = (dvox *)xcalloc(nv, sizeof(dvox));
    d->lc = (dlnk *)xcalloc(nl, sizeof(dlnk));
    d->phase = (double *)xcalloc(nv, sizeof(double));
    d->temp = (double *)xcalloc(nv, sizeof(double));
    d->pos = (v3 *)xcalloc(nv, sizeof(v3));
    d->linMom = (v3 *)xcalloc(nv, sizeof(v3));
    d->angMom = (v3 *)xcalloc(nv, sizeof(v3));
    d->orient = (q4 *)xcalloc(nv, sizeof(q4));
    d->pos2 = (v3 *)xcalloc(nl, sizeof(v3));
    d->a1v = (v3 *)xcalloc(nl, sizeof(v3));
    d->a2v = (v3 *)xcalloc(nl, sizeof(v3));
    d->lflags = (int *)xcalloc(nl, sizeof(int));
    d->fN = (v3 *)xcalloc(nl, sizeof(v3));
    d->fP = (v3 *)xcalloc(nl, sizeof(v3));
    d->mN = (v3 *)xcalloc(nl, sizeof(v3));
    d->mP = (v3 *)xcalloc(nl, sizeof(v3));
    for (int v = 0; v < nv; v++) {
        const vmat *mt = &m->vm[m->vmat_of[v]];
        dvox *c = &d->vc[v];
        double mass = mt->mass, L = mt->nomSize, I = mass * L * L / 6.0;
        c->mInv = 1.0 / mass;
        c->IInv = 1.0 / I;
        c->gDampT = mt->zetaG * 2.0 * sqrt(mass * E[v] * L);
        c->gDampR = mt->zetaG * 2.0 * sqrt(I * E[v] * L * L * L);
        c->cDampT = mt->zetaC * 2.0 * sqrt(mass * E[v] * L);
        c->grav = (double)(float)(-mt->mass * 9.80665f * mt->gravMult);
        c->L = L;
        c->alpha = mt->alphaCTE;
        c->muK = mt->muK;
        c->dmNum = 2.0 * sqrt(mass) * mt->zetaI;
        c->fixed = mt->fixed;
        d->phase[v] = phase ? phase[v] : 0.0;
        d->pos[v] = v3_(m->ix[v] * L, m->iy[v] * L, m->iz[v] * L);
        d->orient[v].w = 1;
    }
    for (int l = 0; l < nl; l++) {
        dlnk *c = &d->lc[l];
        double Ea = E[m->lneg[l]], Eb = E[m->lpos[l]];
        double L = m->lm[l].nomSize;
        c->E = 2.0 * Ea * Eb / (Ea + Eb);
        c->L = L;
        c->a2 = c->E * L * L * L / 12.0;
        c->b1 = c->E * L;
        c->b2 = c->E * L * L / 2.0;
        c->b3 = c->E * L * L * L / 6.0;
        c->sqA1 = sqrt(c->E * L);
        c->sqA2xIp = sqrt(c->a2 * L * L / 6.0);
        c->sqB1 = sqrt(c->b1);
        c->sqB2xFMp = sqrt(c->b2 * L / 2.0);
        c->sqB3xIp = sqrt(c->b3 * L * L / 6.0);
        c->tArea = L * L;
        c->fixedBoth = d->vc[m->lneg[l]].fixed && d->vc[m->lpos[l]].fixed;
        d->lflags[l] = LF_SMALL_ANGLE;
    }
    return d;
}

void vxo_diff_free(vxo_diff *d) {
    if (!d) return;
    free(d->vc); free(d->lc); free(d->phase); free(d->temp); free(d->pos); free(d->linMom); free(d->angMom);
    free(d->orient); free(d->pos2); free(d->a1v); free(d->a2v); free(d->lflags);
    free(d->fN); free(d->fP); free(d->mN); free(d->mP);
    free(d);
}

static void diff_link(vxo_diff *d, int l) {
    const vxo_model *m = d->m;
    const dlnk *c = &d->lc[l];
    if (c->fixedBoth) return;
    int n = m->lneg[l], p = m->lpos[l], axis = m->laxis[l];
    v3 oldPos2 = d->pos2[l], oldA1v = d->a1v[l], oldA2v = d->a2v[l];
    int flags = d->lflags[l];
    const dvox *vn = &d->vc[n], *vp = &d->vc[p];
    double restLen = 0.5 * (vn->L * (1 + d->temp[n] * vn->alpha) + vp->L * (1 + d->temp[p] * vp->alpha));

    v3 pos2 = to_axis_x(vsub(d->pos[p], d->pos[n]), axis);
    q4 angle1 = to_axis_x_q(d->orient[n], axis);
    q4 angle2 = to_axis_x_q(d->orient[p], axis);
    q4 totalRot = qconj(angle1);
    pos2 = qrot(totalRot, pos2);
    angle2 = qmul(totalRot, angle2);
    angle1.w = 1; angle1.x = angle1.y = angle1.z = 0;

    float smallTurn = (float)((fabs(pos2.z) + fabs(pos2.y)) / pos2.x);
    float extendPerc = (float)(fabs(1 - pos2.x / restLen));
    int smallAngle = (flags & LF_SMALL_ANGLE) != 0;
    if (!smallAngle && smallTurn < SA_BOND_BEND_RAD && extendPerc < SA_BOND_EXT_PERC) {
        smallAngle = 1;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    } else if (smallAngle && (smallTurn > HYSTERESIS_FACTOR * SA_BOND_BEND_RAD ||
                              extendPerc > HYSTERESIS_FACTOR * SA_BOND_EXT_PERC)) {
        smallAngle = 0;
        flags &= ~LF_LOCAL_VELOCITY_VALID;
    }
    if (smallAngle) {
        pos2.x -= restLen;
    } else {
        angle1 = q_from_angle_to_posx(pos2, NULL);
        angle2 = qmul(angle1, angle2);
        pos2 = v3_(vlen(pos2) - restLen, 0, 0);
    }
    v3 a1v = q_to_rotvec(angle1, NULL), a2v = q_to_rotvec(angle2, NULL);
    flags = smallAngle ? (flags | LF_SMALL_ANGLE) : (flags & ~LF_SMALL_ANGLE);
    d->pos2[l] = pos2;
    d->a1v[l] = a1v;
    d->a2v[l] = a2v;
    v3 dP = vscale(vsub(pos2, oldPos2), 0.5), dA1 = vscale(vsub(a1v, oldA1v), 0.5), dA2 = vscale(vsub(a2v, oldA2v), 0.5);
    double stress = c->E * (pos2.x / restLen);
    v3 fN = v3_(stress * c->tArea, c->b1 * pos2.y - c->b2 * (a1v.z + a2v.z), c->b1 * pos2.z + c->b2 * (a1v.y + a2v.y));
    v3 fP = vscale(fN, -1);
    v3 mN = v3_(c->a2 * (a2v.x - a1v.x), -c->b2 * pos2.z - c->b3 * (2 * a1v.y + a2v.y),
                c->b2 * pos2.y - c->b3 * (2 * a1v.z + a2v.z));
    v3 mP = v3_(c->a2 * (a1v.x - a2v.x), -c->b2 * pos2.z - c->b3 * (a1v.y + 2 * a2v.y),
                c->b2 * pos2.y - c->b3 * (a1v.z + 2 * a2v.z));
    if (flags & LF_LOCAL_VELOCITY_VALID) {
        v3 pc = v3_(c->sqA1 * dP.x, c->sqB1 * dP.y - c->sqB2xFMp * (dA1.z + dA2.z),
                    c->sqB1 * dP.z + c->sqB2xFMp * (dA1.y + dA2.y));
        double dmN = vn->dmNum / d->prevDt, dmP = vp->dmNum / d->prevDt;
        fN = vadd(fN, vscale(pc, dmN));
        fP = vsub(fP, vscale(pc, dmP));
        v3 tN = v3_(-c->sqA2xIp * (dA2.x - dA1.x), c->sqB2xFMp * dP.z + c->sqB3xIp * (2 * dA1.y + dA2.y),
                    -c->sqB2xFMp * dP.y + c->sqB3xIp * (2 * dA1.z + dA2.z));
        v3 tP = v3_(c->sqA2xIp * (dA2.x - dA1.x), c->sqB2xFMp * dP.z + c->sqB3xIp * (dA1.y + 2 * dA2.y),
                    -c->sqB2xFMp * dP.y + c->sqB3xIp * (dA1.z + 2 * dA2.z));
        mN = vsub(mN, vscale(tN, 0.5 * dmN));
        mP = vsub(mP, vscale(tP, 0.5 * dmP));
    } else {
        flags |= LF_LOCAL_VELOCITY_VALID;
    }
    if (!smallAngle) {
        fN = qrotinv(angle1, fN);
        mN = qrotinv(angle1, mN);
    }
    fP = qrotinv(angle2, fP);
    mP = qrotinv(angle2, mP);
    d->fN[l] = to_axis_original(fN, axis);
    d->fP[l] = to_axis_original(fP, axis);
    d->mN[l] = to_axis_original(mN, axis);
    d->mP[l] = to_axis_original(mP, axis);
    d->lflags[l] = flags;
}

static void diff_voxel(vxo_diff *d, int v, double dt) {
    const vxo_model *m = d->m;
    const dvox *c = &d->vc[v];
    if (c->fixed) return;
    v3 F = v3_(0, 0, 0), M = v3_(0, 0, 0);
    for (int k = 0; k < 6; k++) {
        int l = m->vlinks[v][k];
        if (l < 0) continue;
        F = vadd(F, (k & 1) ? d->fP[l] : d->fN[l]);
        M = vadd(M, (k & 1) ? d->mP[l] : d->mN[l]);
    }
    q4 q = d->orient[v];
    v3 p = d->linMom[v], h = d->angMom[v];
    v3 tf = qrot(q, F);
    v3 vel = vscale(p, c->mInv);
    tf = vsub(tf, vscale(vel, c->gDampT));
    tf.z += c->grav;
    if (d->o.floor_enabled) {
        double L = c->L, b = L * (1 + d->temp[v] * c->alpha);
        double pen = b / 2 - L / 2 - d->pos[v].z;
        double x = pen / d->o.floor_w;
        double N = d->o.floor_k * d->o.floor_w * softplus_d(x);
        tf.z += N - c->cDampT * vel.z * sigmoid_d(x);
        double r = sqrt(vel.x * vel.x + vel.y * vel.y + d->o.fric_v * d->o.fric_v);
        tf.x -= c->muK * N * vel.x / r;
        tf.y -= c->muK * N * vel.y / r;
    }
    p = vadd(p, vscale(tf, dt));
    d->pos[v] = vadd(d->pos[v], vscale(p, dt * c->mInv));
    v3 tm = vsub(qrot(q, M), vscale(h, c->IInv * c->gDampR));
    h = vadd(h, vscale(tm, dt));
    q = qmul(q_from_rotvec(vscale(h, dt * c->IInv)), q);
    double n = 1.0 / sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    q.w *= n; q.x *= n; q.y *= n; q.z *= n;
    d->orient[v] = q;
    d->linMom[v] = p;
    d->angMom[v] = h;
}

void vxo_diff_step(vxo_diff *d, int nsteps) {
    const vxo_model *m = d->m;
    const double twoPiF = (double)(2 * 3.1415926f);
    for (int s = 0; s < nsteps; s++) {
        for (int v = 0; v < m->nv; v++) {
            double T = 0;
            if (d->o.vary_temp && !d->vc[v].fixed) {
                T = d->o.temp_amplitude * sin(twoPiF * (d->time / d->o.temp_period + d->phase[v]));
                if (!d->o.enable_expansion && T > 0) T = 0;
            }
            d->temp[v] = T;
        }
        for (int l = 0; l < m->nl; l++) diff_link(d, l);
        for (int v = 0; v < m->nv; v++) diff_voxel(d, v, d->o.dt);
        d->time += d->o.dt;
        d->prevDt = d->o.dt;
    }
}

void vxo_diff_get(const vxo_diff *d, double *pos3, double *quat4, double *linmom3, double *angmom3) {
    for (int v = 0; v < d->m->nv; v++) {
        if (pos3) { pos3[3 * v] = d->pos[v].x; pos3[3 * v + 1] = d->pos[v].y; pos3[3 * v + 2] = d->pos[v].z; }
        if (quat4) { quat4[4 * v] = d->orient[v].w; quat4[4 * v + 1] = d->orient[v].x; quat4[4 * v + 2] = d->orient[v].y; quat4[4 * v + 3] = d->orient[v].z; }
        if (linmom3) { linmom3[3 * v] = d->linMom[v].x; linmom3[3 * v + 1] = d->linMom[v].y; linmom3[3 * v + 2] = d->linMom[v].z; }
        if (angmom3) { angmom3[3 * v] = d->angMom[v].x; angmom3[3 * v + 1] = d->angMom[v].y; angmom3[3 * v + 2] = d->angMom[v].z; }
    }
}

double vxo_diff_mass(const vxo_diff *d, int v) { return 1.0 / d->vc[v].mInv; }

void vxo_diff_set(vxo_diff *d, const double *pos3, const double *quat4, const double *linmom3, const double *angmom3) {
    for (int v = 0; v < d->m->nv; v++) {
        if (pos3) d->pos[v] = v3_(pos3[3 * v], pos3[3 * v + 1], pos3[3 * v + 2]);
        if (quat4) { q4 q = {quat4[4 * v], quat4[4 * v + 1], quat4[4 * v + 2], quat4[4 * v + 3]}; d->orient[v] = q; }
        if (linmom3) d->linMom[v] = v3_(linmom3[3 * v], linmom3[3 * v + 1], linmom3[3 * v + 2]);
        if (angmom3) d->angMom[v] = v3_(angmom3[3 * v], angmom3[3 * v + 1], angmom3[3 * v + 2]);
    }
}
