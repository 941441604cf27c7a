/*
 * vx3mi_math.h -- double-precision sin, cos and acos that round identically
 * on the host (gcc/clang, x86-64) and on gfx950 (hipcc), provided the including
 * translation unit is built without floating-point contraction
 * (-ffp-contract=off).  The FP64 timestep uses them wherever the reference
 * calls a transcendental (VX3_Quat3D.h:347-432, VX3_VoxelyzeKernel.cu:639), so
 * an FP64 run on the device and an FP64 run on the host agree bit for bit;
 * vendor libm implementations differ in the last ulp, which the timestep's
 * hysteresis and friction switches can amplify.
 *
 * Accuracy (checked against the platform libm): within 2 ulp on the ranges the
 * timestep uses.
 *   sin/cos: Cody-Waite reduction by pi/2 with the three-part split of fdlibm
 *            (|x| < 2^19 * pi/2), then the Taylor series to degree 17/18 on
 *            [-pi/4, pi/4] (truncation below 1e-19 relative).
 *   acos:    acos x = pi/2 - asin x for |x| <= 1/2, 2 asin(sqrt((1-x)/2)) for
 *            x > 1/2, pi - 2 asin(sqrt((1+x)/2)) for x < -1/2, with the asin
 *            Maclaurin series (28 terms, |y| <= 1/2).
 */
#ifndef VX3MI_MATH_H
#define VX3MI_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define VXM_FN static inline __host__ __device__
#else
#define VXM_FN static inline
#endif

/* pi/2 in three parts of 33 bits each (fdlibm's pio2_1, pio2_2, pio2_3) */
#define VXM_PIO2_1 1.57079632673412561417e+00
#define VXM_PIO2_2 6.07710050630396597660e-11
#define VXM_PIO2_3 2.02226624871116645580e-21
#define VXM_PIO2_HI 1.57079632679489655800e+00
#define VXM_PIO2_LO 6.12323399573676603587e-17
#define VXM_PI_HI 3.14159265358979311600e+00
#define VXM_PI_LO 1.22464679914735320717e-16

/* r in [-pi/4, pi/4] */
VXM_FN double vxm_sin_kernel(double r) {
    const double z = r * r;
    double p = 1.0 / 355687428096000.0;  /* 1/17! */
    p = p * z - 1.0 / 1307674368000.0;   /* 1/15! */
    p = p * z + 1.0 / 6227020800.0;      /* 1/13! */
    p = p * z - 1.0 / 39916800.0;        /* 1/11! */
    p = p * z + 1.0 / 362880.0;          /* 1/9!  */
    p = p * z - 1.0 / 5040.0;            /* 1/7!  */
    p = p * z + 1.0 / 120.0;             /* 1/5!  */
    p = p * z - 1.0 / 6.0;               /* 1/3!  */
    return r + (r * z) * p;
}

VXM_FN double vxm_cos_kernel(double r) {
    const double z = r * r;
    double p = -1.0 / 6402373705728000.0; /* 1/18! */
    p = p * z + 1.0 / 20922789888000.0;   /* 1/16! */
    p = p * z - 1.0 / 87178291200.0;      /* 1/14! */
    p = p * z + 1.0 / 479001600.0;        /* 1/12! */
    p = p * z - 1.0 / 3628800.0;          /* 1/10! */
    p = p * z + 1.0 / 40320.0;            /* 1/8!  */
    p = p * z - 1.0 / 720.0;              /* 1/6!  */
    p = p * z + 1.0 / 24.0;               /* 1/4!  */
    /* 1 - z/2 + z^2 p: w = 1 - z/2 rounded, its rounding error recovered exactly */
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * p);
}

/* quadrant q (mod 4) and reduced argument r = x - q*pi/2 */
VXM_FN double vxm_reduce(double x, int *q) {
    const double k = floor(x * 6.36619772367581382433e-01 + 0.5); /* 2/pi */
    *q = (int)(((long long)k) & 3);
    return ((x - k * VXM_PIO2_1) - k * VXM_PIO2_2) - k * VXM_PIO2_3;
}

VXM_FN double vxm_sin(double x) {
    if (x - x != 0.0) return x - x; /* NaN for NaN and +-inf */
    int q;
    const double r = vxm_reduce(x, &q);
    switch (q) {
    case 0: return vxm_sin_kernel(r);
    case 1: return vxm_cos_kernel(r);
    case 2: return -vxm_sin_kernel(r);
    default: return -vxm_cos_kernel(r);
    }
}

VXM_FN double vxm_cos(double x) {
    if (x - x != 0.0) return x - x;
    int q;
    const double r = vxm_reduce(x, &q);
    switch (q) {
    case 0: return vxm_cos_kernel(r);
    case 1: return -vxm_sin_kernel(r);
    case 2: return -vxm_cos_kernel(r);
    default: return vxm_sin_kernel(r);
    }
}

/* asin y for |y| <= 1/2: sum_n (2n)! / (4^n (n!)^2 (2n+1)) y^(2n+1) */
VXM_FN double vxm_asin_small(double y) {
    const double z = y * y;
    double p = 0.0019650336162772837;
    p = p * z + 0.0020776610325181676;
    p = p * z + 0.0022014739737101384;
    p = p * z + 0.002338091892111975;
    p = p * z + 0.0024894486782468836;
    p = p * z + 0.0026578706382072901;
    p = p * z + 0.0028461784011089421;
    p = p * z + 0.0030578216492580306;
    p = p * z + 0.0032970595034734849;
    p = p * z + 0.0035692053938259347;
    p = p * z + 0.0038809645588376691;
    p = p * z + 0.0042409070936793632;
    p = p * z + 0.0046601434869150962;
    p = p * z + 0.0051533096823199046;
    p = p * z + 0.0057400376708419236;
    p = p * z + 0.0064472103118896487;
    p = p * z + 0.0073125258735988454;
    p = p * z + 0.0083903358096168151;
    p = p * z + 0.0097616095291940784;
    p = p * z + 0.011551800896139705;
    p = p * z + 0.013964843750000001;
    p = p * z + 0.017352764423076924;
    p = p * z + 0.022372159090909092;
    p = p * z + 0.030381944444444444;
    p = p * z + 0.044642857142857144;
    p = p * z + 0.074999999999999997;
    p = p * z + 0.16666666666666666;
    return y + (y * z) * p;
}

VXM_FN double vxm_acos(double x) {
    if (x != x) return x;
    if (x >= 1.0) return 0.0;
    if (x <= -1.0) return VXM_PI_HI + VXM_PI_LO;
    if (x > 0.5) return 2.0 * vxm_asin_small(sqrt((1.0 - x) * 0.5));
    if (x < -0.5) return VXM_PI_HI - (2.0 * vxm_asin_small(sqrt((1.0 + x) * 0.5)) - VXM_PI_LO);
    return VXM_PIO2_HI - (vxm_asin_small(x) - VXM_PIO2_LO);
}

#endif
