// fp32_primitives_check.cpp -- host build of the rotation primitives of the
// timestep (vx3mi_physics.h): the FP32 forms (series for small rotations,
// the trigonometry-free align-to-x rotation, the 2 sqrt(2)/sqrt(1+w)
// rotation-vector branch) against the FP64 forms, which follow the
// reference's formulas.  Built by tests/test_math.py with g++.
#include <cmath>
#include <random>

#include "vx3mi_physics.h"

using namespace vx3mi;

namespace {
double qdist(Q4<float> a, Q4<double> b) {
    double d1 = std::fabs(a.w - b.w) + std::fabs(a.x - b.x) + std::fabs(a.y - b.y) + std::fabs(a.z - b.z);
    double d2 = std::fabs(a.w + b.w) + std::fabs(a.x + b.x) + std::fabs(a.y + b.y) + std::fabs(a.z + b.z);
    return d1 < d2 ? d1 : d2;
}
}  // namespace

// worst absolute difference of the float sine of a phase in cycles from the
// double sine of the same float argument 2 * 3.1415926f * c (what sinf of the
// float path's argument approximates), over n phases spread over [0, 1.25) and
// the float neighbours of the quarter points; out[1]: the mean signed
// difference (a bias of the reduction would show here)
extern "C" double check_sin_cycles(int n, double *mean) {
    double worst = 0, sum = 0;
    const float k = 2 * 3.1415926f;
    for (int i = 0; i < n + 10; i++) {
        float c = i < n ? (float)(1.25 * (i + 0.5) / n) : 0.25f * (float)((i - n) / 2);
        if (i >= n && (i - n) % 2) c = std::nextafterf(c, 2.0f);
        const double want = std::sin((double)(k * c));
        const double e = (double)sin_cycles_f32(c) - want;
        worst = std::fmax(worst, std::fabs(e));
        sum += e;
    }
    *mean = sum / (n + 10);
    return worst;
}

// worst relative difference of the float atan2 of the rotation-vector angle
// from the double atan2, over n angles spread over (0, pi] on the unit circle
// (the unit quaternions it sees) and at scales 1e-3 .. 1e3 off it
extern "C" double check_atan2(int n) {
    double worst = 0;
    for (int i = 0; i < n; i++) {
        const double a = M_PI * (i + 1.0) / n;
        for (double sc : {1.0, 1e-3, 1e3}) {
            const float y = (float)(sc * std::sin(a)), x = (float)(sc * std::cos(a));
            if (y < 0 || (x == 0 && y == 0)) continue;
            const double want = std::atan2((double)y, (double)x);
            worst = std::fmax(worst, std::fabs((double)atan2_nonneg_f32(y, x) - want) / want);
        }
    }
    return worst;
}

// worst absolute differences over n random inputs: out[0] from_rotvec,
// out[1] from_angle_to_posx, out[2] to_rotvec (relative to |v|)
extern "C" void check_primitives(unsigned long long seed, int n, double *out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1, 1), lg(-6, 0.5);
    out[0] = out[1] = out[2] = 0;
    for (int i = 0; i < n; i++) {
        // rotation vectors from 1e-6 to ~3 rad
        double sc = std::pow(10.0, lg(rng));
        V3<double> v = {u(rng) * sc, u(rng) * sc, u(rng) * sc};
        V3<float> vf = {(float)v.x, (float)v.y, (float)v.z};
        Q4<double> qd = from_rotvec(v);
        Q4<float> qf = from_rotvec(vf);
        out[0] = std::fmax(out[0], qdist(qf, qd));
        // link vectors: mostly along +x, some far off (beyond the small-angle branch)
        V3<double> r = {1.0 + 0.5 * u(rng), u(rng) * sc, u(rng) * sc};
        V3<float> rf = {(float)r.x, (float)r.y, (float)r.z};
        out[1] = std::fmax(out[1], qdist(from_angle_to_posx(rf), from_angle_to_posx(r)));
        // rotation vector of a unit quaternion
        Q4<double> q = qd;
        Q4<float> qq = {(float)q.w, (float)q.x, (float)q.y, (float)q.z};
        V3<double> a = to_rotvec(q);
        V3<float> af = to_rotvec(qq);
        double na = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
        if (na > 0) {
            double e = std::sqrt((af.x - a.x) * (af.x - a.x) + (af.y - a.y) * (af.y - a.y) + (af.z - a.z) * (af.z - a.z));
            out[2] = std::fmax(out[2], e / na);
        }
    }
}
