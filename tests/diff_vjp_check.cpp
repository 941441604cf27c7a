// Host build of the differentiable-mode equations (vx3mi_diff.h, double) for
// the CPU tests: every hand-written vector-Jacobian product is compared with
// central finite differences of its forward function on random inputs.
// Built by tests/test_diff_host.py with g++; not part of the product.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "vx3mi_diff.h"

using namespace vx3mi;
typedef double D;

static std::mt19937_64 rng;
static D U(D a, D b) { return std::uniform_real_distribution<D>(a, b)(rng); }

static Q4<D> rand_rot(D maxAngle) {
    V3<D> ax = {U(-1, 1), U(-1, 1), U(-1, 1)};
    D n = len(ax);
    D a = U(0, maxAngle);
    Q4<D> q = {cos(a / 2), ax.x / n * sin(a / 2), ax.y / n * sin(a / 2), ax.z / n * sin(a / 2)};
    return q;
}

// Error of an analytic derivative g against central differences of f at step
// h and h/4: |g - fd| relative to |fd| plus ten times the spread between the
// two step sizes (the rounding noise of the difference quotient) plus a floor.
// A correct derivative scores well below 1; a wrong one scores O(1).
struct Err {
    D worst = 0;
    template <typename F> void check(D g, F f, D h, D floor) {
        const D fd1 = (f(h) - f(-h)) / (2 * h), fd2 = (f(h / 4) - f(-h / 4)) / (h / 2);
        const D e = fabs(g - fd2) / (fabs(fd2) + 10 * fabs(fd1 - fd2) + floor);
        if (e > worst) worst = e;
    }
};

// ---- link ------------------------------------------------------------------------
struct LinkIn {
    D x[27];  // du(3) qn(4) qp(4) tN tP P0(3) A10(3) A20(3) E Ln? -> see pack
};

static DLink<D> link_consts(D E, D L, int axis) {
    DLink<D> c;
    c.E = E;
    c.a2 = E * L * L * L / 12.0;
    c.b1 = E * L;
    c.b2 = E * L * L / 2.0;
    c.b3 = E * L * L * L / 6.0;
    c.sqA1 = sqrt(E * L);
    c.sqA2xIp = sqrt(c.a2 * L * L / 6.0);
    c.sqB1 = sqrt(c.b1);
    c.sqB2xFMp = sqrt(c.b2 * L / 2.0);
    c.sqB3xIp = sqrt(c.b3 * L * L / 6.0);
    c.tArea = L * L;
    c.L = L;
    c.nomN = c.nomP = L;
    c.alphaN = 0.01;
    c.alphaP = 0.013;
    c.dmNumN = 0.0063;
    c.dmNumP = 0.0071;
    c.axis = axis;
    c.fixedBoth = 0;
    return c;
}

// unpack 24 inputs: du, qn, qp, tN, tP, P0, A10, A20, logE
static D link_loss(const D *x, int axis, int flags, const D *w, DLinkEval<D> *keep) {
    DLink<D> c = link_consts(exp(x[23]), 0.01, axis);
    V3<D> du = {x[0], x[1], x[2]};
    Q4<D> qn = {x[3], x[4], x[5], x[6]}, qp = {x[7], x[8], x[9], x[10]};
    V3<D> P0 = {x[13], x[14], x[15]}, A10 = {x[16], x[17], x[18]}, A20 = {x[19], x[20], x[21]};
    DLinkEval<D> e;
    d_link_forward(c, du, qn, qp, x[11], x[12], P0, A10, A20, flags, 1.4e-5, e);
    if (keep) *keep = e;
    const V3<D> o[7] = {e.FN, e.FP, e.MN, e.MP, e.P, e.A1v, e.A2v};
    D s = 0;
    for (int k = 0; k < 7; k++) s += w[3 * k] * o[k].x + w[3 * k + 1] * o[k].y + w[3 * k + 2] * o[k].z;
    return s;
}

extern "C" double check_link(int seed, int axis, int flags, double bend, double stretch) {
    rng.seed(seed);
    D x[24];
    const D L = 0.01;
    Q4<D> qn = rand_rot(bend), qp = rand_rot(bend);
    V3<D> du0 = {U(-1, 1) * stretch * L, U(-1, 1) * bend * L, U(-1, 1) * bend * L};
    V3<D> du = to_axis_original(du0, axis);
    D tmp[24] = {du.x, du.y, du.z, qn.w, qn.x, qn.y, qn.z, qp.w, qp.x, qp.y, qp.z, U(-30, 0), U(-30, 0),
                 U(-1e-4, 1e-4), U(-1e-4, 1e-4), U(-1e-4, 1e-4), U(-0.02, 0.02), U(-0.02, 0.02), U(-0.02, 0.02),
                 U(-0.02, 0.02), U(-0.02, 0.02), U(-0.02, 0.02), 0, log(U(1e6, 1e7))};
    for (int i = 0; i < 24; i++) x[i] = tmp[i];
    // force scales differ by orders of magnitude; weight outputs to comparable size
    D w[21];
    for (int i = 0; i < 21; i++) w[i] = U(-1, 1) * (i < 12 ? 1.0 : 1e3);
    DLinkEval<D> e;
    const D base = link_loss(x, axis, flags, w, &e);
    (void)base;
    DLink<D> c = link_consts(exp(x[23]), L, axis);
    V3<D> W[7];
    for (int k = 0; k < 7; k++) W[k] = v3<D>(w[3 * k], w[3 * k + 1], w[3 * k + 2]);
    DLinkBar<D> b;
    d_link_vjp(c, e, Q4<D>{x[3], x[4], x[5], x[6]}, Q4<D>{x[7], x[8], x[9], x[10]}, v3<D>(x[13], x[14], x[15]),
               v3<D>(x[16], x[17], x[18]), v3<D>(x[19], x[20], x[21]), 1.4e-5, W[0], W[1], W[2], W[3], W[4], W[5], W[6],
               b);
    D g[24] = {b.du.x, b.du.y, b.du.z, b.qn.w, b.qn.x, b.qn.y, b.qn.z, b.qp.w, b.qp.x, b.qp.y, b.qp.z, b.tN, b.tP,
               b.P0.x, b.P0.y, b.P0.z, b.A10.x, b.A10.y, b.A10.z, b.A20.x, b.A20.y, b.A20.z, 0, b.Elin};
    Err err;
    bool sameBranch = true;
    for (int i = 0; i < 24; i++) {
        if (i == 22) continue;
        const D h = (i < 3 ? 1e-9 : (i == 11 || i == 12 ? 1e-4 : (i == 23 ? 1e-6 : 1e-8)));
        err.check(g[i], [&](D dx) {
            D xp[24];
            for (int k = 0; k < 24; k++) xp[k] = x[k];
            xp[i] += dx;
            DLinkEval<D> ep;
            const D v = link_loss(xp, axis, flags, w, &ep);
            sameBranch &= ep.small == e.small && ep.lvv == e.lvv;
            return v;
        }, h, 1e-9);
    }
    return sameBranch ? err.worst : -1.0;
}

extern "C" int link_branch(int seed, int axis, int flags, double bend, double stretch) {
    rng.seed(seed);
    const D L = 0.01;
    Q4<D> qn = rand_rot(bend), qp = rand_rot(bend);
    V3<D> du0 = {U(-1, 1) * stretch * L, U(-1, 1) * bend * L, U(-1, 1) * bend * L};
    V3<D> du = to_axis_original(du0, axis);
    DLink<D> c = link_consts(1e6, L, axis);
    DLinkEval<D> e;
    d_link_forward(c, du, qn, qp, -10.0, -20.0, v3<D>(0, 0, 0), v3<D>(0, 0, 0), v3<D>(0, 0, 0), flags, 1.4e-5, e);
    return e.small + 2 * e.lvv;
}

// ---- voxel -----------------------------------------------------------------------
static D vox_loss(const D *x, const D *w, int floorOn, int fixed) {
    // x: u(3) q(4) p(3) h(3) F(3) M(3) temp logE  = 22
    DiffOpts o{1e5, 1e-4, 1e-3, floorOn, 0, 39.0, 0.2, 1};
    const D E = exp(x[21]), L = 0.01, m = 1e-3, I = m * L * L / 6;
    DVox<D> c;
    c.mInv = 1 / m;
    c.IInv = 1 / I;
    c.gDampT = 1.0 * 2 * sqrt(m * E * L);
    c.gDampR = 1.0 * 2 * sqrt(I * E * L * L * L);
    c.cDampT = 0.8 * 2 * sqrt(m * E * L);
    c.grav = -m * 9.81;
    c.L = L;
    c.alpha = 0.01;
    c.muK = 0.5;
    c.dmNum = 0;
    c.iz = 0;
    c.fixed = fixed;
    DVoxState<D> s, out;
    s.u = v3<D>(x[0], x[1], x[2]);
    s.q = Q4<D>{x[3], x[4], x[5], x[6]};
    s.p = v3<D>(x[7], x[8], x[9]);
    s.h = v3<D>(x[10], x[11], x[12]);
    d_voxel_forward(c, o, (D)1.4e-5, x[20], v3<D>(x[13], x[14], x[15]), v3<D>(x[16], x[17], x[18]), s, out);
    const D o13[13] = {out.u.x, out.u.y, out.u.z, out.q.w, out.q.x, out.q.y, out.q.z,
                       out.p.x, out.p.y, out.p.z, out.h.x, out.h.y, out.h.z};
    D r = 0;
    for (int k = 0; k < 13; k++) r += w[k] * o13[k];
    return r;
}

extern "C" double check_voxel(int seed, int floorOn, int fixed, double height) {
    rng.seed(seed);
    Q4<D> q = rand_rot(0.5);
    D x[22] = {U(-1e-3, 1e-3), U(-1e-3, 1e-3), height + U(-2e-4, 2e-4), q.w, q.x, q.y, q.z,
               U(-1e-5, 1e-5), U(-1e-5, 1e-5), U(-1e-5, 1e-5), U(-1e-9, 1e-9), U(-1e-9, 1e-9), U(-1e-9, 1e-9),
               U(-1e-2, 1e-2), U(-1e-2, 1e-2), U(-1e-2, 1e-2), U(-1e-5, 1e-5), U(-1e-5, 1e-5), U(-1e-5, 1e-5),
               0, U(-30, 0), log(U(1e6, 1e7))};
    // weights scaled so every output matters
    D w[13];
    for (int k = 0; k < 13; k++) w[k] = U(-1, 1) * (k < 3 ? 1e3 : (k < 7 ? 1.0 : (k < 10 ? 1e3 : 1e6)));
    const D E = exp(x[21]), L = 0.01, m = 1e-3, I = m * L * L / 6;
    DiffOpts o{1e5, 1e-4, 1e-3, floorOn, 0, 39.0, 0.2, 1};
    DVox<D> c;
    c.mInv = 1 / m; c.IInv = 1 / I;
    c.gDampT = 2 * sqrt(m * E * L); c.gDampR = 2 * sqrt(I * E * L * L * L); c.cDampT = 0.8 * 2 * sqrt(m * E * L);
    c.grav = -m * 9.81; c.L = L; c.alpha = 0.01; c.muK = 0.5; c.dmNum = 0; c.iz = 0; c.fixed = fixed;
    DVoxState<D> s, ob;
    s.u = v3<D>(x[0], x[1], x[2]);
    s.q = Q4<D>{x[3], x[4], x[5], x[6]};
    s.p = v3<D>(x[7], x[8], x[9]);
    s.h = v3<D>(x[10], x[11], x[12]);
    ob.u = v3<D>(w[0], w[1], w[2]);
    ob.q = Q4<D>{w[3], w[4], w[5], w[6]};
    ob.p = v3<D>(w[7], w[8], w[9]);
    ob.h = v3<D>(w[10], w[11], w[12]);
    DVoxBar<D> r;
    d_voxel_vjp(c, o, (D)1.4e-5, x[20], v3<D>(x[13], x[14], x[15]), v3<D>(x[16], x[17], x[18]), s, ob, r);
    D g[22] = {r.s.u.x, r.s.u.y, r.s.u.z, r.s.q.w, r.s.q.x, r.s.q.y, r.s.q.z, r.s.p.x, r.s.p.y, r.s.p.z,
               r.s.h.x, r.s.h.y, r.s.h.z, r.F.x, r.F.y, r.F.z, r.M.x, r.M.y, r.M.z, 0, r.temp, r.Esq * 0.5};
    Err err;
    for (int i = 0; i < 22; i++) {
        if (i == 19) continue;
        const D h = fmax(fabs(x[i]), 1e-9) * 1e-5;
        err.check(g[i], [&](D dx) {
            D xp[22];
            for (int k = 0; k < 22; k++) xp[k] = x[k];
            xp[i] += dx;
            return vox_loss(xp, w, floorOn, fixed);
        }, h, 1e-9);
    }
    return err.worst;
}

// ---- primitives ---------------------------------------------------------------------
extern "C" double check_primitives(int seed) {
    rng.seed(seed);
    Err err;
    const D h = 1e-6;
    auto q4dot = [](Q4<D> a, Q4<D> b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; };
    for (int rep = 0; rep < 20; rep++) {
        Q4<D> q = {U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
        V3<D> f = {U(-1, 1), U(-1, 1), U(-1, 1)}, yb = {U(-1, 1), U(-1, 1), U(-1, 1)};
        for (int inv = 0; inv < 2; inv++) {
            Q4<D> qb = {0, 0, 0, 0};
            V3<D> fb = {0, 0, 0};
            if (inv) qrotinv_vjp(q, f, yb, qb, fb);
            else qrot_vjp(q, f, yb, qb, fb);
            const D g[7] = {qb.w, qb.x, qb.y, qb.z, fb.x, fb.y, fb.z};
            for (int i = 0; i < 7; i++)
                err.check(g[i], [&](D dx) {
                    Q4<D> qq = q;
                    V3<D> ff = f;
                    if (i < 4) (&qq.w)[i] += dx;
                    else (&ff.x)[i - 4] += dx;
                    return dot(inv ? qrotinv(qq, ff) : qrot(qq, ff), yb);
                }, h, 1e-9);
        }
        Q4<D> a = {U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)}, rb = {U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
        Q4<D> ab = {0, 0, 0, 0}, bb = {0, 0, 0, 0};
        qmul_vjp(a, q, rb, ab, bb);
        const D gm[8] = {ab.w, ab.x, ab.y, ab.z, bb.w, bb.x, bb.y, bb.z};
        for (int i = 0; i < 8; i++)
            err.check(gm[i], [&](D dx) {
                Q4<D> aa = a, qq = q;
                if (i < 4) (&aa.w)[i] += dx;
                else (&qq.w)[i - 4] += dx;
                return q4dot(qmul(aa, qq), rb);
            }, h, 1e-9);
        for (int br = 0; br < 2; br++) {
            const Q4<D> u = rand_rot(br ? 2.5 : 0.04);
            Q4<D> qb = {0, 0, 0, 0};
            d_rotvec_vjp(u, yb, qb);
            const D gq[4] = {qb.w, qb.x, qb.y, qb.z};
            for (int i = 0; i < 4; i++)
                err.check(gq[i], [&](D dx) {
                    Q4<D> uu = u;
                    (&uu.w)[i] += dx;
                    return dot(d_rotvec(uu), yb);
                }, 1e-8, 1e-9);
            const D sc = br ? 1.0 : 1e-5;
            const V3<D> in = {U(-1, 1) * sc, U(-1, 1) * sc, U(-1, 1) * sc};
            V3<D> ib = {0, 0, 0};
            from_rotvec_vjp(in, rb, ib);
            const D gi[3] = {ib.x, ib.y, ib.z};
            for (int i = 0; i < 3; i++)
                err.check(gi[i], [&](D dx) {
                    V3<D> ii = in;
                    (&ii.x)[i] += dx;
                    return q4dot(from_rotvec(ii), rb);
                }, 1e-3 * sc, 1e-9);
            const V3<D> rf = br ? v3<D>(U(0.5, 1), U(-1, 1), U(-1, 1)) : v3<D>(U(0.5, 1), U(-1, 1) * 5e-3, U(-1, 1) * 5e-3);
            V3<D> rfb = {0, 0, 0};
            d_angle_to_posx_vjp(rf, rb, rfb);
            const D gr[3] = {rfb.x, rfb.y, rfb.z};
            for (int i = 0; i < 3; i++)
                err.check(gr[i], [&](D dx) {
                    V3<D> rr = rf;
                    (&rr.x)[i] += dx;
                    return q4dot(d_angle_to_posx(rr), rb);
                }, 1e-7, 1e-9);
        }
    }
    return err.worst;
}
