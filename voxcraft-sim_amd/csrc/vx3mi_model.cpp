// vx3mi_model.cpp -- builds a host model from a lattice description the way the
// reference's CVX_Sim::Import (src/VXA/VX_Sim.cpp:59-153) and CVoxelyze
// (src/old/Voxelyze.cpp:422-641) do: same voxel order, same link order and
// orientation, same combined link materials, same float/double rounding of
// the derived material constants (src/old/VX_Material*.cpp).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>

#include "vx3mi_internal.h"

namespace vx3mi {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

namespace {

// Derived quantities of a voxel material (CVX_Material::updateDerived
// src/old/VX_Material.cpp:542-548, CVX_MaterialVoxel::updateDerived
// src/old/VX_MaterialVoxel.cpp:57-79).
struct MatDerived {
    float E, nu, eHat, rho, alpha, muS, muK, zI, zG, zC, gravMult, sigmaFail, epsFail;
    double nomSize;
    float mass, massInv, sqrtMass, inertia, inertiaInv, c2M, c2I;
    int fixed, isMeasured;
    double thermalOn;
    // stress-strain model (CVX_Material strainData / stressData / linear)
    int linear, npts;
    float sd[kMaxCurvePoints], td[kMaxCurvePoints];
};

// CVX_Material::setModelLinear, src/old/VX_Material.cpp:372-401 (arguments
// already validated)
void set_linear(MatDerived &d, float E, float fail) {
    float tf = fail;
    if (tf == -1) tf = 1000000;
    const float tfs = tf / E;
    d.npts = 2;
    d.sd[0] = 0; d.sd[1] = tfs;
    d.td[0] = 0; d.td[1] = tf;
    d.linear = 1;
    d.E = E;
    d.sigmaFail = fail;
    d.epsFail = (fail == -1) ? -1 : tfs;
}

// CVX_Material::setModelBilinear, src/old/VX_Material.cpp:406-451; false
// (model unchanged) for the inputs the reference rejects
bool set_bilinear(MatDerived &d, float E, float plastic, float yield, float fail) {
    if (E <= 0 || plastic <= 0 || plastic >= E || yield <= 0) return false;
    if (fail != -1.0f && fail <= yield) return false;
    const float yieldStrain = yield / E;
    float tf = fail;
    if (tf == -1) tf = 3 * yield;
    const float tM = plastic, tB = yield - tM * yieldStrain, tfs = (tf - tB) / tM;
    d.npts = 3;
    d.sd[0] = 0; d.sd[1] = yieldStrain; d.sd[2] = tfs;
    d.td[0] = 0; d.td[1] = yield; d.td[2] = tf;
    d.linear = 0;
    d.E = E;
    d.sigmaFail = fail;
    d.epsFail = fail == -1.0f ? -1.0f : tfs;
    return true;
}

// CVX_Material::setModel (src/old/VX_Material.cpp:300-367) on float data: a
// leading (0, 0) is dropped; false (model unchanged) for no points, a first
// point not above zero or strains that do not increase.  Decreasing stresses
// only leave a message there, and its slope test compares with 0/0 and never
// fails.  E is the first slope, the last point is the failure point, one point
// is a linear material.
bool set_data(MatDerived &d, int n, const float *strain, const float *stress) {
    if (n > 0 && strain[0] == 0 && stress[0] == 0) { strain++; stress++; n--; }
    if (n <= 0 || n + 1 > kMaxCurvePoints) return false;
    if (strain[0] <= 0 || stress[0] <= 0) return false;
    float sweep = 0.0f;
    for (int i = 0; i < n; i++) {
        if (strain[i] <= sweep) return false;
        sweep = strain[i];
    }
    d.npts = n + 1;
    d.sd[0] = 0; d.td[0] = 0;
    for (int i = 0; i < n; i++) { d.sd[i + 1] = strain[i]; d.td[i + 1] = stress[i]; }
    d.E = d.td[1] / d.sd[1];
    d.sigmaFail = d.td[n];
    d.epsFail = d.sd[n];
    d.linear = n == 1;
    return true;
}

// CVX_Material::modulus / strain (src/old/VX_Material.cpp:214-243)
float mat_modulus(const MatDerived &d, float strain) {
    if (d.epsFail != -1.0f && strain > d.epsFail) return 0.0f;
    if (strain <= d.sd[1] || d.linear) return d.E;
    for (int i = 2; i < d.npts; i++)
        if (strain <= d.sd[i] || i == d.npts - 1) return (d.td[i] - d.td[i - 1]) / (d.sd[i] - d.sd[i - 1]);
    return 0.0f;
}
float mat_strain(const MatDerived &d, float stress) {
    if (stress <= d.td[1] || d.linear) return stress / d.E;
    for (int i = 2; i < d.npts; i++)
        if (stress <= d.td[i] || i == d.npts - 1) {
            const float perc = (stress - d.td[i - 1]) / (d.td[i] - d.td[i - 1]);
            return d.sd[i - 1] + perc * (d.sd[i] - d.sd[i - 1]);
        }
    return 0.0f;
}

void derive(MatDerived &d) {
    d.eHat = d.E / ((1 - 2 * d.nu) * (1 + d.nu));
    double ns = d.nomSize, volume = ns * ns * ns;
    d.mass = (float)(volume * d.rho);
    d.inertia = (float)(d.mass * ns * ns / 6.0f);
    if (volume == 0 || d.mass == 0 || d.inertia == 0) {
        d.massInv = d.sqrtMass = d.inertiaInv = d.c2M = d.c2I = 0.0f;
        return;
    }
    d.massInv = 1.0f / d.mass;
    d.sqrtMass = std::sqrt(d.mass);
    d.inertiaInv = 1.0f / d.inertia;
    d.c2M = (float)(2.0f * std::sqrt((double)(d.mass * d.E) * ns));
    d.c2I = (float)(2.0f * std::sqrt((double)(d.inertia * d.E) * ns * ns * ns));
}

VoxelMat to_voxel_mat(const MatDerived &d) {
    VoxelMat v;
    std::memset(&v, 0, sizeof(v));
    v.nomSize = d.nomSize;
    v.thermalOnAfter = d.thermalOn;
    v.alphaCTE = d.alpha;
    v.mass = d.mass;
    v.massInv = d.massInv;
    v.inertiaInv = d.inertiaInv;
    v.dmNum = 2 * d.sqrtMass * d.zI;
    v.gDampT = d.zG * d.c2M;
    v.gDampR = d.zG * d.c2I;
    v.cDampT = d.zC * d.c2M;
    v.gravF = -d.mass * 9.80665f * d.gravMult;
    v.penK = (float)(2 * d.E * d.nomSize);
    v.penKinv = 1.0f / v.penK;
    v.muS = d.muS;
    v.muK = d.muK;
    v.fixed = d.fixed;
    v.isMeasured = d.isMeasured;
    v.zetaI = d.zI;
    v.zetaG = d.zG;
    v.zetaC = d.zC;
    v.ext = -1;
    return v;
}

// CVX_MaterialLink::updateAll / updateDerived, src/old/VX_MaterialLink.cpp:45-141.
// A pair with a non-linear material gets the springs-in-series curve of the two
// (its points in curve; the caller sets LinkMat::curve).
LinkMat combine(const MatDerived &a, const MatDerived &b, std::vector<CurvePt> &curve) {
    LinkMat l;
    std::memset(&l, 0, sizeof(l));
    curve.clear();
    l.nomSize = 0.5 * (a.nomSize + b.nomSize);
    float f1 = a.sigmaFail, f2 = b.sigmaFail, sf;
    if (f1 == -1.0f) sf = f2;
    else if (f2 == -1.0f) sf = f1;
    else sf = f1 < f2 ? f1 : f2;
    if (a.linear && b.linear) {
        l.E = 2.0f * a.E * b.E / (a.E + b.E);
        l.epsFail = (sf == -1.0f) ? -1.0f : sf / l.E;
    } else {
        // step up through the union of both curves' strain points, each new
        // segment at the series modulus of the two materials just below it
        MatDerived c;
        std::memset(&c, 0, sizeof(c));
        std::vector<float> st(1, 0.0f), ss(1, 0.0f);
        int i1 = 1, i2 = 1;
        while (i1 < a.npts && i2 < b.npts) {
            float e = FLT_MAX;
            if (i1 < a.npts) e = a.sd[i1];
            if (i2 < b.npts && b.sd[i2] < e) e = b.sd[i2];
            if (e == a.sd[i1]) i1++;
            if (e == b.sd[i2]) i2++;
            const float m1 = mat_modulus(a, e - FLT_EPSILON), m2 = mat_modulus(b, e - FLT_EPSILON);
            const float m = 2.0f * m1 * m2 / (m1 + m2);
            const size_t last = st.size() - 1;
            ss.push_back(ss[last] + m * (e - st[last]));
            st.push_back(e);
        }
        // CVX_Material::setModel (VX_Material.cpp:300-370) on these points: the
        // leading (0, 0) is implied; E is the first slope
        const int n = (int)st.size();
        c.npts = std::min(n, kMaxCurvePoints);  // two data materials give at most 2 * 32 + 1 points
        for (int k = 0; k < c.npts; k++) { c.sd[k] = st[k]; c.td[k] = ss[k]; }
        c.E = c.td[1] / c.sd[1];
        c.linear = (n - 1 == 1);
        l.E = c.E;
        l.epsFail = (sf == -1.0f) ? -1.0f : mat_strain(c, sf);
        if (!c.linear)
            for (int k = 0; k < c.npts; k++) curve.push_back(CurvePt{c.sd[k], c.td[k]});
    }
    if (a.nu == 0 && b.nu == 0) l.nu = 0;
    else {
        float tmpEHat = 2 * a.eHat * b.eHat / (a.eHat + b.eHat);
        float c2 = (tmpEHat - l.E) / (2 * tmpEHat) + 0.0625;
        l.nu = (float)(std::sqrt(c2) - 0.25);
    }
    l.eHat = l.E / ((1 - 2 * l.nu) * (1 + l.nu));
    float L = (float)l.nomSize;
    l.a1 = l.E * L;
    l.a2 = l.E * L * L * L / (12.0f * (1 + l.nu));
    l.b1 = l.E * L;
    l.b2 = l.E * L * L / 2.0f;
    l.b3 = l.E * L * L * L / 6.0f;
    l.sqA1 = std::sqrt(l.a1);
    l.sqA2xIp = std::sqrt(l.a2 * L * L / 6.0f);
    l.sqB1 = std::sqrt(l.b1);
    l.sqB2xFMp = std::sqrt(l.b2 * L / 2.0f);
    l.sqB3xIp = std::sqrt(l.b3 * L * L / 6.0f);
    return l;
}

float clamp_nu(float nu) {  // CVX_Material::setPoissonsRatio VX_Material.cpp:488-494
    if (nu < 0) nu = 0;
    if (nu >= 0.5) nu = (float)(0.5 - FLT_EPSILON * 2);
    return nu;
}

}  // namespace

double recommended_dt(const Model &m) {
    double maxFreq2 = 0.0f;
    for (int i = 0; i < m.nl(); i++) {
        const LinkMat &l = m.lmats[m.lmat[i]];
        double m1 = m.vmats[m.vmat[m.lneg[i]]].mass, m2 = m.vmats[m.vmat[m.lpos[i]]].mass;
        float k = l.a1;  // axialStiffness(), VX3_Link.cu:269-278 (nu == 0)
        if (l.nu != 0.0f) {
            float fN = thermal_factor(m.initialTemp, l.alphaN), fP = thermal_factor(m.initialTemp, l.alphaP);
            double rl = 0.5 * (l.nomN * (double)fN + l.nomP * (double)fP);
            k = (float)(l.eHat * l.tArea / ((0 + 1) * rl));
        }
        double f2 = k / (m1 < m2 ? m1 : m2);
        if (f2 > maxFreq2) maxFreq2 = f2;
    }
    if (maxFreq2 <= 0.0f) {
        for (int v = 0; v < m.nv(); v++) {
            const VoxelMat &vm = m.vmats[m.vmat[v]];
            double f2 = m.vmatE[m.vmat[v]] * vm.nomSize / vm.mass;
            if (f2 > maxFreq2) maxFreq2 = f2;
        }
    }
    if (maxFreq2 <= 0.0f) return 0.0f;
    return 1.0f / (6.283185f * std::sqrt(maxFreq2));
}

static int build_model(const vx3mi_lattice *sp, Model &m) {
    if (!sp || sp->nx <= 0 || sp->ny <= 0 || sp->nz <= 0 || !sp->cells || sp->nmat < 0 || (sp->nmat > 0 && !sp->mats)) {
        set_error("vx3mi_model_from_lattice: invalid lattice description");
        return VX3MI_E_ARG;
    }
    if (sp->nx > 1023 || sp->ny > 1023 || sp->nz > 1023) {
        set_error("lattice dimension above 1023 voxels per axis");
        return VX3MI_E_LIMIT;
    }
    const int nx = sp->nx, ny = sp->ny, nz = sp->nz;
    const long ncell = (long)nx * ny * nz;
    m.nx = nx; m.ny = ny; m.nz = nz;
    m.latticeDim = sp->lattice_dim;
    m.dtFrac = sp->dt_frac;
    m.floorEnabled = sp->floor_enabled ? 1 : 0;
    m.enableCollision = sp->enable_collision ? 1 : 0;
    m.enableExpansion = sp->enable_expansion ? 1 : 0;
    m.varyTemp = (sp->vary_temp && sp->temp_period > 0) ? 1 : 0;
    m.tempAmplitude = sp->temp_amplitude;
    m.tempPeriod = sp->temp_period;
    m.tempBase = sp->temp_base;
    m.palette.assign(sp->mats, sp->mats + sp->nmat);
    m.enableCilia = sp->enable_cilia ? 1 : 0;
    m.enableSignals = sp->enable_signals ? 1 : 0;
    m.enableDetach = sp->enable_detach ? 1 : 0;
    m.maxDistPairs = sp->max_dist_pairs;
    m.enableAttach = sp->enable_attach ? 1 : 0;
    m.safetyGuard = sp->safety_guard;
    if (m.enableAttach) {
        if (sp->safety_guard < 0 || sp->safety_guard > kMaxSafetyGuard) {
            set_error("SafetyGuard must lie in 0..65535");
            return VX3MI_E_ARG;
        }
        if (sp->voxel_E) {  // every voxel is a material of its own: nothing could attach
            set_error("EnableAttach with a per-voxel Young's modulus is not supported");
            return VX3MI_E_ARG;
        }
        for (int k = 0; k < sp->nmat; k++)
            if (sp->mats[k].sticky && sp->volume_effects && sp->mats[k].poissons_ratio != 0.0f) {
                // a new link's transverse area would follow its voxels' current strains
                set_error("EnableAttach with a sticky material of non-zero Poisson's ratio under volume effects is not supported");
                return VX3MI_E_ARG;
            }
    }
    m.secondaryExperiment = sp->secondary_experiment ? 1 : 0;

    // UpdateMatTemps (VX_Sim.cpp:467-483) via UpdateCurTemp (VX_Environment.cpp:327-346)
    float t0 = 0.0f;
    if (sp->temp_enabled) {
        float cur;
        if (sp->vary_temp) cur = (sp->temp_period == 0) ? 0.0f : (float)(sp->temp_base + sp->temp_amplitude * std::sin(0.0));
        else cur = (float)(sp->temp_base + sp->temp_amplitude);
        t0 = (float)(cur - sp->temp_base);
    }
    m.initialTemp = t0;

    // palette -> derived voxel materials (CopyMat VX_Sim.cpp:368-420, Import :74-86)
    const float gravMult = (float)(-(float)sp->grav_acc / 9.80665);  // SetGravityAccel VX_Sim.cpp:485
    const float zI = sp->bond_damp_z <= 0 ? 0 : sp->bond_damp_z;
    const float zG = sp->slow_damp_z <= 0 ? 0 : sp->slow_damp_z;
    const float zC = sp->col_damp_z <= 0 ? 0 : sp->col_damp_z;
    std::vector<MatDerived> pal(sp->nmat);
    for (int k = 0; k < sp->nmat; k++) {
        const vx3mi_material &in = sp->mats[k];
        MatDerived &d = pal[k];
        std::memset(&d, 0, sizeof(d));
        d.nomSize = sp->lattice_dim;
        // CVX_Sim::CopyMat (VX_Sim.cpp:390-405): the model of MatModel; inputs
        // the setter rejects leave the addMaterial default (linear, 1e6 Pa)
        set_linear(d, 1e6f, -1.0f);
        const float E = in.elastic_mod, fail = in.fail_stress;
        if (in.mat_model == 0) {
            if (E > 0) set_linear(d, E, -1.0f);
        } else if (in.mat_model == 1) {
            if (E > 0 && (fail == -1.0f || fail > 0)) set_linear(d, E, fail);
        } else if (in.mat_model == 2) {
            set_bilinear(d, E, in.plastic_mod, in.yield_stress, fail);
        } else if (in.mat_model == 3) {
            if (in.ss_points < 0 || in.ss_points > VX3MI_MAX_SS_POINTS) {
                set_error("stress-strain data: 0 to VX3MI_MAX_SS_POINTS points");
                return VX3MI_E_LIMIT;
            }
            set_data(d, in.ss_points, in.ss_strain, in.ss_stress);
        } else {
            set_error("material model must be 0 (linear), 1 (linear with failure), 2 (bilinear) or 3 (data)");
            return VX3MI_E_ARG;
        }
        d.nu = sp->volume_effects ? clamp_nu(in.poissons_ratio) : 0.0f;
        d.rho = in.density <= 0 ? FLT_MIN : in.density;
        d.alpha = in.cte;
        d.muS = in.u_static <= 0 ? 0 : in.u_static;
        d.muK = in.u_dynamic <= 0 ? 0 : in.u_dynamic;
        d.zI = zI; d.zG = zG; d.zC = zC;
        d.gravMult = gravMult;
        d.fixed = in.fixed ? 1 : 0;
        d.isMeasured = in.is_measured ? 1 : 0;
        d.thermalOn = in.thermal_on_after;
        derive(d);
    }

    // voxels in lattice order (x fastest)
    std::vector<int> cellvox(ncell, -1);
    std::vector<MatDerived> vm;  // derived material per voxel-material index
    std::vector<int> palToVm(sp->nmat, -1);
    int nv = 0;
    for (long i = 0; i < ncell; i++) {
        int c = sp->cells[i];
        if (c <= 0 || c > sp->nmat) continue;
        int z = (int)(i / ((long)nx * ny)), y = (int)((i - (long)z * nx * ny) / nx), x = (int)(i - (long)z * nx * ny - (long)y * nx);
        cellvox[i] = nv;
        m.ix.push_back(x); m.iy.push_back(y); m.iz.push_back(z);
        int mi;
        if (sp->voxel_E) {
            MatDerived d = pal[c - 1];
            if (!d.linear) {
                set_error("per-voxel Young's modulus needs linear materials");
                return VX3MI_E_ARG;
            }
            set_linear(d, sp->voxel_E[nv], d.sigmaFail);
            derive(d);
            mi = (int)vm.size();
            vm.push_back(d);
        } else {
            if (palToVm[c - 1] < 0) { palToVm[c - 1] = (int)vm.size(); vm.push_back(pal[c - 1]); }
            mi = palToVm[c - 1];
        }
        m.vmat.push_back(mi);
        m.vpal.push_back(c);
        m.phase.push_back(sp->phase ? sp->phase[nv] : 0.0);
        nv++;
    }
    // per-voxel cilia forces (CVX_Sim::Import copies them to each voxel, VX_Sim.cpp:102-105)
    if (sp->cilia_base) m.ciliaBase.assign(sp->cilia_base, sp->cilia_base + (size_t)3 * nv);
    if (sp->cilia_shift) m.ciliaShift.assign(sp->cilia_shift, sp->cilia_shift + (size_t)3 * nv);
    m.vlinks.assign((size_t)nv * 6, -1);
    // links: created when the later voxel of a pair is added (addVoxel's
    // X_POS..Z_NEG loop only finds lower neighbours), pVNeg = lower coordinate
    std::map<std::pair<int, int>, int> lmatIndex;
    // the link material between voxel materials mn (negative end) and mp
    // (positive end), made on first use
    auto link_mat = [&](int mn, int mp) -> int {
        std::pair<int, int> key(mn, mp);
        auto it = lmatIndex.find(key);
        if (it != lmatIndex.end()) return it->second;
        // combinedMaterial(voxel1 = new voxel, voxel2 = neighbour); the result is symmetric
        std::vector<CurvePt> curve;
        LinkMat l = combine(vm[mp], vm[mn], curve);
        if (!curve.empty()) {
            l.curve = ((int)m.curves.size() << 8) | (int)curve.size();
            m.curves.insert(m.curves.end(), curve.begin(), curve.end());
        }
        const MatDerived &dn = vm[mn], &dp = vm[mp];
        l.nomN = dn.nomSize; l.nomP = dp.nomSize;
        l.alphaN = dn.alpha; l.alphaP = dp.alpha;
        l.dmNumN = 2 * dn.sqrtMass * dn.zI;
        l.dmNumP = 2 * dp.sqrtMass * dp.zI;
        // transverse info frozen at creation (VX3_Link.cu:147-150; all strains zero)
        float sN = (float)dn.nomSize, sP = (float)dp.nomSize;
        l.tArea = 0.5f * (sN * sN + sP * sP);
        l.tStrainSum = 0.0f;
        l.fixedBoth = (dn.fixed && dp.fixed) ? 1 : 0;
        const int lmi = (int)m.lmats.size();
        m.lmats.push_back(l);
        lmatIndex[key] = lmi;
        return lmi;
    };
    for (int v = 0; v < nv; v++) {
        int x = m.ix[v], y = m.iy[v], z = m.iz[v];
        int nb[3] = {x > 0 ? cellvox[(x - 1) + (long)nx * (y + (long)ny * z)] : -1,
                     y > 0 ? cellvox[x + (long)nx * ((y - 1) + (long)ny * z)] : -1,
                     z > 0 ? cellvox[x + (long)nx * (y + (long)ny * (z - 1))] : -1};
        for (int a = 0; a < 3; a++) {
            int n = nb[a];
            if (n < 0) continue;
            int li = m.nl();
            m.lneg.push_back(n);
            m.lpos.push_back(v);
            m.laxis.push_back(a);
            m.vlinks[(size_t)v * 6 + 2 * a + 1] = li;
            m.vlinks[(size_t)n * 6 + 2 * a] = li;
            m.lmat.push_back(link_mat(m.vmat[n], m.vmat[v]));
        }
    }
    // the link material two voxels of one sticky material get when they attach
    // (VX3_Link constructor, VX3_Link.cu:31-57: combinedMaterial of the two ends)
    m.attachLmat.assign(vm.size(), -1);
    if (m.enableAttach)
        for (int c = 0; c < sp->nmat; c++)
            if (sp->mats[c].sticky && palToVm[c] >= 0 && !vm[palToVm[c]].fixed)
                m.attachLmat[palToVm[c]] = link_mat(palToVm[c], palToVm[c]);
    for (const MatDerived &d : vm) {
        m.vmats.push_back(to_voxel_mat(d));
        m.vmatE.push_back(d.E);
    }
    for (int v = 0; v < nv; v++) {
        int c = 0;
        for (int d = 0; d < 6; d++) c += m.vlinks[(size_t)v * 6 + d] >= 0;
        if (c < 6) m.surface.push_back(v);
    }
    return VX3MI_OK;
}

void set_external(Model &m, int v, int dof, const double t[3], const double r[3], const float f[3], const float mo[3]) {
    if (m.vext.empty()) m.vext.assign((size_t)m.nv(), -1);
    if (dof < 0) {
        m.vext[v] = -1;
        return;
    }
    VoxelExt e;
    std::memset(&e, 0, sizeof(e));
    e.dof = dof & DOF_ALL;
    for (int k = 0; k < 3; k++) {
        e.t[k] = t ? t[k] : 0.0;
        e.r[k] = r ? r[k] : 0.0;
        e.f[k] = f ? f[k] : 0.0f;
        e.m[k] = mo ? mo[k] : 0.0f;
    }
    // rotationChanged(), src/old/VX_External.cpp: identity for a zero rotation
    // vector, else Quat3D(extRotation) (FromRotationVector)
    Q4<double> q = {1, 0, 0, 0};
    if (e.r[0] != 0 || e.r[1] != 0 || e.r[2] != 0) q = from_rotvec(V3<double>{e.r[0], e.r[1], e.r[2]});
    e.q[0] = q.w; e.q[1] = q.x; e.q[2] = q.y; e.q[3] = q.z;
    if (m.vext[v] >= 0) m.exts[m.vext[v]] = e;
    else {
        m.vext[v] = (int)m.exts.size();
        m.exts.push_back(e);
    }
}

}  // namespace vx3mi

using namespace vx3mi;

extern "C" {

const char *vx3mi_last_error(void) { return g_error.c_str(); }
const char *vx3mi_version(void) { return "vx3mi 0.1 (gfx950)"; }

int vx3mi_model_from_lattice(const vx3mi_lattice *spec, vx3mi_model **out) {
    if (!out) { set_error("out is NULL"); return VX3MI_E_ARG; }
    *out = nullptr;
    vx3mi_model *h = new (std::nothrow) vx3mi_model();
    if (!h) { set_error("out of host memory"); return VX3MI_E_NOMEM; }
    try {
        int rc = build_model(spec, h->m);
        if (rc != VX3MI_OK) { delete h; return rc; }
    } catch (const std::bad_alloc &) {
        delete h;
        set_error("out of host memory");
        return VX3MI_E_NOMEM;
    }
    *out = h;
    return VX3MI_OK;
}

void vx3mi_model_free(vx3mi_model *m) { delete m; }

int vx3mi_model_counts(const vx3mi_model *m, int *nvox, int *nlink, int *nsurf) {
    if (!m) { set_error("model is NULL"); return VX3MI_E_ARG; }
    if (nvox) *nvox = m->m.nv();
    if (nlink) *nlink = m->m.nl();
    if (nsurf) *nsurf = (int)m->m.surface.size();
    return VX3MI_OK;
}

int vx3mi_model_settings(const vx3mi_model *m, vx3mi_settings *o) {
    if (!m || !o) { set_error("NULL argument"); return VX3MI_E_ARG; }
    const Model &M = m->m;
    o->nx = M.nx; o->ny = M.ny; o->nz = M.nz;
    o->lattice_dim = M.latticeDim;
    o->dt_frac = M.dtFrac;
    o->floor_enabled = M.floorEnabled;
    o->enable_collision = M.enableCollision;
    o->enable_expansion = M.enableExpansion;
    o->vary_temp = M.varyTemp;
    o->temp_base = M.tempBase;
    o->temp_amplitude = M.tempAmplitude;
    o->temp_period = M.tempPeriod;
    o->initial_temp = M.initialTemp;
    o->record_step_size = M.recordStepSize;
    o->record_voxel = M.recordVoxel;
    o->record_link = M.recordLink;
    o->has_stop_formula = !M.stopFormula.empty();
    o->has_fitness_formula = !M.fitnessFormula.empty();
    o->enable_cilia = M.enableCilia;
    o->secondary_experiment = M.secondaryExperiment;
    o->enable_signals = M.enableSignals;
    o->enable_detach = M.enableDetach;
    o->max_dist_pairs = M.maxDistPairs;
    o->enable_attach = M.enableAttach;
    o->safety_guard = M.safetyGuard;
    return VX3MI_OK;
}

const char *vx3mi_model_raw_print(const vx3mi_model *m) { return m ? m->m.rawPrint.c_str() : ""; }
int vx3mi_model_save_positions(const vx3mi_model *m) { return m ? m->m.savePositions : 0; }

int vx3mi_model_palette_index(const vx3mi_model *m, int *out) {
    if (!m || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    for (int v = 0; v < m->m.nv(); v++) out[v] = m->m.vpal[v];
    return VX3MI_OK;
}

int vx3mi_model_cilia(const vx3mi_model *m, double *base3, double *shift3) {
    if (!m) { set_error("NULL argument"); return VX3MI_E_ARG; }
    const Model &M = m->m;
    const size_t n = (size_t)3 * M.nv();
    for (size_t i = 0; i < n; i++) {
        if (base3) base3[i] = M.ciliaBase.empty() ? 0.0 : M.ciliaBase[i];
        if (shift3) shift3[i] = M.ciliaShift.empty() ? 0.0 : M.ciliaShift[i];
    }
    return VX3MI_OK;
}

int vx3mi_model_material(const vx3mi_model *m, int index, vx3mi_material *out) {
    if (!m || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    if (index < 1 || index > (int)m->m.palette.size()) { set_error("palette index out of range"); return VX3MI_E_ARG; }
    *out = m->m.palette[index - 1];
    return VX3MI_OK;
}

int vx3mi_model_phase(const vx3mi_model *m, double *out) {
    if (!m || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    for (int v = 0; v < m->m.nv(); v++) out[v] = m->m.phase[v];
    return VX3MI_OK;
}

int vx3mi_model_recommended_dt(const vx3mi_model *m, double *out) {
    if (!m || !out) { set_error("NULL argument"); return VX3MI_E_ARG; }
    *out = recommended_dt(m->m);
    return VX3MI_OK;
}

int vx3mi_model_topology(const vx3mi_model *m, int *lattice3, int *link_neg, int *link_pos, int *link_axis) {
    if (!m) { set_error("model is NULL"); return VX3MI_E_ARG; }
    const Model &M = m->m;
    for (int v = 0; v < M.nv(); v++)
        if (lattice3) { lattice3[3 * v] = M.ix[v]; lattice3[3 * v + 1] = M.iy[v]; lattice3[3 * v + 2] = M.iz[v]; }
    for (int l = 0; l < M.nl(); l++) {
        if (link_neg) link_neg[l] = M.lneg[l];
        if (link_pos) link_pos[l] = M.lpos[l];
        if (link_axis) link_axis[l] = M.laxis[l];
    }
    return VX3MI_OK;
}

int vx3mi_model_voxel_constants(const vx3mi_model *m, int voxel, float *o) {
    if (!m || !o || voxel < 0 || voxel >= m->m.nv()) { set_error("invalid voxel constants query"); return VX3MI_E_ARG; }
    const VoxelMat &v = m->m.vmats[m->m.vmat[voxel]];
    const float vals[12] = {v.mass, v.massInv, v.inertiaInv, v.dmNum, v.gDampT, v.gDampR,
                            v.cDampT, v.gravF, v.penK, v.muS, v.muK, v.alphaCTE};
    for (int i = 0; i < 12; i++) o[i] = vals[i];
    return VX3MI_OK;
}

int vx3mi_model_link_constants(const vx3mi_model *m, int link, float *o) {
    if (!m || !o || link < 0 || link >= m->m.nl()) { set_error("invalid link constants query"); return VX3MI_E_ARG; }
    const LinkMat &l = m->m.lmats[m->m.lmat[link]];
    const float vals[11] = {l.E, l.a1, l.a2, l.b1, l.b2, l.b3, l.sqA1, l.sqA2xIp, l.sqB1, l.sqB2xFMp, l.sqB3xIp};
    for (int i = 0; i < 11; i++) o[i] = vals[i];
    return VX3MI_OK;
}

int vx3mi_model_link_curve(const vx3mi_model *m, int link, float *strain, float *stress) {
    if (!m || link < 0 || link >= m->m.nl()) {
        set_error("vx3mi_model_link_curve: bad model or link index");
        return VX3MI_E_ARG;
    }
    const LinkMat &l = m->m.lmats[m->m.lmat[link]];
    const int n = curve_points(l), first = curve_first(l);
    for (int k = 0; k < n; k++) {
        if (strain) strain[k] = m->m.curves[first + k].strain;
        if (stress) stress[k] = m->m.curves[first + k].stress;
    }
    return n;
}

int vx3mi_model_set_external(vx3mi_model *m, int voxel, int dof, const double *translation3, const double *rotation3,
                             const double *force3, const double *moment3) {
    if (!m || voxel < 0 || voxel >= m->m.nv()) {
        set_error("vx3mi_model_set_external: bad model or voxel index");
        return VX3MI_E_ARG;
    }
    float f[3] = {0, 0, 0}, mo[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        if (force3) f[k] = (float)force3[k];
        if (moment3) mo[k] = (float)moment3[k];
    }
    set_external(m->m, voxel, dof, translation3, rotation3, f, mo);
    return VX3MI_OK;
}

int vx3mi_model_get_external(const vx3mi_model *m, int voxel, int *dof, double *translation3, double *rotation3,
                             double *force3, double *moment3) {
    if (!m || voxel < 0 || voxel >= m->m.nv()) {
        set_error("vx3mi_model_get_external: bad model or voxel index");
        return VX3MI_E_ARG;
    }
    const int e = m->m.ext_of(voxel);
    if (dof) *dof = e < 0 ? -1 : m->m.exts[e].dof;
    for (int k = 0; k < 3; k++) {
        const VoxelExt *x = e < 0 ? nullptr : &m->m.exts[e];
        if (translation3) translation3[k] = x ? x->t[k] : 0.0;
        if (rotation3) rotation3[k] = x ? x->r[k] : 0.0;
        if (force3) force3[k] = x ? x->f[k] : 0.0;
        if (moment3) moment3[k] = x ? x->m[k] : 0.0;
    }
    return VX3MI_OK;
}

}  // extern "C"
