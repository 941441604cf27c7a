// ref_driver.cpp -- drives the reference's own CPU Voxelyze (src/old, compiled
// from the reference sources by oracle/Makefile into oracle/_ref/) so the FP64
// oracle can be checked against the real reference code.  TEST INFRASTRUCTURE.
//
// The model is assembled the way CVX_Sim::Import does it
// (src/VXA/VX_Sim.cpp:59-153): gravity and floor first, then palette materials
// (CopyMat :368-420 + damping), voxel size, voxels in lattice order, initial
// temperatures, volume effects off.  Each step sets the VX3 thermal actuation
// (src/VX3/VX3_VoxelyzeKernel.cu:625-650) with CVX_Voxel::setTemperature and
// then calls CVoxelyze::doTimeStep (src/old/Voxelyze.cpp:251-284).
//
// stdin (whitespace separated):
//   nx ny nz  lattice_dim  bond col slow  grav_acc floor  temp_amp temp_period vary enable_expansion t0
//   nmat  then per material: E rho cte mu_s mu_k model plastic_mod yield_stress fail_stress
//     (model 0: setModelLinear(E), 1: setModelLinear(E, fail), 2: setModelBilinear,
//      3: followed by n and n strain stress pairs for setModel)
//   nx*ny*nz cell values (0 = empty)
//   nv phase offsets (cycles)
//   dt nsteps record_every
//   [next, then per external: voxel dof tx ty tz rx ry rz fx fy fz mx my mz]
//     applied with CVX_Voxel::external() and CVX_External::setDisplacement /
//     setForce / setMoment (src/old/VX_External.cpp), as CVX_Sim::Import does
//     for boundary-condition regions (src/VXA/VX_Sim.cpp:109-143)
// stdout (binary): per record, nv * (pos3, quat4, vel3, angvel3) doubles,
//   then a trailing timing line on stderr "seconds <s> voxel_steps <n>".
#include "Voxelyze.h"
#include "VX_Voxel.h"
#include "VX_Material.h"
#include "VX_External.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

int main() {
    int nx, ny, nz, floorOn, vary, enableExpansion, nmat;
    double L, tempAmp, tempPeriod;
    float bond, col, slow, gravAcc, t0;
    std::cin >> nx >> ny >> nz >> L >> bond >> col >> slow >> gravAcc >> floorOn >> tempAmp >> tempPeriod >> vary >>
        enableExpansion >> t0 >> nmat;
    CVoxelyze Vx(0.001);
    Vx.setGravity((float)(-gravAcc / 9.80665));
    Vx.enableFloor(floorOn != 0);
    std::vector<CVX_Material *> mats;
    for (int i = 0; i < nmat; i++) {
        float E, rho, cte, mus, muk, plastic, yield, fail;
        int model;
        std::cin >> E >> rho >> cte >> mus >> muk >> model >> plastic >> yield >> fail;
        CVX_Material *pm = Vx.addMaterial();
        if (model == 3) {
            int n = 0;
            std::cin >> n;
            std::vector<float> st(n), ss(n);
            for (int k = 0; k < n; k++) std::cin >> st[k] >> ss[k];
            if (n > 0) pm->setModel(n, st.data(), ss.data());
        } else if (model == 2) pm->setModelBilinear(E, plastic, yield, fail);
        else if (model == 1) pm->setModelLinear(E, fail);
        else pm->setModelLinear(E);
        pm->setPoissonsRatio(0.0f);
        pm->setDensity(rho);
        pm->setCte(cte);
        pm->setStaticFriction(mus);
        pm->setKineticFriction(muk);
        pm->setGlobalDamping(slow);
        pm->setInternalDamping(bond);
        pm->setCollisionDamping(col);
        mats.push_back(pm);
    }
    Vx.setVoxelSize(L);
    for (int i = 0; i < nx * ny * nz; i++) {
        int c;
        std::cin >> c;
        if (c <= 0) continue;
        int z = i / (nx * ny), y = (i - z * nx * ny) / nx, x = i - z * nx * ny - y * nx;
        Vx.setVoxel(mats[c - 1], x, y, z);
    }
    int nv = Vx.voxelCount();
    std::vector<double> phase(nv);
    for (int i = 0; i < nv; i++) std::cin >> phase[i];
    float dt;
    int nsteps, every;
    std::cin >> dt >> nsteps >> every;
    int next = 0;
    if (!(std::cin >> next)) next = 0;
    for (int k = 0; k < next; k++) {
        int v, dof;
        double t[3], r[3];
        float f[3], mo[3];
        std::cin >> v >> dof >> t[0] >> t[1] >> t[2] >> r[0] >> r[1] >> r[2] >> f[0] >> f[1] >> f[2] >> mo[0] >> mo[1] >>
            mo[2];
        CVX_External *e = Vx.voxel(v)->external();
        const dofComponent comp[6] = {X_TRANSLATE, Y_TRANSLATE, Z_TRANSLATE, X_ROTATE, Y_ROTATE, Z_ROTATE};
        for (int b = 0; b < 6; b++)
            if (dof & (1 << b)) e->setDisplacement(comp[b], b < 3 ? t[b] : r[b - 3]);
        e->setForce(f[0], f[1], f[2]);
        e->setMoment(mo[0], mo[1], mo[2]);
    }
    for (int i = 0; i < nv; i++) Vx.voxel(i)->setTemperature(t0);

    std::vector<double> rec((size_t)nv * 13);
    auto emit = [&]() {
        for (int i = 0; i < nv; i++) {
            CVX_Voxel *v = Vx.voxel(i);
            Vec3D<double> p = v->position(), vel = v->velocity(), av = v->angularVelocity();
            Quat3D<double> q = v->orientation();
            double *r = &rec[(size_t)i * 13];
            r[0] = p.x; r[1] = p.y; r[2] = p.z;
            r[3] = q.w; r[4] = q.x; r[5] = q.y; r[6] = q.z;
            r[7] = vel.x; r[8] = vel.y; r[9] = vel.z;
            r[10] = av.x; r[11] = av.y; r[12] = av.z;
        }
        fwrite(rec.data(), sizeof(double), rec.size(), stdout);
    };
    const double twoPiF = (double)(2 * 3.1415926f);
    auto t_start = std::chrono::steady_clock::now();
    double t = 0.0;  // VX3 keeps time in double (VX3_VoxelyzeKernel.cuh:54); CVoxelyze in float
    for (int s = 0; s < nsteps; s++) {
        if (vary && tempPeriod > 0) {
            for (int i = 0; i < nv; i++) {
                double T = tempAmp * sin(twoPiF * (t / tempPeriod + phase[i]));
                if (!enableExpansion && T > 0) T = 0;
                Vx.voxel(i)->setTemperature((float)T);
            }
        }
        if (!Vx.doTimeStep(dt)) {
            fprintf(stderr, "diverged at step %d\n", s);
            return 3;
        }
        t += dt;
        if (every > 0 && (s + 1) % every == 0) emit();
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    fprintf(stderr, "seconds %.6f voxel_steps %.0f\n", secs, (double)nv * nsteps);
    return 0;
}
