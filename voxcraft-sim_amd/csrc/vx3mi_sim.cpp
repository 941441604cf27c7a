// vx3mi_sim.cpp -- the simulation driver around the timestep: the loop of
// CUDA_Simulation (VX3_SimulationManager.cu:11-121) for every robot of a batch.
//
// Per robot, as the reference: the initial centre of mass (measured voxels,
// mass weighted), then up to max_steps iterations of
//     stop condition met? -> stop;  doTimeStep;  diverged? -> stop;
//     every real_stepsize iterations a .history frame;
// where doTimeStep samples the centre of mass every CycleStep = int(TempPeriod/dt)
// steps (VX3_VoxelyzeKernel.cu:314-334: recentAngle from the last three
// samples); at the end the final centre of mass and the fitness formula of
// the displacement.
//
// Robots run together on the device.  The host knows each robot's time
// exactly (t += (double)dt), so it steps the whole batch from event to event:
// the next frame, the next centre-of-mass sample a formula needs, or the
// first step at which the stop formula turns positive.  A robot that has
// stopped is no longer observed (the device keeps integrating it; its results
// were taken when it stopped).  collisionCount (hit) comes from the device;
// targetCloseness and numClosePairs are computed at the CycleStep samples
// (computeTargetCloseness, VX3_VoxelyzeKernel.cu:545-562) from the target
// voxels' positions when MaxDistInVoxelLengthsToCountAsPair is set.  A robot
// whose force fields or attach conditions read the samples has them taken on
// the device (k_cycle_sample), and the runner uses those.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vx3mi_batch.h"

namespace vx3mi {
namespace {

struct RobotRun {
    const Model *m = nullptr;
    double t = 0.0;          // currentTime
    long long j = 0;         // iterations done (doTimeStep calls)
    float dt = 0;
    int cycle = 1;           // CycleStep
    long long realStep = 1;  // real_stepsize
    bool needCycle = false;  // a formula reads the sampled centre of mass or angle
    bool reinitPending = false;  // SecondaryExperiment's one-time re-initialisation
    bool active = true;
    bool diverged = false;
    double com[3] = {0, 0, 0}, comInit[3] = {0, 0, 0}, hist[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double recentAngle = 0;
    int angleSamples = 0;
    int hit = 0;               // collisionCount
    double closeness = 0;      // targetCloseness
    int pairs = 0;             // numClosePairs
    std::vector<int> targets;  // target voxels (registerTargets, VX3_VoxelyzeKernel.cu:536-543)
    std::vector<double> initPos;
};

bool stop_met(const RobotRun &r) {
    if (r.m->stopFormula.empty()) return false;
    const double vars[MV_COUNT] = {r.com[0], r.com[1], r.com[2], (double)r.hit, r.t, r.recentAngle, r.closeness,
                                   (double)r.pairs, (double)r.m->nv()};
    double v = 0;
    return mt_eval(r.m->stopFormula, vars, &v) && v > 0;
}

// CycleStep block of doTimeStep (VX3_VoxelyzeKernel.cu:316-331)
void cycle_sample(RobotRun &r, const double *com) {
    r.angleSamples++;
    std::memcpy(r.hist[0], r.hist[1], sizeof(r.hist[0]));
    std::memcpy(r.hist[1], r.com, sizeof(r.com));
    std::memcpy(r.com, com, sizeof(r.com));
    const double *A = r.hist[0], *B = r.hist[1], *C = r.com;
    const bool bc = B[0] == C[0] && B[1] == C[1] && B[2] == C[2], ab = A[0] == B[0] && A[1] == B[1] && A[2] == B[2];
    if (bc || ab || r.angleSamples < 3) {
        r.recentAngle = 0;
    } else {
        const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, w[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
        const double d = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
        const double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        r.recentAngle = std::acos(d / (lu * lw));
    }
}

// computeTargetCloseness (VX3_VoxelyzeKernel.cu:545-562): over the target pairs,
// the sum of 1 / distance and the number closer than MaxDist voxel lengths
void target_closeness(RobotRun &r, const double *pos) {
    const double R = r.m->maxDistPairs * r.m->latticeDim;
    double ret = 0;
    r.pairs = 0;
    for (size_t i = 0; i < r.targets.size(); i++)
        for (size_t j = i + 1; j < r.targets.size(); j++) {
            const double *a = pos + 3 * r.targets[i], *b = pos + 3 * r.targets[j];
            const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (d < R) r.pairs++;
            ret += 1 / d;
        }
    r.closeness = ret;
}

void appendf(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string &s, const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (n < (int)sizeof(buf)) {
        s.append(buf, n);
    } else {
        std::string big((size_t)n + 1, '\0');
        va_start(ap, fmt);
        vsnprintf(&big[0], big.size(), fmt, ap);
        va_end(ap);
        s.append(big.data(), n);
    }
}

// .history header (VX3_SimulationManager.cu:40-51)
std::string history_header(const Model &m) {
    std::string s = "\n{{{setting}}}<rescale>0.001</rescale>\n";
    for (size_t k = 0; k < m.palette.size(); k++) {
        const vx3mi_material &p = m.palette[k];
        auto ci = [](float c) { int v = (int)(c * 255); return v > 255 ? 255 : (v < 0 ? 0 : v); };
        appendf(s, "{{{setting}}}<matcolor><id>%d</id><r>%.2f</r><g>%.2f</g><b>%.2f</b><a>%.2f</a></matcolor>\n",
                (int)k + 1, ci(p.red) / 255., ci(p.green) / 255., ci(p.blue) / 255., ci(p.alpha) / 255.);
    }
    appendf(s, "\n{{{setting}}}<voxel_size>%f</voxel_size>\n", m.latticeDim);
    return s;
}

// one .history frame (VX3_SimulationManager.cu:61-111) from the robot's
// positions, orientations, temperatures and link strains.  With removal
// (vrem / lrem: the step count at whose end a voxel / link went, 0 present;
// stampNow: the step count after the frame's step) or attach (att: the links
// attach made that are present, with the step count at whose end each was
// made) the voxels listed are those of the surface list regenerated at the
// start of the frame's step that are still present, removed links are left
// out and attached links follow the lattice links.
std::string history_frame(const Model &m, long long j, double t, const double *pos, const double *q,
                          const float *temp, const float *strain, const float *maxStrain, const int *vrem = nullptr,
                          const int *lrem = nullptr, long long stampNow = 0, const double *signal = nullptr,
                          const std::vector<int4> *att = nullptr) {
    std::string s;
    const double vs = 1 / 0.001;
    auto link_at = [&](int v, int d) {  // the voxel's link in direction d now, -1 none
        const int l = m.vlinks[(size_t)v * 6 + d];
        return (l >= 0 && lrem && lrem[l]) ? -1 : l;
    };
    std::vector<int> listed;
    if (vrem || att) {
        std::vector<int> extra((size_t)m.nv(), 0);  // slots attach had filled when the frame's step began
        if (att)
            for (const int4 &a : *att)
                if (a.w < stampNow) { extra[a.x]++; extra[a.y]++; }
        for (int v = 0; v < m.nv(); v++) {
            if (vrem && vrem[v]) continue;
            int filled = extra[v];
            for (int d = 0; d < 6; d++) {
                const int l = m.vlinks[(size_t)v * 6 + d];
                filled += l >= 0 && !(lrem && lrem[l] != 0 && lrem[l] < stampNow);
            }
            if (filled < 6) listed.push_back(v);
        }
    }
    const std::vector<int> &voxels = (vrem || att) ? listed : m.surface;
    if (m.recordVoxel) {
        appendf(s, "<<<Step%lld Time:%f>>>", j, t);
        for (int v : voxels) {
            appendf(s, "%.1f,%.1f,%.1f,", pos[3 * v] * vs, pos[3 * v + 1] * vs, pos[3 * v + 2] * vs);
            const double w = q[4 * v];
            const double angle = 2.0 * std::acos(w > 1 ? 1 : w) * 57.29577951308232;  // AngleDegrees
            appendf(s, "%.1f,%.2f,%.2f,%.2f,", angle, q[4 * v + 1], q[4 * v + 2], q[4 * v + 3]);
            // cornerOffset(NNN) and cornerOffset(PPP) (VX3_Voxel.cu:147-159)
            const VoxelMat &vm = m.vmats[m.vmat[v]];
            const double half = 0.5 * (vm.nomSize * (double)(1 + temp[v] * vm.alphaCTE));
            float off[2][3];
            for (int corner = 0; corner < 2; corner++) {
                const bool posLink = corner == 1;
                for (int i = 0; i < 3; i++) {
                    const int l = link_at(v, 2 * i + (posLink ? 0 : 1));
                    const float epsFail = l >= 0 ? m.lmats[m.lmat[l]].epsFail : -1.0f;
                    double st;
                    if (l >= 0 && !(epsFail != -1.0f && maxStrain[l] > epsFail)) {  // present, not failed
                        const float ratio = m.vmatE[m.vmat[m.lpos[l]]] / m.vmatE[m.vmat[m.lneg[l]]];
                        const float a = posLink ? 2.0f * strain[l] * ratio / (1.0f + ratio) : 2.0f * strain[l] / (1.0f + ratio);
                        st = (1 + a) * (posLink ? 1 : -1);
                    } else {
                        st = posLink ? 1.0 : -1.0;
                    }
                    off[corner][i] = (float)(half * st);
                }
            }
            appendf(s, "%.1f,%.1f,%.1f,%.1f,%.1f,%.1f,", off[0][0] * vs, off[0][1] * vs, off[0][2] * vs, off[1][0] * vs,
                    off[1][1] * vs, off[1][2] * vs);
            appendf(s, "%d,", m.vpal[v]);
            appendf(s, "%.1f,", signal ? signal[v] : 0.0);  // localSignal
            s += ";";
        }
        s += "<<<>>>";
    }
    if (m.recordLink) {
        appendf(s, "|[[[%lld]]]", j);
        for (int l = 0; l < m.nl(); l++) {
            if (lrem && lrem[l]) continue;
            const int a = m.lpos[l], b = m.lneg[l];
            appendf(s, "%.4f,%.4f,%.4f,", pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
            appendf(s, "%.4f,%.4f,%.4f,", pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]);
            s += ";";
        }
        if (att)
            for (const int4 &l : *att) {
                appendf(s, "%.4f,%.4f,%.4f,", pos[3 * l.y], pos[3 * l.y + 1], pos[3 * l.y + 2]);
                appendf(s, "%.4f,%.4f,%.4f,", pos[3 * l.x], pos[3 * l.x + 1], pos[3 * l.x + 2]);
                s += ";";
            }
        s += "[[[]]]";
    }
    s += "\n";
    return s;
}

}  // namespace
}  // namespace vx3mi

using namespace vx3mi;

extern "C" int vx3mi_run(vx3mi_batch *bh, vx3mi_model *const *models, int n, long long max_steps, int mode,
                         vx3mi_text_fn history, void *user, vx3mi_result *results, double *const *final_positions) {
    if (!bh || !bh->b || !models || !results) { set_error("vx3mi_run: NULL argument"); return VX3MI_E_ARG; }
    BatchBase &b = *bh->b;
    if (n != b.R) { set_error("vx3mi_run: models do not match the batch"); return VX3MI_E_ARG; }
    for (int r = 0; r < n; r++)
        if (!models[r] || models[r]->m.nv() != b.hrp[r].vcount || models[r]->m.nl() != b.hrp[r].lcount) {
            set_error("vx3mi_run: models do not match the batch");
            return VX3MI_E_ARG;
        }
    if (max_steps <= 0) max_steps = 1000000;  // CUDA_Simulation's loop bound
    std::vector<RobotRun> run((size_t)n);
    std::vector<double> com((size_t)3 * n);
    std::vector<double> pos, quat;
    std::vector<float> temp, strain, maxStrain;
    std::vector<int> vrem, lrem;
    std::vector<int4> attached;
    std::vector<double> signal;
    int rc = b.com(com.data());
    if (rc) return rc;
    for (int r = 0; r < n; r++) {
        RobotRun &x = run[r];
        const Model &m = models[r]->m;
        x.m = &m;
        x.dt = b.defaultDt[r];
        x.cycle = std::max(1, (int)(m.tempPeriod / (double)x.dt));
        const double rec = recommended_dt(m);
        x.realStep = (long long)(m.recordStepSize / (10000 * rec * m.dtFrac)) + 1;
        const unsigned su = mt_vars_used(m.stopFormula), fu = mt_vars_used(m.fitnessFormula);
        x.reinitPending = m.secondaryExperiment != 0;
        const unsigned sampled = (1u << MV_ANGLE) | (1u << MV_CLOSENESS) | (1u << MV_PAIRS);
        x.needCycle = (su & ((1u << MV_X) | (1u << MV_Y) | (1u << MV_Z) | sampled)) || (fu & sampled) ||
                      x.reinitPending || m.maxDistPairs != 0;
        for (int v = 0; v < m.nv(); v++)
            if (m.palette[m.vpal[v] - 1].is_target) x.targets.push_back(v);
        std::memcpy(x.com, &com[3 * r], sizeof(x.com));
        std::memcpy(x.comInit, x.com, sizeof(x.com));
        x.initPos.resize((size_t)3 * m.nv());
        for (int v = 0; v < m.nv(); v++) {
            const double L = m.vmats[m.vmat[v]].nomSize;
            x.initPos[3 * v] = m.ix[v] * L;
            x.initPos[3 * v + 1] = m.iy[v] * L;
            x.initPos[3 * v + 2] = m.iz[v] * L;
        }
        if (history && m.recordStepSize) {
            std::string h = history_header(m);
            history(user, r, h.data(), h.size());
        }
        std::memset(&results[r], 0, sizeof(vx3mi_result));
        results[r].real_stepsize = (int)x.realStep;
        results[r].recommended_dt = rec;
    }

    auto finish = [&](int r) -> int {
        RobotRun &x = run[r];
        const Model &m = *x.m;
        x.active = false;
        pos.resize((size_t)3 * m.nv());
        int e = b.frame(r, pos.data(), nullptr, nullptr, nullptr);
        if (e) return e;
        if (final_positions && final_positions[r]) std::memcpy(final_positions[r], pos.data(), sizeof(double) * pos.size());
        vx3mi_result &o = results[r];
        o.current_time = x.t;
        o.steps = x.j;
        o.diverged = x.diverged;
        o.angle_sample_times = x.needCycle ? x.angleSamples : (int)(x.j / x.cycle);
        o.num_voxel = m.nv();
        o.num_close_pairs = x.pairs;
        o.vox_size = m.latticeDim;
        for (int k = 0; k < 3; k++) {
            o.initial_com[k] = x.comInit[k];
            o.current_com[k] = com[3 * r + k];
        }
        for (int v = 0; v < m.nv(); v++) {
            if (!m.vmats[m.vmat[v]].isMeasured) continue;
            o.num_measured_voxel++;
            const double dx = pos[3 * v] - x.initPos[3 * v], dy = pos[3 * v + 1] - x.initPos[3 * v + 1],
                         dz = pos[3 * v + 2] - x.initPos[3 * v + 2];
            o.total_distance_of_all_voxels += std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        o.fitness_score = 0;
        if (!m.fitnessFormula.empty()) {
            const double vars[MV_COUNT] = {o.current_com[0] - x.comInit[0], o.current_com[1] - x.comInit[1],
                                           o.current_com[2] - x.comInit[2], (double)x.hit, x.t, x.recentAngle,
                                           x.closeness, (double)x.pairs, (double)m.nv()};
            mt_eval(m.fitnessFormula, vars, &o.fitness_score);
        }
        return VX3MI_OK;
    };

    std::vector<RobotTime> rt((size_t)n);
    HIP_TRY(hipMemcpy(rt.data(), b.d_rt, sizeof(RobotTime) * n, hipMemcpyDeviceToHost));
    std::vector<int> steps0((size_t)n);
    for (int r = 0; r < n; r++) steps0[r] = rt[r].steps;
    for (;;) {
        // stop checks at the current iteration, then the next event of every robot
        long long seg = -1;
        bool any = false;
        for (int r = 0; r < n; r++) {
            RobotRun &x = run[r];
            if (!x.active) continue;
            if (x.j >= max_steps || stop_met(x)) {
                if ((rc = finish(r))) return rc;
                continue;
            }
            any = true;
            long long k = max_steps - x.j;                                   // loop bound
            if (x.m->recordStepSize && history) {
                const long long nextRec = ((x.j + x.realStep - 1) / x.realStep) * x.realStep;  // iteration index
                k = std::min(k, nextRec - x.j + 1);
            }
            if (x.needCycle) k = std::min(k, (long long)(x.cycle - (x.j % x.cycle)));
            if (x.reinitPending) {
                // doTimeStep re-initialises after the step that starts past the time
                // (VX3_VoxelyzeKernel.cu:336-349)
                double tt = x.t;
                for (long long s = 1; s <= k; s++) {
                    if (x.m->reinitInitialPositionAfter < tt) { k = s; break; }
                    tt += (double)x.dt;
                }
            }
            // the stop formula can only change with t (and hit, every step) between samples
            if (!x.m->stopFormula.empty() && (mt_vars_used(x.m->stopFormula) & (1u << MV_HIT))) k = 1;
            if (!x.m->stopFormula.empty() && (mt_vars_used(x.m->stopFormula) & (1u << MV_T))) {
                RobotRun probe = x;
                for (long long s = 1; s < k; s++) {
                    probe.t += (double)x.dt;
                    if (stop_met(probe)) { k = s; break; }
                }
            }
            if (seg < 0 || k < seg) seg = k;
        }
        if (!any) break;
        rc = b.step((int)seg, -1.0f, mode);
        if (rc != VX3MI_OK && rc != VX3MI_DIVERGED) return rc;
        rc = b.com(com.data());
        if (rc) return rc;
        HIP_TRY(hipMemcpy(rt.data(), b.d_rt, sizeof(RobotTime) * n, hipMemcpyDeviceToHost));
        for (int r = 0; r < n; r++) {
            RobotRun &x = run[r];
            if (!x.active) continue;
            if (rt[r].diverged) {
                // the diverging doTimeStep returns before currentTime advances: the
                // robot's device time and step count stop at the last good step
                x.t = std::max(rt[r].time[0], rt[r].time[1]);  // the later of the two step-parity slots
                x.j = (long long)(rt[r].steps - steps0[r]) + 1;
                x.diverged = true;
                if ((rc = finish(r))) return rc;
                continue;
            }
            double tLast = x.t;  // time at the start of the segment's last step
            for (long long s = 0; s < seg; s++) {
                tLast = x.t;
                x.t += (double)x.dt;
            }
            x.j += seg;
            x.hit = rt[r].collisionCount;
            if (b.hrp[r].sample && x.j % x.cycle == 0) {
                // the device took this sample (its force fields or attach conditions read it)
                const RobotTime &t = rt[r];
                std::memcpy(x.hist[0], t.shist[0], sizeof(x.hist[0]));
                std::memcpy(x.hist[1], t.shist[1], sizeof(x.hist[1]));
                std::memcpy(x.com, t.scom, sizeof(x.com));
                x.recentAngle = t.recentAngle;
                x.angleSamples = t.angleSamples;
                if (x.m->maxDistPairs != 0) {
                    x.closeness = t.closeness;
                    x.pairs = t.closePairs;
                }
            } else if (x.needCycle && x.j % x.cycle == 0) {
                cycle_sample(x, &com[3 * r]);
                if (x.m->maxDistPairs != 0) {
                    pos.resize((size_t)3 * x.m->nv());
                    if ((rc = b.frame(r, pos.data(), nullptr, nullptr, nullptr))) return rc;
                    target_closeness(x, pos.data());
                }
            }
            if (x.reinitPending && x.m->reinitInitialPositionAfter < tLast) {
                x.reinitPending = false;
                std::memcpy(x.comInit, x.com, sizeof(x.com));  // InitializeCenterOfMass: the sampled centre of mass
                if ((rc = b.frame(r, x.initPos.data(), nullptr, nullptr, nullptr))) return rc;  // saveInitialPosition
            }
            if (x.m->recordStepSize && history && (x.j - 1) % x.realStep == 0) {
                const Model &m = *x.m;
                pos.resize((size_t)3 * m.nv());
                quat.resize((size_t)4 * m.nv());
                temp.resize((size_t)m.nv());
                strain.resize((size_t)std::max(1, m.nl()));
                maxStrain.resize(strain.size());
                if ((rc = b.frame(r, pos.data(), quat.data(), temp.data(), strain.data(), maxStrain.data()))) return rc;
                const int *vr = nullptr, *lr = nullptr;
                if (b.anyRemoval && m.secondaryExperiment) {
                    vrem.resize((size_t)m.nv());
                    lrem.resize((size_t)std::max(1, m.nl()));
                    if ((rc = b.get_removed(r, vrem.data(), lrem.data()))) return rc;
                    vr = vrem.data();
                    lr = lrem.data();
                }
                const double *sg = nullptr;
                if (m.enableSignals) {
                    signal.resize((size_t)m.nv());
                    if ((rc = b.get_signals(r, signal.data()))) return rc;
                    sg = signal.data();
                }
                const std::vector<int4> *at = nullptr;
                if (b.anyAttach && m.enableAttach) {
                    if ((rc = b.attach_links(r, attached))) return rc;
                    at = &attached;
                }
                const std::string f = history_frame(m, x.j - 1, x.t, pos.data(), quat.data(), temp.data(), strain.data(),
                                                    maxStrain.data(), vr, lr, (long long)steps0[r] + x.j, sg, at);
                history(user, r, f.data(), f.size());
            }
        }
    }
    return VX3MI_OK;
}
