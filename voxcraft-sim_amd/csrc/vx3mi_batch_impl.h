// vx3mi_batch_impl.h -- device side of the VX3 timestep for gfx950 (templated on
// the state precision; instantiated by vx3mi_batch_fp32.hip / _fp64.hip).
//
// Data layout (structure of arrays, one entry per voxel / link of the whole
// batch; robots are CSR-packed by voxel and link offsets):
//   voxel  pq0 = (ux, uy, uz, qw)  pq1 = (qx, qy, qz, T)   u: displacement from
//          mo0 = (px, py, pz, hx)  mo1 = (hy, hz, flags, 0)  the lattice point
//          lat (ix | iy<<10 | iz<<20), vmat, vrob, phase, vlinks[6]
//   link   ls0..2 = pos2, angle1v, angle2v, maxStrain, flags, strain
//          lf0..2 = forceNeg, momentNeg, forcePos, momentPos (voxel frames)
//          lends = (pVNeg, pVPos), lmeta = (axis | linkMat<<2, robot)
// T is the temperature of the voxel for the step about to run; the voxel
// kernel of step n writes the temperature of step n+1 (gpu_update_temperature,
// VX3_VoxelyzeKernel.cu:625-650, evaluated one step ahead).  Per-robot time and
// previous dt are double-buffered by step parity so that no thread reads a
// value another thread of the same launch writes.
//
// Kernels per step: k_links (one work-item per link, writes both end forces;
// no atomics), k_voxels (one work-item per voxel, gathers its <=6 link ends in
// direction order X_POS..Z_NEG, integrates).  k_persistent runs many steps with
// one workgroup per robot and workgroup barriers between the two phases.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vx3mi_batch.h"

namespace vx3mi {

template <typename T> struct Dev {
    T4<T> *pq0, *pq1, *mo0, *mo1;
    const int *lat, *vmat, *vrob;
    const T *phase;
    int *vlinks;           // edited by removal, detach and attach
    int2 *lends, *lmeta;   // written for the link entries attach hands out
    T4<T> *ls0, *ls1, *ls2;
    T4<T> *lf0, *lf1, *lf2;
    const VoxelMat *vmats;
    const VoxelExt *exts;  // voxel externals (VoxelMat::ext indexes it)
    const LinkMat *lmats;
    const CurvePt *curves; // stress-strain curves of non-linear link materials
    const MtToken *formulas; // force-field postfix programs (RobotParams::ffOff / ffLen)
    T4<T> *field;            // per-voxel force-field force of the step, or nullptr without force fields
    const RobotParams *rp;
    RobotTime *rt;
    const T4<T> *contact;  // per-voxel contact force, or nullptr without collisions
    // collision broadphase (candidate lists, see the collision section)
    const int *surf;       // surface voxel ids, CSR by robot
    const int *vsurf;      // voxel -> its index in surf, -1 when not listed
    T4<T> *bpos;           // per voxel: position field at the last listing (surface voxels)
    int *anyStale;         // per step parity: some robot's lists are stale
    int *ctick;            // per robot: workgroups done with their part of the count pass (zero between listings)
    int4 *scell;           // per surface voxel: grid cell and bucket at the last listing
    int *srank;            // per surface voxel: rank within its bucket (rebuild)
    int *ccnt;             // bucket counts (rebuild; zero between rebuilds)
    int *cstart;           // bucket starts within the robot's records, mask + 2 per robot
    int4 *grec;            // a robot's surface voxels in bucket order, CSR by robot: cell at the listing, voxel id
    T4<T> *gpos;           // with grec: position field at the listing, lattice site in the fourth word
    int *clist;            // per surface voxel: kMaxCand candidate surface indices, ascending
    int *ccount;           // per surface voxel: candidates (> kMaxCand: search the grid)
    int *tcand;            // collision tiles (k_tile): packed candidates, kMaxCand x KV x NT words per tile
    T4<T> *contactW;
    int2 *pairRec;         // when recording: pairs in contact (voxel ids, smaller surface index first)
    int *pairN;
    int pairCap;
    const VoxelCilia *cilia;  // per-voxel cilia, or nullptr when no robot has EnableCilia
    // removal (SecondaryExperiment): step count at whose end a voxel / link was
    // removed (0 present), or nullptr when no robot can lose voxels
    int *vrem, *lrem;
    const int *remCand;       // voxels of materials with RemoveFromSimulationAfterThisManySeconds > 0
    const double *remAt;      // that time, per candidate
    int nRemCand;
    const int *detCand;       // links of EnableDetach robots that can fail
    int nDetCand;
    VoxelSignal *sig;         // per-voxel signal state, or nullptr when no robot has EnableSignals
    const SignalMat *sigMat;  // per-voxel signal constants
    const int *vtarget;       // per-voxel isTarget, or nullptr when no voxel is a target
    // attach (nullptr when no robot has EnableAttach): per voxel the palette
    // material it attaches with (-1: never) and the link material of such a
    // link, and its enableAttach flag; per surface voxel this step's candidates
    const int2 *vatt;
    int *vflagA;
    int *acand;               // kAttachPerVoxel partner voxels per surface voxel, ascending
    int *acount;              // candidates found (beyond kAttachPerVoxel: not kept)
    int *lmade;               // per link entry: the step count at whose end attach made it (0: lattice)
    const int *targets;       // target voxels of robots with device samples, CSR by robot (RobotParams::tgOff)
};

// ------------------------------------------------------------------ device
// absolute position in double of a voxel's position field u
template <typename T> __device__ __forceinline__ V3<double> position_d(int lat, double nomSize, V3<T> u) {
    if constexpr (kAbsolutePositions<T>) return V3<double>{u.x, u.y, u.z};
    return V3<double>{(lat & 1023) * nomSize + (double)u.x, ((lat >> 10) & 1023) * nomSize + (double)u.y,
                      (lat >> 20) * nomSize + (double)u.z};
}

// lattice offset of w from v
__device__ __forceinline__ int3 lat_offset(int pkv, int pkw) {
    return int3{(pkw & 1023) - (pkv & 1023), ((pkw >> 10) & 1023) - ((pkv >> 10) & 1023), (pkw >> 20) - (pkv >> 20)};
}

template <typename T> __device__ __forceinline__ void link_update(const Dev<T> &a, int l, int parity) {
    // the link's own words and its ends' poses are loaded before the checks
    // (each check returns early, and a load after it would wait for it)
    const int2 meta = a.lmeta[l];
    const int2 e = a.lends[l];
    const T4<T> s0 = a.ls0[l], s1 = a.ls1[l], s2 = a.ls2[l];
    const T4<T> n0 = a.pq0[e.x], n1 = a.pq1[e.x], p0 = a.pq0[e.y], p1 = a.pq1[e.y];
    RobotTime *rt = a.rt + meta.y;
    if (rt->diverged) return;
    if (a.lrem && a.lrem[l]) return;  // removed link (gpu_update_links :570-571)
    const LinkMat &lm = a.lmats[meta.x >> 2];
    if (lm.fixedBoth) return;  // gpu_update_links VX3_VoxelyzeKernel.cu:572-573
    const int axis = meta.x & 3;
    V3<T> du = {p0.x - n0.x, p0.y - n0.y, p0.z - n0.z};
    Q4<T> qn = {n0.w, n1.x, n1.y, n1.z}, qp = {p0.w, p1.x, p1.y, p1.z};
    LinkState<T> st;
    st.pos2 = {s0.x, s0.y, s0.z};
    st.a1v = {s0.w, s1.x, s1.y};
    st.a2v = {s1.z, s1.w, s2.x};
    st.maxStrain = (float)s2.y;
    st.flags = (int)s2.z;
    if constexpr (!kAbsolutePositions<T>) {
        // the displacement form takes the ends one voxel apart along the axis;
        // an attached link adds the rest of its ends' lattice offset
        if (st.flags & LF_ATTACHED) {
            const int3 lo = lat_offset(a.lat[e.x], a.lat[e.y]);
            const T L = (T)lm.nomSize;
            du.x += (T)(lo.x - (axis == 0)) * L;
            du.y += (T)(lo.y - (axis == 1)) * L;
            du.z += (T)(lo.z - (axis == 2)) * L;
        }
    }
    LinkOut<T> o;
    link_forces<T>(lm, a.curves, axis, du, qn, qp, (float)n1.w, (float)p1.w, rt->prevDt[parity], st, o);
    // isNewLink (VX3_Link.cu:206-213): the first SafetyGuard updates of a link
    // made by attach are scaled by 0.01, a failed link's zero forces aside
    if ((st.flags >> LF_NEW_SHIFT) > 0 && !(lm.epsFail != -1.0f && st.maxStrain > lm.epsFail)) {
        const T k = (T)0.01;
        o.fN = o.fN * k; o.fP = o.fP * k; o.mN = o.mN * k; o.mP = o.mP * k;
        st.flags -= 1 << LF_NEW_SHIFT;
    }
    a.ls0[l] = T4<T>{st.pos2.x, st.pos2.y, st.pos2.z, st.a1v.x};
    a.ls1[l] = T4<T>{st.a1v.y, st.a1v.z, st.a2v.x, st.a2v.y};
    a.ls2[l] = T4<T>{st.a2v.z, (T)st.maxStrain, (T)st.flags, (T)o.strain};
    a.lf0[l] = T4<T>{o.fN.x, o.fN.y, o.fN.z, o.mN.x};
    a.lf1[l] = T4<T>{o.mN.y, o.mN.z, o.fP.x, o.fP.y};
    a.lf2[l] = T4<T>{o.fP.z, o.mP.x, o.mP.y, o.mP.z};
    // divergence (VX3_VoxelyzeKernel.cu:271-280): any link, not one sampled link;
    // kept apart from `diverged` until the voxel update, so that work running
    // beside the links in the same step (the contact phase) sees the flag of
    // the earlier steps only
    if (o.strain > 100.0f) atomicOr(&rt->divergedNow, 1);
}

template <typename T> __device__ __forceinline__ void col_track(const Dev<T> &a, int v, const V3<T> &u, RobotTime *rt,
                                                                const RobotParams &rp, int parity);

// an empty asm that takes x as it stands here: a load issued earlier stays in
// flight until this point, and the arithmetic on x is not moved above it (the
// FP32 code contracts multiplies and adds within a basic block, so the blocks
// the arithmetic sits in decide its rounding)
template <typename T> __device__ __forceinline__ void pin(T &x) { asm volatile("" : "+v"(x)); }
template <typename T> __device__ __forceinline__ void pin(T4<T> &x) {
    pin(x.x);
    pin(x.y);
    pin(x.z);
    pin(x.w);
}

template <typename T> __device__ __forceinline__ void voxel_update(const Dev<T> &a, int v, int parity) {
    // the loads that need only v (and the link forces, once the link ids are in)
    // are issued before the robot's checks: each check returns early, and a load
    // after it would wait for the loads before it.  The values are pinned where
    // the arithmetic starts, which keeps it in the blocks it had with the loads
    // there (the same rounding as the tiled kernel's voxel update)
    const int r = a.vrob[v];
    const int mat = a.vmat[v];
    int vl[6];
#pragma unroll
    for (int d = 0; d < 6; d++) vl[d] = a.vlinks[6 * v + d];
    T4<T> q0 = a.pq0[v], q1 = a.pq1[v], m0 = a.mo0[v], m1 = a.mo1[v];
    T4<T> lfA[6], lfB[6];  // force and moment words of each present link at this voxel's end
#pragma unroll
    for (int d = 0; d < 6; d++) {
        const int l = vl[d];
        if (l < 0) continue;
        if (d & 1) {  // link in a NEG direction: this voxel is its pVPos end
            lfA[d] = a.lf1[l];
            lfB[d] = a.lf2[l];
        } else {
            lfA[d] = a.lf0[l];
            lfB[d] = a.lf1[l];
        }
    }
    RobotTime *rt = a.rt + r;
    {  // this step's listing flags: the lists are built and read by now
        const RobotParams &rp = a.rp[r];
        if (v == rp.voff && rp.scount > 0 && rt->stale[parity]) {
            rt->listings += 1;
            rt->stale[parity] = 0;
        }
        if (v == 0 && a.anyStale) a.anyStale[parity] = 0;
    }
    if (rt->diverged) return;
    if (rt->divergedNow) {  // doTimeStep returns before the voxels move (:271-280)
        rt->diverged = 1;
        return;
    }
    const RobotParams &rp = a.rp[r];
    const float dt = rt->dt;
    const double tnext = rt->time[parity] + (double)dt;  // currentTime += dt (:352)
    if (v == rp.voff) {
        rt->time[parity ^ 1] = tnext;
        rt->prevDt[parity ^ 1] = dt;
        rt->steps += 1;
    }
    if (a.vrem && a.vrem[v]) return;  // removed voxel (gpu_update_voxels :586-587)
    const VoxelMat &vm = a.vmats[mat];
    if (vm.fixed) return;  // gpu_update_voxels :588-589

    V3<T> fs = {0, 0, 0}, ms = {0, 0, 0};
    bool surface = false;  // fewer than six links (VX3_Voxel::updateSurface, VX3_Voxel.cu:515-524)
#pragma unroll
    for (int d = 0; d < 6; d++) {
        if (vl[d] < 0) { surface = true; continue; }
        pin(lfA[d]);
        pin(lfB[d]);
        if (d & 1) {
            fs = fs + V3<T>{lfA[d].z, lfA[d].w, lfB[d].x};
            ms = ms + V3<T>{lfB[d].y, lfB[d].z, lfB[d].w};
        } else {
            fs = fs + V3<T>{lfA[d].x, lfA[d].y, lfA[d].z};
            ms = ms + V3<T>{lfA[d].w, lfB[d].x, lfB[d].y};
        }
    }
    pin(q0);
    pin(q1);
    pin(m0);
    pin(m1);
    VoxelState<T> s;
    s.u = {q0.x, q0.y, q0.z};
    s.q = {q0.w, q1.x, q1.y, q1.z};
    s.p = {m0.x, m0.y, m0.z};
    s.h = {m0.w, m1.x, m1.y};
    s.flags = (int)m1.z;
    float temp = (float)q1.w;
    V3<T> contact = {0, 0, 0};
    if (a.contact) {
        const T4<T> c = a.contact[v];
        contact = {c.x, c.y, c.z};
    }
    V3<T> field = {0, 0, 0};
    if (a.field) {
        const T4<T> f = a.field[v];
        field = {f.x, f.y, f.z};
    }
    V3<T> cil = {0, 0, 0};
    if (a.cilia) {
        // gpu_update_cilia_force (VX3_VoxelyzeKernel.cu:846-858): surface voxels of a
        // cilia material once TurnOnCiliaAfterThisManySeconds has passed, with the
        // orientation and the signal entering the step
        const VoxelCilia &c = a.cilia[v];
        if (surface && c.coef != 0 && !(c.onAfter > rt->time[parity])) {
            const double ls = a.sig ? a.sig[v].local : 0.0;  // localSignal
            const V3<T> b = {(T)(c.base[0] + ls * c.shift[0]), (T)(c.base[1] + ls * c.shift[1]),
                             (T)(c.base[2] + ls * c.shift[2])};
            cil = qrot(s.q, b) * (T)c.coef;
        }
    }
    voxel_step<T>(vm, a.exts, dt, a.lat[v], temp, fs, ms, contact, field, s, a.cilia ? &cil : nullptr);
    if (rp.varyTemp && !(vm.thermalOnAfter > tnext)) {
        if constexpr (std::is_same<T, float>::value)
            temp = actuation_temperature_f32(rp.tempAmp, rp.tempInvPeriod, tnext, (double)a.phase[v], rp.enableExpansion);
        else
            temp = actuation_temperature<T>(rp.tempAmp, rp.tempPeriod, tnext, (double)a.phase[v], rp.enableExpansion);
    }
    a.pq0[v] = T4<T>{s.u.x, s.u.y, s.u.z, s.q.w};
    a.pq1[v] = T4<T>{s.q.x, s.q.y, s.q.z, (T)temp};
    a.mo0[v] = T4<T>{s.p.x, s.p.y, s.p.z, s.h.x};
    a.mo1[v] = T4<T>{s.h.y, s.h.z, (T)s.flags, (T)0};
    if (rp.collisions || rp.attach) col_track(a, v, s.u, rt, rp, parity);  // attach uses the candidate lists too
    // the voxel's own signal section ends a full timeStep (not one of a voxel
    // with all six degrees of freedom fixed, VX3_Voxel.cu:166-172)
    if (a.sig && rp.signals && dt != 0.0f && !(vm.ext >= 0 && a.exts[vm.ext].dof == DOF_ALL)) {
        VoxelSignal sg = a.sig[v];
        signal_own(sg, a.sigMat[v], rt->time[parity]);
        a.sig[v] = sg;
    }
}

// the receive phase of signals: each present voxel takes the signals its linked
// neighbours sent in this step, in its link direction order (the own section of
// every voxel precedes what it receives)
template <typename T> __global__ __launch_bounds__(256) void k_signal_deliver(Dev<T> a, int v0, int n) {
    const int v = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= v0 + n) return;
    if (!a.rp[a.vrob[v]].signals || a.rt[a.vrob[v]].diverged || (a.vrem && a.vrem[v])) return;
    const SignalMat sm = a.sigMat[v];
    for (int d = 0; d < 6; d++) {
        const int l = a.vlinks[6 * v + d];
        if (l < 0) continue;
        const int o = (d & 1) ? a.lends[l].x : a.lends[l].y;
        const double value = a.sig[o].outValue;
        if (value != 0) {
            VoxelSignal &me = a.sig[v];
            signal_receive(me, sm, value, a.sig[o].outTime, false);
        }
    }
}

// ---- force fields (VX3_Voxel.cu:177-183) ---------------------------------------
// The force-field force of each voxel for the coming voxel update, from the
// position entering the step and the time before it (a kernel of its own: the
// formula evaluator stays out of the update kernels).
template <typename T> __global__ __launch_bounds__(256) void k_force_field(Dev<T> a, int v0, int nv, int parity) {
    const int v = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= v0 + nv) return;
    const int r = a.vrob[v];
    const RobotParams &rp = a.rp[r];
    const RobotTime *rt = a.rt + r;
    V3<T> f = {0, 0, 0};
    if (!rt->diverged && (rp.ffLen[0] | rp.ffLen[1] | rp.ffLen[2])) {
        const T4<T> q0 = a.pq0[v];
        const V3<double> p = position_d<T>(a.lat[v], a.vmats[a.vmat[v]].nomSize, V3<T>{q0.x, q0.y, q0.z});
        f = force_field<T>(a.formulas, rp.ffOff, rp.ffLen, p.x, p.y, p.z, rt->time[parity], rp.vcount, rt->collisionCount,
                           rt->recentAngle, rt->closeness, rt->closePairs);
    }
    a.field[v] = T4<T>{f.x, f.y, f.z, (T)0};
}

// ---- collisions: candidate lists from a counting-sort grid --------------------
// Replaces the all-pairs surface loop (updateAttach VX3_VoxelyzeKernel.cu:448-458,
// gpu_update_attach :833-843).  Every surface voxel keeps the ascending list of
// its robot's surface voxels within listR (the largest contact envelope plus a
// skin of kSkin or, for small robots, kSkinSmall voxel lengths).  The lists are built from a uniform grid of
// cell edge listR by a counting sort -- count per bucket, prefix sum, scatter
// in bucket order -- whenever a surface voxel of the robot has moved more than
// half the skin since the last listing, and each step the contact phase tests
// every listed candidate against the envelope exactly.  A pair inside the
// envelope now was within listR when the lists were built (each end moved less
// than half the skin since), so the partners are exactly those of the all-pairs
// loop, in ascending order, whenever the lists were built: results do not
// depend on the rebuild schedule.
__device__ __forceinline__ unsigned cell_hash(int cx, int cy, int cz) {
    return ((unsigned)cx * 73856093u) ^ ((unsigned)cy * 19349663u) ^ ((unsigned)cz * 83492791u);
}

// a voxel of the surface list that currently takes part in collisions: with
// removal the list also holds voxels that become surface when a neighbour goes
// (regenerateSurfaceVoxels VX3_VoxelyzeKernel.cu:495-513 keeps surface voxels
// that are not removed)
template <typename T> __device__ __forceinline__ bool col_active(const Dev<T> &a, int v) {
    if (!a.vrem) return true;
    if (a.vrem[v]) return false;
    bool surface = false;
    for (int d = 0; d < 6; d++) surface |= a.vlinks[6 * v + d] < 0;
    return surface;
}

// v and w are directly linked (is_neighbor depth 1, VX3_VoxelyzeKernel.cu:651-680,
// :701-703): a link joins lattice neighbours only, so a unit lattice offset in
// direction d (X_POS..Z_NEG) is linked when v's link slot d is present
template <typename T> __device__ __forceinline__ bool col_linked(const Dev<T> &a, int v, int lx, int ly, int lz) {
    if (abs(lx) + abs(ly) + abs(lz) != 1) return false;
    const int d = lx ? (lx > 0 ? 0 : 1) : (ly ? (ly > 0 ? 2 : 3) : (lz > 0 ? 4 : 5));
    return a.vlinks[6 * v + d] >= 0;
}

// v and w share a link in any slot: the test for robots whose links can join
// voxels that are not lattice neighbours (attach)
template <typename T> __device__ __forceinline__ bool col_linked_any(const Dev<T> &a, int v, int w) {
    for (int d = 0; d < 6; d++) {
        const int l = a.vlinks[6 * v + d];
        if (l >= 0) {
            const int2 e = a.lends[l];
            if (e.x == w || e.y == w) return true;
        }
    }
    return false;
}


// a lattice site carried in the fourth word of a position record: the bits of
// the packed site in float, its value in double (exact below 2^53)
template <typename T> __device__ __forceinline__ T lat_word(int lat) {
    if constexpr (sizeof(T) == 4) return __int_as_float(lat);
    else return (T)lat;
}
template <typename T> __device__ __forceinline__ int word_lat(T w) {
    if constexpr (sizeof(T) == 4) return __float_as_int(w);
    else return (int)w;
}

// pos_v - pos_w: absolute positions in double; in float the lattice offset and
// the displacement difference are kept apart
template <typename T> __device__ __forceinline__ V3<T> col_diff(const T4<T> &pv, const T4<T> &pw, int3 o, T L) {
    if constexpr (kAbsolutePositions<T>) return V3<T>{pv.x - pw.x, pv.y - pw.y, pv.z - pw.z};
    return V3<T>{(T)(-o.x) * L + (pv.x - pw.x), (T)(-o.y) * L + (pv.y - pw.y), (T)(-o.z) * L + (pv.z - pw.z)};
}

// rebuild pass 1 (robots whose lists are stale): the cell of each surface voxel,
// its rank within its bucket, its position field at the listing
template <typename T> __device__ __forceinline__ void col_count(const Dev<T> &a, int s, int parity) {
    const int v = a.surf[s];
    const int r = a.vrob[v];
    const RobotTime *rt = a.rt + r;
    if (!rt->stale[parity] || rt->diverged) return;
    const RobotParams &rp = a.rp[r];
    const T4<T> q0 = a.pq0[v];
    const V3<double> p = position_d<T>(a.lat[v], a.vmats[a.vmat[v]].nomSize, V3<T>{q0.x, q0.y, q0.z});
    const double h = rp.listR;
    const int cx = (int)floor(p.x / h), cy = (int)floor(p.y / h), cz = (int)floor(p.z / h);
    const int key = (int)(cell_hash(cx, cy, cz) & (unsigned)rp.mask);
    a.scell[s] = int4{cx, cy, cz, key};
    a.srank[s] = atomicAdd(&a.ccnt[rp.toff + key], 1);
    a.bpos[v] = q0;
}

// rebuild pass 2, one workgroup of blockDim threads (a multiple of 64, at most
// 1024) per robot: exclusive prefix sum of the robot's bucket counts into cstart
// (mask + 2 entries, the last the total), counts zeroed for the next rebuild;
// wsum holds 16 ints of LDS.  WIDE: the buckets go in tiles of four per thread
// (the table of a robot starts on a multiple of four entries and holds a power
// of two, at least 64), so that a wave reads and writes consecutive 16-byte
// words -- a robot of any size; otherwise each thread takes a run of
// consecutive buckets, fewer registers for tables of a few buckets per thread
template <bool WIDE, typename T> __device__ __forceinline__ void col_scan(const Dev<T> &a, int r, int *wsum) {
    const RobotParams &rp = a.rp[r];
    const int nt = blockDim.x, tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, nw = (nt + 63) >> 6;
    // inclusive scan of x over the wave
    auto wave_scan = [&](int x) {
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        return x;
    };
    if constexpr (WIDE) {
        const int n4 = (rp.mask + 1) >> 2;
        int4 *cnt = reinterpret_cast<int4 *>(a.ccnt + rp.toff);
        int4 *start = reinterpret_cast<int4 *>(a.cstart + rp.toff);
        int carry = 0;
        for (int t0 = 0; t0 < n4; t0 += nt) {
            const int i = t0 + tid;
            const int4 c = i < n4 ? cnt[i] : int4{0, 0, 0, 0};
            const int sum = c.x + c.y + c.z + c.w, incl = wave_scan(sum);
            if (lane == 63 || tid == nt - 1) wsum[wv] = incl;
            __syncthreads();
            int e = carry + incl - sum;  // the wave totals before this wave, in order
            for (int w = 0; w < nw; w++) {
                const int x = wsum[w];
                if (w < wv) e += x;
                carry += x;
            }
            if (i < n4) {
                start[i] = int4{e, e + c.x, e + c.x + c.y, e + c.x + c.y + c.z};
                cnt[i] = int4{0, 0, 0, 0};
            }
            __syncthreads();  // wsum is rewritten by the next tile
        }
        if (tid == 0) start[n4].x = carry;
    } else {
        const int n = rp.mask + 1;
        const int chunk = (n + nt - 1) / nt, b0 = min(n, tid * chunk), b1 = min(n, b0 + chunk);
        int *cnt = a.ccnt + rp.toff, *start = a.cstart + rp.toff;
        int sum = 0;
        for (int b = b0; b < b1; b++) sum += cnt[b];
        const int incl = wave_scan(sum);
        if (lane == 63 || tid == nt - 1) wsum[wv] = incl;
        __syncthreads();
        if (tid == 0) {  // exclusive prefix of the wave totals
            int acc = 0;
            for (int i = 0; i < nw; i++) { const int x = wsum[i]; wsum[i] = acc; acc += x; }
            start[n] = acc;
        }
        __syncthreads();
        int acc = wsum[wv] + incl - sum;
        for (int b = b0; b < b1; b++) {
            const int c = cnt[b];
            start[b] = acc;
            acc += c;
            cnt[b] = 0;
        }
    }
}

// rebuild pass 3: surface indices (robot-relative) in bucket order
template <typename T> __device__ __forceinline__ void col_scatter(const Dev<T> &a, int s, int parity) {
    const int v = a.surf[s];
    const int r = a.vrob[v];
    const RobotTime *rt = a.rt + r;
    if (!rt->stale[parity] || rt->diverged) return;
    const RobotParams &rp = a.rp[r];
    const int4 c = a.scell[s];
    const int k = rp.soff + a.cstart[rp.toff + c.w] + a.srank[s];
    const T4<T> q = a.bpos[v];
    a.grec[k] = int4{c.x, c.y, c.z, v};
    a.gpos[k] = T4<T>{q.x, q.y, q.z, lat_word<T>(a.lat[v])};
}

// calls f(w, pw) for every surface voxel w of the robot in the 27 grid cells
// around surface voxel s (voxel v), from the grid of the last listing; pw is
// w's record of that listing (position field, lattice site), loaded with POS
template <bool POS, typename T, typename F>
__device__ __forceinline__ void col_grid_each(const Dev<T> &a, const RobotParams &rp, int s, int v, F &&f, int c0 = 0,
                                              int cstep = 1) {
    const int4 cs = a.scell[s];
    for (int c = c0; c < 27; c += cstep) {  // the cells c0, c0 + cstep, ...
        const int cx = cs.x + c % 3 - 1, cy = cs.y + (c / 3) % 3 - 1, cz = cs.z + c / 9 - 1;
        const int key = (int)(cell_hash(cx, cy, cz) & (unsigned)rp.mask);
        const int b0 = a.cstart[rp.toff + key], b1 = a.cstart[rp.toff + key + 1];
        for (int k = rp.soff + b0; k < rp.soff + b1; k++) {
            const int4 g = a.grec[k];
            T4<T> pw{};
            if constexpr (POS) pw = a.gpos[k];  // in flight with the cell
            if (g.w == v || g.x != cx || g.y != cy || g.z != cz) continue;  // itself, another cell of the bucket
            f(g.w, pw);
        }
    }
}

// rebuild pass 4: the candidate list of surface voxel s, ascending (voxel ids:
// the surface list of a robot is in voxel order, so this is the surface order
// of the reference's pair loop): surface voxels of its robot in the 27 cells
// around it within listR, not both fixed, not directly linked unless the robot
// can lose links or voxels (the contact phase tests the current links then);
// more than kMaxCand: ccount = their number and the voxel searches the grid
// itself
// The candidates of surface voxel s (voxel v) in the cells c0, c0 + cstep, ...
// of the 27, in ascending order into out[0 .. cap); returns their number (also
// beyond cap).  out is global or LDS memory.
template <typename T>
__device__ __forceinline__ int col_list_cells(const Dev<T> &a, const RobotParams &rp, int s, int v, int c0, int cstep,
                                              int *out, int cap) {
    const VoxelMat &mi = a.vmats[a.vmat[v]];
    const T4<T> pv = a.pq0[v];
    const int pkv = a.lat[v];
    const T L = (T)mi.nomSize, R = (T)rp.listR;
    int n = 0;
    col_grid_each<true>(
        a, rp, s, v,
        [&](int w, const T4<T> &pw) {  // pw: w's position now
            const int3 o = lat_offset(pkv, word_lat<T>(pw.w));
            if (!rp.listLinked && col_linked(a, v, o.x, o.y, o.z)) return;
            if (mi.fixed && a.vmats[a.vmat[w]].fixed) return;
            const V3<T> d = col_diff<T>(pv, pw, o, L);
            if (d.x * d.x + d.y * d.y + d.z * d.z > R * R) return;
            if (n < cap) {  // insertion in ascending order
                int k = n;
                while (k > 0 && out[k - 1] > w) { out[k] = out[k - 1]; k--; }
                out[k] = w;
            }
            n++;
        },
        c0, cstep);
    return n;
}
template <typename T> __device__ __forceinline__ void col_list(const Dev<T> &a, int s, int parity) {
    const int v = a.surf[s];
    const int r = a.vrob[v];
    const RobotTime *rt = a.rt + r;
    if (!rt->stale[parity] || rt->diverged) return;
    a.ccount[s] = col_list_cells(a, a.rp[r], s, v, 0, 1, a.clist + (long)s * kMaxCand, kMaxCand);
}

// The contact force of surface voxel s from the partners inside its envelope,
// summed in ascending partner index (the order in which the serialized pair loop
// of the reference accumulates this voxel's contact force), with the target
// hits, the forced signal, and the pair record when one is kept.  kContactLanes
// consecutive work-items (lane 0..) share a voxel: each evaluates every
// kContactLanes-th candidate, and the forces are added in list order.  LIST: in
// a step whose listing passes have run (the grid is rebuilt), lane 0 builds the
// voxel's candidate list first and then takes every candidate itself (the list
// is its own writes; a listing comes every ~100 steps).  The voxel update of the
// step clears the listing flags.
constexpr int kContactLanes = 4;

// The list of surface voxel s (voxel v) built by its kContactLanes contact lanes
// (lane = 0 ..; seg: kMaxCand ints of LDS for the voxel): lane j walks the cells
// j, j + kContactLanes, ... into its part of seg, in ascending order; then each
// lane ranks its candidates among all of them and writes them to their places
// in the list row -- the list and count col_list makes.  When a lane finds more
// than its part holds, lane 0 makes the whole list alone.  Returns the count to
// every lane.
template <typename T>
__device__ __forceinline__ int col_list_lanes(const Dev<T> &a, const RobotParams &rp, int s, int v, int lane, int *seg) {
    constexpr int PART = kMaxCand / kContactLanes;
    int *list = a.clist + (long)s * kMaxCand;
    const int own = col_list_cells(a, rp, s, v, lane, kContactLanes, seg + lane * PART, PART);
    int cnt[kContactLanes], total = 0;
    bool over = false;
#pragma unroll
    for (int j = 0; j < kContactLanes; j++) {
        cnt[j] = __shfl(own, j, kContactLanes);
        total += cnt[j];
        over |= cnt[j] > PART;
    }
    if (over) {  // the same for the voxel's lanes
        if (lane == 0) total = col_list_cells(a, rp, s, v, 0, 1, list, kMaxCand);
        total = __shfl(total, 0, kContactLanes);
    } else {
        __threadfence_block();  // the parts of the other lanes
        for (int k = 0; k < own; k++) {
            const int e = seg[lane * PART + k];
            int rank = k;
#pragma unroll
            for (int j = 0; j < kContactLanes; j++) {
                if (j == lane) continue;
                for (int m = 0; m < cnt[j]; m++) rank += seg[j * PART + m] < e;
            }
            list[rank] = e;
        }
    }
    if (lane == 0) a.ccount[s] = total;
    __threadfence_block();  // the list, for the other lanes
    return total;
}

template <bool LIST, typename T> __device__ __forceinline__ void col_contact(const Dev<T> &a, int s, int parity, int lane, int *seg) {
    // the loads that need only s or v are issued before the robot's checks (each
    // returns early, and a load after it would wait for it): the voxel's pose,
    // and its list as it stands with this lane's first candidate (a listing step
    // rebuilds the list below, and lane 0 then reads it again)
    const int v = a.surf[s];
    const int ncPre = a.ccount[s];
    const int *list = a.clist + (long)s * kMaxCand;
    const int wFirst = list[lane];
    const int r = a.vrob[v];
    const int matI = a.vmat[v], pkv = a.lat[v];
    const T4<T> qi0 = a.pq0[v], qi1 = a.pq1[v], mi0 = a.mo0[v];
    RobotTime *rt = a.rt + r;
    const RobotParams &rp = a.rp[r];
    const bool listed = LIST && rt->stale[parity] && !rt->diverged;  // the same for the lanes of a voxel
    const int nc = listed ? col_list_lanes(a, rp, s, v, lane, seg) : ncPre;
    if (!rp.collisions || rt->diverged) return;
    if (!col_active(a, v)) {
        if (lane == 0) a.contactW[v] = T4<T>{0, 0, 0, 0};
        return;
    }
    const VoxelMat &mi = a.vmats[matI];
    const T L = (T)mi.nomSize;
    const T bsaI = base_size_average<T>(mi.nomSize, (float)qi1.w, mi.alphaCTE);
    const V3<T> velI = V3<T>{mi0.x, mi0.y, mi0.z} * (T)mi.massInv;
    const int ti = a.vtarget ? a.vtarget[v] : 0;
    int hit = 0;
    // partner w when it is in contact now, its force in fw (links: whether to
    // test the current links; lists leave out linked neighbours unless the
    // robot can lose links)
    auto partner = [&](int w, bool links, V3<T> &fw) -> int {
        // w's words are loaded before the tests that return early (and pinned
        // where the arithmetic on them starts, which keeps its rounding)
        const int latW = a.lat[w], matW = a.vmat[w];
        T4<T> qj0 = a.pq0[w], mj0 = a.mo0[w];
        T tj = a.pq1[w].w;
        if (!col_active(a, w)) return 0;
        const int3 o = lat_offset(pkv, latW);
        if (links && (rp.attach ? col_linked_any(a, v, w) : col_linked(a, v, o.x, o.y, o.z))) return 0;
        const VoxelMat &mj = a.vmats[matW];
        if (mi.fixed && mj.fixed) return 0;
        pin(qj0);
        const V3<T> diff = col_diff<T>(qi0, qj0, o, L);
        pin(tj);
        const T bsaJ = base_size_average<T>(mj.nomSize, (float)tj, mj.alphaCTE);
        if (!in_contact_envelope(diff, bsaI, bsaJ)) return 0;
        if (a.vtarget && a.vtarget[w] != ti) {  // a target touching a non-target (:711-726)
            if (v < w) atomicAdd(&rt->collisionCount, 1);
            hit = 1;
        }
        if (a.pairRec && v < w) {
            const int k = atomicAdd(a.pairN, 1);
            if (k < a.pairCap) a.pairRec[k] = int2{v, w};
        }
        pin(mj0);
        const V3<T> velJ = V3<T>{mj0.x, mj0.y, mj0.z} * (T)mj.massInv;
        fw = contact_force(diff, bsaI, bsaJ, velI, velJ, mi.penKinv, mj.penKinv, mi.cDampT, mj.cDampT);
        return 1;
    };
    V3<T> f = {0, 0, 0};
    if (nc <= kMaxCand) {
        const bool links = rp.listLinked != 0;
        int w = listed ? list[lane] : wFirst;  // a listing step has just written the list
        for (int k0 = 0; k0 < nc; k0 += kContactLanes) {
            const int kn = k0 + kContactLanes + lane;
            const int wn = kn < nc ? list[kn] : 0;  // the next batch in flight
            V3<T> fk = {0, 0, 0};
            int in = 0;
            if (k0 + lane < nc) in = partner(w, links, fk);
            w = wn;
            for (int j = 0; j < kContactLanes && k0 + j < nc; j++) {  // list order
                const int inj = __shfl(in, j, kContactLanes);
                const V3<T> fj = {__shfl(fk.x, j, kContactLanes), __shfl(fk.y, j, kContactLanes),
                                  __shfl(fk.z, j, kContactLanes)};
                if (inj) f = f + fj;
            }
        }
    } else if (lane == 0) {
        // more candidates than a list holds: ascending partners straight from the
        // grid, which is rebuilt every step while this lasts
        rt->stale[parity ^ 1] = 1;
        if (a.anyStale) a.anyStale[parity ^ 1] = 1;
        int prev = -1;
        for (;;) {
            int next = INT_MAX;
            col_grid_each<false>(a, rp, s, v, [&](int w, const T4<T> &) {
                if (w > prev && w < next) next = w;
            });
            if (next == INT_MAX) break;
            V3<T> fw;
            if (partner(next, true, fw)) f = f + fw;
            prev = next;
        }
    }
    for (int o = 1; o < kContactLanes; o <<= 1) hit |= __shfl_xor(hit, o, kContactLanes);
    if (lane != 0) return;
    // with signals a non-target voxel touching a target receives a forced 100
    if (hit && !ti && a.sig && rp.signals) {
        VoxelSignal sg = a.sig[v];
        signal_receive(sg, a.sigMat[v], 100, rt->time[parity], true);
        a.sig[v] = sg;
    }
    a.contactW[v] = T4<T>{f.x, f.y, f.z, 0};
}

// a moved surface voxel marks its robot's lists stale for the next step
template <typename T> __device__ __forceinline__ void col_track(const Dev<T> &a, int v, const V3<T> &u, RobotTime *rt,
                                                                const RobotParams &rp, int parity) {
    if (!a.vsurf || a.vsurf[v] < 0) return;
    const T4<T> b = a.bpos[v];
    const T dx = u.x - b.x, dy = u.y - b.y, dz = u.z - b.z;
    if (dx * dx + dy * dy + dz * dz > (T)rp.staleD2) {
        rt->stale[parity ^ 1] = 1;
        a.anyStale[parity ^ 1] = 1;
    }
}

// rebuild kernels for robots whose lists are stale (all = 0: they return at
// once unless some robot is).  The grid: the workgroups of a robot count its
// surface voxels into buckets, kCountPerGroup each, and the last of them to
// finish takes the prefix sum and scatters (one launch fewer per step than a
// separate count pass; a listing comes every ~100 steps, the launch every step)
// passes 1 and 3 for one robot known to be listing (the checks of col_count /
// col_scatter done once), four surface voxels per thread at a time so that
// their loads are in flight together: the count of surface voxels [s0, s1), the
// scatter of the whole robot (one workgroup scatters a robot of any size, 40
// voxels per thread for a 64x64x16 body).  The same cells and records as the
// per-voxel passes (the ranks inside a bucket may differ, which the ascending
// lists do not see).
constexpr int kCountPerGroup = 4096;  // four surface voxels per thread of a 1024-thread workgroup
template <typename T> __device__ __forceinline__ void col_count_robot(const Dev<T> &a, const RobotParams &rp, int s0, int s1) {
    const double h = rp.listR;
    const int end = s1, nt = blockDim.x;
    for (int i0 = s0 + (int)threadIdx.x; i0 < end; i0 += 4 * nt) {
        int v[4], pk[4];
        T4<T> q[4];
        double L[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = i0 + k * nt < end ? a.surf[i0 + k * nt] : -1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (v[k] < 0) continue;
            q[k] = a.pq0[v[k]];
            pk[k] = a.lat[v[k]];
            L[k] = a.vmats[a.vmat[v[k]]].nomSize;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (v[k] < 0) continue;
            const int s = i0 + k * nt;
            const V3<double> p = position_d<T>(pk[k], L[k], V3<T>{q[k].x, q[k].y, q[k].z});
            const int cx = (int)floor(p.x / h), cy = (int)floor(p.y / h), cz = (int)floor(p.z / h);
            const int key = (int)(cell_hash(cx, cy, cz) & (unsigned)rp.mask);
            a.scell[s] = int4{cx, cy, cz, key};
            a.srank[s] = atomicAdd(&a.ccnt[rp.toff + key], 1);
            a.bpos[v[k]] = q[k];
        }
    }
}
template <typename T> __device__ __forceinline__ void col_scatter_robot(const Dev<T> &a, const RobotParams &rp) {
    const int end = rp.soff + rp.scount, nt = blockDim.x;
    for (int i0 = rp.soff + (int)threadIdx.x; i0 < end; i0 += 4 * nt) {
        int4 c[4];
        int rank[4], v[4], lat[4], at[4];
        T4<T> q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i0 + k * nt >= end) continue;
            c[k] = a.scell[i0 + k * nt];
            rank[k] = a.srank[i0 + k * nt];
            v[k] = a.surf[i0 + k * nt];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i0 + k * nt >= end) continue;
            at[k] = rp.soff + a.cstart[rp.toff + c[k].w] + rank[k];
            q[k] = a.bpos[v[k]];
            lat[k] = a.lat[v[k]];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i0 + k * nt >= end) continue;
            a.grec[at[k]] = int4{c[k].x, c[k].y, c[k].z, v[k]};
            a.gpos[at[k]] = T4<T>{q[k].x, q[k].y, q[k].z, lat_word<T>(lat[k])};
        }
    }
}

// blockIdx.x: robot r0 + x; blockIdx.y: its part of the count pass
template <typename T> __global__ __launch_bounds__(1024) void k_col_scan(Dev<T> a, int r0, int parity, int all) {
    __shared__ int wsum[16];  // wave totals of col_scan
    __shared__ int lastPart;
    if (!all && !a.anyStale[parity]) return;
    const int r = r0 + blockIdx.x;
    const RobotParams &rp = a.rp[r];
    if (!a.rt[r].stale[parity] || a.rt[r].diverged || rp.scount == 0) return;
    const int parts = (rp.scount + kCountPerGroup - 1) / kCountPerGroup, part = blockIdx.y;
    if (part >= parts) return;
    col_count_robot(a, rp, rp.soff + part * kCountPerGroup, rp.soff + min(rp.scount, (part + 1) * kCountPerGroup));
    if (parts > 1) {
        // the workgroup that finishes its part last goes on with the whole
        // robot: every workgroup releases its cells, ranks and positions before
        // taking its ticket, the last one acquires them after
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) lastPart = atomicAdd(&a.ctick[r], 1) == parts - 1;
        __syncthreads();
        if (!lastPart) return;
        __threadfence();
        if (threadIdx.x == 0) a.ctick[r] = 0;  // for the next listing (a later launch)
    }
    __syncthreads();
    col_scan<true>(a, r, wsum);
    __syncthreads();
    col_scatter_robot(a, rp);
}
template <typename T> __global__ __launch_bounds__(256) void k_col_contact(Dev<T> a, int s0, int n, int parity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // kContactLanes per surface voxel
    __shared__ int seg[256 / kContactLanes * kMaxCand];  // a listing step's parts of each voxel's list
    if (i < n * kContactLanes)
        col_contact<true>(a, s0 + i / kContactLanes, parity, i % kContactLanes, seg + threadIdx.x / kContactLanes * kMaxCand);
}

// the contact phase and the link updates of one step side by side in one launch
// (they read the state entering the step and write disjoint outputs): blocks
// [0, nbl) update links, the rest compute contact forces
template <typename T>
__global__ __launch_bounds__(256) void k_contact_links(Dev<T> a, int nl, int nbl, int ns, int parity) {
    __shared__ int seg[256 / kContactLanes * kMaxCand];  // a listing step's parts of each voxel's list
    if ((int)blockIdx.x < nbl) {
        const int l = blockIdx.x * blockDim.x + threadIdx.x;
        if (l < nl) link_update(a, l, parity);
    } else {
        const int i = (blockIdx.x - nbl) * blockDim.x + threadIdx.x;
        if (i < ns * kContactLanes)
            col_contact<true>(a, i / kContactLanes, parity, i % kContactLanes, seg + threadIdx.x / kContactLanes * kMaxCand);
    }
}

// ---- attach (handle_collision_attachment VX3_VoxelyzeKernel.cu:729-831) -------
// Each step, after the contact forces: every surface voxel lists its attach
// candidates among its partners of lower index (k_attach_find, in parallel,
// from the topology entering the step); then one workgroup per robot walks the
// candidates in the order of the reference's pair loop and makes the links
// whose slots are still free, one at a time (k_attach_walk).  The reference
// runs these tests and link creations concurrently with unsynchronised slot
// writes; this is the interleaving in which every test precedes every creation
// and creations happen in pair order.

// is_neighbor(from, to, NULL, 5) (:651-680): `to` is reached from `from` over at
// most five links without stepping straight back over the link just taken;
// depth first in slot order with an explicit stack
template <typename T> __device__ bool linked_within5(const Dev<T> &a, int from, int to) {
    if (from == to) return true;
    int node[5], inc[5], dir[5];
    int top = 0;
    node[0] = from;
    inc[0] = -1;
    dir[0] = 0;
    while (top >= 0) {
        if (dir[top] == 6) {
            top--;
            continue;
        }
        const int l = a.vlinks[6 * node[top] + dir[top]];
        dir[top]++;
        if (l < 0 || l == inc[top]) continue;
        const int2 e = a.lends[l];
        const int o = e.x == node[top] ? e.y : e.x;
        if (o == to) return true;
        if (top < 4) {
            top++;
            node[top] = o;
            inc[top] = l;
            dir[top] = 0;
        }
    }
    return false;
}

// the contact force on v from w, the term col_contact adds for the pair
template <typename T> __device__ V3<T> pair_contact(const Dev<T> &a, int v, int w) {
    const VoxelMat &mi = a.vmats[a.vmat[v]], &mj = a.vmats[a.vmat[w]];
    const T4<T> mi0 = a.mo0[v], mj0 = a.mo0[w];
    const T bsaI = base_size_average<T>(mi.nomSize, (float)a.pq1[v].w, mi.alphaCTE);
    const T bsaJ = base_size_average<T>(mj.nomSize, (float)a.pq1[w].w, mj.alphaCTE);
    const V3<T> diff = col_diff<T>(a.pq0[v], a.pq0[w], lat_offset(a.lat[v], a.lat[w]), (T)mi.nomSize);
    const V3<T> velI = V3<T>{mi0.x, mi0.y, mi0.z} * (T)mi.massInv;
    const V3<T> velJ = V3<T>{mj0.x, mj0.y, mj0.z} * (T)mj.massInv;
    return contact_force(diff, bsaI, bsaJ, velI, velJ, mi.penKinv, mj.penKinv, mi.cDampT, mj.cDampT);
}

// The attach candidates of surface voxel v = surf[s] as voxel1 (:684-746):
// partners w < v inside the contact envelope, not linked, both with
// enableAttach, of v's sticky material (neither fixed), and with no path of at
// most five links between them; ascending, the first kAttachPerVoxel kept.
template <typename T> __device__ void attach_find(const Dev<T> &a, int s) {
    const int v = a.surf[s];
    const int r = a.vrob[v];
    const RobotParams &rp = a.rp[r];
    int n = 0;
    if (rp.attach && !a.rt[r].diverged && a.vflagA[v] && a.vatt[v].x >= 0 && col_active(a, v)) {
        const int mat = a.vatt[v].x;
        const VoxelMat &mi = a.vmats[a.vmat[v]];
        const T4<T> qi0 = a.pq0[v];
        const int pkv = a.lat[v];
        const T L = (T)mi.nomSize;
        const T bsaI = base_size_average<T>(mi.nomSize, (float)a.pq1[v].w, mi.alphaCTE);
        auto test = [&](int w) {
            if (w >= v || !a.vflagA[w] || a.vatt[w].x != mat || !col_active(a, w)) return;
            const VoxelMat &mj = a.vmats[a.vmat[w]];
            const V3<T> diff = col_diff<T>(qi0, a.pq0[w], lat_offset(pkv, a.lat[w]), L);
            const T bsaJ = base_size_average<T>(mj.nomSize, (float)a.pq1[w].w, mj.alphaCTE);
            if (!in_contact_envelope(diff, bsaI, bsaJ)) return;
            if (col_linked_any(a, v, w) || linked_within5(a, v, w)) return;
            if (n < kAttachPerVoxel) a.acand[(long)s * kAttachPerVoxel + n] = w;
            n++;
        };
        const int nc = a.ccount[s];
        if (nc <= kMaxCand) {
            const int *list = a.clist + (long)s * kMaxCand;
            for (int k = 0; k < nc; k++) test(list[k]);
        } else {  // ascending partners from the grid, as the contact phase takes them
            int prev = -1;
            for (;;) {
                int next = INT_MAX;
                col_grid_each<false>(a, rp, s, v, [&](int w, const T4<T> &) {
                    if (w > prev && w < next) next = w;
                });
                if (next == INT_MAX) break;
                test(next);
                prev = next;
            }
        }
    }
    a.acount[s] = n;
}

template <typename T> __global__ __launch_bounds__(256) void k_attach_find(Dev<T> a, int s0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) attach_find(a, s0 + i);
}

// The link of candidate pair (v1, v2) when both slots are free (:748-830, the
// VX3_Link constructor and reset VX3_Link.cu:31-76): the axis is the dominant
// component of voxel2's offset in voxel1's frame, voxel1 takes the slot facing
// voxel2 and voxel2 the opposite one, and the voxel with the NEG-direction slot
// is the positive end.  The entry comes from the robot's pool (entries of links
// that left are reused once every entry was handed out); the link starts from
// the reset state with isNewLink = SafetyGuard; the pair's contact force is
// given back (:819-825).
template <typename T>
__device__ void attach_pair(const Dev<T> &a, const RobotParams &rp, RobotTime *rt, int r, int v1, int v2) {
    const T4<T> p1 = a.pq0[v1], p1b = a.pq1[v1];
    const Q4<T> q1 = {p1.w, p1b.x, p1b.y, p1b.z};
    const V3<T> e = col_diff<T>(p1, a.pq0[v2], lat_offset(a.lat[v1], a.lat[v2]), (T)a.vmats[a.vmat[v1]].nomSize);
    const V3<T> ea = qrotinv(q1, V3<T>{-e.x, -e.y, -e.z});
    const T ax = mabs(ea.x), ay = mabs(ea.y), az = mabs(ea.z);
    int axis;
    bool neg1;  // voxel1 takes its NEG-direction slot
    if (ax >= ay && ax >= az) { axis = 0; neg1 = ea.x < 0; }
    else if (ay >= ax && ay >= az) { axis = 1; neg1 = ea.y < 0; }
    else { axis = 2; neg1 = ea.z < 0; }
    const int d1 = 2 * axis + (neg1 ? 1 : 0), d2 = 2 * axis + (neg1 ? 0 : 1);
    if (a.vlinks[6 * v1 + d1] >= 0 || a.vlinks[6 * v2 + d2] >= 0) return;
    int l = -1;
    if (rt->attachUsed < rp.pcount) {
        l = rp.poff + rt->attachUsed;
        rt->attachUsed += 1;
    } else {
        for (int k = 0; k < rp.pcount && l < 0; k++)
            if (a.lrem[rp.poff + k]) l = rp.poff + k;
    }
    if (l < 0) {
        rt->attachOverflow = 1;
        return;
    }
    a.lends[l] = int2{neg1 ? v2 : v1, neg1 ? v1 : v2};
    a.lmeta[l] = int2{axis | (a.vatt[v1].y << 2), r};
    const T4<T> z4 = {0, 0, 0, 0};
    a.ls0[l] = z4;
    a.ls1[l] = z4;
    a.ls2[l] = T4<T>{0, 0, (T)(LF_SMALL_ANGLE | LF_ATTACHED | (rp.safetyGuard << LF_NEW_SHIFT)), 0};
    a.lf0[l] = z4;
    a.lf1[l] = z4;
    a.lf2[l] = z4;
    a.lrem[l] = 0;
    a.lmade[l] = rt->steps + 1;  // the count the voxel update of this step leaves
    a.vlinks[6 * v1 + d1] = l;
    a.vlinks[6 * v2 + d2] = l;
    if (rp.collisions) {
        const V3<T> f = pair_contact(a, v1, v2);
        const T4<T> c1 = a.contactW[v1], c2 = a.contactW[v2];
        a.contactW[v1] = T4<T>{c1.x - f.x, c1.y - f.y, c1.z - f.z, c1.w};
        a.contactW[v2] = T4<T>{c2.x - (-f.x), c2.y - (-f.y), c2.z - (-f.z), c2.w};
    }
}

// one workgroup per robot: its surface voxels' candidates in ascending voxel1
// (then voxel2) order, made by one work-item
template <typename T> __global__ __launch_bounds__(256) void k_attach_walk(Dev<T> a, int r0) {
    const int r = r0 + blockIdx.x;
    const RobotParams &rp = a.rp[r];
    RobotTime *rt = a.rt + r;
    if (!rp.attach || rp.scount == 0 || rt->diverged) return;
    const int end = rp.soff + rp.scount;
    for (int base = rp.soff; base < end; base += 256) {
        const int s = base + (int)threadIdx.x;
        const int c = s < end ? a.acount[s] : 0;
        if (!__syncthreads_or(c > 0)) continue;
        if (threadIdx.x == 0) {
            for (int i = base; i < end && i < base + 256; i++) {
                const int n = a.acount[i];
                if (n > kAttachPerVoxel) rt->attachOverflow = 1;
                for (int k = 0; k < n && k < kAttachPerVoxel; k++)
                    attach_pair(a, rp, rt, r, a.surf[i], a.acand[(long)i * kAttachPerVoxel + k]);
            }
        }
        __syncthreads();
    }
}

// gpu_update_voxels (:607-621): after the voxel update every present, non-fixed
// voxel of an attach robot re-evaluates enableAttach, true when each
// AttachCondition given is > 0, at its new position with the step's start time,
// the collision count and the last CycleStep samples
template <typename T> __global__ __launch_bounds__(256) void k_attach_flags(Dev<T> a, int v0, int n, int parity) {
    const int v = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= v0 + n) return;
    const int r = a.vrob[v];
    const RobotParams &rp = a.rp[r];
    const RobotTime *rt = a.rt + r;
    if (!rp.attach || rt->diverged || a.vrem[v]) return;
    const VoxelMat &vm = a.vmats[a.vmat[v]];
    if (vm.fixed) return;
    const T4<T> q0 = a.pq0[v];
    const V3<double> p = position_d<T>(a.lat[v], vm.nomSize, V3<T>{q0.x, q0.y, q0.z});
    const double vars[MV_COUNT] = {p.x, p.y, p.z, (double)rt->collisionCount, rt->time[parity], rt->recentAngle,
                                   rt->closeness, (double)rt->closePairs, (double)rp.vcount};
    int ok = 1;
    for (int k = 0; k < 5; k++)
        if (rp.acLen[k] && !(formula_eval(a.formulas + rp.acOff[k], rp.acLen[k], vars) > 0)) ok = 0;
    a.vflagA[v] = ok;
}

// The CycleStep block of doTimeStep (VX3_VoxelyzeKernel.cu:314-334) for robots
// whose force fields or attach conditions read its samples: every CycleStep =
// int(TempPeriod / dt) steps (at least 1; the reference divides by zero below)
// the centre of mass is sampled (updateCurrentCenterOfMass :477-493, summed in
// voxel order as the reference's loop), recentAngle is the angle between the
// last two displacements of the sampled centre of mass (0 for the first two
// samples or a repeated point), and with a pair radius targetCloseness sums
// 1 / distance over the target pairs and numClosePairs counts those closer
// (computeTargetCloseness :545-562).  One work-item per robot.
template <typename T> __global__ __launch_bounds__(64) void k_cycle_sample(Dev<T> a, int r0, int nr) {
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r0 + nr) return;
    const RobotParams &rp = a.rp[r];
    RobotTime *rt = a.rt + r;
    if (!rp.sample || rt->diverged) return;
    int cycle = (int)(rp.tempPeriod / (double)rt->dt);
    if (cycle < 1) cycle = 1;
    if (rt->steps % cycle != 0) return;  // the step count includes this step (CurStepCount)
    rt->angleSamples += 1;
    double s[3] = {0, 0, 0}, total = 0;
    for (int v = rp.voff; v < rp.voff + rp.vcount; v++) {
        const VoxelMat &vm = a.vmats[a.vmat[v]];
        if (!vm.isMeasured) continue;
        const T4<T> q = a.pq0[v];
        const V3<double> p = position_d<T>(a.lat[v], vm.nomSize, V3<T>{q.x, q.y, q.z});
        const double w = vm.mass;
        s[0] = s[0] + p.x * w;
        s[1] = s[1] + p.y * w;
        s[2] = s[2] + p.z * w;
        total += w;
    }
    double A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) {
        A[k] = rt->shist[1][k];
        B[k] = rt->scom[k];
        C[k] = total == 0 ? 0.0 : s[k] / total;
        rt->shist[0][k] = A[k];
        rt->shist[1][k] = B[k];
        rt->scom[k] = C[k];
    }
    const bool bc = B[0] == C[0] && B[1] == C[1] && B[2] == C[2], ab = A[0] == B[0] && A[1] == B[1] && A[2] == B[2];
    if (bc || ab || rt->angleSamples < 3) {
        rt->recentAngle = 0;
    } else {
        const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, w[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
        const double d = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
        const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double lw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        rt->recentAngle = vxm_acos(d / (lu * lw));
    }
    if (rp.pairR == 0) return;
    double ret = 0;
    int pairs = 0;
    for (int i = 0; i < rp.tgCount; i++) {
        const int vi = a.targets[rp.tgOff + i];
        const T4<T> qi = a.pq0[vi];
        const V3<double> pi = position_d<T>(a.lat[vi], a.vmats[a.vmat[vi]].nomSize, V3<T>{qi.x, qi.y, qi.z});
        for (int j = i + 1; j < rp.tgCount; j++) {
            const int vj = a.targets[rp.tgOff + j];
            const T4<T> qj = a.pq0[vj];
            const V3<double> pj = position_d<T>(a.lat[vj], a.vmats[a.vmat[vj]].nomSize, V3<T>{qj.x, qj.y, qj.z});
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < rp.pairR) pairs++;
            ret += 1 / dist;
        }
    }
    rt->closeness = ret;
    rt->closePairs = pairs;
}

template <typename T> __global__ __launch_bounds__(256) void k_links(Dev<T> a, int l0, int n, int parity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) link_update(a, l0 + i, parity);
}

template <typename T> __global__ __launch_bounds__(256) void k_voxels(Dev<T> a, int v0, int n, int parity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) voxel_update(a, v0 + i, parity);
}

// VX3_VoxelyzeKernel::removeVoxels (VX3_VoxelyzeKernel.cu:365-399), at the end
// of a step before currentTime advances (:336-343): a voxel whose material's
// RemoveFromSimulationAfterThisManySeconds lies below the step's start time is
// removed with its links, which both ends forget.  Two linked voxels removed in
// the same step write the same values, in either order.
template <typename T> __global__ __launch_bounds__(256) void k_removal(Dev<T> a, int parity) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nRemCand) return;
    const int v = a.remCand[i];
    const int r = a.vrob[v];
    const RobotTime *rt = a.rt + r;
    if (!a.rp[r].secondary || rt->diverged || rt->dt == 0.0f || a.vrem[v]) return;
    if (!(a.remAt[i] < rt->time[parity])) return;
    const int stamp = rt->steps;
    a.vrem[v] = stamp;
    for (int d = 0; d < 6; d++) {
        const int l = a.vlinks[6 * v + d];
        if (l < 0) continue;
        a.lrem[l] = stamp;
        const int o = (d & 1) ? a.lends[l].x : a.lends[l].y;
        a.vlinks[6 * o + (d ^ 1)] = -1;
        a.vlinks[6 * v + d] = -1;
    }
}

// gpu_update_detach (VX3_VoxelyzeKernel.cu:946-968): a link whose maxStrain has
// passed its failure strain (VX3_Link::isFailed, VX3_Link.cu:78) leaves both
// voxels' link lists.  The reference detaches after the step's collisions; the
// surface, cilia and collisions see it from the next step, the step's voxel
// update sums the failed link's zero forces either way, and the step's signals
// already go without it: so it runs after the voxel update, before the signal
// delivery.
template <typename T> __global__ __launch_bounds__(256) void k_detach(Dev<T> a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nDetCand) return;
    const int l = a.detCand[i];
    const int2 meta = a.lmeta[l];
    const RobotTime *rt = a.rt + meta.y;
    if (rt->diverged || rt->dt == 0.0f || a.lrem[l]) return;
    const float epsFail = a.lmats[meta.x >> 2].epsFail;
    if (!(epsFail != -1.0f && (float)a.ls2[l].y > epsFail)) return;
    a.lrem[l] = rt->steps;
    const int axis = meta.x & 3;
    const int2 e = a.lends[l];
    a.vlinks[6 * e.x + 2 * axis] = -1;      // the link is pVNeg's POS-direction link
    a.vlinks[6 * e.y + 2 * axis + 1] = -1;  // and pVPos's NEG-direction link
}

template <typename T = int> __global__ __launch_bounds__(1) void k_zero_int(int *p) { *p = 0; }

template <typename T> __global__ __launch_bounds__(256) void k_mark_stale(RobotTime *rt, int R, int *anyStale) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) rt[r].stale[0] = rt[r].stale[1] = 1;
    if (r == 0 && anyStale) anyStale[0] = anyStale[1] = 1;
}

// nsteps steps with dt = 0: the step count advances, time and the previous dt
// stay (an odd count carries them, and pending listings, to the other parity)
template <typename T> __global__ __launch_bounds__(256) void k_zero_steps(RobotTime *rt, int R, int nsteps, int parity, int *anyStale) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0 && anyStale && (nsteps & 1) && anyStale[parity]) anyStale[parity ^ 1] = 1;
    if (r >= R || rt[r].diverged) return;
    RobotTime &t = rt[r];
    if (nsteps & 1) {
        t.time[parity ^ 1] = t.time[parity];
        t.prevDt[parity ^ 1] = t.prevDt[parity];
        t.stale[parity ^ 1] |= t.stale[parity];
    }
    t.steps += nsteps;
}

// One workgroup per robot; nsteps steps with barrier-separated phases (the
// collision listing when it is stale, the contact forces, links, voxels).
template <typename T, int NT> __global__ __launch_bounds__(NT) void k_persistent(Dev<T> a, int nsteps, int parity0) {
    __shared__ int wsum[16];  // wave totals of col_scan
    const int r = blockIdx.x;
    const RobotParams rp = a.rp[r];
    const bool col = rp.collisions && rp.scount > 0 && a.contactW != nullptr;
    for (int s = 0; s < nsteps; s++) {
        const int parity = (parity0 + s) & 1;
        if (col) {
            if (a.rt[r].stale[parity] && !a.rt[r].diverged) {
                for (int i = rp.soff + threadIdx.x; i < rp.soff + rp.scount; i += NT) col_count(a, i, parity);
                __syncthreads();
                col_scan<true>(a, r, wsum);
                __syncthreads();
                for (int i = rp.soff + threadIdx.x; i < rp.soff + rp.scount; i += NT) col_scatter(a, i, parity);
                __syncthreads();
                for (int i = rp.soff + threadIdx.x; i < rp.soff + rp.scount; i += NT) col_list(a, i, parity);
            }
            __syncthreads();
            for (int i = threadIdx.x; i < rp.scount * kContactLanes; i += NT)
                col_contact<false>(a, rp.soff + i / kContactLanes, parity, i % kContactLanes, nullptr);  // listed above
        }
        for (int l = rp.loff + threadIdx.x; l < rp.loff + rp.lcount; l += NT) link_update(a, l, parity);
        __syncthreads();
        for (int v = rp.voff + threadIdx.x; v < rp.voff + rp.vcount; v += NT) voxel_update(a, v, parity);
        __syncthreads();
    }
}

}  // namespace vx3mi
#include "vx3mi_tile.h"
namespace vx3mi {

// recommendedTimeStep (VX3_VoxelyzeKernel.cu:184-217) as a max-reduction:
// per-lane maxima, a 64-lane butterfly, then one value per wavefront in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_recommended_dt(const int2 *lends, const int2 *lmeta, const int *vmat,
                                                        const VoxelMat *vmats, const LinkMat *lmats,
                                                        const float *vmatE, const RobotParams *rp, double *out) {
    __shared__ double wmax[4];
    const RobotParams p = rp[blockIdx.x];
    double mx = 0.0;
    for (int l = p.loff + threadIdx.x; l < p.loff + p.lcount; l += blockDim.x) {
        const LinkMat &lm = lmats[lmeta[l].x >> 2];
        const int2 e = lends[l];
        const double m1 = vmats[vmat[e.x]].mass, m2 = vmats[vmat[e.y]].mass;
        const double f2 = lm.a1 / (m1 < m2 ? m1 : m2);
        mx = f2 > mx ? f2 : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) wmax[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) mx = wmax[w] > mx ? wmax[w] : mx;
        if (mx <= 0.0) {  // no links: voxel stiffness E*L/m (:204-211)
            for (int v = p.voff; v < p.voff + p.vcount; v++) {
                const VoxelMat &vm = vmats[vmat[v]];
                const double f2 = vmatE[vmat[v]] * vm.nomSize / vm.mass;
                mx = f2 > mx ? f2 : mx;
            }
        }
        out[blockIdx.x] = mx <= 0.0 ? 0.0 : 1.0f / (6.283185f * sqrt(mx));
    }
}

// mass-weighted centre of mass of measured voxels, one workgroup per robot
template <typename T>
__global__ __launch_bounds__(256) void k_com(const T4<T> *pq0, const int *lat, const int *vmat, const VoxelMat *vmats,
                                             const RobotParams *rp, double *out) {
    __shared__ double red[4][256];
    const RobotParams p = rp[blockIdx.x];
    double sx = 0, sy = 0, sz = 0, sm = 0;
    for (int v = p.voff + threadIdx.x; v < p.voff + p.vcount; v += blockDim.x) {
        const VoxelMat &vm = vmats[vmat[v]];
        if (!vm.isMeasured) continue;
        const int pk = lat[v];
        const double L = kAbsolutePositions<T> ? 0.0 : vm.nomSize, w = vm.mass;
        const T4<T> q = pq0[v];
        sx += ((pk & 1023) * L + (double)q.x) * w;
        sy += (((pk >> 10) & 1023) * L + (double)q.y) * w;
        sz += ((pk >> 20) * L + (double)q.z) * w;
        sm += w;
    }
    red[0][threadIdx.x] = sx; red[1][threadIdx.x] = sy; red[2][threadIdx.x] = sz; red[3][threadIdx.x] = sm;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tm = red[3][0];
        out[3 * blockIdx.x + 0] = tm == 0 ? 0.0 : red[0][0] / tm;
        out[3 * blockIdx.x + 1] = tm == 0 ? 0.0 : red[1][0] / tm;
        out[3 * blockIdx.x + 2] = tm == 0 ? 0.0 : red[2][0] / tm;
    }
}

// ------------------------------------------------------------------- host
template <typename T> struct Batch : BatchBase {
    Dev<T> d{};
    struct TilePlan {
        bool ok = false;
        int ntiles = 0;
        int4 *d_tiles = nullptr;  // robot begin, robot count, voxels, links
        int *d_tileRobots = nullptr;
    } plans[2];                   // large and small tile configuration
    bool tileOK = false;          // some tile plan exists
    bool anyForceField = false;   // some robot has a force field (plain and flat modes only)
    bool anyCilia = false;        // some robot has EnableCilia (not in the tiles)
    bool anySignals = false;      // some robot has EnableSignals
    bool anySample = false;       // some robot's formulas read the CycleStep samples (angle, closeness, pairs)
    // plain and flat modes only: force fields (their own kernel), removal and
    // signals (kernels between steps)
    bool plainFlatOnly() const { return anyForceField || anyRemoval || anySignals || anyDetach || anyAttach; }
    // launch configuration used by VX3MI_MODE_AUTO (set by the autotuner)
    int autoMode = 0, tileSel = 0, persistThreads = 256;
    int nLinkMats = 0;  // entries of the device link material table
    RobotParams *d_rp = nullptr;
    VoxelMat *d_vmats = nullptr;
    VoxelExt *d_exts = nullptr;
    CurvePt *d_curves = nullptr;
    MtToken *d_formulas = nullptr;
    LinkMat *d_lmats = nullptr;
    float *d_vmatE = nullptr;
    double *d_out = nullptr;

    int create(vx3mi_model *const *models, int n) {
        R = n;
        std::vector<int> lat, vmat, vrob, vlinks;
        std::vector<T> phase;
        std::vector<int2> lends, lmeta;
        std::vector<LinkMat> lmats;
        std::vector<float> vmatE;
        std::vector<VoxelExt> exts;
        std::vector<CurvePt> curves;
        std::vector<MtToken> formulas;
        std::vector<T4<T>> pq0, pq1;
        std::vector<VoxelCilia> cilia;
        std::vector<SignalMat> sigMat;
        std::vector<int> vtarget;
        std::vector<int2> vatt;          // attach material and link material per voxel
        std::vector<int> targets;        // target voxels of robots with device samples
        std::vector<char> poolDetach;    // per robot: its attach links can fail and detach
        bool anyTarget = false;
        std::vector<int> remCand, detCand;
        std::vector<double> remAt;
        std::map<std::string, int> vmIndex, lmIndex, curveIndex;
        // materials are deduplicated across robots (byte-identical constants);
        // a voxel with an external gets an entry of its own material whose
        // ext field points at the external
        auto vm_index = [&](const VoxelMat &vmIn, float E, int ext) {
            VoxelMat vm = vmIn;
            vm.ext = ext;
            std::string key((const char *)&vm, sizeof(VoxelMat));
            key.append((const char *)&E, sizeof(float));
            auto it = vmIndex.find(key);
            if (it == vmIndex.end()) {
                it = vmIndex.emplace(key, (int)hvmats.size()).first;
                hvmats.push_back(vm);
                vmatE.push_back(E);
            }
            return it->second;
        };
        hrp.resize(n);
        for (int r = 0; r < n; r++) {
            const Model &m = models[r]->m;
            const int voff = (int)lat.size(), loff = (int)lends.size();
            RobotParams &p = hrp[r];
            std::memset(&p, 0, sizeof(p));
            p.tempAmp = m.tempAmplitude;
            p.tempPeriod = m.tempPeriod;
            p.tempInvPeriod = 1.0 / m.tempPeriod;
            p.varyTemp = m.varyTemp;
            p.enableExpansion = m.enableExpansion;
            p.collisions = m.enableCollision;
            p.secondary = m.secondaryExperiment;
            p.detach = m.enableDetach;
            p.signals = m.enableSignals;
            p.voff = voff; p.vcount = m.nv(); p.loff = loff; p.lcount = m.nl();
            if (m.enableCilia) anyCilia = true;
            if (m.enableSignals) anySignals = true;
            for (int ax = 0; ax < 3; ax++) {
                std::vector<MtToken> post;
                if (!m.forceField[ax].empty() && !mt_postfix(m.forceField[ax], post)) {
                    set_error("malformed force-field formula");
                    return VX3MI_E_ARG;
                }
                p.ffOff[ax] = (int)formulas.size();
                p.ffLen[ax] = (int)post.size();
                if (!post.empty()) anyForceField = true;
                formulas.insert(formulas.end(), post.begin(), post.end());
            }
            p.attach = 0;
            p.safetyGuard = m.safetyGuard;
            if (m.enableAttach)
                for (int lmi : m.attachLmat) p.attach |= lmi >= 0;
            for (int k = 0; k < 5; k++) {
                std::vector<MtToken> post;
                if (!m.attachCond[k].empty() && !mt_postfix(m.attachCond[k], post)) {
                    set_error("malformed attach condition");
                    return VX3MI_E_ARG;
                }
                p.acOff[k] = (int)formulas.size();
                p.acLen[k] = p.attach ? (int)post.size() : 0;
                if (p.attach) formulas.insert(formulas.end(), post.begin(), post.end());
            }
            if (p.attach) anyAttach = true;
            // device CycleStep samples when a force field or attach condition reads them
            const unsigned sampledVars = (1u << MV_ANGLE) | (1u << MV_CLOSENESS) | (1u << MV_PAIRS);
            unsigned used = 0;
            for (int ax = 0; ax < 3; ax++) used |= mt_vars_used(m.forceField[ax]);
            for (int k = 0; p.attach && k < 5; k++) used |= mt_vars_used(m.attachCond[k]);
            p.sample = (used & sampledVars) ? 1 : 0;
            p.pairR = m.maxDistPairs * m.latticeDim;
            p.tgOff = (int)targets.size();
            if (p.sample)
                for (int v = 0; v < m.nv(); v++)
                    if (m.palette[m.vpal[v] - 1].is_target) targets.push_back(voff + v);
            p.tgCount = (int)targets.size() - p.tgOff;
            if (p.sample) anySample = true;
            std::vector<int> vmRemap(m.vmats.size()), lmRemap(m.lmats.size());
            for (size_t k = 0; k < m.vmats.size(); k++) vmRemap[k] = vm_index(m.vmats[k], m.vmatE[k], -1);
            const int extOff = (int)exts.size();
            exts.insert(exts.end(), m.exts.begin(), m.exts.end());
            for (size_t k = 0; k < m.lmats.size(); k++) {
                LinkMat lm = m.lmats[k];
                if (lm.curve) {  // stress-strain curves are shared by content too
                    const int first = curve_first(lm), np = curve_points(lm);
                    std::string ck((const char *)&m.curves[first], sizeof(CurvePt) * np);
                    auto ci = curveIndex.find(ck);
                    if (ci == curveIndex.end()) {
                        ci = curveIndex.emplace(ck, (int)curves.size()).first;
                        curves.insert(curves.end(), m.curves.begin() + first, m.curves.begin() + first + np);
                    }
                    lm.curve = (ci->second << 8) | np;
                }
                std::string key((const char *)&lm, sizeof(LinkMat));
                auto it = lmIndex.find(key);
                if (it == lmIndex.end()) {
                    it = lmIndex.emplace(key, (int)lmats.size()).first;
                    lmats.push_back(lm);
                }
                lmRemap[k] = it->second;
            }
            for (int v = 0; v < m.nv(); v++) {
                lat.push_back(m.ix[v] | (m.iy[v] << 10) | (m.iz[v] << 20));
                const int e = m.ext_of(v);
                vmat.push_back(e < 0 ? vmRemap[m.vmat[v]] : vm_index(m.vmats[m.vmat[v]], m.vmatE[m.vmat[v]], extOff + e));
                vrob.push_back(r);
                phase.push_back((T)m.phase[v]);
                for (int k = 0; k < 6; k++) {
                    int l = m.vlinks[(size_t)v * 6 + k];
                    vlinks.push_back(l < 0 ? -1 : loff + l);
                }
                const VoxelMat &vm = m.vmats[m.vmat[v]];
                float t0 = m.initialTemp;
                if (m.varyTemp && !vm.fixed && !(vm.thermalOnAfter > 0.0))
                    t0 = actuation_temperature<T>(m.tempAmplitude, m.tempPeriod, 0.0, m.phase[v], m.enableExpansion);
                const double L0 = vm.nomSize;
                if (kAbsolutePositions<T>) pq0.push_back(T4<T>{(T)(m.ix[v] * L0), (T)(m.iy[v] * L0), (T)(m.iz[v] * L0), 1});
                else pq0.push_back(T4<T>{0, 0, 0, 1});
                pq1.push_back(T4<T>{0, 0, 0, (T)t0});
                const vx3mi_material &pm = m.palette[m.vpal[v] - 1];
                VoxelCilia c{};
                for (int k = 0; k < 3; k++) {
                    c.base[k] = m.ciliaBase.empty() ? 0.0 : m.ciliaBase[(size_t)3 * v + k];
                    c.shift[k] = m.ciliaShift.empty() ? 0.0 : m.ciliaShift[(size_t)3 * v + k];
                }
                c.coef = m.enableCilia ? pm.cilia : 0.0;
                c.onAfter = pm.cilia_on_after;
                cilia.push_back(c);
                SignalMat sm{};
                sm.period = pm.pacemaker_period;
                sm.decay = pm.signal_value_decay;
                sm.delay = pm.signal_time_delay;
                sm.inactive = pm.inactive_period;
                sm.pacemaker = m.enableSignals && pm.is_pacemaker;
                sigMat.push_back(sm);
                vtarget.push_back(pm.is_target ? 1 : 0);
                anyTarget |= pm.is_target != 0 && m.enableCollision;
                if (m.secondaryExperiment && pm.remove_after > 0) {
                    remCand.push_back(voff + v);
                    remAt.push_back(pm.remove_after);
                }
                const int alm = p.attach ? m.attachLmat[m.vmat[v]] : -1;
                vatt.push_back(alm < 0 ? int2{-1, 0} : int2{m.vpal[v], lmRemap[alm]});
            }
            // attach pool: a voxel holds six links, so a robot holds at most
            // 3 nv links at once and at most 3 nv - nl made by attach
            p.pcount = p.attach ? std::max(0, 3 * m.nv() - m.nl()) : 0;
            bool poolCanFail = false;
            for (int lmi : m.attachLmat)
                if (lmi >= 0 && m.lmats[lmi].epsFail != -1.0f) poolCanFail = true;
            poolDetach.push_back(p.attach && m.enableDetach && poolCanFail);
            for (int l = 0; l < m.nl(); l++) {
                lends.push_back(int2{voff + m.lneg[l], voff + m.lpos[l]});
                lmeta.push_back(int2{m.laxis[l] | (lmRemap[m.lmat[l]] << 2), r});
                if (m.enableDetach && m.lmats[m.lmat[l]].epsFail != -1.0f) detCand.push_back(loff + l);
            }
            float dflt = (float)(m.dtFrac * std::max(recommended_dt(m), 1e-10));
            defaultDt.push_back(dflt);
        }
        V = (int)lat.size();
        L = (int)lends.size();
        // the attach pools follow the lattice links of every robot; an entry
        // not handed out yet is marked removed (-1)
        LP = L;
        for (int r = 0; r < n; r++) {
            RobotParams &p = hrp[r];
            p.poff = LP;
            LP += p.pcount;
            int lm0 = 0;  // any valid link material for the entries not handed out
            for (int v = p.voff; p.attach && v < p.voff + p.vcount; v++)
                if (vatt[v].x >= 0) { lm0 = vatt[v].y; break; }
            for (int k = 0; k < p.pcount; k++) {
                lends.push_back(int2{p.voff, p.voff});
                lmeta.push_back(int2{lm0 << 2, r});
                if (poolDetach[r]) detCand.push_back(p.poff + k);
            }
        }
        nLinkMats = (int)lmats.size();
        anyRemoval = !remCand.empty();
        anyDetach = !detCand.empty();
        hlat = lat;
        hvmat = vmat;

        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        T4<T> *pq0d, *pq1d, *mo0d, *mo1d, *ls0, *ls1, *ls2, *lf0, *lf1, *lf2;
        int *latd, *vmatd, *vrobd, *vlinksd;
        T *phased;
        int2 *lendsd, *lmetad;
        int rc = VX3MI_OK;
        if ((rc = alloc(&pq0d, V)) || (rc = alloc(&pq1d, V)) || (rc = alloc(&mo0d, V)) || (rc = alloc(&mo1d, V)) ||
            (rc = alloc(&latd, V)) || (rc = alloc(&vmatd, V)) || (rc = alloc(&vrobd, V)) || (rc = alloc(&phased, V)) ||
            (rc = alloc(&vlinksd, (size_t)6 * V)) || (rc = alloc(&lendsd, LP)) || (rc = alloc(&lmetad, LP)) ||
            (rc = alloc(&ls0, LP)) || (rc = alloc(&ls1, LP)) || (rc = alloc(&ls2, LP)) || (rc = alloc(&lf0, LP)) ||
            (rc = alloc(&lf1, LP)) || (rc = alloc(&lf2, LP)) || (rc = alloc(&d_vmats, hvmats.size())) ||
            (rc = alloc(&d_exts, exts.size())) || (rc = alloc(&d_curves, curves.size())) ||
            (rc = alloc(&d_formulas, formulas.size())) ||
            (rc = alloc(&d_lmats, lmats.size())) || (rc = alloc(&d_vmatE, vmatE.size())) || (rc = alloc(&d_rp, R)) ||
            (rc = alloc(&d_rt, R)) || (rc = alloc(&d_out, (size_t)3 * R)))
            return rc;
        std::vector<T4<T>> zeros((size_t)std::max(V, LP), T4<T>{0, 0, 0, 0});
        std::vector<T4<T>> mo1((size_t)V, T4<T>{0, 0, 0, 0});
        for (int r = 0; r < n; r++) {
            const Model &m = models[r]->m;
            const int fl = VF_FLOOR_STATIC_FRICTION | (m.floorEnabled ? VF_FLOOR_ENABLED : 0);
            for (int v = 0; v < m.nv(); v++) mo1[hrp[r].voff + v].z = (T)fl;
        }
        std::vector<T4<T>> ls2h((size_t)LP, T4<T>{0, 0, (T)LF_SMALL_ANGLE, 0});
        HIP_TRY(hipMemcpy(pq0d, pq0.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(pq1d, pq1.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(mo0d, zeros.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(mo1d, mo1.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(latd, lat.data(), sizeof(int) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(vmatd, vmat.data(), sizeof(int) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(vrobd, vrob.data(), sizeof(int) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(phased, phase.data(), sizeof(T) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(vlinksd, vlinks.data(), sizeof(int) * 6 * V, hipMemcpyHostToDevice));
        if (LP) {
            HIP_TRY(hipMemcpy(lendsd, lends.data(), sizeof(int2) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(lmetad, lmeta.data(), sizeof(int2) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(ls0, zeros.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(ls1, zeros.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(ls2, ls2h.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(lf0, zeros.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(lf1, zeros.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(lf2, zeros.data(), sizeof(T4<T>) * LP, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d_lmats, lmats.data(), sizeof(LinkMat) * lmats.size(), hipMemcpyHostToDevice));
        }
        HIP_TRY(hipMemcpy(d_vmats, hvmats.data(), sizeof(VoxelMat) * hvmats.size(), hipMemcpyHostToDevice));
        if (!exts.empty())
            HIP_TRY(hipMemcpy(d_exts, exts.data(), sizeof(VoxelExt) * exts.size(), hipMemcpyHostToDevice));
        if (!curves.empty())
            HIP_TRY(hipMemcpy(d_curves, curves.data(), sizeof(CurvePt) * curves.size(), hipMemcpyHostToDevice));
        if (!formulas.empty())
            HIP_TRY(hipMemcpy(d_formulas, formulas.data(), sizeof(MtToken) * formulas.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_vmatE, vmatE.data(), sizeof(float) * vmatE.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_rp, hrp.data(), sizeof(RobotParams) * R, hipMemcpyHostToDevice));
        std::vector<RobotTime> rt((size_t)R);
        for (int r = 0; r < R; r++) {
            std::memset(&rt[r], 0, sizeof(RobotTime));
            rt[r].dt = defaultDt[r];
            rt[r].stale[0] = rt[r].stale[1] = 1;  // candidate lists are built by the first step
            // InitializeCenterOfMass: the sampled centre of mass starts at the initial one
            const Model &m = models[r]->m;
            double sc[3] = {0, 0, 0}, total = 0;
            for (int v = 0; v < m.nv(); v++) {
                const VoxelMat &vm = m.vmats[m.vmat[v]];
                if (!vm.isMeasured) continue;
                const double L0 = vm.nomSize, w = vm.mass;
                sc[0] = sc[0] + (m.ix[v] * L0) * w;
                sc[1] = sc[1] + (m.iy[v] * L0) * w;
                sc[2] = sc[2] + (m.iz[v] * L0) * w;
                total += w;
            }
            for (int k = 0; k < 3; k++) rt[r].scom[k] = total == 0 ? 0.0 : sc[k] / total;
        }
        curDt = defaultDt;
        HIP_TRY(hipMemcpy(d_rt, rt.data(), sizeof(RobotTime) * R, hipMemcpyHostToDevice));

        // collision broadphase: surface list, per-robot hash tables
        std::vector<int> surf;
        long buckets = 0;
        for (int r = 0; r < n; r++) {
            const Model &m = models[r]->m;
            RobotParams &p = hrp[r];
            p.soff = (int)surf.size();
            // with removal the list also holds the voxels linked to a removable
            // one (they become surface when it goes), in voxel order
            std::vector<int> rs;
            if (m.enableCollision || p.attach) {  // attach finds its pairs through the same lists
                std::vector<char> in((size_t)m.nv(), 0);
                for (int v : m.surface) in[v] = 1;
                if (m.enableDetach)  // the ends of links that can fail
                    for (int l = 0; l < m.nl(); l++)
                        if (m.lmats[m.lmat[l]].epsFail != -1.0f) in[m.lneg[l]] = in[m.lpos[l]] = 1;
                if (m.secondaryExperiment)
                    for (int v = 0; v < m.nv(); v++) {
                        if (!(m.palette[m.vpal[v] - 1].remove_after > 0)) continue;
                        for (int k = 0; k < 6; k++) {
                            const int l = m.vlinks[(size_t)v * 6 + k];
                            if (l >= 0) in[(k & 1) ? m.lneg[l] : m.lpos[l]] = 1;
                        }
                    }
                for (int v = 0; v < m.nv(); v++)
                    if (in[v]) rs.push_back(v);
            }
            p.scount = (int)rs.size();
            if (p.scount) {
                for (int v : rs) surf.push_back(p.voff + v);
                int hsz = 64;
                while (hsz < p.scount) hsz <<= 1;
                p.toff = (int)buckets;
                p.mask = hsz - 1;
                buckets += hsz + 4;  // the total after the last bucket; tables on 16 bytes
                // the largest base size: thermal factor 1 + T * CTE over the
                // temperatures the robot takes (the import temperature and, with
                // actuation, -A .. A or -A .. 0 without EnableExpansion)
                double tlo = m.initialTemp, thi = m.initialTemp;
                if (m.varyTemp) {
                    tlo = std::min(tlo, -std::fabs(m.tempAmplitude));
                    thi = std::max(thi, m.enableExpansion ? std::fabs(m.tempAmplitude) : 0.0);
                }
                double fmax = 1.0;
                for (const VoxelMat &vm : m.vmats)
                    fmax = std::max({fmax, 1.0 + (double)vm.alphaCTE * tlo, 1.0 + (double)vm.alphaCTE * thi});
                const double L0 = m.latticeDim;
                const double skin = m.nv() <= kSmallRobotVoxels ? kSkinSmall : kSkin;
                p.listR = (2 * kCOLLISION_ENVELOPE_RADIUS * L0 * fmax + skin * L0) * 1.0001;
                p.staleD2 = (float)((0.49 * skin * L0) * (0.49 * skin * L0));
                p.listLinked = (m.enableDetach || m.secondaryExperiment || p.attach) ? 1 : 0;
                anyCollisions = true;
            }
        }
        S = (int)surf.size();
        tableBuckets = buckets;
        HIP_TRY(hipMemcpy(d_rp, hrp.data(), sizeof(RobotParams) * R, hipMemcpyHostToDevice));
        d.contact = nullptr;
        d.contactW = nullptr;
        d.vsurf = nullptr;
        d.anyStale = nullptr;
        d.ctick = nullptr;
        d.pairRec = nullptr;
        d.pairN = nullptr;
        d.pairCap = 0;
        if (anyCollisions) {
            int *surfd, *vsurfd, *anyd, *srankd, *ccntd, *cstartd, *clistd, *ccountd, *tickd;
            int4 *scd, *grecd;
            T4<T> *cw, *bpd, *gposd;
            if ((rc = alloc(&surfd, S)) || (rc = alloc(&vsurfd, V)) || (rc = alloc(&anyd, 2)) || (rc = alloc(&scd, S)) ||
                (rc = alloc(&srankd, S)) || (rc = alloc(&ccntd, buckets)) || (rc = alloc(&cstartd, buckets)) ||
                (rc = alloc(&grecd, S)) || (rc = alloc(&gposd, S)) || (rc = alloc(&clistd, (size_t)S * kMaxCand)) || (rc = alloc(&ccountd, S)) ||
                (rc = alloc(&cw, V)) || (rc = alloc(&bpd, V)) || (rc = alloc(&tickd, R)))
                return rc;
            std::vector<int> vs((size_t)V, -1);
            for (int i = 0; i < S; i++) vs[surf[i]] = i;
            const int two[2] = {1, 1};
            HIP_TRY(hipMemcpy(surfd, surf.data(), sizeof(int) * S, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(vsurfd, vs.data(), sizeof(int) * V, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(anyd, two, sizeof(two), hipMemcpyHostToDevice));
            HIP_TRY(hipMemset(ccntd, 0, sizeof(int) * buckets));
            HIP_TRY(hipMemset(cstartd, 0, sizeof(int) * buckets));
            HIP_TRY(hipMemset(ccountd, 0, sizeof(int) * S));
            HIP_TRY(hipMemset(cw, 0, sizeof(T4<T>) * V));
            HIP_TRY(hipMemset(bpd, 0, sizeof(T4<T>) * V));
            HIP_TRY(hipMemset(tickd, 0, sizeof(int) * R));
            HIP_TRY(hipDeviceSynchronize());
            d.surf = surfd; d.vsurf = vsurfd; d.anyStale = anyd; d.ctick = tickd; d.scell = scd; d.srank = srankd;
            d.ccnt = ccntd; d.cstart = cstartd; d.grec = grecd; d.gpos = gposd; d.clist = clistd; d.ccount = ccountd; d.bpos = bpd;
            d.contactW = cw;
            d.contact = cw;
        }

        d.pq0 = pq0d; d.pq1 = pq1d; d.mo0 = mo0d; d.mo1 = mo1d;
        d.lat = latd; d.vmat = vmatd; d.vrob = vrobd; d.phase = phased; d.vlinks = vlinksd;
        d.lends = lendsd; d.lmeta = lmetad;
        d.ls0 = ls0; d.ls1 = ls1; d.ls2 = ls2; d.lf0 = lf0; d.lf1 = lf1; d.lf2 = lf2;
        d.field = nullptr;
        if (anyForceField) {
            T4<T> *fd;
            if ((rc = alloc(&fd, V))) return rc;
            HIP_TRY(hipMemcpy(fd, zeros.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
            d.field = fd;
        }
        d.vmats = d_vmats; d.exts = d_exts; d.lmats = d_lmats; d.curves = d_curves; d.formulas = d_formulas;
        d.rp = d_rp; d.rt = d_rt;
        d.cilia = nullptr;
        if (anyCilia) {
            VoxelCilia *cd;
            if ((rc = alloc(&cd, V))) return rc;
            HIP_TRY(hipMemcpy(cd, cilia.data(), sizeof(VoxelCilia) * V, hipMemcpyHostToDevice));
            d.cilia = cd;
        }
        d.vrem = d.lrem = nullptr;
        d.remCand = nullptr;
        d.remAt = nullptr;
        d.nRemCand = 0;
        d.detCand = nullptr;
        d.nDetCand = 0;
        if (anyRemoval || anyDetach || anyAttach) {
            int *vr, *lr, *rc0, *dc;
            double *ra;
            const int nc = (int)remCand.size(), nd = (int)detCand.size();
            if ((rc = alloc(&vr, V)) || (rc = alloc(&lr, LP)) || (rc = alloc(&rc0, nc)) || (rc = alloc(&ra, nc)) ||
                (rc = alloc(&dc, nd)))
                return rc;
            HIP_TRY(hipMemset(vr, 0, sizeof(int) * V));
            std::vector<int> lr0((size_t)std::max(LP, 1), 0);
            std::fill(lr0.begin() + L, lr0.begin() + LP, -1);  // pool entries not handed out
            HIP_TRY(hipMemcpy(lr, lr0.data(), sizeof(int) * std::max(LP, 1), hipMemcpyHostToDevice));
            if (nc) HIP_TRY(hipMemcpy(rc0, remCand.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
            if (nc) HIP_TRY(hipMemcpy(ra, remAt.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
            if (nd) HIP_TRY(hipMemcpy(dc, detCand.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
            HIP_TRY(hipDeviceSynchronize());
            d.vrem = vr; d.lrem = lr; d.remCand = rc0; d.remAt = ra; d.nRemCand = nc;
            d.detCand = dc; d.nDetCand = nd;
        }
        d.vtarget = nullptr;
        if (anyTarget) {
            int *td;
            if ((rc = alloc(&td, V))) return rc;
            HIP_TRY(hipMemcpy(td, vtarget.data(), sizeof(int) * V, hipMemcpyHostToDevice));
            d.vtarget = td;
        }
        d.targets = nullptr;
        if (anySample && !targets.empty()) {
            int *tg;
            if ((rc = alloc(&tg, targets.size()))) return rc;
            HIP_TRY(hipMemcpy(tg, targets.data(), sizeof(int) * targets.size(), hipMemcpyHostToDevice));
            d.targets = tg;
        }
        d.vatt = nullptr;
        d.vflagA = d.acand = d.acount = d.lmade = nullptr;
        if (anyAttach) {
            int2 *vad;
            int *vfd, *acd, *acn, *lmd;
            if ((rc = alloc(&vad, V)) || (rc = alloc(&vfd, V)) || (rc = alloc(&acd, (size_t)std::max(S, 1) * kAttachPerVoxel)) ||
                (rc = alloc(&acn, std::max(S, 1))) || (rc = alloc(&lmd, LP)))
                return rc;
            HIP_TRY(hipMemset(lmd, 0, sizeof(int) * std::max(LP, 1)));
            const std::vector<int> ones((size_t)V, 1);  // enableAttach is true until the first voxel update
            HIP_TRY(hipMemcpy(vad, vatt.data(), sizeof(int2) * V, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(vfd, ones.data(), sizeof(int) * V, hipMemcpyHostToDevice));
            HIP_TRY(hipMemset(acn, 0, sizeof(int) * std::max(S, 1)));
            d.vatt = vad; d.vflagA = vfd; d.acand = acd; d.acount = acn; d.lmade = lmd;
        }
        d.sig = nullptr;
        d.sigMat = nullptr;
        if (anySignals) {
            VoxelSignal *sd;
            SignalMat *smd;
            if ((rc = alloc(&sd, V)) || (rc = alloc(&smd, V))) return rc;
            HIP_TRY(hipMemset(sd, 0, sizeof(VoxelSignal) * V));
            HIP_TRY(hipMemcpy(smd, sigMat.data(), sizeof(SignalMat) * V, hipMemcpyHostToDevice));
            d.sig = sd;
            d.sigMat = smd;
        }
        return plan_tiles();
    }

    // Longest-processing-time bin packing of robots into tiles (k_tile): at
    // least one tile per workgroup slot of the device (compute units x
    // workgroups per unit), robots sorted by link count, each placed in the
    // least-loaded tile that still has room.
    template <int SMALL, int COL> int plan_tiles_cfg() {
        using C = typename TileSel<T, SMALL, COL>::C;
        TilePlan &P = plans[SMALL];
        P.ok = false;
        if (anyForceField || anyCilia || plainFlatOnly() || R < 2) return VX3MI_OK;
        if (anyCollisions && SMALL) return VX3MI_OK;  // collision tiles: the large configuration
        long totL = 0, totV = 0;
        for (const RobotParams &p : hrp) {
            if (p.vcount > C::VMAX || p.lcount > C::LMAX) return VX3MI_OK;
            totL += p.lcount;
            totV += p.vcount;
        }
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
        }
        const long slots = (long)cus * (SMALL ? 2 : 1);
        long want = std::max<long>({slots, (totL + C::LMAX - 1) / C::LMAX, (totV + C::VMAX - 1) / C::VMAX});
        want = std::min<long>(want, R);
        std::vector<int> order(R);
        for (int r = 0; r < R; r++) order[r] = r;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return hrp[x].lcount > hrp[y].lcount; });
        struct Bin { long l = 0, v = 0; std::vector<int> robots; };
        std::vector<Bin> bins((size_t)want);
        for (int r : order) {
            const RobotParams &p = hrp[r];
            int best = -1;
            for (int b = 0; b < (int)bins.size(); b++) {
                const Bin &x = bins[b];
                if (x.l + p.lcount > C::LMAX || x.v + p.vcount > C::VMAX || (int)x.robots.size() >= C::RMAX) continue;
                if (best < 0 || x.l + x.v < bins[best].l + bins[best].v) best = b;
            }
            if (best < 0) { bins.emplace_back(); best = (int)bins.size() - 1; }
            bins[best].l += p.lcount;
            bins[best].v += p.vcount;
            bins[best].robots.push_back(r);
        }
        std::vector<int4> tiles;
        std::vector<int> trobots;
        for (Bin &b : bins) {
            if (b.robots.empty()) continue;
            std::sort(b.robots.begin(), b.robots.end());
            tiles.push_back(int4{(int)trobots.size(), (int)b.robots.size(), (int)b.v, (int)b.l});
            for (int r : b.robots) trobots.push_back(r);
        }
        P.ntiles = (int)tiles.size();
        int rc;
        if ((rc = alloc(&P.d_tiles, tiles.size())) || (rc = alloc(&P.d_tileRobots, trobots.size()))) return rc;
        if (COL && (rc = alloc(&d.tcand, tiles.size() * kMaxCand * C::KV * C::NT))) return rc;
        HIP_TRY(hipMemcpy(P.d_tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P.d_tileRobots, trobots.data(), sizeof(int) * trobots.size(), hipMemcpyHostToDevice));
        P.ok = true;
        return VX3MI_OK;
    }

    int plan_tiles() {
        int rc;
        if (anyCollisions) rc = plan_tiles_cfg<0, 1>();
        else rc = plan_tiles_cfg<0, 0>();
        if (rc || (rc = plan_tiles_cfg<1, 0>())) return rc;
        tileOK = plans[0].ok || plans[1].ok;
        tileSel = plans[0].ok ? 0 : 1;
        autoMode = tileOK ? VX3MI_MODE_TILED
                          : ((R > 1 && V / R <= 4096 && !plainFlatOnly()) ? VX3MI_MODE_PERSISTENT : VX3MI_MODE_FLAT);
        return VX3MI_OK;
    }

    void launch_tiled(int k, int par) {
        const int nlm = nLinkMats, nvm = (int)hvmats.size();
        const int cap = tileSel ? TileMatCache<1>::N : TileMatCache<0>::N;
        const bool cache = nlm <= cap && nvm <= cap;
        const TilePlan &P = plans[tileSel];
        if (anyCollisions) {
            constexpr int NTC = TileColCfg<T>::NT;
            if (cache) k_tile<T, 0, 1, 1><<<P.ntiles, NTC, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
            else k_tile<T, 0, 0, 1><<<P.ntiles, NTC, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
        } else if (tileSel == 0) {
            if (cache) k_tile<T, 0, 1, 0><<<P.ntiles, TileCfg<T, 0>::NT, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
            else k_tile<T, 0, 0, 0><<<P.ntiles, TileCfg<T, 0>::NT, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
        } else {
            if (cache) k_tile<T, 1, 1, 0><<<P.ntiles, TileCfg<T, 1>::NT, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
            else k_tile<T, 1, 0, 0><<<P.ntiles, TileCfg<T, 1>::NT, 0, stream>>>(d, P.d_tiles, P.d_tileRobots, k, par, nlm, nvm);
        }
    }
    void launch_persistent(int k, int par) {
        if (persistThreads == 128) k_persistent<T, 128><<<R, 128, 0, stream>>>(d, k, par);
        else if (persistThreads == 512) k_persistent<T, 512><<<R, 512, 0, stream>>>(d, k, par);
        else k_persistent<T, 256><<<R, 256, 0, stream>>>(d, k, par);
    }

    // upload per-robot dt if it changed (robot time/prevDt stay on the device)
    int set_dt(float dt) {
        std::vector<float> want = defaultDt;
        if (dt >= 0) std::fill(want.begin(), want.end(), dt);
        if (want == curDt) return VX3MI_OK;
        for (int r = 0; r < R; r++)
            HIP_TRY(hipMemcpyAsync(&d_rt[r].dt, &want[r], sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        curDt = want;
        return VX3MI_OK;
    }

    // the listing passes of robots [r0, r0 + nr) -- count, prefix sum and
    // scatter; they return at once unless some robot's lists are stale (all:
    // check the robots' own flags only) -- the lists themselves are built by the
    // contact kernels that follow
    void launch_listing(int r0, int nr, int parity, int all) {
        int parts = 1;  // count-pass workgroups of the largest robot
        for (int r = r0; r < r0 + nr; r++) parts = std::max(parts, (hrp[r].scount + kCountPerGroup - 1) / kCountPerGroup);
        k_col_scan<T><<<dim3(nr, parts), 1024, 0, stream>>>(d, r0, parity, all);
    }
    // the collision phase of one step for surface voxels [s0, s0 + n) (robots
    // [r0, r0 + nr)): the listing passes, then the lists and contact forces
    void launch_collisions(int s0, int n, int r0, int nr, int parity, int all) {
        launch_listing(r0, nr, parity, all);
        k_col_contact<T><<<(n * kContactLanes + 255) / 256, 256, 0, stream>>>(d, s0, n, parity);
    }
    static constexpr int kCollisionLaunches = 2;  // plain mode; the flat step runs the contacts beside the links

    void launch_flat_step(int parity) {
        if (anyCollisions) {
            launch_listing(0, R, parity, 0);
            const int nbl = (LP + 255) / 256, nbc = (S * kContactLanes + 255) / 256;
            k_contact_links<T><<<nbl + nbc, 256, 0, stream>>>(d, LP, nbl, S, parity);
        } else if (LP) {
            k_links<T><<<(LP + 255) / 256, 256, 0, stream>>>(d, 0, LP, parity);
        }
        if (anyAttach) launch_attach(0, S, 0, R);
        if (anyForceField) k_force_field<T><<<(V + 255) / 256, 256, 0, stream>>>(d, 0, V, parity);
        k_voxels<T><<<(V + 255) / 256, 256, 0, stream>>>(d, 0, V, parity);
        if (anyAttach) k_attach_flags<T><<<(V + 255) / 256, 256, 0, stream>>>(d, 0, V, parity);
        if (anySample) k_cycle_sample<T><<<(R + 63) / 64, 64, 0, stream>>>(d, 0, R);
        if (anyDetach) k_detach<T><<<(d.nDetCand + 255) / 256, 256, 0, stream>>>(d);
        if (anySignals) k_signal_deliver<T><<<(V + 255) / 256, 256, 0, stream>>>(d, 0, V);
        if (anyRemoval) k_removal<T><<<(d.nRemCand + 255) / 256, 256, 0, stream>>>(d, parity);
    }
    int kernels_per_flat_step() const {
        return 2 + (anyCollisions ? kCollisionLaunches - 1 : 0) + (anyForceField ? 1 : 0) + (anyRemoval ? 1 : 0) + (anySignals ? 1 : 0) +
               (anyDetach ? 1 : 0) + (anyAttach ? 3 : 0) + (anySample ? 1 : 0);
    }
    // the attach phase of surface voxels [s0, s0 + n) (robots [r0, r0 + nr)):
    // candidates, then the links made robot by robot
    void launch_attach(int s0, int n, int r0, int nr) {
        if (n) k_attach_find<T><<<(n + 255) / 256, 256, 0, stream>>>(d, s0, n);
        k_attach_walk<T><<<nr, 256, 0, stream>>>(d, r0);
    }

    int step(int nsteps, float dt, int mode) override {
        if (nsteps <= 0) return VX3MI_OK;
        if (dt == 0.0f) {  // doTimeStep(0) only advances the step counter (VX3_VoxelyzeKernel.cu:241-243)
            k_zero_steps<T><<<(R + 255) / 256, 256, 0, stream>>>(d_rt, R, nsteps, parity, d.anyStale);
            HIP_TRY(hipGetLastError());
            parity = (parity + nsteps) & 1;
            return VX3MI_OK;
        }
        int rc = set_dt(dt);
        if (rc) return rc;
        if (mode == VX3MI_MODE_TILED && !plans[tileSel].ok) {
            set_error("tiled mode needs force fields, cilia, signals, detach and removal off and every robot within one tile");
            return VX3MI_E_ARG;
        }
        if (mode == VX3MI_MODE_AUTO) mode = autoMode;
        if (mode == VX3MI_MODE_PERSISTENT && plainFlatOnly()) {
            set_error("force fields, material removal, detach and signals run in the plain and flat modes only");
            return VX3MI_E_ARG;
        }
        if (d.pairRec) {  // each call records its own steps (a fill kernel: no memset node under capture)
            k_zero_int<<<1, 1, 0, stream>>>(d.pairN);
            HIP_TRY(hipGetLastError());
        }
        if (graphStale) {
            drop_graphs();
            graphStale = false;
        }
        HIP_TRY(hipEventRecord(ev0, stream));
        launches = 0;
        lastMode = mode;
        if (mode == VX3MI_MODE_PLAIN) {
            for (int s = 0; s < nsteps; s++) {
                for (int r = 0; r < R; r++) {
                    const RobotParams &p = hrp[r];
                    if (p.scount) launch_collisions(p.soff, p.scount, r, 1, parity, 1);
                    if (p.lcount) k_links<T><<<(p.lcount + 255) / 256, 256, 0, stream>>>(d, p.loff, p.lcount, parity);
                    if (p.pcount) k_links<T><<<(p.pcount + 255) / 256, 256, 0, stream>>>(d, p.poff, p.pcount, parity);
                    if (p.attach) launch_attach(p.soff, p.scount, r, 1);
                    if (anyForceField) k_force_field<T><<<(p.vcount + 255) / 256, 256, 0, stream>>>(d, p.voff, p.vcount, parity);
                    k_voxels<T><<<(p.vcount + 255) / 256, 256, 0, stream>>>(d, p.voff, p.vcount, parity);
                    if (p.attach) k_attach_flags<T><<<(p.vcount + 255) / 256, 256, 0, stream>>>(d, p.voff, p.vcount, parity);
                    if (p.sample) k_cycle_sample<T><<<1, 64, 0, stream>>>(d, r, 1);
                    launches += 2 + (p.scount ? kCollisionLaunches : 0) + (anyForceField ? 1 : 0) + (p.pcount ? 1 : 0) +
                                (p.attach ? 3 : 0) + (p.sample ? 1 : 0);
                }
                // the end-of-step phases for the whole batch (each touches only its
                // robot's voxels and links)
                if (anyDetach) {
                    k_detach<T><<<(d.nDetCand + 255) / 256, 256, 0, stream>>>(d);
                    launches++;
                }
                if (anySignals) {
                    k_signal_deliver<T><<<(V + 255) / 256, 256, 0, stream>>>(d, 0, V);
                    launches++;
                }
                if (anyRemoval) {
                    k_removal<T><<<(d.nRemCand + 255) / 256, 256, 0, stream>>>(d, parity);
                    launches++;
                }
                parity ^= 1;
            }
        } else if (mode == VX3MI_MODE_TILED) {
            for (int s = 0; s < nsteps; s += persistentChunk) {
                const int k = std::min(persistentChunk, nsteps - s);
                launch_tiled(k, parity);
                parity = (parity + k) & 1;
                launches++;
            }
        } else if (mode == VX3MI_MODE_PERSISTENT) {
            // launches of at most persistentChunk steps keep each kernel short
            for (int s = 0; s < nsteps; s += persistentChunk) {
                const int k = std::min(persistentChunk, nsteps - s);
                launch_persistent(k, parity);
                parity = (parity + k) & 1;
                launches++;
            }
        } else {
            // 1000-step graphs while at least that many steps remain, then 64-step
            // ones, then single steps
            int s = 0;
            for (int g = 0; g < 2; g++) {
                const int chunk = kGraphChunks[g];
                if (nsteps - s < chunk) continue;
                if (!graphs[g] || graphParity[g] != parity) {
                    if (graphs[g]) { (void)hipGraphExecDestroy(graphs[g]); graphs[g] = nullptr; }
                    hipGraph_t cg;
                    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                    for (int k = 0; k < chunk; k++) launch_flat_step((parity + k) & 1);
                    HIP_TRY(hipStreamEndCapture(stream, &cg));
                    const hipError_t e = hipGraphInstantiate(&graphs[g], cg, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(cg);
                    HIP_TRY(e);
                    graphParity[g] = parity;
                }
                for (; s + chunk <= nsteps; s += chunk) {
                    HIP_TRY(hipGraphLaunch(graphs[g], stream));
                    launches += (long)kernels_per_flat_step() * chunk;
                }
            }
            for (; s < nsteps; s++) {
                launch_flat_step(parity);
                parity ^= 1;
                launches += kernels_per_flat_step();
            }
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ev1, stream));
        return VX3MI_OK;
    }

    int fetch(std::vector<T4<T>> &a0, std::vector<T4<T>> &a1, std::vector<T4<T>> &b0, std::vector<T4<T>> &b1) {
        a0.resize(V); a1.resize(V); b0.resize(V); b1.resize(V);
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(a0.data(), d.pq0, sizeof(T4<T>) * V, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(a1.data(), d.pq1, sizeof(T4<T>) * V, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b0.data(), d.mo0, sizeof(T4<T>) * V, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b1.data(), d.mo1, sizeof(T4<T>) * V, hipMemcpyDeviceToHost));
        return VX3MI_OK;
    }

    void range(int robot, int &v0, int &v1) {
        if (robot < 0) { v0 = 0; v1 = V; }
        else { v0 = hrp[robot].voff; v1 = v0 + hrp[robot].vcount; }
    }

    int get_state(int robot, double *pos, double *quat, double *lm, double *am) override {
        std::vector<T4<T>> a0, a1, b0, b1;
        int rc = fetch(a0, a1, b0, b1);
        if (rc) return rc;
        int v0, v1;
        range(robot, v0, v1);
        for (int v = v0; v < v1; v++) {
            const int i = v - v0, pk = hlat[v];
            const double Ls = kAbsolutePositions<T> ? 0.0 : hvmats[hvmat[v]].nomSize;
            if (pos) {
                pos[3 * i] = (pk & 1023) * Ls + (double)a0[v].x;
                pos[3 * i + 1] = ((pk >> 10) & 1023) * Ls + (double)a0[v].y;
                pos[3 * i + 2] = (pk >> 20) * Ls + (double)a0[v].z;
            }
            if (quat) { quat[4 * i] = a0[v].w; quat[4 * i + 1] = a1[v].x; quat[4 * i + 2] = a1[v].y; quat[4 * i + 3] = a1[v].z; }
            if (lm) { lm[3 * i] = b0[v].x; lm[3 * i + 1] = b0[v].y; lm[3 * i + 2] = b0[v].z; }
            if (am) { am[3 * i] = b0[v].w; am[3 * i + 1] = b1[v].x; am[3 * i + 2] = b1[v].y; }
        }
        return VX3MI_OK;
    }

    int set_state(int robot, const double *pos, const double *quat, const double *lm, const double *am) override {
        std::vector<T4<T>> a0, a1, b0, b1;
        int rc = fetch(a0, a1, b0, b1);
        if (rc) return rc;
        int v0, v1;
        range(robot, v0, v1);
        for (int v = v0; v < v1; v++) {
            const int i = v - v0, pk = hlat[v];
            const double Ls = kAbsolutePositions<T> ? 0.0 : hvmats[hvmat[v]].nomSize;
            if (pos) {
                a0[v].x = (T)(pos[3 * i] - (pk & 1023) * Ls);
                a0[v].y = (T)(pos[3 * i + 1] - ((pk >> 10) & 1023) * Ls);
                a0[v].z = (T)(pos[3 * i + 2] - (pk >> 20) * Ls);
            }
            if (quat) { a0[v].w = (T)quat[4 * i]; a1[v].x = (T)quat[4 * i + 1]; a1[v].y = (T)quat[4 * i + 2]; a1[v].z = (T)quat[4 * i + 3]; }
            if (lm) { b0[v].x = (T)lm[3 * i]; b0[v].y = (T)lm[3 * i + 1]; b0[v].z = (T)lm[3 * i + 2]; }
            if (am) { b0[v].w = (T)am[3 * i]; b1[v].x = (T)am[3 * i + 1]; b1[v].y = (T)am[3 * i + 2]; }
        }
        HIP_TRY(hipMemcpy(d.pq0, a0.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.pq1, a1.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.mo0, b0.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.mo1, b1.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice));
        return mark_stale();
    }

    // the candidate lists of every robot are rebuilt by the next step (after the
    // state was replaced)
    int mark_stale() {
        if (!anyCollisions) return VX3MI_OK;
        k_mark_stale<T><<<(R + 255) / 256, 256, 0, stream>>>(d_rt, R, d.anyStale);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream));
        return VX3MI_OK;
    }

    // ---- autotuner: candidates are the large / small tile, persistent
    // workgroups of 128 / 256 / 512 threads, and flat graph launches
    static constexpr int kCandidates = 6;
    int autotune(int trialSteps, double *ms, int *choice) override {
        if (trialSteps <= 0) trialSteps = 100;
        // snapshot of everything a step changes
        struct Buf { void *p; size_t bytes; };
        std::vector<Buf> bufs = {{d.pq0, sizeof(T4<T>) * V}, {d.pq1, sizeof(T4<T>) * V}, {d.mo0, sizeof(T4<T>) * V},
                                 {d.mo1, sizeof(T4<T>) * V}, {d.ls0, sizeof(T4<T>) * LP}, {d.ls1, sizeof(T4<T>) * LP},
                                 {d.ls2, sizeof(T4<T>) * LP}, {d.lf0, sizeof(T4<T>) * LP}, {d.lf1, sizeof(T4<T>) * LP},
                                 {d.lf2, sizeof(T4<T>) * LP}, {d_rt, sizeof(RobotTime) * R}};
        // (collision lists need no snapshot: they are rebuilt after each restore,
        // and a step's partners do not depend on when the lists were built)
        if (anySignals) bufs.push_back({d.sig, sizeof(VoxelSignal) * (size_t)V});
        if (anyRemoval || anyDetach || anyAttach) {
            bufs.push_back({d.vlinks, sizeof(int) * 6 * (size_t)V});
            bufs.push_back({d.vrem, sizeof(int) * (size_t)V});
            bufs.push_back({d.lrem, sizeof(int) * (size_t)LP});
        }
        if (anyAttach) {
            bufs.push_back({d.lends, sizeof(int2) * (size_t)LP});
            bufs.push_back({d.lmeta, sizeof(int2) * (size_t)LP});
            bufs.push_back({d.vflagA, sizeof(int) * (size_t)V});
            if (d.contactW) bufs.push_back({d.contactW, sizeof(T4<T>) * (size_t)V});
            bufs.push_back({d.lmade, sizeof(int) * (size_t)LP});
        }
        size_t total = 0;
        for (const Buf &b : bufs) total += (b.bytes + 255) & ~(size_t)255;
        char *snap = nullptr;
        HIP_TRY(hipMalloc((void **)&snap, total));
        auto copyAll = [&](bool save) -> int {
            size_t off = 0;
            for (const Buf &b : bufs) {
                if (b.bytes) {
                    void *dst = save ? (void *)(snap + off) : b.p;
                    const void *src = save ? b.p : (const void *)(snap + off);
                    HIP_TRY(hipMemcpyAsync(dst, src, b.bytes, hipMemcpyDeviceToDevice, stream));
                }
                off += (b.bytes + 255) & ~(size_t)255;
            }
            HIP_TRY(hipStreamSynchronize(stream));
            return save ? VX3MI_OK : mark_stale();
        };
        const std::vector<float> curDt0 = curDt;  // the dt the snapshot holds
        const int parity0 = parity, tile0 = tileSel, threads0 = persistThreads, mode0 = autoMode;
        const int modes[kCandidates] = {VX3MI_MODE_TILED, VX3MI_MODE_TILED, VX3MI_MODE_PERSISTENT,
                                        VX3MI_MODE_PERSISTENT, VX3MI_MODE_PERSISTENT, VX3MI_MODE_FLAT};
        const int tiles[kCandidates] = {0, 1, 0, 0, 0, 0}, threads[kCandidates] = {0, 0, 128, 256, 512, 0};
        int rc = copyAll(true), best = -1;
        double bestMs = 0;
        for (int c = 0; rc == VX3MI_OK && c < kCandidates; c++) {
            ms[c] = -1;
            if (modes[c] == VX3MI_MODE_TILED && !plans[tiles[c]].ok) continue;
            if (modes[c] == VX3MI_MODE_PERSISTENT && plainFlatOnly()) continue;
            tileSel = tiles[c];
            if (threads[c]) persistThreads = threads[c];
            if ((rc = step(std::min(trialSteps, 64), -1.0f, modes[c])) < 0) break;  // code load, graph capture
            if ((rc = copyAll(false))) break;
            parity = parity0;
            if ((rc = step(trialSteps, -1.0f, modes[c])) < 0) break;
            HIP_TRY(hipStreamSynchronize(stream));
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, ev0, ev1));
            ms[c] = t;
            if (best < 0 || t < bestMs) { best = c; bestMs = t; }
            if ((rc = copyAll(false))) break;
            parity = parity0;
        }
        if (rc == VX3MI_DIVERGED) rc = VX3MI_OK;
        (void)hipFree(snap);
        curDt = curDt0;
        if (rc != VX3MI_OK) {
            tileSel = tile0; persistThreads = threads0; autoMode = mode0; parity = parity0;
            return rc;
        }
        autoMode = modes[best];
        tileSel = modes[best] == VX3MI_MODE_TILED ? tiles[best] : (plans[0].ok ? 0 : 1);
        persistThreads = threads[best] ? threads[best] : 256;
        parity = parity0;
        drop_graphs();
        choice[0] = autoMode;
        choice[1] = tileSel;
        choice[2] = persistThreads;
        return VX3MI_OK;
    }

    int2 *pairBuf = nullptr;
    int *pairNBuf = nullptr;
    int pairAlloc = 0;
    int record_contacts(int cap) override {
        if (cap < 0) { set_error("negative capacity"); return VX3MI_E_ARG; }
        if (cap > 0 && !anyCollisions) { set_error("no robot of the batch has collisions"); return VX3MI_E_ARG; }
        HIP_TRY(hipStreamSynchronize(stream));
        if (cap > pairAlloc) {
            int rc;
            if ((rc = alloc(&pairBuf, cap)) || (!pairNBuf && (rc = alloc(&pairNBuf, 1)))) return rc;
            pairAlloc = cap;
        }
        return record_contacts_at(cap ? pairBuf : nullptr, cap ? pairNBuf : nullptr, cap);
    }

    int record_contacts_at(int2 *rec, int *counter, int cap) override {
        if (cap < 0 || (cap > 0 && (!rec || !counter))) { set_error("invalid contact record"); return VX3MI_E_ARG; }
        if (cap > 0 && !anyCollisions) { set_error("no robot of the batch has collisions"); return VX3MI_E_ARG; }
        d.pairRec = cap ? rec : nullptr;
        d.pairN = cap ? counter : nullptr;
        d.pairCap = cap;
        graphStale = true;  // the flat graph holds the launch arguments
        if (cap) {
            k_zero_int<<<1, 1, 0, stream>>>(counter);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipStreamSynchronize(stream));
        return VX3MI_OK;
    }

    int attach_links(int robot, std::vector<int4> &out) override {
        out.clear();
        const RobotParams &p = hrp[robot];
        if (!p.attach) return VX3MI_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        RobotTime rt;
        HIP_TRY(hipMemcpy(&rt, d_rt + robot, sizeof(RobotTime), hipMemcpyDeviceToHost));
        const int used = rt.attachUsed;
        if (!used) return VX3MI_OK;
        std::vector<int2> e((size_t)used), mt((size_t)used);
        std::vector<int> rem((size_t)used), made((size_t)used);
        HIP_TRY(hipMemcpy(e.data(), d.lends + p.poff, sizeof(int2) * used, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(mt.data(), d.lmeta + p.poff, sizeof(int2) * used, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(rem.data(), d.lrem + p.poff, sizeof(int) * used, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(made.data(), d.lmade + p.poff, sizeof(int) * used, hipMemcpyDeviceToHost));
        for (int k = 0; k < used; k++)
            if (!rem[k]) out.push_back(int4{e[k].x - p.voff, e[k].y - p.voff, mt[k].x & 3, made[k]});
        return VX3MI_OK;
    }

    int attach_state(int robot, int *vflags, int *neg, int *pos, int *axis, int *newLeft, int cap,
                     int *overflow) override {
        const RobotParams &p = hrp[robot];
        HIP_TRY(hipStreamSynchronize(stream));
        if (vflags) {
            if (p.attach) HIP_TRY(hipMemcpy(vflags, d.vflagA + p.voff, sizeof(int) * p.vcount, hipMemcpyDeviceToHost));
            else std::fill(vflags, vflags + p.vcount, 1);
        }
        if (overflow) *overflow = 0;
        if (!p.attach) return 0;
        RobotTime rt;
        HIP_TRY(hipMemcpy(&rt, d_rt + robot, sizeof(RobotTime), hipMemcpyDeviceToHost));
        if (overflow) *overflow = rt.attachOverflow;
        const int used = rt.attachUsed;
        std::vector<int2> e((size_t)std::max(used, 1)), mt((size_t)std::max(used, 1));
        std::vector<T4<T>> s2((size_t)std::max(used, 1));
        std::vector<int> rem((size_t)std::max(used, 1));
        if (used) {
            HIP_TRY(hipMemcpy(e.data(), d.lends + p.poff, sizeof(int2) * used, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(mt.data(), d.lmeta + p.poff, sizeof(int2) * used, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(s2.data(), d.ls2 + p.poff, sizeof(T4<T>) * used, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(rem.data(), d.lrem + p.poff, sizeof(int) * used, hipMemcpyDeviceToHost));
        }
        int n = 0;  // present links in pool order (the order made, unless entries were reused)
        for (int k = 0; k < used; k++) {
            if (rem[k]) continue;
            if (n < cap) {
                if (neg) neg[n] = e[k].x - p.voff;
                if (pos) pos[n] = e[k].y - p.voff;
                if (axis) axis[n] = mt[k].x & 3;
                if (newLeft) newLeft[n] = (int)s2[k].z >> LF_NEW_SHIFT;
            }
            n++;
        }
        return n;
    }

    int contact_pairs(int robot, int *out, int cap) override {
        if (!d.pairRec) { set_error("contacts are not being recorded"); return VX3MI_E_ARG; }
        HIP_TRY(hipStreamSynchronize(stream));
        int n = 0;
        HIP_TRY(hipMemcpy(&n, d.pairN, sizeof(int), hipMemcpyDeviceToHost));
        if (n > d.pairCap) { set_error("more contact pairs than the record holds"); return VX3MI_E_LIMIT; }
        std::vector<int2> h((size_t)std::max(n, 1));
        if (n) HIP_TRY(hipMemcpy(h.data(), d.pairRec, sizeof(int2) * n, hipMemcpyDeviceToHost));
        int v0, v1;
        range(robot, v0, v1);
        std::vector<std::pair<int, int>> pr;
        for (int k = 0; k < n; k++) {
            const int a = h[k].x, b = h[k].y;
            if (a < v0 || a >= v1) continue;
            pr.emplace_back(std::min(a, b) - v0, std::max(a, b) - v0);
        }
        std::sort(pr.begin(), pr.end());
        for (int k = 0; k < (int)pr.size() && k < cap; k++) {
            out[2 * k] = pr[k].first;
            out[2 * k + 1] = pr[k].second;
        }
        return (int)pr.size();
    }

    int configure(int mode, int tile, int threads) override {
        if (mode == VX3MI_MODE_TILED && (tile < 0 || tile > 1 || !plans[tile].ok)) {
            set_error("this tile variant cannot hold every robot of the batch (batches with collisions: the large tile)");
            return VX3MI_E_ARG;
        }
        if (threads != 0 && threads != 128 && threads != 256 && threads != 512) {
            set_error("threads per robot must be 128, 256 or 512");
            return VX3MI_E_ARG;
        }
        if (mode < VX3MI_MODE_PLAIN || mode > VX3MI_MODE_TILED) { set_error("invalid mode"); return VX3MI_E_ARG; }
        if (mode == VX3MI_MODE_PERSISTENT && plainFlatOnly()) {
            set_error("force fields, material removal, detach and signals run in the plain and flat modes only");
            return VX3MI_E_ARG;
        }
        autoMode = mode;
        if (mode == VX3MI_MODE_TILED) tileSel = tile;
        if (threads) persistThreads = threads;
        return VX3MI_OK;
    }

    int frame(int robot, double *pos, double *quat, float *temp, float *strain, float *maxStrain) override {
        if (robot < 0 || robot >= R) { set_error("robot index out of range"); return VX3MI_E_ARG; }
        const RobotParams &p = hrp[robot];
        std::vector<T4<T>> a0((size_t)p.vcount), a1((size_t)p.vcount), l2((size_t)p.lcount);
        HIP_TRY(hipStreamSynchronize(stream));
        if (p.vcount) {
            HIP_TRY(hipMemcpy(a0.data(), d.pq0 + p.voff, sizeof(T4<T>) * p.vcount, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(a1.data(), d.pq1 + p.voff, sizeof(T4<T>) * p.vcount, hipMemcpyDeviceToHost));
        }
        if (p.lcount) HIP_TRY(hipMemcpy(l2.data(), d.ls2 + p.loff, sizeof(T4<T>) * p.lcount, hipMemcpyDeviceToHost));
        for (int i = 0; i < p.vcount; i++) {
            const int v = p.voff + i, pk = hlat[v];
            const double Ls = kAbsolutePositions<T> ? 0.0 : hvmats[hvmat[v]].nomSize;
            if (pos) {
                pos[3 * i] = (pk & 1023) * Ls + (double)a0[i].x;
                pos[3 * i + 1] = ((pk >> 10) & 1023) * Ls + (double)a0[i].y;
                pos[3 * i + 2] = (pk >> 20) * Ls + (double)a0[i].z;
            }
            if (quat) { quat[4 * i] = a0[i].w; quat[4 * i + 1] = a1[i].x; quat[4 * i + 2] = a1[i].y; quat[4 * i + 3] = a1[i].z; }
            if (temp) temp[i] = (float)a1[i].w;
        }
        for (int l = 0; l < p.lcount; l++) {
            if (strain) strain[l] = (float)l2[l].w;
            if (maxStrain) maxStrain[l] = (float)l2[l].y;
        }
        return VX3MI_OK;
    }

    int get_flags(int robot, int *flags) override {
        std::vector<T4<T>> a0, a1, b0, b1;
        int rc = fetch(a0, a1, b0, b1);
        if (rc) return rc;
        int v0, v1;
        range(robot, v0, v1);
        for (int v = v0; v < v1; v++) flags[v - v0] = (int)b1[v].z;
        return VX3MI_OK;
    }

    int get_link_flags(int robot, int *flags, float *strain) override {
        int l0 = 0, l1 = L;
        if (robot >= 0) { l0 = hrp[robot].loff; l1 = l0 + hrp[robot].lcount; }
        std::vector<T4<T>> s2((size_t)std::max(L, 1));
        HIP_TRY(hipStreamSynchronize(stream));
        if (L) HIP_TRY(hipMemcpy(s2.data(), d.ls2, sizeof(T4<T>) * L, hipMemcpyDeviceToHost));
        for (int l = l0; l < l1; l++) {
            if (flags) flags[l - l0] = (int)s2[l].z;
            if (strain) strain[l - l0] = (float)s2[l].w;
        }
        return VX3MI_OK;
    }

    int get_removed(int robot, int *vstep, int *lstep) override {
        int v0 = 0, v1 = V, l0 = 0, l1 = L;
        if (robot >= 0) {
            v0 = hrp[robot].voff; v1 = v0 + hrp[robot].vcount;
            l0 = hrp[robot].loff; l1 = l0 + hrp[robot].lcount;
        }
        HIP_TRY(hipStreamSynchronize(stream));
        if (!anyRemoval && !anyDetach) {
            if (vstep) std::fill(vstep, vstep + (v1 - v0), 0);
            if (lstep) std::fill(lstep, lstep + (l1 - l0), 0);
            return VX3MI_OK;
        }
        if (vstep && v1 > v0) HIP_TRY(hipMemcpy(vstep, d.vrem + v0, sizeof(int) * (v1 - v0), hipMemcpyDeviceToHost));
        if (lstep && l1 > l0) HIP_TRY(hipMemcpy(lstep, d.lrem + l0, sizeof(int) * (l1 - l0), hipMemcpyDeviceToHost));
        return VX3MI_OK;
    }

    int get_signals(int robot, double *local) override {
        int v0 = 0, v1 = V;
        if (robot >= 0) { v0 = hrp[robot].voff; v1 = v0 + hrp[robot].vcount; }
        HIP_TRY(hipStreamSynchronize(stream));
        if (!anySignals) {
            std::fill(local, local + (v1 - v0), 0.0);
            return VX3MI_OK;
        }
        std::vector<VoxelSignal> h((size_t)(v1 - v0));
        if (v1 > v0) HIP_TRY(hipMemcpy(h.data(), d.sig + v0, sizeof(VoxelSignal) * (v1 - v0), hipMemcpyDeviceToHost));
        for (int v = 0; v < v1 - v0; v++) local[v] = h[v].local;
        return VX3MI_OK;
    }

    int eval_links(double *out) override {
        if (L) k_links<T><<<(L + 255) / 256, 256, 0, stream>>>(d, 0, L, parity);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream));
        std::vector<T4<T>> f0(L), f1(L), f2(L);
        if (L) {
            HIP_TRY(hipMemcpy(f0.data(), d.lf0, sizeof(T4<T>) * L, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(f1.data(), d.lf1, sizeof(T4<T>) * L, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(f2.data(), d.lf2, sizeof(T4<T>) * L, hipMemcpyDeviceToHost));
        }
        for (int l = 0; l < L; l++) {
            double *o = out + 12 * l;
            // forceNeg, forcePos, momentNeg, momentPos
            o[0] = f0[l].x; o[1] = f0[l].y; o[2] = f0[l].z;
            o[3] = f1[l].z; o[4] = f1[l].w; o[5] = f2[l].x;
            o[6] = f0[l].w; o[7] = f1[l].x; o[8] = f1[l].y;
            o[9] = f2[l].y; o[10] = f2[l].z; o[11] = f2[l].w;
        }
        return VX3MI_OK;
    }

    int com(double *out) override {
        k_com<T><<<R, 256, 0, stream>>>(d.pq0, d.lat, d.vmat, d_vmats, d_rp, d_out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, d_out, sizeof(double) * 3 * R, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return VX3MI_OK;
    }

    int recommended(double *out) override {
        k_recommended_dt<T><<<R, 256, 0, stream>>>(d.lends, d.lmeta, d.vmat, d_vmats, d_lmats, d_vmatE, d_rp, d_out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, d_out, sizeof(double) * R, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return VX3MI_OK;
    }
};

template <typename T> BatchBase *make_batch(vx3mi_model *const *models, int n, int *rc) {
    Batch<T> *b = new (std::nothrow) Batch<T>();
    if (!b) { *rc = VX3MI_E_NOMEM; set_error("out of host memory"); return nullptr; }
    b->precision = (int)(8 * sizeof(T));
    try {
        *rc = b->create(models, n);
    } catch (const std::bad_alloc &) {
        *rc = VX3MI_E_NOMEM;
        set_error("out of host memory");
    }
    if (*rc != VX3MI_OK) { delete b; return nullptr; }
    return b;
}

}  // namespace vx3mi
