// vx3mi_diff_impl.h -- GPU kernels and host driver of the differentiable
// timestep (instantiated for float and double by vx3mi_diff_fp32.hip /
// vx3mi_diff_fp64.hip).  The per-item work is in vx3mi_diff_steps.h.
//
// Forward: one workgroup per robot runs many steps (link phase, barrier, voxel
// phase, barrier), writing a checkpoint of the full state every K steps.
// Backward, per segment of K steps from the last: the segment is recomputed
// from its checkpoint into a tape (state entering each step and the link force
// sums of each voxel), then swept in reverse.  The recompute of segment s-1
// runs on a second stream into the other of two tapes while segment s is swept
// (the two touch disjoint arrays: forward state vs adjoint state), so the
// backward costs about one reverse sweep plus one segment's recompute.
// A reverse step is two
// barrier-separated phases:
//   voxel phase  gather the link adjoints of step n+1 (deterministic, <=6 per
//                voxel, direction order), then the voxel adjoint of step n;
//   link phase   recompute the link forward of step n from the tape and apply
//                its adjoint; each link writes its end contributions to its
//                own slots (no atomics).
// When the whole run's tape fits the memory budget, the forward writes it and
// the backward needs no recompute.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vx3mi_batch.h"
#include "vx3mi_diff.h"
#include "vx3mi_diff_batch.h"
#include "vx3mi_diff_steps.h"

namespace vx3mi {

template <typename T> __global__ __launch_bounds__(256) void k_diff_init(DD<T> d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.V) diff_init_voxel(d, i);
    if (i < d.L) diff_init_link(d, i);
}

// nsteps steps from step n0 with one workgroup per robot; a checkpoint every
// ckEvery steps and/or a tape of the steps [tapeStart, tapeStart + nsteps)
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_diff_forward(DD<T> d, const RobotParams *rps, int n0, int nsteps, int ckEvery,
                                                      T4<T> *ckV, T4<T> *ckL, T4<T> *tapeV, T4<T> *tapeL, int tapeStart) {
    const RobotParams rp = rps[blockIdx.x];
    const int vb = rp.voff, ve = rp.voff + rp.vcount, lb = rp.loff, le = rp.loff + rp.lcount;
    for (int s = 0; s < nsteps; s++) {
        const int n = n0 + s;
        const size_t ts = (size_t)(n - tapeStart);
        if (ckV && n % ckEvery == 0) {
            const size_t slot = (size_t)(n / ckEvery);
            for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) diff_save_voxel(d, v, ckV + (slot * d.V + v) * 4);
            for (int l = lb + threadIdx.x; l < le; l += blockDim.x) diff_save_link(d, l, ckL + (slot * d.L + l) * 3);
        }
        // the link state entering the step, saved by the thread that updates it
        if (tapeL)
            for (int l = lb + threadIdx.x; l < le; l += blockDim.x) diff_save_link(d, l, tapeL + (ts * d.L + l) * 3);
        const T prevDt = n > 0 ? d.dt : (T)0;
        for (int l = lb + threadIdx.x; l < le; l += blockDim.x) diff_link_fwd(d, l, prevDt);
        __syncthreads();
        const double tNext = d.times[n + 1];
        for (int v = vb + threadIdx.x; v < ve; v += blockDim.x)
            diff_voxel_fwd(d, v, tNext, tapeV ? tapeV + (ts * d.V + v) * 5 : (T4<T> *)nullptr);
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_diff_restore(DD<T> d, const T4<T> *ckV, const T4<T> *ckL, int slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.V) diff_restore_voxel(d, i, ckV + ((size_t)slot * d.V + i) * 4);
    if (i < d.L) diff_restore_link(d, i, ckL + ((size_t)slot * d.L + i) * 3);
}

// reverse sweep of the steps [segStart, segStart + segLen) from their tape
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_diff_reverse(DD<T> d, const RobotParams *rps, int segStart, int segLen,
                                                      int gatherFirst, const T4<T> *tapeV, const T4<T> *tapeL) {
    const RobotParams rp = rps[blockIdx.x];
    const int vb = rp.voff, ve = rp.voff + rp.vcount, lb = rp.loff, le = rp.loff + rp.lcount;
    for (int n = segStart + segLen - 1; n >= segStart; n--) {
        const size_t ts = (size_t)(n - segStart);
        const bool gather = n < segStart + segLen - 1 || gatherFirst;
        for (int v = vb + threadIdx.x; v < ve; v += blockDim.x) {
            if (gather) diff_gather_adjoint(d, v, d.times[n + 1]);
            diff_voxel_rev(d, v, tapeV + (ts * d.V + v) * 5);
        }
        __syncthreads();
        const T prevDt = n > 0 ? d.dt : (T)0;
        for (int l = lb + threadIdx.x; l < le; l += blockDim.x)
            diff_link_rev(d, l, prevDt, tapeV + ts * d.V * 5, tapeL + (ts * d.L + l) * 3);
        __syncthreads();
    }
}

// the same reverse step n over every robot at once, one item per thread, two
// launches in stream order instead of the workgroup barriers: the voxel phase
// (gather of step n + 1, voxel adjoint of step n) ...
template <typename T>
__global__ __launch_bounds__(256) void k_diff_rev_voxels(DD<T> d, int n, int segStart, int gather, const T4<T> *tapeV) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.V) return;
    if (gather) diff_gather_adjoint(d, v, d.times[n + 1]);
    diff_voxel_rev(d, v, tapeV + ((size_t)(n - segStart) * d.V + v) * 5);
}
// ... and the link phase (forward recompute of the link from the tape, its adjoint)
template <typename T>
__global__ __launch_bounds__(256) void k_diff_rev_links(DD<T> d, int n, int segStart, const T4<T> *tapeV,
                                                         const T4<T> *tapeL) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.L) return;
    const size_t ts = (size_t)(n - segStart);
    diff_link_rev(d, l, n > 0 ? d.dt : (T)0, tapeV + ts * d.V * 5, tapeL + (ts * d.L + l) * 3);
}

template <typename T> __global__ __launch_bounds__(256) void k_diff_finish(DD<T> d, T *gE) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < d.V) gE[v] = diff_finish_voxel(d, v);
}

// ------------------------------------------------------------------- host
// threads per robot workgroup: as many as the registers of each kernel allow
// (forward FP32 84 VGPRs, FP64 ~230; reverse FP32 166, FP64 ~370 with
// accumulation registers; the optional wide FP32 reverse holds 128 with a
// small spill).  The recompute beside the reverse sweep uses fewer
// threads so that, in FP32, two of its waves and two reverse waves share a SIMD
// (2 x 88 + 2 x 168 = 512 VGPRs).  The thread count does not change results:
// every item is computed by one thread between barriers.
template <typename T> struct DiffThreads;
template <> struct DiffThreads<float> { static constexpr int FWD = 1024, RECOMP = 512, REV = 512, REV_WIDE = 1024; };
template <> struct DiffThreads<double> { static constexpr int FWD = 512, RECOMP = 512, REV = 256, REV_WIDE = 256; };

template <typename T> struct DiffBatch : DiffBase {
    DD<T> d{};
    RobotParams *d_rp = nullptr;
    DVox<T> *d_vc = nullptr;
    DLink<T> *d_lc = nullptr;
    double *d_phase = nullptr, *d_times = nullptr;
    T *d_E = nullptr, *d_gE = nullptr;
    T4<T> *ckV = nullptr, *ckL = nullptr, *tapeV = nullptr, *tapeL = nullptr;
    T4<T> *tapeV2 = nullptr, *tapeL2 = nullptr;  // second tape: recompute beside the reverse sweep
    int ckSlots = 0, tapeSlots = 0, tape2Slots = 0, timesCap = 0;
    std::vector<VoxelMat> hvm;           // material of each voxel
    T4<T> *fin = nullptr;                // final state of the last forward (su, sq, sp, sh)
    std::vector<int> hiz, hvl;           // lattice z; 6 link ids per voxel
    std::vector<int2v> hends;
    std::vector<int> haxis;
    std::vector<double> hnom;            // link nominal length

    int create(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts) {
        R = n;
        optsIn = opts;
        std::vector<int> vl;
        std::vector<int2> ends;
        hrp.resize(n);
        double maxF2 = 0;
        double dtFrac = 0.9;
        const Model &m0 = models[0]->m;
        for (int r = 0; r < n; r++) {
            const Model &m = models[r]->m;
            if (m.varyTemp != m0.varyTemp || m.tempAmplitude != m0.tempAmplitude || m.tempPeriod != m0.tempPeriod ||
                m.enableExpansion != m0.enableExpansion || m.floorEnabled != m0.floorEnabled) {
                set_error("differentiable batches need one actuation and floor setting for every robot");
                return VX3MI_E_ARG;
            }
            if (!m.forceField[0].empty() || !m.forceField[1].empty() || !m.forceField[2].empty()) {
                set_error("differentiable mode: force fields are not supported");
                return VX3MI_E_ARG;
            }
            for (int v = 0; v < m.nv(); v++)
                if (m.ext_of(v) >= 0) {
                    set_error("differentiable mode: voxel externals (boundary conditions) are not supported");
                    return VX3MI_E_ARG;
                }
            bool removal = false;
            for (const vx3mi_material &pm : m.palette) removal |= m.secondaryExperiment && pm.remove_after > 0;
            if (m.enableCilia || removal || m.enableSignals || m.enableDetach || m.enableAttach) {
                set_error("differentiable mode: cilia, signals, detach, attach and material removal are not supported");
                return VX3MI_E_ARG;
            }
            RobotParams &p = hrp[r];
            std::memset(&p, 0, sizeof(p));
            p.voff = (int)hvm.size();
            p.vcount = m.nv();
            p.loff = (int)hends.size();
            p.lcount = m.nl();
            for (int v = 0; v < m.nv(); v++) {
                const VoxelMat &vm = m.vmats[m.vmat[v]];
                if (vm.thermalOnAfter > 0.0) {
                    set_error("differentiable mode: TurnOnThermalExpansionAfterThisManySeconds is not supported");
                    return VX3MI_E_ARG;
                }
                hvm.push_back(vm);
                hiz.push_back(m.iz[v]);
                hlat.push_back(m.ix[v] | (m.iy[v] << 10) | (m.iz[v] << 20));
                hphase0.push_back(m.phase[v]);
                hE0.push_back(m.vmatE[m.vmat[v]]);
                for (int k = 0; k < 6; k++) {
                    const int l = m.vlinks[(size_t)v * 6 + k];
                    hvl.push_back(l < 0 ? -1 : p.loff + l);
                }
            }
            for (int l = 0; l < m.nl(); l++) {
                if (m.lmats[m.lmat[l]].nu != 0.0f) {
                    set_error("differentiable mode: volume effects (Poisson's ratio) are not supported");
                    return VX3MI_E_ARG;
                }
                if (m.lmats[m.lmat[l]].curve) {
                    set_error("differentiable mode: non-linear (bilinear) materials are not supported");
                    return VX3MI_E_ARG;
                }
                hends.push_back(int2v{p.voff + m.lneg[l], p.voff + m.lpos[l]});
                haxis.push_back(m.laxis[l]);
                hnom.push_back(m.lmats[m.lmat[l]].nomSize);
                const double m1 = m.vmats[m.vmat[m.lneg[l]]].mass, m2 = m.vmats[m.vmat[m.lpos[l]]].mass;
                const double k = (double)(float)(opts.e_max * m.lmats[m.lmat[l]].nomSize);
                maxF2 = std::max(maxF2, k / std::min(m1, m2));
            }
            if (m.nl() == 0)
                for (int v = 0; v < m.nv(); v++) {
                    const VoxelMat &vm = m.vmats[m.vmat[v]];
                    maxF2 = std::max(maxF2, opts.e_max * vm.nomSize / vm.mass);
                }
            dtFrac = m.dtFrac;
        }
        V = (int)hvm.size();
        L = (int)hends.size();
        o.floorK = opts.floor_stiffness;
        o.floorW = opts.floor_width;
        o.fricV = opts.friction_speed;
        o.floorEnabled = m0.floorEnabled;
        o.enableExpansion = m0.enableExpansion;
        o.tempAmp = m0.tempAmplitude;
        o.tempPeriod = m0.tempPeriod;
        o.varyTemp = m0.varyTemp;
        if (opts.dt > 0) dt = (float)opts.dt;
        else dt = (float)(dtFrac * (maxF2 > 0 ? 1.0f / (6.283185f * std::sqrt(maxF2)) : 1e-10));

        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&rstream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        int rc;
        int *vld;
        int2v *endsd;
        if ((rc = alloc(&d_rp, R)) || (rc = alloc(&vld, (size_t)6 * V)) || (rc = alloc(&endsd, L)) ||
            (rc = alloc(&d_vc, V)) || (rc = alloc(&d_lc, L)) || (rc = alloc(&d_phase, V)) || (rc = alloc(&d_E, V)) ||
            (rc = alloc(&d_gE, V)))
            return rc;
        T4<T> *arr[27];
        const size_t cnt[27] = {(size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)L, (size_t)L, (size_t)L,
                                (size_t)L, (size_t)L, (size_t)L, (size_t)V, (size_t)V, (size_t)V, (size_t)V,
                                (size_t)V, (size_t)V, (size_t)L, (size_t)L, (size_t)L, (size_t)L, (size_t)L,
                                (size_t)L, (size_t)L, 1, 1, 1, 1};
        for (int k = 0; k < 23; k++)
            if ((rc = alloc(&arr[k], cnt[k]))) return rc;
        T *gph;
        if ((rc = alloc(&gph, V)) || (rc = alloc(&fin, (size_t)4 * V))) return rc;
        HIP_TRY(hipMemcpy(d_rp, hrp.data(), sizeof(RobotParams) * R, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(vld, hvl.data(), sizeof(int) * 6 * V, hipMemcpyHostToDevice));
        if (L) HIP_TRY(hipMemcpy(endsd, hends.data(), sizeof(int2v) * L, hipMemcpyHostToDevice));
        d.vlinks = vld; d.lends = endsd; d.vc = d_vc; d.lc = d_lc; d.phase = d_phase; d.E = d_E;
        d.o = o; d.dt = (T)dt; d.V = V; d.L = L;
        d.su = arr[0]; d.sq = arr[1]; d.sp = arr[2]; d.sh = arr[3];
        d.sl0 = arr[4]; d.sl1 = arr[5]; d.sl2 = arr[6];
        d.lf0 = arr[7]; d.lf1 = arr[8]; d.lf2 = arr[9];
        d.au = arr[10]; d.aq = arr[11]; d.ap = arr[12]; d.ah = arr[13]; d.lamF0 = arr[14]; d.lamF1 = arr[15];
        d.al0 = arr[16]; d.al1 = arr[17]; d.al2 = arr[18];
        d.c0 = arr[19]; d.c1 = arr[20]; d.c2 = arr[21]; d.c3 = arr[22];
        d.gPhase = gph;
        return set_params(hphase0.data(), hE0.data());
    }

    int set_params(const double *phase, const double *E) override {
        std::vector<DVox<T>> vc((size_t)V);
        std::vector<DLink<T>> lc((size_t)std::max(L, 1));
        std::vector<T> Ev((size_t)V);
        for (int v = 0; v < V; v++) {
            if (!(E[v] > 0)) {
                set_error("Young's modulus must be positive");
                return VX3MI_E_ARG;
            }
            Ev[v] = (T)E[v];
        }
        hphase.assign(phase, phase + V);
        hE.assign(E, E + V);
        diff_constants<T>(hvm, hends.data(), haxis.data(), hnom.data(), hiz.data(), V, L, E, vc.data(), lc.data());
        HIP_TRY(hipMemcpy(d_vc, vc.data(), sizeof(DVox<T>) * V, hipMemcpyHostToDevice));
        if (L) HIP_TRY(hipMemcpy(d_lc, lc.data(), sizeof(DLink<T>) * L, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_phase, phase, sizeof(double) * V, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_E, Ev.data(), sizeof(T) * V, hipMemcpyHostToDevice));
        steps = 0;
        return VX3MI_OK;
    }

    size_t bytes_ck_slot() const override { return sizeof(T4<T>) * ((size_t)4 * V + (size_t)3 * L); }
    size_t bytes_tape_slot() const override { return sizeof(T4<T>) * ((size_t)5 * V + (size_t)3 * L); }

    int ensure_buffers(int nsteps) {
        // checkpoint interval: as given, else the whole run if its tape fits the
        // budget (no recompute), else the interval minimising checkpoint + tape
        // bytes (two tapes once there are several segments)
        int kk = optsIn.checkpoint_every;
        if (kk <= 0) {
            const double budget = optsIn.hbm_budget;
            const double full = (double)bytes_tape_slot() * nsteps + (double)bytes_ck_slot();
            if (full <= budget) kk = nsteps;
            else {
                kk = (int)std::ceil(std::sqrt((double)nsteps * bytes_ck_slot() / (2.0 * bytes_tape_slot())));
                kk = std::max(1, std::min(kk, nsteps));
                const double need =
                    (double)bytes_ck_slot() * ((nsteps + kk - 1) / kk) + 2.0 * (double)bytes_tape_slot() * kk;
                if (need > budget) {
                    set_error("checkpoints and tape of this run exceed the memory budget");
                    return VX3MI_E_LIMIT;
                }
            }
        }
        K = std::max(1, std::min(kk, nsteps));
        const int needCk = (nsteps + K - 1) / K, needTape = K, needTape2 = needCk > 1 ? K : 0;
        int rc;
        if (needCk > ckSlots) {
            if (ckV) { free_one(ckV); free_one(ckL); }
            if ((rc = alloc(&ckV, (size_t)needCk * 4 * V)) || (rc = alloc(&ckL, (size_t)needCk * 3 * std::max(L, 1))))
                return rc;
            ckSlots = needCk;
        }
        if (needTape > tapeSlots) {
            if (tapeV) { free_one(tapeV); free_one(tapeL); }
            if ((rc = alloc(&tapeV, (size_t)needTape * 5 * V)) ||
                (rc = alloc(&tapeL, (size_t)needTape * 3 * std::max(L, 1))))
                return rc;
            tapeSlots = needTape;
        }
        if (needTape2 > tape2Slots) {
            if (tapeV2) { free_one(tapeV2); free_one(tapeL2); }
            if ((rc = alloc(&tapeV2, (size_t)needTape2 * 5 * V)) ||
                (rc = alloc(&tapeL2, (size_t)needTape2 * 3 * std::max(L, 1))))
                return rc;
            tape2Slots = needTape2;
        }
        if (nsteps + 1 > timesCap) {
            if (d_times) free_one(d_times);
            if ((rc = alloc(&d_times, (size_t)nsteps + 1))) return rc;
            timesCap = nsteps + 1;
        }
        ckBytes = (double)bytes_ck_slot() * needCk;
        tapeBytes = (double)bytes_tape_slot() * (needTape + needTape2);
        return VX3MI_OK;
    }

    int forward(int nsteps) override {
        if (nsteps <= 0) { set_error("nsteps must be positive"); return VX3MI_E_ARG; }
        int rc = ensure_buffers(nsteps);
        if (rc) return rc;
        std::vector<double> times((size_t)nsteps + 1);
        times[0] = 0.0;
        for (int n = 0; n < nsteps; n++) times[n + 1] = times[n] + (double)dt;  // currentTime += dt
        HIP_TRY(hipMemcpyAsync(d_times, times.data(), sizeof(double) * (nsteps + 1), hipMemcpyHostToDevice, stream));
        d.times = d_times;
        HIP_TRY(hipEventRecord(ev0, stream));
        const int nb = (std::max(V, L) + 255) / 256;
        k_diff_init<T><<<std::max(nb, 1), 256, 0, stream>>>(d);
        // the forward keeps checkpoints; with one segment it also writes the tape
        const bool oneSeg = K >= nsteps;
        for (int s0 = 0; s0 < nsteps; s0 += chunk) {
            const int k = std::min(chunk, nsteps - s0);
            k_diff_forward<T, DiffThreads<T>::FWD><<<R, DiffThreads<T>::FWD, 0, stream>>>(
                d, d_rp, s0, k, K, ckV, ckL, oneSeg ? tapeV : nullptr, oneSeg ? tapeL : nullptr, 0);
        }
        HIP_TRY(hipEventRecord(ev1, stream));
        T4<T> *src[4] = {d.su, d.sq, d.sp, d.sh};
        for (int k = 0; k < 4; k++)
            HIP_TRY(hipMemcpyAsync(fin + (size_t)k * V, src[k], sizeof(T4<T>) * V, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventElapsedTime(&fwdMs, ev0, ev1));
        steps = nsteps;
        tapeValid = oneSeg;
        return VX3MI_OK;
    }

    int get_state(int robot, double *pos, double *quat, double *lm, double *am) override {
        int v0 = 0, nv = V;
        if (robot >= 0) {
            if (robot >= R) { set_error("robot index out of range"); return VX3MI_E_ARG; }
            v0 = hrp[robot].voff;
            nv = hrp[robot].vcount;
        }
        std::vector<T4<T>> a((size_t)nv), b((size_t)nv), c((size_t)nv), e((size_t)nv);
        if (steps <= 0) { set_error("no forward run yet"); return VX3MI_E_ARG; }
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(a.data(), fin + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b.data(), fin + (size_t)V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(c.data(), fin + (size_t)2 * V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(e.data(), fin + (size_t)3 * V + v0, sizeof(T4<T>) * nv, hipMemcpyDeviceToHost));
        for (int i = 0; i < nv; i++) {
            const int v = v0 + i, pk = hlat[v];
            const double Lv = hvm[v].nomSize;
            if (pos) {
                pos[3 * i] = (pk & 1023) * Lv + (double)a[i].x;
                pos[3 * i + 1] = ((pk >> 10) & 1023) * Lv + (double)a[i].y;
                pos[3 * i + 2] = (pk >> 20) * Lv + (double)a[i].z;
            }
            if (quat) { quat[4 * i] = b[i].x; quat[4 * i + 1] = b[i].y; quat[4 * i + 2] = b[i].z; quat[4 * i + 3] = b[i].w; }
            if (lm) { lm[3 * i] = c[i].x; lm[3 * i + 1] = c[i].y; lm[3 * i + 2] = c[i].z; }
            if (am) { am[3 * i] = e[i].x; am[3 * i + 1] = e[i].y; am[3 * i + 2] = e[i].z; }
        }
        return VX3MI_OK;
    }

    int backward(const double *gpos, const double *gquat, const double *glm, const double *gam, double *gphase,
                 double *gE) override {
        if (steps <= 0) { set_error("backward needs a forward run first"); return VX3MI_E_ARG; }
        // adjoint of the final state
        std::vector<T4<T>> au((size_t)V, T4<T>{0, 0, 0, 0}), aq = au, ap = au, ah = au;
        for (int v = 0; v < V; v++) {
            if (gpos) au[v] = T4<T>{(T)gpos[3 * v], (T)gpos[3 * v + 1], (T)gpos[3 * v + 2], 0};
            if (gquat) aq[v] = T4<T>{(T)gquat[4 * v], (T)gquat[4 * v + 1], (T)gquat[4 * v + 2], (T)gquat[4 * v + 3]};
            if (glm) ap[v] = T4<T>{(T)glm[3 * v], (T)glm[3 * v + 1], (T)glm[3 * v + 2], 0};
            if (gam) ah[v] = T4<T>{(T)gam[3 * v], (T)gam[3 * v + 1], (T)gam[3 * v + 2], 0};
        }
        HIP_TRY(hipEventRecord(ev0, stream));
        HIP_TRY(hipMemcpyAsync(d.au, au.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.aq, aq.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.ap, ap.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d.ah, ah.data(), sizeof(T4<T>) * V, hipMemcpyHostToDevice, stream));
        if (L) {
            HIP_TRY(hipMemsetAsync(d.al0, 0, sizeof(T4<T>) * L, stream));
            HIP_TRY(hipMemsetAsync(d.al1, 0, sizeof(T4<T>) * L, stream));
            HIP_TRY(hipMemsetAsync(d.al2, 0, sizeof(T4<T>) * L, stream));
        }
        HIP_TRY(hipMemsetAsync(d.gPhase, 0, sizeof(T) * V, stream));
        const int nseg = (steps + K - 1) / K;
        const int nb = (std::max(V, L) + 255) / 256;
        while ((int)segEv.size() < 4 * nseg) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            segEv.push_back(e);
        }
        // segment s uses tape s % 2; its recompute (on rstream) waits until the
        // reverse sweep of segment s + 2, the last reader of that tape, is done
        T4<T> *bufV[2] = {tapeV, tapeV2}, *bufL[2] = {tapeL, tapeL2};
        const bool recompute = !(tapeValid && nseg == 1);
        // flat: every robot's reverse step at once over the whole device (a robot
        // larger than the others no longer sets the pace of the sweep)
        const bool flat = optsIn.reverse_threads < 0;
        const bool wide = flat || (optsIn.reverse_threads == DiffThreads<T>::REV_WIDE &&
                                   DiffThreads<T>::REV_WIDE != DiffThreads<T>::REV);
        hipStream_t rs = rstream;
        HIP_TRY(hipStreamWaitEvent(rs, ev0, 0));
        auto recompute_seg = [&](int seg) -> int {
            const int s0 = seg * K, len = std::min(K, steps - s0);
            if (seg + 2 < nseg) HIP_TRY(hipStreamWaitEvent(rs, segEv[4 * (seg + 2) + 3], 0));
            HIP_TRY(hipEventRecord(segEv[4 * seg], rs));
            k_diff_restore<T><<<std::max(nb, 1), 256, 0, rs>>>(d, ckV, ckL, seg);
            for (int c = 0; c < len; c += chunk) {
                const int k = std::min(chunk, len - c);
                if (wide)  // a wide reverse fills the compute units: recompute at full width too
                    k_diff_forward<T, DiffThreads<T>::FWD><<<R, DiffThreads<T>::FWD, 0, rs>>>(
                        d, d_rp, s0 + c, k, K, nullptr, nullptr, bufV[seg & 1], bufL[seg & 1], s0);
                else
                    k_diff_forward<T, DiffThreads<T>::RECOMP><<<R, DiffThreads<T>::RECOMP, 0, rs>>>(
                        d, d_rp, s0 + c, k, K, nullptr, nullptr, bufV[seg & 1], bufL[seg & 1], s0);
            }
            HIP_TRY(hipEventRecord(segEv[4 * seg + 1], rs));
            return VX3MI_OK;
        };
        int rc;
        if (recompute && (rc = recompute_seg(nseg - 1))) return rc;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int s0 = seg * K, len = std::min(K, steps - s0);
            if (recompute) HIP_TRY(hipStreamWaitEvent(stream, segEv[4 * seg + 1], 0));
            HIP_TRY(hipEventRecord(segEv[4 * seg + 2], stream));
            const int gatherFirst = seg != nseg - 1;
            if (flat) {
                for (int n = s0 + len - 1; n >= s0; n--) {
                    const int gather = n < s0 + len - 1 || gatherFirst;
                    k_diff_rev_voxels<T><<<(V + 255) / 256, 256, 0, stream>>>(d, n, s0, gather, bufV[seg & 1]);
                    if (L) k_diff_rev_links<T><<<(L + 255) / 256, 256, 0, stream>>>(d, n, s0, bufV[seg & 1], bufL[seg & 1]);
                }
            } else if (wide)
                k_diff_reverse<T, DiffThreads<T>::REV_WIDE><<<R, DiffThreads<T>::REV_WIDE, 0, stream>>>(
                    d, d_rp, s0, len, gatherFirst, bufV[seg & 1], bufL[seg & 1]);
            else
                k_diff_reverse<T, DiffThreads<T>::REV><<<R, DiffThreads<T>::REV, 0, stream>>>(
                    d, d_rp, s0, len, gatherFirst, bufV[seg & 1], bufL[seg & 1]);
            HIP_TRY(hipEventRecord(segEv[4 * seg + 3], stream));
            if (recompute && seg > 0 && (rc = recompute_seg(seg - 1))) return rc;
        }
        k_diff_finish<T><<<(V + 255) / 256, 256, 0, stream>>>(d, d_gE);
        HIP_TRY(hipEventRecord(ev1, stream));
        std::vector<T> gp((size_t)V), ge((size_t)V);
        HIP_TRY(hipMemcpyAsync(gp.data(), d.gPhase, sizeof(T) * V, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(ge.data(), d_gE, sizeof(T) * V, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventElapsedTime(&bwdMs, ev0, ev1));
        recomputeMs = reverseMs = 0;
        for (int seg = 0; seg < nseg; seg++) {
            float a = 0, b = 0;
            if (recompute) HIP_TRY(hipEventElapsedTime(&a, segEv[4 * seg], segEv[4 * seg + 1]));
            HIP_TRY(hipEventElapsedTime(&b, segEv[4 * seg + 2], segEv[4 * seg + 3]));
            recomputeMs += a;
            reverseMs += b;
        }
        for (int v = 0; v < V; v++) {
            if (gphase) gphase[v] = (double)gp[v];
            if (gE) gE[v] = (double)ge[v];
        }
        tapeValid = tapeValid && nseg == 1;  // a recompute leaves the tape of segment 0
        return VX3MI_OK;
    }
};

template <typename T> DiffBase *make_diff(vx3mi_model *const *models, int n, const vx3mi_diff_opts &opts, int *rc) {
    DiffBatch<T> *b = new (std::nothrow) DiffBatch<T>();
    if (!b) { *rc = VX3MI_E_NOMEM; set_error("out of host memory"); return nullptr; }
    b->precision = sizeof(T) == 4 ? 32 : 64;
    *rc = b->create(models, n, opts);
    if (*rc != VX3MI_OK) { delete b; return nullptr; }
    return b;
}

}  // namespace vx3mi
