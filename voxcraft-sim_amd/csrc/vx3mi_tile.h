// vx3mi_tile.h -- the tiled persistent timestep for batches of small robots.
//
// Robots are bin-packed (longest processing time first, by link count) into
// tiles of at most LMAX links and VMAX voxels; one workgroup of NT threads owns
// a tile for a whole launch of many steps.  Each thread owns up to KL links and
// KV voxels and keeps their state in registers across steps; the only data
// exchanged between threads are the voxel poses (position field, quaternion,
// temperature) and the link-end forces, both in LDS as structure of arrays.
// Global memory is read at the start of a launch and written at its end.
// The per-link and per-voxel arithmetic is the same as k_links / k_voxels
// (link_forces / voxel_step, link ends summed in direction order), so every
// launch mode produces the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace vx3mi {

// Tile configurations: NT threads, up to KL links and KV voxels per thread,
// MINW the minimum waves per SIMD the register budget must allow (the second
// __launch_bounds__ argument).  The large tile
// uses ~158 KB of LDS (one workgroup per compute unit); the FP32 one runs 768
// threads, 12 waves, so each SIMD holds three waves of at most 168 VGPRs (the
// kernel is VALU-issue bound, and a third wave hides the LDS and barrier
// waits: +33 % over 512 threads holding four links each at two waves).  The
// small tile stays within 78 KB of LDS, so two workgroups fit a compute unit's
// 160 KB (one computes while the other waits at a barrier), 384 threads each,
// again three waves per SIMD.  The batch's autotuner picks one.
template <typename T, int SMALL> struct TileCfg;
template <> struct TileCfg<float, 0> {
    static constexpr int NT = 768, KL = 3, KV = 2, LMAX = 2048, VMAX = 1536, RMAX = 128, MINW = 1;
};
template <> struct TileCfg<float, 1> {
    static constexpr int NT = 384, KL = 3, KV = 2, LMAX = 960, VMAX = 768, RMAX = 64, MINW = 3;
};
template <> struct TileCfg<double, 0> {
    static constexpr int NT = 512, KL = 2, KV = 2, LMAX = 1024, VMAX = 768, RMAX = 64, MINW = 2;
};
template <> struct TileCfg<double, 1> {
    static constexpr int NT = 256, KL = 2, KV = 2, LMAX = 512, VMAX = 384, RMAX = 32, MINW = 2;
};

// Collision tiles (large configuration): the partner momenta of the contact
// phase sit in LDS beside the poses, so the capacities are smaller (every
// robot of the batched workload still fits: at most 1,861 links, 1,077 voxels).
template <typename T> struct TileColCfg;
template <> struct TileColCfg<float> {
    static constexpr int NT = 768, KL = 3, KV = 2, LMAX = 1920, VMAX = 1248, RMAX = 128, MINW = 1;
};
template <> struct TileColCfg<double> {
    static constexpr int NT = 512, KL = 2, KV = 2, LMAX = 896, VMAX = 640, RMAX = 64, MINW = 2;
};
template <typename T, int SMALL, int COL> struct TileSel { using C = TileCfg<T, SMALL>; };
template <typename T, int SMALL> struct TileSel<T, SMALL, 1> { using C = TileColCfg<T>; };

// LDS robot table of a tile
template <int RMAX> struct TileRobots {
    int rid[RMAX];
    int voffG[RMAX], loffG[RMAX];
    int vpre[RMAX + 1], lpre[RMAX + 1];
    double time[2][RMAX];
    float prevDt[2][RMAX];
    float dt[RMAX];
    int diverged[RMAX];
    int steps[RMAX];
    double tempAmp[RMAX], tempPeriod[RMAX];  // actuation of each robot (FP32: 1 / period)
    int tempFlags[RMAX];                     // varyTemp | enableExpansion << 1
    int col[RMAX];                           // the robot has collisions (surface voxels in the lists)
    int spre[RMAX + 1];                      // prefix of the robots' surface voxel counts (collision tiles)
    int anyStale;                            // some robot of the tile lists its candidates this step
};

// material tables of at most this many entries (large / small tile) are
// copied into LDS at launch; the small tile keeps within half the LDS
template <int SMALL> struct TileMatCache { static constexpr int N = SMALL ? 12 : 24; };

// largest j with pre[j] <= x (pre ascending, pre[0] = 0, pre[n] = total)
__device__ __forceinline__ int tile_slot(const int *pre, int n, int x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// LDSMAT: the batch has at most kTileMatCache voxel and link materials, which
// are read from LDS in the step loop instead of global memory
// COL: collision tiles.  Each step, robots whose candidate lists are stale
// list them first (the tile publishes its poses to global memory and runs the
// listing passes of the flat path); the contact forces are computed at the
// start of the voxel phase, from the poses and momenta entering the step (LDS)
// and the lists, summed in list order as in k_col_contact.
template <typename T, int SMALL, int LDSMAT, int COL>
__global__ __launch_bounds__((TileSel<T, SMALL, COL>::C::NT), (TileSel<T, SMALL, COL>::C::MINW)) void k_tile(
    Dev<T> a, const int4 *tiles, const int *tileRobots, int nsteps, int parity0, int nLinkMats, int nVoxMats) {
    using C = typename TileSel<T, SMALL, COL>::C;
    constexpr int NT = C::NT, KL = C::KL, KV = C::KV;
    constexpr int LMAX = C::LMAX, VMAX = C::VMAX, RMAX = C::RMAX;
    static_assert(KL * NT >= LMAX && KV * NT >= VMAX && RMAX <= NT && NT <= 3 * LMAX,
                  "a tile's threads must cover its capacity");
    __shared__ T sPose[8][VMAX];    // position field xyz, qw, qx, qy, qz, temperature
    __shared__ T sForce[12][LMAX];  // forceNeg, momentNeg, forcePos, momentPos
    __shared__ T sMom[COL ? 3 : 1][COL ? VMAX : 1];  // linear momentum entering the step (contacts)
    __shared__ TileRobots<RMAX> tr;
    __shared__ LinkMat sLM[LDSMAT ? TileMatCache<SMALL>::N : 1];
    __shared__ VoxelMat sVM[LDSMAT ? TileMatCache<SMALL>::N : 1];

    const int tid = threadIdx.x;
    const int4 info = tiles[blockIdx.x];  // robot begin, robot count, voxels, links
    const int nr = info.y, nv = info.z, nl = info.w;

    if (tid < nr) {
        const int r = tileRobots[info.x + tid];
        const RobotParams &p = a.rp[r];
        const RobotTime &t = a.rt[r];
        tr.rid[tid] = r;
        tr.voffG[tid] = p.voff;
        tr.loffG[tid] = p.loff;
        tr.vpre[tid + 1] = p.vcount;
        tr.lpre[tid + 1] = p.lcount;
        tr.time[0][tid] = t.time[0];
        tr.time[1][tid] = t.time[1];
        tr.prevDt[0][tid] = t.prevDt[0];
        tr.prevDt[1][tid] = t.prevDt[1];
        tr.dt[tid] = t.dt;
        tr.diverged[tid] = t.diverged | t.divergedNow;
        tr.steps[tid] = t.steps;
        tr.tempAmp[tid] = p.tempAmp;
        tr.tempPeriod[tid] = std::is_same<T, float>::value ? p.tempInvPeriod : p.tempPeriod;
        tr.tempFlags[tid] = (p.varyTemp ? 1 : 0) | (p.enableExpansion ? 2 : 0);
        tr.col[tid] = COL && p.collisions && p.scount > 0;
        tr.spre[tid + 1] = COL ? p.scount : 0;
    }
    if (LDSMAT) {
        for (int i = tid; i < nLinkMats; i += NT) sLM[i] = a.lmats[i];
        for (int i = tid; i < nVoxMats; i += NT) sVM[i] = a.vmats[i];
    }
    __syncthreads();
    if (tid == 0) {
        tr.vpre[0] = 0;
        tr.lpre[0] = 0;
        tr.spre[0] = 0;
        for (int j = 0; j < nr; j++) {
            tr.vpre[j + 1] += tr.vpre[j];
            tr.lpre[j + 1] += tr.lpre[j];
            tr.spre[j + 1] += tr.spre[j];
        }
    }
    __syncthreads();

    // ---- owned links: static description and state in registers
    int lGlob[KL], lEnds[KL], lMeta[KL], lMat[KL];  // global id; neg | pos << 16 (local); axis | slot << 2
    float lStrain[KL];
    LinkState<T> ls[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) {
        const int ll = tid + k * NT;
        lGlob[k] = -1;
        if (ll < nl) {
            const int j = tile_slot(tr.lpre, nr, ll);
            const int gl = tr.loffG[j] + (ll - tr.lpre[j]);
            const int2 e = a.lends[gl];
            const int base = tr.vpre[j] - tr.voffG[j];
            lGlob[k] = gl;
            lEnds[k] = (e.x + base) | ((e.y + base) << 16);
            lMeta[k] = (a.lmeta[gl].x & 3) | (j << 2);
            lMat[k] = a.lmeta[gl].x >> 2;
            const T4<T> s0 = a.ls0[gl], s1 = a.ls1[gl], s2 = a.ls2[gl];
            lStrain[k] = (float)s2.w;
            ls[k].pos2 = {s0.x, s0.y, s0.z};
            ls[k].a1v = {s0.w, s1.x, s1.y};
            ls[k].a2v = {s1.z, s1.w, s2.x};
            ls[k].maxStrain = (float)s2.y;
            ls[k].flags = (int)s2.z;
        }
    }
    // ---- owned voxels
    int vGlob[KV], vSlot[KV], vMat[KV], vLat[KV];
    int vLinks[KV][3];  // local link ids, two 16-bit fields each (0xffff = none)
    T vPhase[KV];
    VoxelState<T> vs[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int lv = tid + k * NT;
        vGlob[k] = -1;
        if (lv < nv) {
            const int j = tile_slot(tr.vpre, nr, lv);
            const int gv = tr.voffG[j] + (lv - tr.vpre[j]);
            vGlob[k] = gv;
            vSlot[k] = j;
            vMat[k] = a.vmat[gv];
            vLat[k] = a.lat[gv];
            vPhase[k] = a.phase[gv];
            const int lbase = tr.lpre[j] - tr.loffG[j];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int l0 = a.vlinks[6 * gv + 2 * d], l1 = a.vlinks[6 * gv + 2 * d + 1];
                vLinks[k][d] = (l0 < 0 ? 0xffff : l0 + lbase) | ((l1 < 0 ? 0xffff : l1 + lbase) << 16);
            }
            const T4<T> q0 = a.pq0[gv], q1 = a.pq1[gv], m0 = a.mo0[gv], m1 = a.mo1[gv];
            vs[k].u = {q0.x, q0.y, q0.z};
            vs[k].q = {q0.w, q1.x, q1.y, q1.z};
            vs[k].p = {m0.x, m0.y, m0.z};
            vs[k].h = {m0.w, m1.x, m1.y};
            vs[k].flags = (int)m1.z;
            sPose[0][lv] = q0.x; sPose[1][lv] = q0.y; sPose[2][lv] = q0.z; sPose[3][lv] = q0.w;
            sPose[4][lv] = q1.x; sPose[5][lv] = q1.y; sPose[6][lv] = q1.z; sPose[7][lv] = q1.w;
            if constexpr (COL) { sMom[0][lv] = m0.x; sMom[1][lv] = m0.y; sMom[2][lv] = m0.z; }
        }
    }
    __syncthreads();

    // surface voxel i of the tile (collision robots in tile order): its global
    // surface index and its robot's slot
    auto tile_surface = [&](int i, int &j) {
        j = tile_slot(tr.spre, nr, i);
        return a.rp[tr.rid[j]].soff + (i - tr.spre[j]);
    };
    // the contact force on owned voxel gv (tile index lv, robot slot j), from the
    // listed candidates in contact now: col_contact's partner test and force with
    // the poses and momenta of the tile (the robots of a tile lose neither links
    // nor voxels, and have no signals)
    auto contact = [&](int gv, int lv, int j, int pkv, const VoxelMat &mi, int parity) -> V3<T> {
        V3<T> f = {0, 0, 0};
        const int s = a.vsurf[gv];
        if (s < 0) return f;
        const int r = tr.rid[j];
        RobotTime *rt = a.rt + r;
        const RobotParams &rp = a.rp[r];
        if (s == rp.soff && rt->stale[parity]) {
            rt->listings += 1;
            rt->stale[parity] = 0;
        }
        const int lbase = tr.vpre[j] - tr.voffG[j];  // tile index of a voxel of the robot = global id + lbase
        const T4<T> qi0 = {sPose[0][lv], sPose[1][lv], sPose[2][lv], sPose[3][lv]};
        const T L = (T)mi.nomSize;
        const T bsaI = base_size_average<T>(mi.nomSize, (float)sPose[7][lv], mi.alphaCTE);
        const V3<T> velI = V3<T>{sMom[0][lv], sMom[1][lv], sMom[2][lv]} * (T)mi.massInv;
        const int ti = a.vtarget ? a.vtarget[gv] : 0;
        auto partner = [&](int w, bool links, V3<T> &fw) -> bool {
            const int lw = w + lbase;
            const int3 o = lat_offset(pkv, a.lat[w]);
            if (links && col_linked(a, gv, o.x, o.y, o.z)) return false;
            const VoxelMat &mj = LDSMAT ? sVM[a.vmat[w]] : a.vmats[a.vmat[w]];
            if (mi.fixed && mj.fixed) return false;
            const T4<T> qj0 = {sPose[0][lw], sPose[1][lw], sPose[2][lw], sPose[3][lw]};
            const V3<T> diff = col_diff<T>(qi0, qj0, o, L);
            const T bsaJ = base_size_average<T>(mj.nomSize, (float)sPose[7][lw], mj.alphaCTE);
            if (!in_contact_envelope(diff, bsaI, bsaJ)) return false;
            if (a.vtarget && a.vtarget[w] != ti && gv < w) atomicAdd(&rt->collisionCount, 1);
            if (a.pairRec && gv < w) {
                const int k = atomicAdd(a.pairN, 1);
                if (k < a.pairCap) a.pairRec[k] = int2{gv, w};
            }
            const V3<T> velJ = V3<T>{sMom[0][lw], sMom[1][lw], sMom[2][lw]} * (T)mj.massInv;
            fw = contact_force(diff, bsaI, bsaJ, velI, velJ, mi.penK, mj.penK, mi.cDampT, mj.cDampT);
            return true;
        };
        const int nc = a.ccount[s];
        if (nc <= kMaxCand) {
            const int *list = a.clist + (long)s * kMaxCand;
            for (int k = 0; k < nc; k++) {
                V3<T> fw;
                if (partner(list[k], false, fw)) f = f + fw;
            }
        } else {  // more candidates than a list holds: the grid of this step, ascending
            rt->stale[parity ^ 1] = 1;
            int prev = -1;
            for (;;) {
                int next = INT_MAX;
                col_grid_each(a, rp, s, [&](int sj) {
                    if (sj > prev && sj < next) next = sj;
                });
                if (next == INT_MAX) break;
                V3<T> fw;
                if (partner(a.surf[next], true, fw)) f = f + fw;
                prev = next;
            }
        }
        return f;
    };

    for (int s = 0; s < nsteps; s++) {
        const int parity = (parity0 + s) & 1;
        if constexpr (COL) {
            // ---- candidate lists of the robots whose lists are stale (rare)
            if (tid == 0) tr.anyStale = 0;
            __syncthreads();
            if (tid < nr && tr.col[tid] && !tr.diverged[tid] && a.rt[tr.rid[tid]].stale[parity]) tr.anyStale = 1;
            __syncthreads();
            if (tr.anyStale) {
                auto listing = [&](int j) { return tr.col[j] && !tr.diverged[j] && a.rt[tr.rid[j]].stale[parity]; };
#pragma unroll
                for (int k = 0; k < KV; k++) {  // the listing passes read poses from global memory
                    const int lv = tid + k * NT;
                    if (vGlob[k] < 0 || !listing(vSlot[k])) continue;
                    a.pq0[vGlob[k]] = T4<T>{sPose[0][lv], sPose[1][lv], sPose[2][lv], sPose[3][lv]};
                    a.pq1[vGlob[k]] = T4<T>{sPose[4][lv], sPose[5][lv], sPose[6][lv], sPose[7][lv]};
                }
                __syncthreads();
                for (int i = tid; i < tr.spre[nr]; i += NT) {
                    int j;
                    const int si = tile_surface(i, j);
                    if (listing(j)) col_count(a, si, parity);
                }
                __syncthreads();
                int *wsum = reinterpret_cast<int *>(&sForce[0][0]);  // free until the link phase
                for (int j = 0; j < nr; j++) {
                    if (!listing(j)) continue;  // (the same for every thread)
                    const int r = tr.rid[j];
                    col_scan(a, r, wsum);
                    __syncthreads();
                    const RobotParams &rp = a.rp[r];
                    for (int i = rp.soff + tid; i < rp.soff + rp.scount; i += NT) col_scatter(a, i, parity);
                    __syncthreads();
                }
                for (int i = tid; i < tr.spre[nr]; i += NT) {
                    int j;
                    const int si = tile_surface(i, j);
                    if (listing(j)) col_list(a, si, parity);
                }
                __syncthreads();
            }
        }
        // ---- links
#pragma unroll
        for (int k = 0; k < KL; k++) {
            const int ll = tid + k * NT;
            if (lGlob[k] < 0) continue;
            const int j = lMeta[k] >> 2;
            if (tr.diverged[j]) continue;
            const LinkMat &lm = LDSMAT ? sLM[lMat[k]] : a.lmats[lMat[k]];
            if (lm.fixedBoth) continue;
            const int n = lEnds[k] & 0xffff, p = lEnds[k] >> 16;
            const V3<T> du = {sPose[0][p] - sPose[0][n], sPose[1][p] - sPose[1][n], sPose[2][p] - sPose[2][n]};
            const Q4<T> qn = {sPose[3][n], sPose[4][n], sPose[5][n], sPose[6][n]};
            const Q4<T> qp = {sPose[3][p], sPose[4][p], sPose[5][p], sPose[6][p]};
            LinkOut<T> o;
            link_forces<T>(lm, a.curves, lMeta[k] & 3, du, qn, qp, (float)sPose[7][n], (float)sPose[7][p], tr.prevDt[parity][j],
                           ls[k], o);
            sForce[0][ll] = o.fN.x; sForce[1][ll] = o.fN.y; sForce[2][ll] = o.fN.z;
            sForce[3][ll] = o.mN.x; sForce[4][ll] = o.mN.y; sForce[5][ll] = o.mN.z;
            sForce[6][ll] = o.fP.x; sForce[7][ll] = o.fP.y; sForce[8][ll] = o.fP.z;
            sForce[9][ll] = o.mP.x; sForce[10][ll] = o.mP.y; sForce[11][ll] = o.mP.z;
            lStrain[k] = o.strain;
            if (o.strain > 100.0f) tr.diverged[j] = 1;
        }
        __syncthreads();
        // ---- voxels (contact forces first, from the poses entering the step)
        V3<T> cf[KV];
#pragma unroll
        for (int k = 0; k < KV; k++) {
            cf[k] = V3<T>{0, 0, 0};
            if constexpr (COL) {
                const int lv = tid + k * NT;
                if (vGlob[k] < 0 || !tr.col[vSlot[k]] || tr.diverged[vSlot[k]]) continue;
                const VoxelMat &vm = LDSMAT ? sVM[vMat[k]] : a.vmats[vMat[k]];
                cf[k] = contact(vGlob[k], lv, vSlot[k], vLat[k], vm, parity);
            }
        }
        if constexpr (COL) __syncthreads();
#pragma unroll
        for (int k = 0; k < KV; k++) {
            const int lv = tid + k * NT;
            if (vGlob[k] < 0) continue;
            const int j = vSlot[k];
            if (tr.diverged[j]) continue;
            const float dt = tr.dt[j];
            const double tnext = tr.time[parity][j] + (double)dt;
            if (lv == tr.vpre[j]) {
                tr.time[parity ^ 1][j] = tnext;
                tr.prevDt[parity ^ 1][j] = dt;
                tr.steps[j] += 1;
            }
            const VoxelMat &vm = LDSMAT ? sVM[vMat[k]] : a.vmats[vMat[k]];
            if (vm.fixed) continue;
            V3<T> fs = {0, 0, 0}, ms = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const int l = (vLinks[k][d >> 1] >> (16 * (d & 1))) & 0xffff;
                if (l == 0xffff) continue;
                const int o = (d & 1) ? 6 : 0;  // NEG direction: this voxel is the pVPos end
                fs = fs + V3<T>{sForce[o + 0][l], sForce[o + 1][l], sForce[o + 2][l]};
                ms = ms + V3<T>{sForce[o + 3][l], sForce[o + 4][l], sForce[o + 5][l]};
            }
            float temp = (float)sPose[7][lv];
            const int tf = tr.tempFlags[j];
            voxel_step<T>(vm, a.exts, dt, vLat[k], temp, fs, ms, cf[k], V3<T>{0, 0, 0}, vs[k]);
            if ((tf & 1) && !(vm.thermalOnAfter > tnext)) {
                if constexpr (std::is_same<T, float>::value)
                    temp = actuation_temperature_f32(tr.tempAmp[j], tr.tempPeriod[j], tnext, (double)vPhase[k], tf >> 1);
                else
                    temp = actuation_temperature<T>(tr.tempAmp[j], tr.tempPeriod[j], tnext, (double)vPhase[k], tf >> 1);
            }
            sPose[0][lv] = vs[k].u.x; sPose[1][lv] = vs[k].u.y; sPose[2][lv] = vs[k].u.z; sPose[3][lv] = vs[k].q.w;
            sPose[4][lv] = vs[k].q.x; sPose[5][lv] = vs[k].q.y; sPose[6][lv] = vs[k].q.z; sPose[7][lv] = (T)temp;
            if constexpr (COL) {
                sMom[0][lv] = vs[k].p.x; sMom[1][lv] = vs[k].p.y; sMom[2][lv] = vs[k].p.z;
                if (tr.col[j]) col_track(a, vGlob[k], vs[k].u, a.rt + tr.rid[j], a.rp[tr.rid[j]], parity);
            }
        }
        __syncthreads();
    }

    // ---- write back
#pragma unroll
    for (int k = 0; k < KL; k++) {
        if (lGlob[k] < 0) continue;
        const int gl = lGlob[k];
        a.ls0[gl] = T4<T>{ls[k].pos2.x, ls[k].pos2.y, ls[k].pos2.z, ls[k].a1v.x};
        a.ls1[gl] = T4<T>{ls[k].a1v.y, ls[k].a1v.z, ls[k].a2v.x, ls[k].a2v.y};
        a.ls2[gl] = T4<T>{ls[k].a2v.z, (T)ls[k].maxStrain, (T)ls[k].flags, (T)lStrain[k]};
    }
#pragma unroll
    for (int k = 0; k < KV; k++) {
        if (vGlob[k] < 0) continue;
        const int gv = vGlob[k], lv = tid + k * NT;
        a.pq0[gv] = T4<T>{vs[k].u.x, vs[k].u.y, vs[k].u.z, vs[k].q.w};
        a.pq1[gv] = T4<T>{vs[k].q.x, vs[k].q.y, vs[k].q.z, sPose[7][lv]};
        a.mo0[gv] = T4<T>{vs[k].p.x, vs[k].p.y, vs[k].p.z, vs[k].h.x};
        a.mo1[gv] = T4<T>{vs[k].h.y, vs[k].h.z, (T)vs[k].flags, (T)0};
    }
    if (tid < nr) {
        RobotTime &t = a.rt[tr.rid[tid]];
        t.time[0] = tr.time[0][tid];
        t.time[1] = tr.time[1][tid];
        t.prevDt[0] = tr.prevDt[0][tid];
        t.prevDt[1] = tr.prevDt[1][tid];
        t.diverged = tr.diverged[tid];
        t.steps = tr.steps[tid];
    }
    // the flat path's hint that some robot may be stale (per-robot flags decide)
    if (COL && blockIdx.x == 0 && tid == 0) a.anyStale[0] = a.anyStale[1] = 1;
}

}  // namespace vx3mi
