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
// The FP32 tile holds fewer robots (64) to keep 1,984 links and 1,280 voxels:
// the batched workload then packs into 256 tiles, one per compute unit, where
// 1,920 links made 259 and three compute units ran a second tile after the
// first (the launch took twice as long).
template <typename T> struct TileColCfg;
template <> struct TileColCfg<float> {
    static constexpr int NT = 768, KL = 3, KV = 2, LMAX = 1984, VMAX = 1280, RMAX = 64, MINW = 1;
};
template <> struct TileColCfg<double> {
    static constexpr int NT = 512, KL = 2, KV = 2, LMAX = 896, VMAX = 640, RMAX = 64, MINW = 2;
};
// the thinner list skin goes to the robots the FP32 collision tile can hold
static_assert(kSmallRobotVoxels == TileColCfg<float>::VMAX, "small-robot skin threshold");
template <typename T, int SMALL, int COL> struct TileSel { using C = TileCfg<T, SMALL>; };
template <typename T, int SMALL> struct TileSel<T, SMALL, 1> { using C = TileColCfg<T>; };

// LDS robot table of a tile
template <int RMAX, int COL> struct TileRobots {
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
    // collision tiles: the robot lists its candidates at the start of the next
    // step (set by the tracking of the voxel phase), its stale distance, a link
    // of this step passed the divergence strain (merged into diverged once the
    // contact and link phases that read diverged are done), and per step parity
    // whether some robot of the tile is stale
    int stale[COL ? RMAX : 1];
    float staleD2[COL ? RMAX : 1];
    int divNow[RMAX];
    int anyStale[2];
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
// listing passes of the flat path, then packs each voxel's list for the tile);
// each thread computes the contact forces of its voxels after its links, from
// the poses and momenta entering the step (LDS) and the packed lists, summed in
// list order as in k_col_contact.  Tracking and the stale flags stay in LDS
// during a launch.
template <typename T, int SMALL, int LDSMAT, int COL>
__global__ __launch_bounds__((TileSel<T, SMALL, COL>::C::NT), (TileSel<T, SMALL, COL>::C::MINW)) void k_tile(
    Dev<T> a, const int4 *tiles, const int *tileRobots, int nsteps, int parity0, int nLinkMats, int nVoxMats) {
    using C = typename TileSel<T, SMALL, COL>::C;
    constexpr int NT = C::NT, KL = C::KL, KV = C::KV;
    constexpr int LMAX = C::LMAX, VMAX = C::VMAX, RMAX = C::RMAX;
    static_assert(KL * NT >= LMAX && KV * NT >= VMAX && RMAX <= NT && NT <= 3 * LMAX,
                  "a tile's threads must cover its capacity");
    static_assert(RMAX <= 128 && VMAX <= 2048, "robot slots take 7 bits, packed candidates 11");
    __shared__ T sPose[8][VMAX];    // position field xyz, qw, qx, qy, qz, temperature
    __shared__ T sForce[12][LMAX];  // forceNeg, momentNeg, forcePos, momentPos (in the link frame)
    __shared__ T sMom[COL ? 3 : 1][COL ? VMAX : 1];  // linear momentum entering the step (contacts)
    __shared__ TileRobots<RMAX, COL> tr;
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
        if constexpr (COL) {
            tr.stale[tid] = tr.col[tid] && t.stale[parity0 & 1];
            tr.staleD2[tid] = p.staleD2;
        }
        tr.divNow[tid] = 0;
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
        int any = 0;
        for (int j = 0; j < nr; j++) {
            tr.vpre[j + 1] += tr.vpre[j];
            tr.lpre[j + 1] += tr.lpre[j];
            tr.spre[j + 1] += tr.spre[j];
            if constexpr (COL) any |= tr.stale[j] && !tr.diverged[j];
        }
        tr.anyStale[parity0 & 1] = any;
        tr.anyStale[(parity0 & 1) ^ 1] = 0;
    }
    __syncthreads();

    // ---- owned links: static description and state in registers
    int lEnds[KL], lInfo[KL];  // neg | pos << 16 (tile indices); axis | robot slot << 2 | material << 9
    float lStrain[KL];
    LinkState<T> ls[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) {
        const int ll = tid + k * NT;
        if (ll < nl) {
            const int j = tile_slot(tr.lpre, nr, ll);
            const int gl = tr.loffG[j] + (ll - tr.lpre[j]);
            const int2 e = a.lends[gl];
            const int base = tr.vpre[j] - tr.voffG[j];
            lEnds[k] = (e.x + base) | ((e.y + base) << 16);
            lInfo[k] = (a.lmeta[gl].x & 3) | (j << 2) | ((a.lmeta[gl].x >> 2) << 9);
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
    int vInfo[KV], vLat[KV];  // robot slot | material << 7; lattice position
    int vLinks[KV][3];  // local link ids, two 16-bit fields each (0xffff = none)
    T vPhase[KV];
    VoxelState<T> vs[KV];
    // collision tiles: the voxel's candidates (kNoCand: none, not a surface voxel
    // of a robot with collisions; kWideCand: read the candidate list itself; else
    // the number of packed entries) and its position field at the last listing
    constexpr int kNoCand = -2, kWideCand = -1;
    int vCand[KV];
    V3<T> vBp[KV];
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int lv = tid + k * NT;
        vCand[k] = kNoCand;
        vBp[k] = V3<T>{0, 0, 0};
        if (lv < nv) {
            const int j = tile_slot(tr.vpre, nr, lv);
            const int gv = tr.voffG[j] + (lv - tr.vpre[j]);
            vInfo[k] = j | (a.vmat[gv] << 7);
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

    // owned voxel k: its robot's slot, its material, its global id
    auto vSlot = [&](int k) { return vInfo[k] & 127; };
    auto vMat = [&](int k) { return vInfo[k] >> 7; };
    auto vGlob = [&](int k) {
        const int j = vInfo[k] & 127;
        return tr.voffG[j] + (tid + k * NT - tr.vpre[j]);
    };
    // surface voxel i of the tile (collision robots in tile order): its global
    // surface index and its robot's slot
    auto tile_surface = [&](int i, int &j) {
        j = tile_slot(tr.spre, nr, i);
        return a.rp[tr.rid[j]].soff + (i - tr.spre[j]);
    };
    // packed candidates of the tile's voxels: one word per candidate holding the
    // partner's tile index (11 bits), its material (5 bits; the materials are in
    // LDS) and its lattice offset from the voxel (5 signed bits per axis); four
    // candidates of a voxel per int4, laid out by (rank / 4, owned voxel slot k,
    // thread), so the threads of a wave read consecutive int4s, and a tile's
    // entries stay in its XCD's L2
    int4 *const tc4 = reinterpret_cast<int4 *>(a.tcand);
    const size_t cbase = COL ? (size_t)blockIdx.x * (kMaxCand / 4) * KV * NT : 0;
    auto cand4_at = [&](int c4, int k) -> size_t { return cbase + ((size_t)c4 * KV + k) * NT + tid; };
    // owned voxel k's packed entries from its candidate list, and its position
    // field at the listing (at the start of a launch and after each listing)
    auto pack = [&](int k) {
        vCand[k] = kNoCand;
        if (tid + k * NT >= nv || !tr.col[vSlot(k)]) return;
        const int gv = vGlob(k);
        const int s = a.vsurf[gv];
        if (s < 0) return;
        const T4<T> b = a.bpos[gv];
        vBp[k] = V3<T>{b.x, b.y, b.z};
        vCand[k] = kWideCand;
        const int n = a.ccount[s];
        if (!LDSMAT || n > kMaxCand) return;
        const int *list = a.clist + (long)s * kMaxCand;
        const int lbase = tr.vpre[vSlot(k)] - tr.voffG[vSlot(k)];
        int4 q = {0, 0, 0, 0};
        for (int c = 0; c < n; c++) {
            const int w = list[c];
            const int3 o = lat_offset(vLat[k], a.lat[w]);
            if (o.x < -16 || o.x > 15 || o.y < -16 || o.y > 15 || o.z < -16 || o.z > 15) return;
            const int e = (w + lbase) | (a.vmat[w] << 11) | ((o.x & 31) << 16) | ((o.y & 31) << 21) | ((o.z & 31) << 26);
            if ((c & 3) == 0) q.x = e;
            else if ((c & 3) == 1) q.y = e;
            else if ((c & 3) == 2) q.z = e;
            else q.w = e;
            if ((c & 3) == 3 || c == n - 1) {
                tc4[cand4_at(c >> 2, k)] = q;
                q = int4{0, 0, 0, 0};
            }
        }
        vCand[k] = n;
    };
    // the contact force on owned voxel k (tile index lv) from its candidates in
    // contact now, summed in candidate order: col_contact's partner test and
    // force with the poses and momenta of the tile (the robots of a tile lose
    // neither links nor voxels, and have no signals)
    auto contact = [&](int k, int lv, const VoxelMat &mi, int parity) -> V3<T> {
        V3<T> f = {0, 0, 0};
        const int j = vSlot(k), gv = vGlob(k);
        RobotTime *rt = a.rt + tr.rid[j];
        const int lbase = tr.vpre[j] - tr.voffG[j];  // tile index of a voxel of the robot = global id + lbase
        const T4<T> qi0 = {sPose[0][lv], sPose[1][lv], sPose[2][lv], sPose[3][lv]};
        const T L = (T)mi.nomSize;
        const T bsaI = base_size_average<T>(mi.nomSize, (float)sPose[7][lv], mi.alphaCTE);
        const V3<T> velI = V3<T>{sMom[0][lv], sMom[1][lv], sMom[2][lv]} * (T)mi.massInv;
        const int ti = a.vtarget ? a.vtarget[gv] : 0;
        // the partner of tile index lw at lattice offset o when in contact now,
        // its force in fw
        auto partner = [&](int lw, int3 o, const VoxelMat &mj, V3<T> &fw) -> bool {
            if (mi.fixed && mj.fixed) return false;
            const T4<T> qj0 = {sPose[0][lw], sPose[1][lw], sPose[2][lw], sPose[3][lw]};
            const V3<T> diff = col_diff<T>(qi0, qj0, o, L);
            const T bsaJ = base_size_average<T>(mj.nomSize, (float)sPose[7][lw], mj.alphaCTE);
            if (!in_contact_envelope(diff, bsaI, bsaJ)) return false;
            const int w = lw - lbase;
            if (a.vtarget && a.vtarget[w] != ti && gv < w) atomicAdd(&rt->collisionCount, 1);
            if (a.pairRec && gv < w) {
                const int q = atomicAdd(a.pairN, 1);
                if (q < a.pairCap) a.pairRec[q] = int2{gv, w};
            }
            const V3<T> velJ = V3<T>{sMom[0][lw], sMom[1][lw], sMom[2][lw]} * (T)mj.massInv;
            fw = contact_force(diff, bsaI, bsaJ, velI, velJ, mi.penKinv, mj.penKinv, mi.cDampT, mj.cDampT);
            return true;
        };
        // a partner by voxel id (lists read in place); links: test the current
        // links first
        auto partner_id = [&](int w, bool links, V3<T> &fw) -> bool {
            const int3 o = lat_offset(vLat[k], a.lat[w]);
            if (links && col_linked(a, gv, o.x, o.y, o.z)) return false;
            return partner(w + lbase, o, LDSMAT ? sVM[a.vmat[w]] : a.vmats[a.vmat[w]], fw);
        };
        if constexpr (LDSMAT) {
            const int n = vCand[k];
            if (n >= 0) {
                // four candidates per load; the first twelve requested at once,
                // each later load two batches ahead of its use (three registers
                // in rotation, so no load result is copied before it arrives)
                auto batch = [&](const int4 &E, int c0) {
                    for (int q = 0; q < 4 && c0 + q < n; q++) {
                        const int e = q == 0 ? E.x : q == 1 ? E.y : q == 2 ? E.z : E.w;
                        const unsigned u = (unsigned)e;
                        const int3 o = {(int)(u << 11) >> 27, (int)(u << 6) >> 27, (int)(u << 1) >> 27};
                        V3<T> fw;
                        if (partner(e & 2047, o, sVM[(e >> 11) & 31], fw)) f = f + fw;
                    }
                };
                int4 R0 = {0, 0, 0, 0}, R1 = {0, 0, 0, 0}, R2 = {0, 0, 0, 0};
                if (n > 0) R0 = tc4[cand4_at(0, k)];
                if (n > 4) R1 = tc4[cand4_at(1, k)];
                if (n > 8) R2 = tc4[cand4_at(2, k)];
                for (int c0 = 0; c0 < n; c0 += 12) {
                    batch(R0, c0);
                    if (c0 + 12 < n) R0 = tc4[cand4_at((c0 >> 2) + 3, k)];
                    batch(R1, c0 + 4);
                    if (c0 + 16 < n) R1 = tc4[cand4_at((c0 >> 2) + 4, k)];
                    batch(R2, c0 + 8);
                    if (c0 + 20 < n) R2 = tc4[cand4_at((c0 >> 2) + 5, k)];
                }
                return f;
            }
        }
        const int s = a.vsurf[gv];
        const int nc = a.ccount[s];
        if (nc <= kMaxCand) {
            const int *list = a.clist + (long)s * kMaxCand;
            for (int c = 0; c < nc; c++) {
                V3<T> fw;
                if (partner_id(list[c], false, fw)) f = f + fw;
            }
        } else {  // more candidates than a list holds: the grid of this step, ascending
            tr.stale[j] = 1;  // rebuilt every step while this lasts
            tr.anyStale[parity ^ 1] = 1;
            const RobotParams &rp = a.rp[tr.rid[j]];
            int prev = -1;
            for (;;) {
                int next = INT_MAX;
                col_grid_each<false>(a, rp, s, -1, [&](int w, const T4<T> &) {
                    if (w > prev && w < next && w != gv) next = w;
                });
                if (next == INT_MAX) break;
                V3<T> fw;
                if (partner_id(next, true, fw)) f = f + fw;
                prev = next;
            }
        }
        return f;
    };
    if constexpr (COL) {
#pragma unroll
        for (int k = 0; k < KV; k++) pack(k);
    }

    for (int s = 0; s < nsteps; s++) {
        const int parity = (parity0 + s) & 1;
        if constexpr (COL) {
            // ---- candidate lists of the robots whose lists are stale (rare)
            if (tr.anyStale[parity]) {
                auto listing = [&](int j) { return tr.col[j] && !tr.diverged[j] && tr.stale[j]; };
                if (tid < nr && listing(tid)) a.rt[tr.rid[tid]].stale[parity] = 1;  // the flag the listing passes read
#pragma unroll
                for (int k = 0; k < KV; k++) {  // the listing passes read poses from global memory
                    const int lv = tid + k * NT;
                    if (lv >= nv || !listing(vSlot(k))) continue;
                    a.pq0[vGlob(k)] = T4<T>{sPose[0][lv], sPose[1][lv], sPose[2][lv], sPose[3][lv]};
                    a.pq1[vGlob(k)] = T4<T>{sPose[4][lv], sPose[5][lv], sPose[6][lv], sPose[7][lv]};
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
                    col_scan<false>(a, r, wsum);  // at most VMAX surface voxels: a few buckets per thread
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
#pragma unroll
                for (int k = 0; k < KV; k++)
                    if (tid + k * NT < nv && listing(vSlot(k))) pack(k);
                __syncthreads();  // every thread has read the stale flags
                if (tid < nr && listing(tid)) {
                    RobotTime &t = a.rt[tr.rid[tid]];
                    t.stale[parity] = 0;
                    t.listings += 1;
                    tr.stale[tid] = 0;
                }
                __syncthreads();  // before the contact phase may set them again
            }
        }
        // ---- links; a link past the divergence strain stops its robot before
        // the voxel phase (diverged itself changes only after the barrier, so
        // every link and contact of the step sees the same flag)
#pragma unroll
        for (int k = 0; k < KL; k++) {
            const int ll = tid + k * NT;
            if (ll >= nl) continue;
            const int j = (lInfo[k] >> 2) & 127;
            if (tr.diverged[j]) continue;
            const LinkMat &lm = LDSMAT ? sLM[lInfo[k] >> 9] : a.lmats[lInfo[k] >> 9];
            if (lm.fixedBoth) continue;
            const int n = lEnds[k] & 0xffff, p = lEnds[k] >> 16;
            const V3<T> du = {sPose[0][p] - sPose[0][n], sPose[1][p] - sPose[1][n], sPose[2][p] - sPose[2][n]};
            const Q4<T> qn = {sPose[3][n], sPose[4][n], sPose[5][n], sPose[6][n]};
            const Q4<T> qp = {sPose[3][p], sPose[4][p], sPose[5][p], sPose[6][p]};
            LinkOut<T> o;
            link_forces<T, true>(lm, a.curves, lInfo[k] & 3, du, qn, qp, (float)sPose[7][n], (float)sPose[7][p], tr.prevDt[parity][j],
                           ls[k], o);
            sForce[0][ll] = o.fN.x; sForce[1][ll] = o.fN.y; sForce[2][ll] = o.fN.z;
            sForce[3][ll] = o.mN.x; sForce[4][ll] = o.mN.y; sForce[5][ll] = o.mN.z;
            sForce[6][ll] = o.fP.x; sForce[7][ll] = o.fP.y; sForce[8][ll] = o.fP.z;
            sForce[9][ll] = o.mP.x; sForce[10][ll] = o.mP.y; sForce[11][ll] = o.mP.z;
            lStrain[k] = o.strain;
            if (o.strain > 100.0f) tr.divNow[j] = 1;
        }
        // ---- contact forces, from the poses and momenta entering the step
        V3<T> cf[KV];
#pragma unroll
        for (int k = 0; k < KV; k++) {
            cf[k] = V3<T>{0, 0, 0};
            if constexpr (COL) {
                if (vCand[k] == kNoCand || tr.diverged[vSlot(k)]) continue;
                const VoxelMat &vm = LDSMAT ? sVM[vMat(k)] : a.vmats[vMat(k)];
                cf[k] = contact(k, tid + k * NT, vm, parity);
            }
        }
        __syncthreads();
        if (tid < nr && tr.divNow[tid]) tr.diverged[tid] = 1;  // (the voxel phase reads both)
        if (COL && tid == 0) tr.anyStale[parity] = 0;            // every thread has read it
        // ---- voxels
#pragma unroll
        for (int k = 0; k < KV; k++) {
            const int lv = tid + k * NT;
            if (lv >= nv) continue;
            const int j = vSlot(k);
            if (tr.diverged[j] || tr.divNow[j]) continue;
            const float dt = tr.dt[j];
            const double tnext = tr.time[parity][j] + (double)dt;
            if (lv == tr.vpre[j]) {
                tr.time[parity ^ 1][j] = tnext;
                tr.prevDt[parity ^ 1][j] = dt;
                tr.steps[j] += 1;
            }
            const VoxelMat &vm = LDSMAT ? sVM[vMat(k)] : a.vmats[vMat(k)];
            if (vm.fixed) continue;
            V3<T> fs = {0, 0, 0}, ms = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const int l = (vLinks[k][d >> 1] >> (16 * (d & 1))) & 0xffff;
                if (l == 0xffff) continue;
                const int o = (d & 1) ? 6 : 0;  // NEG direction: this voxel is the pVPos end
                // link-frame end vectors back to the lattice frame (axis d / 2, known here)
                fs = fs + to_axis_original(V3<T>{sForce[o + 0][l], sForce[o + 1][l], sForce[o + 2][l]}, d >> 1);
                ms = ms + to_axis_original(V3<T>{sForce[o + 3][l], sForce[o + 4][l], sForce[o + 5][l]}, d >> 1);
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
                if (vCand[k] != kNoCand) {  // a moved surface voxel: its robot lists again next step
                    const T dx = vs[k].u.x - vBp[k].x, dy = vs[k].u.y - vBp[k].y, dz = vs[k].u.z - vBp[k].z;
                    if (dx * dx + dy * dy + dz * dz > (T)tr.staleD2[j]) {
                        tr.stale[j] = 1;
                        tr.anyStale[parity ^ 1] = 1;
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---- write back
#pragma unroll
    for (int k = 0; k < KL; k++) {
        const int ll = tid + k * NT;
        if (ll >= nl) continue;
        const int j = (lInfo[k] >> 2) & 127;
        const int gl = tr.loffG[j] + (ll - tr.lpre[j]);
        a.ls0[gl] = T4<T>{ls[k].pos2.x, ls[k].pos2.y, ls[k].pos2.z, ls[k].a1v.x};
        a.ls1[gl] = T4<T>{ls[k].a1v.y, ls[k].a1v.z, ls[k].a2v.x, ls[k].a2v.y};
        a.ls2[gl] = T4<T>{ls[k].a2v.z, (T)ls[k].maxStrain, (T)ls[k].flags, (T)lStrain[k]};
    }
#pragma unroll
    for (int k = 0; k < KV; k++) {
        const int lv = tid + k * NT;
        if (lv >= nv) continue;
        const int gv = vGlob(k);
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
        t.diverged = tr.diverged[tid] | tr.divNow[tid];
        t.steps = tr.steps[tid];
        if constexpr (COL) {
            if (tr.col[tid]) {  // lists to build at the next step, as the flat path's tracking leaves them
                const int pe = (parity0 + nsteps) & 1;
                t.stale[pe] = tr.stale[tid];
                t.stale[pe ^ 1] = 0;
            }
        }
    }
    // the flat path's hint that some robot may be stale (per-robot flags decide)
    if (COL && blockIdx.x == 0 && tid == 0) a.anyStale[0] = a.anyStale[1] = 1;
}

}  // namespace vx3mi
