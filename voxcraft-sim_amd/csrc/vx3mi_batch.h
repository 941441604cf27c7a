// vx3mi_batch.h -- host-side batch interface shared by the FP32 and FP64
// device translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vx3mi_internal.h"

namespace vx3mi {

struct RobotParams {
    double tempAmp, tempPeriod;
    double tempInvPeriod;  // 1 / tempPeriod (the FP32 path multiplies)
    int varyTemp, enableExpansion, collisions;
    int secondary;           // SecondaryExperiment: materials are removed after their time
    int detach;              // EnableDetach: failed links leave both voxels
    int signals;             // EnableSignals (the signal arrays are the batch's)
    int ffOff[3], ffLen[3];  // force-field postfix programs in the formula table (length 0: none)
    int voff, vcount, loff, lcount;
    // collision broadphase: surface voxels [soff, soff+scount) of the surface
    // list; the rebuild grid's bucket table of (mask+1) buckets at toff (mask+2
    // prefix entries), its cell edge = the candidate radius listR; candidate
    // lists are rebuilt when a surface voxel moved more than sqrt(staleD2)
    int soff, scount, toff, mask;
    int listLinked;    // lists keep linked lattice neighbours (the robot can lose links or voxels)
    float staleD2;
    double listR;
    // attach (EnableAttach): SafetyGuard, the AttachCondition programs in the
    // formula table (length 0: no condition), and the robot's pool of link
    // entries [poff, poff + pcount) for links made at run time
    int attach, safetyGuard;
    int acOff[5], acLen[5];
    int poff, pcount;
    // CycleStep samples on the device (a force field or attach condition reads
    // angle, targetCloseness or numClosePairs): the robot's target voxels
    // [tgOff, tgOff + tgCount) of the target list and the pair radius
    // MaxDistInVoxelLengthsToCountAsPair * voxel size (0: closeness not computed)
    int sample, tgOff, tgCount;
    double pairR;
};

// Candidate lists: every surface voxel lists (ascending) the surface voxels of
// its robot within listR = the largest contact envelope + a skin of kSkin voxel
// lengths, found through a counting-sort grid whenever a surface voxel of the
// robot has moved more than half the skin (less a margin) since the last
// listing.  A voxel with more than kMaxCand candidates searches the grid itself,
// and its robot rebuilds the grid every step while it does.  Robots of at most
// kSmallRobotVoxels voxels (those a collision tile holds) take the thinner skin
// kSkinSmall: their listing runs inside the tile and costs little, and every
// step's contact phase tests fewer candidates.  The partners, and so the
// results, do not depend on the skin.
constexpr int kMaxCand = 48;
constexpr double kSkin = 0.25;
constexpr double kSkinSmall = 0.15;
constexpr int kSmallRobotVoxels = 1280;

struct RobotTime {
    double time[2];
    float prevDt[2];
    float dt;
    int diverged;
    int steps;
    int stale[2];        // candidate lists to be rebuilt before the contact phase of the step of this parity
    int collisionCount;  // target / non-target contacts summed over the steps (VX3_VoxelyzeKernel.cu:711-726)
    int listings;        // steps whose collision phase rebuilt the robot's candidate lists
    int divergedNow;     // a link of this step passed the divergence strain (the voxel update sets diverged)
    // the CycleStep samples (VX3_VoxelyzeKernel.cu:314-334, 545-562): centre of
    // mass now and at the last two samples, recentAngle, targetCloseness,
    // numClosePairs, angleSampleTimes
    double scom[3], shist[2][3];
    double recentAngle, closeness;
    int closePairs, angleSamples;
    int attachUsed;      // pool entries handed out so far (attach)
    int attachOverflow;  // a voxel had more attach candidates than it keeps (kAttachPerVoxel), or the pool ran out
};

// attach candidates a surface voxel keeps per step (partners of lower index)
constexpr int kAttachPerVoxel = 8;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
            return VX3MI_E_HIP;                                                                \
        }                                                                                      \
    } while (0)

struct BatchBase {
    int precision = 32;
    int R = 0, V = 0, L = 0;
    int parity = 0;
    int persistentChunk = 1000;  // steps per persistent launch
    long launches = 0;           // kernel launches of the last step call
    int lastMode = 0;
    int S = 0;                   // surface voxels (collision candidates) of the batch
    bool anyCollisions = false;
    bool anyRemoval = false;      // some robot can lose voxels (SecondaryExperiment)
    bool anyDetach = false;       // some robot can lose links (EnableDetach with failing materials)
    bool anyAttach = false;       // some robot can make links (EnableAttach with a sticky material)
    int LP = 0;                   // link entries: the L lattice links, then the attach pools
    long tableBuckets = 0;         // rebuild-grid prefix entries of the batch
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void *> allocs;
    std::vector<RobotParams> hrp;
    std::vector<float> defaultDt;  // (float)(DtFrac * recommendedTimeStep) per robot
    std::vector<float> curDt;      // dt uploaded to the device
    std::vector<int> hlat, hvmat;
    std::vector<VoxelMat> hvmats;
    // graphs of flat steps: chunks of kGraphChunks[k] steps (even, so the step
    // parity returns to its start), each captured for the parity it starts at
    static constexpr int kGraphChunks[2] = {1000, 64};
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    int graphParity[2] = {-1, -1};
    bool graphStale = false;  // the launch arguments changed: capture the flat graphs again
    void drop_graphs() {
        for (int k = 0; k < 2; k++) {
            if (graphs[k]) (void)hipGraphExecDestroy(graphs[k]);
            graphs[k] = nullptr;
            graphParity[k] = -1;
        }
    }
    virtual ~BatchBase() {
        drop_graphs();
        for (void *p : allocs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <typename P> int alloc(P **p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void **)p, n * sizeof(P));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return VX3MI_E_HIP;
        }
        allocs.push_back((void *)*p);
        return VX3MI_OK;
    }
    virtual int step(int nsteps, float dt, int mode) = 0;
    virtual int get_state(int robot, double *pos, double *quat, double *lm, double *am) = 0;
    virtual int set_state(int robot, const double *pos, const double *quat, const double *lm, const double *am) = 0;
    virtual int get_flags(int robot, int *flags) = 0;
    virtual int get_link_flags(int robot, int *flags, float *strain) = 0;
    // removal state: per voxel / link of the robot the step count (the robot's
    // steps since the batch was created) at whose end it was removed, 0 while
    // present
    virtual int get_removed(int robot, int *vstep, int *lstep) = 0;
    // each voxel's localSignal (zeros without signals)
    virtual int get_signals(int robot, double *local) = 0;
    virtual int eval_links(double *out) = 0;
    virtual int com(double *out) = 0;
    virtual int recommended(double *out) = 0;
    // one robot's absolute positions, orientations, temperatures and link
    // strains (robot voxel / link order)
    virtual int frame(int robot, double *pos, double *quat, float *temp, float *strain, float *maxStrain = nullptr) = 0;
    // times the launch configurations on trial steps and keeps the fastest for
    // VX3MI_MODE_AUTO; the batch state is left as it was
    virtual int autotune(int trialSteps, double *ms, int *choice) = 0;
    // fixes the launch configuration VX3MI_MODE_AUTO uses
    virtual int configure(int mode, int tile, int threads) = 0;
    // contact pairs of the collision phases of the next step calls (cap pairs
    // per call; 0 stops), and those of the last call for one robot
    virtual int record_contacts(int cap) = 0;
    // record into the given device queue (pairs) and counter, or stop (nullptr);
    // allocates nothing
    virtual int record_contacts_at(int2 *rec, int *counter, int cap) = 0;
    virtual int contact_pairs(int robot, int *out, int cap) = 0;
    // attach state of one robot: per voxel the enableAttach flag, and the links
    // made so far in the order they were made (robot voxel indices; per link the
    // negative end, positive end, axis and the updates left at 0.01 force);
    // returns the number of links made (at most cap written), *overflow set when
    // a candidate was lost
    virtual int attach_state(int robot, int *vflags, int *neg, int *pos, int *axis, int *newLeft, int cap,
                             int *overflow) = 0;
    // the robot's present attached links: negative end, positive end (robot
    // voxel indices), axis, and the step count at whose end each was made
    virtual int attach_links(int robot, std::vector<int4> &out) = 0;
    RobotTime *d_rt = nullptr;
};

BatchBase *make_batch_fp32(vx3mi_model *const *models, int n, int *rc);
BatchBase *make_batch_fp64(vx3mi_model *const *models, int n, int *rc);

}  // namespace vx3mi

struct vx3mi_batch {
    vx3mi::BatchBase *b = nullptr;
};
