// vx3mi_internal.h -- host-side model representation shared by the VXA reader,
// the model builder and the batch (device) code.
#pragma once

#include <string>
#include <vector>

#include "vx3mi.h"
#include "vx3mi_physics.h"

namespace vx3mi {

static_assert(kMaxCurvePoints == VX3MI_MAX_CURVE_POINTS, "curve capacity of the ABI and the model differ");

void set_error(const std::string &msg);

// Host model: topology, materials and initial state of one robot, in the
// reference's voxel and link order.
// MathTree formulas (MtOp, MtToken) are declared in vx3mi_physics.h.  A VXA
// formula is read as tokens in reversed breadth-first order of the XML subtree,
// MT_END last, evaluated as a FIFO of values (an operator consumes the oldest
// unconsumed values).
// false when the token list is malformed (an operator without operands, no MT_END)
bool mt_eval(const std::vector<MtToken> &f, const double vars[MV_COUNT], double *out);
// The same formula as a postfix program for formula_eval (operands before
// their operator, left before right); false when malformed or when evaluating
// it needs more than kMaxFormulaStack values at once.
bool mt_postfix(const std::vector<MtToken> &f, std::vector<MtToken> &post);
// bit v set when variable v occurs in the formula
unsigned mt_vars_used(const std::vector<MtToken> &f);

struct Model {
    int nx = 0, ny = 0, nz = 0;
    double latticeDim = 0.001;
    // voxels (voxel order)
    std::vector<int> ix, iy, iz;
    std::vector<int> vmat;          // index into vmats
    std::vector<int> vpal;          // palette index (1-based, the reference's matid)
    std::vector<double> phase;      // cycles
    std::vector<int> vlinks;        // 6 per voxel, X_POS..Z_NEG, -1 none
    std::vector<int> surface;       // surface voxel indices
    std::vector<int> vext;          // index into exts per voxel, -1 none (empty: no externals)
    std::vector<VoxelExt> exts;     // externals (CVX_Voxel::external())
    // links (link order)
    std::vector<int> lneg, lpos, laxis, lmat;
    // materials
    std::vector<VoxelMat> vmats;
    std::vector<LinkMat> lmats;
    std::vector<CurvePt> curves;    // stress-strain curves of non-linear link materials (LinkMat::curve)
    std::vector<float> vmatE;       // Young's modulus per voxel material (host only)
    std::vector<vx3mi_material> palette;  // display/metadata of palette entries
    // settings
    double dtFrac = 0.9;
    int floorEnabled = 1, enableCollision = 1, enableExpansion = 0;
    int varyTemp = 0;               // VaryTempEnabled && TempPeriod > 0
    double tempAmplitude = 0, tempPeriod = 0.1, tempBase = 25;
    float initialTemp = 0;          // UpdateMatTemps result (src/VXA/VX_Sim.cpp:467-483)
    // settings carried for the drop-in runner (VX3_SimulationManager.cu:323-376)
    std::vector<MtToken> stopFormula, fitnessFormula;
    std::vector<MtToken> forceField[3];  // Simulator/ForceField/{x,y,z}_forcefield (empty: none)
    int recordStepSize = 0, recordVoxel = 1, recordLink = 0;
    std::string rawPrint;           // VXA.RawPrint
    int savePositions = 0;          // Simulator/SavePositionOfAllVoxels
    int secondaryExperiment = 0;
    double reinitInitialPositionAfter = 0.0;
    int enableCilia = 0;             // Simulator/EnableCilia
    int enableSignals = 0;           // Simulator/EnableSignals
    int enableDetach = 0;            // Simulator/AttachDetach/EnableDetach
    int enableAttach = 0;            // Simulator/AttachDetach/EnableAttach
    int safetyGuard = 500;           // Simulator/AttachDetach/SafetyGuard
    std::vector<MtToken> attachCond[5];  // AttachCondition/Condition_0..4 (empty: none)
    // per voxel material: the link material (lmats index) a link between two
    // voxels of it gets when they attach, -1 when they cannot (not sticky, fixed)
    std::vector<int> attachLmat;
    double maxDistPairs = 0.0;       // Simulator/MaxDistInVoxelLengthsToCountAsPair
    std::vector<double> ciliaBase, ciliaShift;  // 3 per voxel (empty: zeros)
    int nv() const { return (int)ix.size(); }
    int ext_of(int v) const { return vext.empty() ? -1 : vext[v]; }
    int nl() const { return (int)lneg.size(); }
};

// CVX_Voxel::external() + CVX_External::setDisplacement / setForce / setMoment
// for voxel v (dof < 0 removes its external)
void set_external(Model &m, int v, int dof, const double t[3], const double r[3], const float f[3], const float mo[3]);

// recommendedTimeStep, VX3_VoxelyzeKernel.cu:184-217 (double arithmetic on
// the float material constants)
double recommended_dt(const Model &m);

}  // namespace vx3mi

struct vx3mi_model {
    vx3mi::Model m;
};
