/*
 * vx3mi.h -- C ABI of the MI355X-native VX3 timestep.
 *
 * This is the drop-in boundary for the reference's VX3 kernel interface
 * (class VX3_VoxelyzeKernel, src/VX3/VX3_VoxelyzeKernel.cuh:21-49) and its
 * host-side construction path (VX3_SimulationManager::readVXD,
 * src/VX3/VX3_SimulationManager.cu:277-381).  Plain pointers and sizes only;
 * no exception crosses this boundary; every call returns an int status and
 * vx3mi_last_error() describes the last failure on the calling thread.
 *
 * Ownership: a vx3mi_model is a host object; a vx3mi_batch owns every device
 * buffer it uses (allocated once in vx3mi_batch_create, freed by
 * vx3mi_batch_destroy) and one HIP stream.  A handle is not re-entrant.
 */
#ifndef VX3MI_H
#define VX3MI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define VX3MI_OK 0
#define VX3MI_E_ARG (-1)      /* invalid argument */
#define VX3MI_E_PARSE (-2)    /* VXA/VXD could not be read */
#define VX3MI_E_HIP (-3)      /* HIP runtime failure */
#define VX3MI_E_NOMEM (-4)    /* allocation failure */
#define VX3MI_E_LIMIT (-5)    /* model exceeds a documented limit */
#define VX3MI_DIVERGED 1      /* at least one robot diverged (see vx3mi_batch_diverged) */

/* precision of the device state */
#define VX3MI_FP32 32
#define VX3MI_FP64 64

/* step modes (vx3mi_step) */
#define VX3MI_MODE_AUTO 0        /* tiled if possible, else persistent for small robots, else flat */
#define VX3MI_MODE_PLAIN 1       /* one robot per launch sequence (reference-style) */
#define VX3MI_MODE_FLAT 2        /* one link launch + one voxel launch per step for the whole batch */
#define VX3MI_MODE_PERSISTENT 3  /* one workgroup per robot, many steps per launch */
#define VX3MI_MODE_TILED 4       /* robots bin-packed into register/LDS-resident tiles (collisions off) */

#define VX3MI_MAX_SS_POINTS 32    /* stress-strain data points of a MatModel 3 material */
#define VX3MI_MAX_CURVE_POINTS 72 /* points of a link's combined stress-strain curve */

/* One palette material: VXA Palette/Material (src/VXA/VX_Object.cpp:1373-1475). */
typedef struct {
    float elastic_mod;   /* Elastic_Mod (Pa) */
    float density;       /* Density (kg/m^3) */
    float poissons_ratio;/* Poissons_Ratio (inert unless volume_effects) */
    float cte;           /* CTE (1/degC) */
    float u_static;      /* uStatic */
    float u_dynamic;     /* uDynamic */
    float fail_stress;   /* Fail_Stress (mat_model 1, 2; -1 = none) */
    float plastic_mod;   /* Plastic_Mod (mat_model 2) */
    float yield_stress;  /* Yield_Stress (mat_model 2) */
    int mat_model;       /* MatModel: 0 linear, 1 linear with failure, 2 bilinear, 3 stress-strain data */
    int fixed;           /* Fixed */
    int is_measured;     /* isMeasured */
    int is_target;       /* isTarget: contacts with non-targets count as hits (collisionCount) */
    double thermal_on_after; /* TurnOnThermalExpansionAfterThisManySeconds */
    float red, green, blue, alpha; /* Display */
    double cilia;            /* Cilia: multiplier of the voxel's cilia force (with enable_cilia) */
    double cilia_on_after;   /* TurnOnCiliaAfterThisManySeconds */
    double remove_after;     /* RemoveFromSimulationAfterThisManySeconds (with secondary_experiment; 0 never) */
    int ss_points;           /* mat_model 3: SSData points as given (a leading 0, 0 is allowed) */
    float ss_strain[VX3MI_MAX_SS_POINTS], ss_stress[VX3MI_MAX_SS_POINTS]; /* SSData StrainData / StressData */
    int is_pacemaker;        /* isPaceMaker (with enable_signals) */
    double pacemaker_period; /* PaceMakerPeriod (s) */
    double signal_value_decay; /* signalValueDecay (a VXA without it: 0.9) */
    double signal_time_delay;  /* signalTimeDelay (s; a VXA without it: 0.03) */
    double inactive_period;    /* inactivePeriod (s; a VXA without it: 0.03) */
    int sticky;                /* Sticky: touching voxels of this material attach (with enable_attach) */
} vx3mi_material;

/* Lattice description: the information content of a merged VXA (base.vxa +
 * robot.vxd), plus the per-voxel arrays a binding can supply directly. */
typedef struct {
    int nx, ny, nz;                /* Structure X_Voxels, Y_Voxels, Z_Voxels */
    const signed char *cells;      /* nx*ny*nz palette indices, x fastest; <=0 empty */
    const double *phase;           /* per-voxel PhaseOffset (cycles), voxel order, or NULL */
    const float *voxel_E;          /* per-voxel Young's modulus override, or NULL */
    int nmat;                      /* palette size (materials 1..nmat) */
    const vx3mi_material *mats;
    double lattice_dim;            /* Lattice/Lattice_Dim */
    double dt_frac;                /* Simulator/Integration/DtFrac */
    float bond_damp_z, col_damp_z, slow_damp_z; /* Simulator/Damping */
    int volume_effects;            /* Simulator/Features/VolumeEffectsEnabled */
    int enable_collision;          /* Simulator/AttachDetach/EnableCollision */
    int enable_expansion;          /* Simulator/EnableExpansion */
    float grav_acc;                /* Environment/Gravity/GravAcc */
    int floor_enabled;             /* Environment/Gravity/FloorEnabled */
    int temp_enabled, vary_temp;   /* Environment/Thermal */
    double temp_base, temp_amplitude, temp_period;
    int enable_cilia;              /* Simulator/EnableCilia */
    int secondary_experiment;      /* Simulator/SecondaryExperiment (material removal, re-initialisation) */
    const double *cilia_base;      /* per-voxel BaseCiliaForce (N, 3 per voxel, voxel order) or NULL */
    const double *cilia_shift;     /* per-voxel ShiftCiliaForce (scaled by the voxel's signal) or NULL */
    int enable_signals;            /* Simulator/EnableSignals */
    int enable_detach;             /* Simulator/AttachDetach/EnableDetach: failed links leave their voxels */
    double max_dist_pairs;         /* Simulator/MaxDistInVoxelLengthsToCountAsPair (0: no close-pair counting) */
    int enable_attach;             /* Simulator/AttachDetach/EnableAttach: touching sticky voxels link up */
    int safety_guard;              /* Simulator/AttachDetach/SafetyGuard (a VXA without it: 500), 0..65535 */
} vx3mi_lattice;

/* Settings a model carries besides topology and materials, read back by
 * vx3mi_model_settings (the VXA keys they come from are named). */
typedef struct {
    int nx, ny, nz;                /* Structure dims */
    double lattice_dim;            /* Lattice_Dim * X_Dim_Adj */
    double dt_frac;                /* Integration/DtFrac */
    int floor_enabled, enable_collision, enable_expansion;
    int vary_temp;                 /* VaryTempEnabled && TempPeriod > 0 */
    double temp_base, temp_amplitude, temp_period;
    float initial_temp;            /* temperature set at import (UpdateMatTemps) */
    int record_step_size, record_voxel, record_link; /* Simulator/RecordHistory */
    int has_stop_formula, has_fitness_formula;
    int enable_cilia, secondary_experiment, enable_signals, enable_detach;
    double max_dist_pairs;         /* MaxDistInVoxelLengthsToCountAsPair */
    int enable_attach, safety_guard; /* AttachDetach/EnableAttach, SafetyGuard */
} vx3mi_settings;

typedef struct vx3mi_model vx3mi_model;
typedef struct vx3mi_batch vx3mi_batch;

const char *vx3mi_last_error(void);
const char *vx3mi_version(void);

/* ---- models (host) -------------------------------------------------------
 * Replaces VX3_SimulationManager::readVXD (VX3_SimulationManager.cu:277-381)
 * + CVX_Sim::ReadVXA/Import (src/VXA/VX_Sim.cpp:59-183) + the
 * VX3_VoxelyzeKernel(CVX_Sim*) constructor (VX3_VoxelyzeKernel.cu:27-105). */
int vx3mi_model_from_vxa(const char *base_vxa_path, const char *vxd_path, vx3mi_model **out);
int vx3mi_model_from_xml(const char *base_vxa_text, const char *vxd_text, vx3mi_model **out);
int vx3mi_model_from_lattice(const vx3mi_lattice *spec, vx3mi_model **out);
void vx3mi_model_free(vx3mi_model *m);
/* settings read back (see vx3mi_settings) */
int vx3mi_model_settings(const vx3mi_model *m, vx3mi_settings *out);
/* per-voxel PhaseOffset in cycles, voxel order (nvox doubles) */
int vx3mi_model_phase(const vx3mi_model *m, double *out);
/* Evaluate the model's stop-condition or fitness MathTree (VXA
 * Simulator/StopCondition/StopConditionFormula, Simulator/FitnessFunction;
 * VX3_MathTree::eval, src/Utils/VX3_MathTree.h:50-190) on the host.
 * vars (9): x, y, z, hit, t, angle, targetCloseness, numClosePairs, num_voxel.
 * VX3MI_E_ARG when the model has no such formula. */
#define VX3MI_FORMULA_STOP 0
#define VX3MI_FORMULA_FITNESS 1
int vx3mi_model_eval_formula(const vx3mi_model *m, int which, const double *vars9, double *out);
/* counts: voxels, links, surface voxels */
int vx3mi_model_counts(const vx3mi_model *m, int *nvox, int *nlink, int *nsurf);
/* recommendedTimeStep (VX3_VoxelyzeKernel.cu:184-217), host evaluation */
int vx3mi_model_recommended_dt(const vx3mi_model *m, double *out);
/* topology in voxel / link order: lattice (3 per voxel), link ends and axis */
int vx3mi_model_topology(const vx3mi_model *m, int *lattice3, int *link_neg, int *link_pos, int *link_axis);
/* derived material constants as the device uses them (float, the reference's
 * rounding; VX3_MaterialVoxel.h:29-44, VX3_MaterialLink.h:27-36).
 * voxel (12): mass, 1/mass, 1/inertia, 2*sqrt(mass)*zetaInternal,
 *   globalDampingTranslateC, globalDampingRotateC, collisionDampingTranslateC,
 *   gravityForce, penetrationStiffness, muStatic, muKinetic, CTE
 * link (11): E, a1, a2, b1, b2, b3, sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp */
int vx3mi_model_voxel_constants(const vx3mi_model *m, int voxel, float *out12);
int vx3mi_model_link_constants(const vx3mi_model *m, int link, float *out11);
/* stress-strain curve of a link's material when it is non-linear (a bilinear
 * or data palette material at either end): the springs-in-series combination
 * of CVX_MaterialLink::updateAll (src/old/VX_MaterialLink.cpp:66-101), starting
 * at (0, 0).  Returns the number of points (at most VX3MI_MAX_CURVE_POINTS; 0
 * for a linear material) and writes them to strain / stress when not NULL. */
int vx3mi_model_link_curve(const vx3mi_model *m, int link, float *strain, float *stress);

/* External of a voxel: replaces CVX_Voxel::external() with
 * CVX_External::setDisplacement / setForce / setMoment (src/old/VX_Voxel.h:75,
 * src/old/VX_External.h:59-80), which the reference copies to the device as
 * VX3_External (src/VX3/VX3_External.h:27-31, VX3_Voxel.cu:39-45).
 * dof: bits 1,2,4 fix x,y,z translation and 8,16,32 fix x,y,z rotation;
 * dof < 0 removes the external (any dof >= 0, even 0, makes one: the force and
 * moment still apply).  translation3 (m) / rotation3 (rotation vector, rad):
 * where the fixed degrees of freedom are held, relative to the lattice
 * position; force3 (N) / moment3 (N m) are stored as float, as the reference
 * does.  NULL arrays mean zeros.  Boundary-condition regions of a VXA
 * (Environment/Boundary_Conditions) are applied through the same call. */
int vx3mi_model_set_external(vx3mi_model *m, int voxel, int dof, const double *translation3,
                             const double *rotation3, const double *force3, const double *moment3);
/* Force field of one axis (0 x, 1 y, 2 z): the MathTree XML a VXA's
 * Simulator/ForceField/{x,y,z}_forcefield element holds (e.g.
 * "<mtMUL><mtCONST>0.01</mtCONST><mtVAR>t</mtVAR></mtMUL>"; empty or NULL
 * removes it), added to every voxel's force as VX3_Voxel::timeStep does
 * (src/VX3/VX3_Voxel.cu:177-183, parsed as VX3_SimulationManager::ParseMathTree,
 * VX3_SimulationManager.cu:157-275).  Variables x, y, z (the voxel's position),
 * t, hit (collisionCount), num_voxel, and angle, targetCloseness and
 * numClosePairs from the CycleStep samples of doTimeStep (every
 * int(TempPeriod / dt) steps, VX3_VoxelyzeKernel.cu:314-334, 545-562), taken on
 * the device.  Batches with force fields run in the plain and flat modes. */
int vx3mi_model_set_force_field(vx3mi_model *m, int axis, const char *formula_xml);
/* AttachCondition Condition_k (k = 0..4, VX3_SimulationManager.cu:334-343): a
 * MathTree XML as for force fields (empty or NULL removes it).  After each voxel
 * update a voxel may attach when every condition given is > 0
 * (VX3_VoxelyzeKernel.cu:607-621); variables x, y, z (the voxel's new position),
 * hit (collisionCount), t (the step's start time), num_voxel, and angle,
 * targetCloseness, numClosePairs from the last CycleStep samples. */
int vx3mi_model_set_attach_condition(vx3mi_model *m, int k, const char *formula_xml);
/* the voxel's external (*dof = -1 and zeros when it has none) */
int vx3mi_model_get_external(const vx3mi_model *m, int voxel, int *dof, double *translation3,
                             double *rotation3, double *force3, double *moment3);

/* ---- batches (device) -----------------------------------------------------
 * A batch packs n models (CSR: voxel and link offsets per robot) into one
 * set of device arrays.  precision: VX3MI_FP32 or VX3MI_FP64. */
int vx3mi_batch_create(vx3mi_model *const *models, int n, int precision, vx3mi_batch **out);
void vx3mi_batch_destroy(vx3mi_batch *b);
int vx3mi_batch_counts(const vx3mi_batch *b, int *nrobot, int *nvox, int *nlink);

/* recommendedTimeStep of every robot, computed on the device with a
 * wavefront max-reduction (replaces the serial loop at
 * VX3_VoxelyzeKernel.cu:195-203).  out: n doubles. */
int vx3mi_batch_recommended_dt(vx3mi_batch *b, double *out);

/* doTimeStep (VX3_VoxelyzeKernel.cu:237-359) nsteps times for every robot.
 * dt < 0: each robot uses (float)(DtFrac * recommendedTimeStep), as
 * doTimeStep(-1).  dt >= 0: every robot uses dt (dt = 0 only counts the
 * steps, as doTimeStep(0) does).  A robot whose link strain passes 100 stops
 * (doTimeStep returns false, VX3_VoxelyzeKernel.cu:271-280) and is not
 * advanced further; vx3mi_batch_diverged reports it.  Returns VX3MI_OK or an
 * error.  Asynchronous on the batch stream except for the return code of
 * launch errors; vx3mi_batch_sync waits. */
int vx3mi_step(vx3mi_batch *b, int nsteps, float dt, int mode);
int vx3mi_batch_sync(vx3mi_batch *b);

/* State access in voxel order (absolute positions, as VX3_Voxel::pos). robot
 * < 0 selects all robots.  Any output pointer may be NULL. */
int vx3mi_get_state(vx3mi_batch *b, int robot, double *pos3, double *quat4, double *linmom3, double *angmom3);
int vx3mi_set_state(vx3mi_batch *b, int robot, const double *pos3, const double *quat4, const double *linmom3,
                    const double *angmom3);
/* per-voxel flag words (floor / static friction bits as the reference) */
int vx3mi_get_voxel_flags(vx3mi_batch *b, int robot, int *flags);
/* Material removal (VX3_VoxelyzeKernel::removeVoxels, VX3_VoxelyzeKernel.cu:365-399,
 * with SecondaryExperiment) and detach (gpu_update_detach :946-968, with
 * EnableDetach: a link past its failure strain leaves both voxels): per voxel
 * and link of the robot, the number of the step (counted from batch creation)
 * at whose end it was removed or detached, 0 while present.  Removed voxels
 * keep their last state; their links are gone from both ends.  Batches that can
 * lose voxels or links run in the plain and flat modes. */
int vx3mi_batch_removed(vx3mi_batch *b, int robot, int *voxel_step, int *link_step);
/* Signals (EnableSignals; VX3_Voxel::propagateSignal / packMaker /
 * localSignalDecay / receiveSignal, src/VX3/VX3_Voxel.cu:271-348): each voxel's
 * localSignal, which scales its ShiftCiliaForce.  The reference writes a
 * neighbour's signal from inside the parallel voxel update; here each voxel runs
 * its own signal section and then receives its neighbours' signals of that
 * step in link direction order (X_POS..Z_NEG).  Batches with signals run in the
 * plain and flat modes. */
int vx3mi_batch_signals(vx3mi_batch *b, int robot, double *local);
/* collisionCount of every robot: contacts between a target (isTarget) and a
 * non-target voxel summed over the steps (VX3_VoxelyzeKernel.cu:711-726), the
 * MathTree variable "hit".  out: n ints. */
int vx3mi_batch_hits(vx3mi_batch *b, int *out);
/* per-link flag words (bit 0 local velocity valid, bit 1 small-angle
 * formulation, VX3_Link.cu:117-131, 172-191) and axial strain of the last step,
 * link order; either pointer may be NULL */
int vx3mi_get_link_flags(vx3mi_batch *b, int robot, int *flags, float *strain);
/* per-robot simulated time (currentTime), diverged flag, step count */
int vx3mi_batch_time(vx3mi_batch *b, double *out);
int vx3mi_batch_diverged(vx3mi_batch *b, int *out);
/* Contact pairs, a diagnostic of the collision phase (the pairs the
 * reference's all-pairs loop finds in contact, VX3_VoxelyzeKernel.cu:448-458,
 * 682-729): with cap > 0 every later vx3mi_step call records the pairs of
 * surface voxels inside each other's contact envelope in its steps (at most
 * cap pairs per call); cap 0 stops recording.  Error when no robot has
 * collisions. */
int vx3mi_batch_record_contacts(vx3mi_batch *b, int cap);
/* No counterpart in the reference: the same contact record for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops).
 * The record lives in the library's static per-GPU queue of at most 1,048,576
 * pairs (cap beyond that: VX3MI_E_LIMIT); the call allocates nothing.  Call it
 * with the batch's device current.  Every batch recording this way on one GPU
 * shares that GPU's queue: step calls of such batches that overlap race, and
 * each step call starts the record afresh (the queue keeps the last call's
 * pairs until then).  Flat-mode steps capture their graph again after it. */
int vx3mi_batch_record_contacts_static(vx3mi_batch *b, int cap);
/* per robot: the steps whose collision phase rebuilt the robot's candidate
 * lists (every surface voxel lists the surface voxels within its largest
 * contact envelope plus a skin of a quarter voxel; they are rebuilt, through a
 * counting-sort grid, once some surface voxel has moved more than half the
 * skin).  out: n ints. */
int vx3mi_batch_listings(vx3mi_batch *b, int *out);
/* The pairs recorded by the last vx3mi_step call for one robot (robot < 0:
 * all, batch voxel indices): voxel index pairs, the smaller first, sorted.
 * Returns the number of pairs (at most cap are written) or VX3MI_E_LIMIT when
 * the call found more pairs than the record holds. */
int vx3mi_batch_contact_pairs(vx3mi_batch *b, int robot, int *pairs2, int cap);
/* Attach state of one robot (replaces reading VX3_Voxel::enableAttach,
 * VX3_Voxel.h:302, and the links handle_collision_attachment appends to
 * d_v_links, VX3_VoxelyzeKernel.cu:729-831): voxel_flags (robot voxel count
 * ints, may be NULL) gets each voxel's enableAttach flag; for the links made so
 * far, in the order they were made, link_neg / link_pos / link_axis get the
 * negative end, positive end (robot voxel indices) and axis, new_left the
 * updates left whose forces are scaled by 0.01 (isNewLink); at most cap links
 * are written (the arrays may be NULL with cap 0).  overflow (may be NULL) is
 * set when a voxel had more than 8 attach candidates of lower index in one step
 * or the robot's pool of link entries ran out; the candidates beyond were not
 * considered.  Returns the number of links made. */
int vx3mi_batch_attach_state(vx3mi_batch *b, int robot, int *voxel_flags, int *link_neg, int *link_pos,
                             int *link_axis, int *new_left, int cap, int *overflow);
/* updateCurrentCenterOfMass (VX3_VoxelyzeKernel.cu:477-493) per robot: out 3n */
int vx3mi_center_of_mass(vx3mi_batch *b, double *out);

/* Link forces of the current state (VX3_Link::updateForces) without moving
 * voxels: updates the link state as the reference does and writes 12 values
 * per link (forceNeg, forcePos, momentNeg, momentPos in voxel frames). */
int vx3mi_eval_link_forces(vx3mi_batch *b, double *out12);

/* Device timing of the last vx3mi_step call, from HIP events recorded on the
 * batch stream around its launches (ms), and the number of kernel launches
 * that call made. */
int vx3mi_last_step_ms(vx3mi_batch *b, float *ms);
int vx3mi_last_step_launches(vx3mi_batch *b, long long *launches);

/* Autotuner: times trial_steps (<= 0: 100) of each launch configuration on
 * this batch -- [0] tiled, 512-thread tiles; [1] tiled, 256-thread tiles (two
 * workgroups per compute unit); [2..4] one workgroup of 128 / 256 / 512
 * threads per robot; [5] flat launches in captured graphs -- writes the
 * milliseconds to ms6 (-1: not applicable), keeps the fastest for
 * VX3MI_MODE_AUTO and reports it in choice3 (mode, tile variant, threads per
 * robot).  The batch state is restored afterwards. */
int vx3mi_batch_autotune(vx3mi_batch *b, int trial_steps, double *ms6, int *choice3);
/* fixes the configuration VX3MI_MODE_AUTO uses: mode (PLAIN..TILED), tile
 * variant (0 large, 1 small) and threads per robot for PERSISTENT (128, 256,
 * 512; 0 keeps the current) */
int vx3mi_batch_configure(vx3mi_batch *b, int mode, int tile, int threads);

/* ---- simulation driver --------------------------------------------------------
 * CUDA_Simulation (VX3_SimulationManager.cu:11-121) for every robot of a batch:
 * iterate {stop condition met -> stop; doTimeStep; diverged -> stop; history
 * frame every real_stepsize iterations} at most max_steps times (<= 0: the
 * reference's 1000000), then the final centre of mass and fitness.  The
 * models must be the ones the batch was created from.  history (may be NULL)
 * receives the .history text of each robot (header and frames, the format of
 * VX3_SimulationManager.cu:40-111) in order per robot. */
typedef struct {
    double current_time;                  /* currentTime when the robot stopped */
    double fitness_score;                 /* FitnessFunction of the CoM displacement (0 without one) */
    double initial_com[3], current_com[3];
    double total_distance_of_all_voxels;  /* sum over measured voxels of |pos - initial pos| */
    double vox_size;
    double recommended_dt;                /* recommendedTimeStep() */
    int num_voxel, num_measured_voxel, num_close_pairs;
    int diverged;
    int angle_sample_times;
    int real_stepsize;
    long long steps;                      /* doTimeStep calls */
} vx3mi_result;
typedef void (*vx3mi_text_fn)(void *user, int robot, const char *text, size_t len);
int vx3mi_run(vx3mi_batch *b, vx3mi_model *const *models, int n, long long max_steps, int mode, vx3mi_text_fn history,
              void *user, vx3mi_result *results, double *const *final_positions);
/* final_positions (may be NULL, or hold NULL entries): per robot 3*nvox doubles
 * receiving the absolute voxel positions when that robot stopped */
/* palette index (1-based, the reference's matid) of every voxel, voxel order */
int vx3mi_model_palette_index(const vx3mi_model *m, int *out);
/* palette entry index (1..nmat) as the model holds it */
int vx3mi_model_material(const vx3mi_model *m, int index, vx3mi_material *out);
/* per-voxel BaseCiliaForce / ShiftCiliaForce (3 per voxel, voxel order; zeros
 * when the VXA has none), as CVX_Sim::Import gives them to the voxels
 * (src/VXA/VX_Sim.cpp:102-105) */
int vx3mi_model_cilia(const vx3mi_model *m, double *base3, double *shift3);
/* Simulator/SavePositionOfAllVoxels of a model read from XML (0 otherwise) */
int vx3mi_model_save_positions(const vx3mi_model *m);
/* VXA.RawPrint of a model read from XML ("" otherwise); valid while the model lives */
const char *vx3mi_model_raw_print(const vx3mi_model *m);

/* ---- differentiable mode ----------------------------------------------------
 * No counterpart in the reference: a differentiable variant of doTimeStep
 * (same link equations, VX3_Link.cu:90-218, and integrator,
 * VX3_Voxel.cu:162-277) with a softplus floor penalty, kinetic-only friction
 * with a regularised direction, collisions off, per-voxel actuation phase and
 * Young's modulus as parameters, and a hand-written adjoint with
 * checkpointing every K steps and recomputation. */
typedef struct vx3mi_diff vx3mi_diff;
typedef struct {
    double floor_stiffness;  /* softplus floor stiffness k (N/m) */
    double floor_width;      /* softplus width w (m) */
    double friction_speed;   /* friction regularisation speed (m/s) */
    double dt;               /* timestep; <= 0: DtFrac x recommendedTimeStep with every modulus at e_max */
    double e_max;            /* modulus bound for the default timestep (Pa) */
    int checkpoint_every;    /* K; <= 0: chosen under hbm_budget */
    double hbm_budget;       /* device bytes for checkpoints + tape when K is chosen */
    int reverse_threads;     /* threads per robot of the reverse sweep: FP32 512 or 1024, FP64 256; 0 = 512 / 256;
                                -1 = flat: each reverse step of every robot at once, one voxel or link per
                                thread over the device (batches whose robots differ much in size) */
} vx3mi_diff_opts;
/* defaults: k 1e5, w 1e-4, friction speed 1e-3, dt from e_max 1e7, K chosen, budget 16 GiB,
 * reverse threads 0 */
void vx3mi_diff_default_opts(vx3mi_diff_opts *o);
/* every model must have one actuation / floor setting, no volume effects and
 * no TurnOnThermalExpansion delay; collisions are not simulated */
int vx3mi_diff_create(vx3mi_model *const *models, int n, int precision, const vx3mi_diff_opts *opts, vx3mi_diff **out);
void vx3mi_diff_destroy(vx3mi_diff *d);
int vx3mi_diff_counts(const vx3mi_diff *d, int *nrobot, int *nvox, int *nlink);
/* per-voxel parameters in batch voxel order: phase (cycles) and Young's modulus (Pa) */
int vx3mi_diff_set_params(vx3mi_diff *d, const double *phase, const double *E);
/* nsteps from the initial state (voxels at rest on their lattice points),
 * keeping checkpoints for the adjoint */
int vx3mi_diff_forward(vx3mi_diff *d, int nsteps);
/* final state of the last forward; robot -1 = whole batch; absolute positions */
int vx3mi_diff_get_state(vx3mi_diff *d, int robot, double *pos, double *quat, double *linmom, double *angmom);
/* adjoint of the last forward: from dL/d(final state) (each may be NULL = 0),
 * dL/dphase (per cycle) and dL/dE per voxel */
int vx3mi_diff_backward(vx3mi_diff *d, const double *grad_pos, const double *grad_quat, const double *grad_linmom,
                        const double *grad_angmom, double *grad_phase, double *grad_E);
/* out (10): K, checkpoint bytes, tape bytes, dt, forward ms, backward ms, steps, precision,
 * recompute ms and reverse-sweep ms of the backward */
int vx3mi_diff_info(const vx3mi_diff *d, double *out10);

#ifdef __cplusplus
}
#endif
#endif
