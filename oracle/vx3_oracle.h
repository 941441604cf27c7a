/*
 * vx3_oracle.h -- FP64 CPU restatement of the VX3 timestep, used ONLY as the
 * checker for the HIP implementation (tests, smoke check, CPU baseline leg of
 * the benchmark).  Nothing in the product links against this file.
 *
 * The oracle rebuilds a voxel model from a lattice description exactly the way
 * the reference does (voxel order, link order, combined link materials,
 * material constants with the reference's float/double rounding) and advances
 * it with the reference's update equations:
 *
 *   temperature  src/VX3/VX3_VoxelyzeKernel.cu:625-650, VX3_Voxel.cu:107-113,
 *                VX3_Link.cu:80-83, VX3_Voxel.h:96-100
 *   dt           src/VX3/VX3_VoxelyzeKernel.cu:184-217, 237-254
 *   links        src/VX3/VX3_Link.cu:90-260, VX3_Link.h:97-191,
 *                VX3_Material.cu:90-124, src/Utils/VX3_Quat3D.h
 *   collisions   src/VX3/VX3_VoxelyzeKernel.cu:682-843, VX3_Collision.cu:3-31
 *   voxels       src/VX3/VX3_Voxel.cu:162-277, 350-426, VX3_Voxel.h:185-208
 *   cilia        src/VX3/VX3_VoxelyzeKernel.cu:296-302, 846-858, VX3_Voxel.cu:376-377
 *   removal      src/VX3/VX3_VoxelyzeKernel.cu:336-399, 495-513 (SecondaryExperiment)
 *   signals      src/VX3/VX3_Voxel.cu:271-348 (EnableSignals)
 *   detach       src/VX3/VX3_VoxelyzeKernel.cu:461-475, 946-968 (EnableDetach)
 *   attach       src/VX3/VX3_VoxelyzeKernel.cu:593-621, 651-831, VX3_Link.cu:31-76, 206-213 (EnableAttach)
 *   CoM          src/VX3/VX3_VoxelyzeKernel.cu:477-493
 *
 * Deliberate (documented) deviations from the racy reference:
 *   - both endpoint temperatures used for a link rest length are the new ones;
 *   - contact forces are summed serially over surface pairs (first > second);
 *   - divergence is "any link axial strain > 100" instead of one random link;
 *   - signals: VX3_Voxel::propagateSignal writes a neighbour's signal state from
 *     inside the parallel voxel update, so the reference's result depends on
 *     thread order.  Here every voxel first runs its own signal section
 *     (propagate, pacemaker, decay), then receives the signals its neighbours
 *     sent in that step, in its link direction order X_POS..Z_NEG: the
 *     interleaving in which a voxel's own section precedes what it receives.
 *   - attach (VX3_VoxelyzeKernel.cu:729-831): every surface pair first adds its
 *     contact force and is tested against the topology entering the step (the
 *     depth-5 neighbour search included); then, serially in the all-pairs order
 *     (first > second), a candidate whose two link slots are still free gets
 *     its link and gives its contact force back.  The reference runs these
 *     threads concurrently with unsynchronised slot writes; this is the
 *     interleaving in which every test precedes every creation and creations
 *     happen one at a time in pair order.  Attach needs EnableAttach (the
 *     reference's pair loop also runs for EnableCollision alone, with per-voxel
 *     flags whose formulas default to an uninitialised value), and a condition
 *     that is not given does not constrain.
 */
#ifndef VX3_ORACLE_H
#define VX3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define VXO_MAX_SS_POINTS 32   /* stress-strain data points of a MatModel 3 material */
#define VXO_MAX_CURVE_POINTS 72 /* points of a link's combined curve (two data materials: 65) */

/* One palette material (VXA Palette/Material/Mechanical, MatModel 0-3). */
typedef struct {
    float E;            /* Elastic_Mod, Pa */
    float rho;          /* Density, kg/m^3 */
    float nu;           /* Poissons_Ratio (only used with volume effects) */
    float cte;          /* CTE, 1/degC */
    float mu_s, mu_k;   /* uStatic, uDynamic */
    float fail_stress;  /* -1 = no failure */
    float plastic_mod, yield_stress; /* Plastic_Mod, Yield_Stress (model 2) */
    int model;          /* MatModel: 0 linear, 1 linear with failure, 2 bilinear, 3 data */
    int fixed;          /* Fixed */
    int is_measured;    /* isMeasured */
    double thermal_on_after; /* TurnOnThermalExpansionAfterThisManySeconds */
    double cilia;            /* Cilia: multiplier of the voxel's cilia force */
    double cilia_on_after;   /* TurnOnCiliaAfterThisManySeconds */
    double remove_after;     /* RemoveFromSimulationAfterThisManySeconds (0: never) */
    int ss_points;           /* model 3: SSData points (strain, stress), as given */
    float ss_strain[VXO_MAX_SS_POINTS], ss_stress[VXO_MAX_SS_POINTS];
    int is_pacemaker;        /* isPaceMaker */
    double pacemaker_period; /* PaceMakerPeriod (s) */
    double signal_decay;     /* signalValueDecay */
    double signal_delay;     /* signalTimeDelay (s) */
    double inactive_period;  /* inactivePeriod (s) */
    int is_target;           /* isTarget */
    int sticky;              /* Sticky: touching voxels of this material may attach (with enable_attach) */
} vxo_material;

/* Lattice + simulator + environment description. */
typedef struct {
    int nx, ny, nz;
    const signed char *cells;   /* nx*ny*nz palette indices, x fastest; <=0 empty */
    const double *phase;        /* per-voxel phase offset in cycles (voxel order) or NULL */
    const float *voxel_E;       /* per-voxel Young's modulus override (voxel order) or NULL */
    int nmat;                   /* palette entries 1..nmat stored at mats[0..nmat-1] */
    const vxo_material *mats;
    double lattice_dim;         /* Lattice_Dim (m) */
    double dt_frac;             /* Integration/DtFrac */
    float bond_damp_z, col_damp_z, slow_damp_z; /* Damping */
    int volume_effects;         /* Features/VolumeEffectsEnabled */
    int enable_collision;       /* AttachDetach/EnableCollision */
    int enable_expansion;       /* Simulator/EnableExpansion */
    float grav_acc;             /* Environment/Gravity/GravAcc */
    int floor_enabled;          /* Environment/Gravity/FloorEnabled */
    int temp_enabled, vary_temp;                   /* Thermal */
    double temp_base, temp_amplitude, temp_period; /* Thermal */
    int enable_cilia;           /* Simulator/EnableCilia */
    int secondary_experiment;   /* Simulator/SecondaryExperiment (material removal) */
    const double *cilia_base;   /* per-voxel BaseCiliaForce (3 per voxel, voxel order) or NULL */
    const double *cilia_shift;  /* per-voxel ShiftCiliaForce or NULL */
    int enable_signals;         /* Simulator/EnableSignals */
    int enable_detach;          /* Simulator/AttachDetach/EnableDetach */
    int enable_attach;          /* Simulator/AttachDetach/EnableAttach */
    int safety_guard;           /* Simulator/AttachDetach/SafetyGuard: updates a new link's forces are scaled by 0.01 */
    double max_dist_pairs;      /* Simulator/MaxDistInVoxelLengthsToCountAsPair (0: targetCloseness not computed) */
} vxo_spec;

typedef struct vxo_model vxo_model;

/* flags for vxo_step */
#define VXO_CPU_SEMANTICS 1  /* CVoxelyze (src/old) integrator: no angMom reset under static friction */
#define VXO_NO_TEMPERATURE 2 /* skip the temperature update (caller drives temperatures) */
#define VXO_ALL_PAIRS 4      /* collisions by the literal all-pairs loop (VX3_VoxelyzeKernel.cu:448-458), O(S^2) */

vxo_model *vxo_build(const vxo_spec *spec);
void vxo_free(vxo_model *m);
int vxo_num_voxels(const vxo_model *m);
int vxo_num_links(const vxo_model *m);
int vxo_num_surface(const vxo_model *m);
double vxo_time(const vxo_model *m);
double vxo_recommended_dt(const vxo_model *m);
float vxo_step_dt(vxo_model *m);            /* the float dt doTimeStep(-1) uses */

/* Advance nsteps steps with dt (<0: DtFrac*recommended).  Returns the number of
 * steps completed; fewer than nsteps means the model diverged. */
int vxo_step(vxo_model *m, float dt, int nsteps, int flags);

/* Link forces for the current state without advancing: the temperature
 * update and the link phase of the next doTimeStep run; link state (pos2,
 * angle1v, angle2v, smallAngle, local-velocity flag) IS updated, exactly as
 * VX3_Link::updateForces does.  out: 12 doubles per link (forceNeg, forcePos,
 * momentNeg, momentPos), each in its voxel's local frame. */
void vxo_update_links(vxo_model *m, double *out);

/* external of a voxel (CVX_Voxel::external(), VX3_External.h:27-31): dof bits
 * 1,2,4 = X,Y,Z translation fixed, 8,16,32 = X,Y,Z rotation fixed (dof < 0
 * removes the external); translation (m) and rotation vector (rad) of the
 * fixed degrees of freedom; force (N) and moment (N m).  NULL = zeros. */
void vxo_set_external(vxo_model *m, int voxel, int dof, const double *t3, const double *r3, const float *f3,
                      const float *m3);

/* MathTree force field (VXA Simulator/ForceField/{x,y,z}_forcefield): the
 * tokens in the reference's order (reversed breadth-first order of the XML
 * subtree, END last) with these op codes (VX3_MathTreeOperator order); a VAR's
 * value is the variable index (x, y, z, hit, t, angle, targetCloseness,
 * numClosePairs, num_voxel). */
enum { VXO_MT_END, VXO_MT_CONST, VXO_MT_E, VXO_MT_PI, VXO_MT_VAR, VXO_MT_ADD, VXO_MT_SUB, VXO_MT_MUL, VXO_MT_DIV,
       VXO_MT_POW, VXO_MT_SQRT, VXO_MT_SIN, VXO_MT_COS, VXO_MT_TAN, VXO_MT_ATAN, VXO_MT_LOG, VXO_MT_INT, VXO_MT_ABS,
       VXO_MT_NOT, VXO_MT_GT, VXO_MT_LT, VXO_MT_AND, VXO_MT_OR, VXO_MT_NORMALCDF };
void vxo_set_force_field(vxo_model *m, int axis, int n, const int *ops, const double *values);
/* AttachCondition Condition_k (k = 0..4, VX3_SimulationManager.cu:334-343), tokens
 * as for force fields.  A voxel may attach when every condition it has is > 0
 * (a condition never set does not constrain). */
void vxo_set_attach_condition(vxo_model *m, int k, int n, const int *ops, const double *values);
/* attach state: per voxel the enableAttach flag of the last voxel update (1
 * before the first), per link the updates left whose forces are scaled by 0.01
 * (isNewLink) */
void vxo_get_attach(const vxo_model *m, int *voxel_flags, int *link_new);
/* the CycleStep samples of doTimeStep (VX3_VoxelyzeKernel.cu:314-334, 545-562),
 * which force fields and attach conditions read as angle, targetCloseness and
 * numClosePairs: out7 = recentAngle, targetCloseness, numClosePairs,
 * angleSampleTimes, the sampled centre of mass (3) */
void vxo_get_samples(const vxo_model *m, double *out7);

/* stress-strain curve of a non-linear link material (n <= VXO_MAX_CURVE_POINTS
 * points, 0 when linear) */
int vxo_get_link_curve(const vxo_model *m, int link, float *strain, float *stress);

/* state access (voxel order) */
void vxo_get_voxels(const vxo_model *m, double *pos3, double *quat4, double *linmom3, double *angmom3);
void vxo_set_voxels(vxo_model *m, const double *pos3, const double *quat4, const double *linmom3, const double *angmom3);
void vxo_get_flags(const vxo_model *m, int *voxel_flags, int *link_flags);
/* removal (SecondaryExperiment) and detach (EnableDetach): per voxel / link the
 * number of the step (1-based, counted over vxo_step calls) at whose end it was
 * removed or detached, 0 if present */
void vxo_get_removed(const vxo_model *m, int *voxel_step, int *link_step);
/* signals: each voxel's localSignal */
void vxo_get_signals(const vxo_model *m, double *local);
/* collisionCount: contacts between a target and a non-target voxel, summed over
 * the steps (VX3_VoxelyzeKernel.cu:711-726) */
long vxo_collision_count(const vxo_model *m);
/* the pairs in contact in the last step's collision phase, in the order their
 * effects were applied: (voxel1, voxel2) per pair, at most cap pairs written;
 * returns the number of pairs */
int vxo_contact_pairs(const vxo_model *m, int *v1v2, int cap);
void vxo_get_link_state(const vxo_model *m, double *state9, float *strain);
void vxo_get_temperature(const vxo_model *m, float *temp, double *rest_length);
void vxo_set_temperature(vxo_model *m, const float *temp);
void vxo_get_topology(const vxo_model *m, int *lattice3, int *link_neg, int *link_pos, int *link_axis, int *surface);
void vxo_center_of_mass(const vxo_model *m, double *com3);
void vxo_get_contact(const vxo_model *m, double *contact3);

/* material constants as the reference stores them (for builder parity tests)
 * voxel: 12 floats, link: 11 floats (see vx3_oracle.c for the order) */
void vxo_get_voxel_material(const vxo_model *m, int voxel, float *out12);
void vxo_get_link_material(const vxo_model *m, int link, float *out11);

/* elementary functions used by the oracle (include/vx3mi_math.h) */
double vxo_sin(double x);
double vxo_cos(double x);
double vxo_acos(double x);

/* ---- differentiable mode (softplus floor, per-voxel E and phase) ---- */
typedef struct {
    double floor_k, floor_w, fric_v;  /* softplus stiffness, width; friction regularisation speed */
    int floor_enabled, vary_temp, enable_expansion;
    double temp_amplitude, temp_period;
    double dt;
} vxo_diff_opts;
typedef struct vxo_diff vxo_diff;
/* the model m supplies topology and materials and must outlive the result;
 * phase (cycles) and E (Pa) are per voxel */
vxo_diff *vxo_diff_new(const vxo_model *m, const double *phase, const double *E, const vxo_diff_opts *o);
void vxo_diff_free(vxo_diff *d);
void vxo_diff_step(vxo_diff *d, int nsteps);
void vxo_diff_get(const vxo_diff *d, double *pos3, double *quat4, double *linmom3, double *angmom3);
double vxo_diff_mass(const vxo_diff *d, int v);
void vxo_diff_set(vxo_diff *d, const double *pos3, const double *quat4, const double *linmom3, const double *angmom3);

#ifdef __cplusplus
}
#endif
#endif
