/*
 * mpm.h — C ABI of the MI355X MPM timestep engine (libmpm_hip.so).
 *
 * The reference (Matter, twmeehan/matter-tests) has no FFI: its seam is the C++ class
 * `Simulation` (src/simulation/simulation.hpp:23-225), whose `simulate()` frame loop
 * (src/simulation/simulation.cpp:168-201) calls `advanceStep()` (src/simulation/simulation.cpp:219-273).
 * Each entry point below states the reference interface it replaces.
 *
 * Conventions
 *   - Plain C types only. Host buffers are owned by the caller; device state is owned by the context.
 *   - Every call returns 0 on success or a bitwise OR of MPM_ERR_* codes.
 *     A nonzero return from mpm_step is the reference's `exit = 1` (src/simulation/simulation.hpp:28).
 *   - One host thread per context. Work is issued on one HIP stream per context.
 *   - Vectors are `dim` doubles per particle, interleaved (the memory image of std::vector<TV>);
 *     matrices are dim*dim doubles per particle in column-major order (the memory image of
 *     std::vector<TM> with Eigen's default storage), so a caller can pass `&particles.F[0](0,0)`.
 */
#ifndef MATTER_MPM_H
#define MATTER_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 2

/* Enumerations. Integer values follow the declaration order of the reference enums
 * (src/tools.hpp:53-57), so `static_cast<int32_t>(PlasticModel::DP)` is a valid value. */
enum mpm_plate_type { MPM_PLATE_TOP = 0, MPM_PLATE_BOTTOM, MPM_PLATE_LEFT, MPM_PLATE_RIGHT,
                      MPM_PLATE_FRONT, MPM_PLATE_BACK };
enum mpm_elastic_model { MPM_ELASTIC_HENCKY = 0, MPM_ELASTIC_NEOHOOKEAN };
enum mpm_plastic_model { MPM_PLASTIC_NONE = 0, MPM_PLASTIC_VM, MPM_PLASTIC_DP, MPM_PLASTIC_DPSOFT,
                         MPM_PLASTIC_MCC, MPM_PLASTIC_VMVISC, MPM_PLASTIC_DPVISC, MPM_PLASTIC_MCCVISC,
                         MPM_PLASTIC_DPMUI, MPM_PLASTIC_MCCMUI, MPM_PLASTIC_SNOW };
enum mpm_hardening_law { MPM_HARD_NONE = 0, MPM_HARD_EXPO_EXPL, MPM_HARD_EXPO_IMPL, MPM_HARD_SINH_EXPL,
                         MPM_HARD_SINH_IMPL };
enum mpm_bc { MPM_BC_NOSLIP = 0, MPM_BC_SLIPSTICK, MPM_BC_SLIPFREE };

/* Analytic boundary objects: the closed set of ObjectGeneral subclasses (src/objects/object_*.hpp). */
enum mpm_object_kind {
    MPM_OBJECT_BOX_ROTATED = 1, /* ObjectBoxRotated  src/objects/object_box.hpp:6-44    */
    MPM_OBJECT_BUMP        = 2, /* ObjectBump        src/objects/object_bump.hpp:8-42   */
    MPM_OBJECT_GROUND      = 3, /* ObjectGround      src/objects/object_ground.hpp:8-38 */
    MPM_OBJECT_RAMP        = 4, /* ObjectRamp        src/objects/object_ramp.hpp:8-41   */
    MPM_OBJECT_GATE        = 5, /* ObjectGate        src/objects/object_gate.hpp:8-43   */
    MPM_OBJECT_SILO        = 6, /* ObjectSilo (3D)   src/objects/object_silo.hpp:8-56   */
    MPM_OBJECT_CURVE       = 7, /* ObjectCurve       src/objects/object_curve.hpp:8-39  */
    MPM_OBJECT_SDF         = 8  /* ObjectVdb         src/objects/object_vdb.hpp:17-94 (mpm_add_sdf_object) */
};

/* Arithmetic form of the step. */
enum mpm_precision {
    MPM_PRECISION_FP64  = 0, /* plain form: every quantity in binary64, the drop-in */
    MPM_PRECISION_SPLIT = 1  /* transfers and grid in binary32, return mapping and reductions in binary64 */
};

/* Status / error bits. */
#define MPM_OK                    0u
#define MPM_ERR_INVALID_ARGUMENT  (1u << 0)  /* bad pointer, size or enum value                     */
#define MPM_ERR_UNSUPPORTED       (1u << 1)  /* option outside this engine (Snow, linear spline)       */
#define MPM_ERR_HIP               (1u << 2)  /* a HIP runtime call failed                             */
#define MPM_ERR_OUT_OF_MEMORY     (1u << 3)
#define MPM_ERR_PERZYNA_NEGATIVE  (1u << 4)  /* plasticity.cpp:190-193,286-289,546-553 */
#define MPM_ERR_NEWTON_MAXITER    (1u << 5)  /* plasticity.cpp:200-203,297-300,562-565 */
#define MPM_ERR_NEWTON_SMALL_DERIV (1u << 6) /* plasticity.cpp:219-222,313-316,578-581 */
#define MPM_ERR_MCC_NO_STEP       (1u << 7)  /* mcc_rma_implicit_*_onevar.hpp (limited search failed, sinh max iter) */
#define MPM_ERR_GRID_BOUNDS       (1u << 8)  /* a stencil left the grid (undefined behaviour in the reference) */
#define MPM_ERR_NONFINITE         (1u << 9)  /* NaN/Inf reached the particle state */
#define MPM_ERR_CONSERVATION      (1u << 10) /* simulation.cpp:312-316,328-332 */
#define MPM_ERR_NOT_READY         (1u << 11) /* step before particles were set */
#define MPM_ERR_SCRATCH_LIMIT     (1u << 12) /* mpm_step_noalloc: the grid does not fit the device scratch */
#define MPM_WARN_SUPERSONIC       (1u << 16) /* update_dt.cpp:14-17 (warning only; not an exit) */

#define MPM_ERR_EXIT_MASK 0x0000FFFFu /* bits that map to the reference's `exit = 1` */

/* Flags for mpm_step. */
#define MPM_STEP_STOP_AT_FRAME_END  1u /* return after the step that completes a frame */
#define MPM_STEP_CHECK_CONSERVATION 2u /* run checkMassConservation/checkMomentumConservation after P2G */

/*
 * Simulation parameters. One field per public member of the reference's Simulation
 * (src/simulation/simulation.hpp:28-106), with the same names and meanings, plus the
 * compile-time switches of src/tools.hpp:19-21 (dim, spline_degree) and the arithmetic form.
 * Call mpm_config_defaults() first: it fills the reference's default member values.
 */
typedef struct mpm_config {
    int32_t dim;            /* 2 or 3              (tools.hpp:20 THREEDIM)  */
    int32_t spline_degree;  /* 2 (quadratic) or 3 (cubic)  (tools.hpp:21)   */
    int32_t precision;      /* enum mpm_precision                           */
    int32_t n_threads;      /* OpenMP threads of the reference; ignored by the GPU engine */

    int32_t elastic_model;  /* enum mpm_elastic_model */
    int32_t plastic_model;  /* enum mpm_plastic_model */
    int32_t hardening_law;  /* enum mpm_hardening_law */
    int32_t use_pradhana;
    int32_t use_mises_q;
    int32_t use_mibf;
    int32_t use_musl;
    int32_t gravity_special;
    int32_t pbc;                       /* periodic in x over [0, Lx): pbc.cpp, remesh.cpp:14-36 */
    int32_t change_particle_positions; /* position_update.cpp:38-43 */
    uint32_t end_frame;
    uint32_t reserved0;

    double grid_reference_point[3];
    double gravity[3];
    double fps, cfl, cfl_elastic, flip_ratio, rho, gravity_time;
    double Lx, Ly, Lz;
    double particle_mass, particle_volume, dx;
    double E, nu, stress_tolerance, xi;
    double q_max, q_min, p_min;           /* von Mises        */
    double M, q_cohesion;                 /* Drucker-Prager   */
    double perzyna_exp, perzyna_visc;     /* Perzyna          */
    double beta, p0;                      /* Modified Cam-Clay */
    double rho_s, grain_diameter, I_ref, mu_1, mu_2; /* mu(I) */
} mpm_config;

/* Axis-aligned plate: ObjectPlate (src/objects/object_plate.hpp:13-28 3D, :83-96 2D). */
typedef struct mpm_plate {
    double pos_object;
    int32_t plate_type;   /* enum mpm_plate_type */
    int32_t bc;           /* enum mpm_bc         */
    double friction;
    double pos_lower, pos_upper;
    double v_object[3];   /* vx_object, vy_object, vz_object */
    double vmin_factor, load_factor;
} mpm_plate;

/* Analytic object (tagged union over the ObjectGeneral subclasses). */
typedef struct mpm_object {
    int32_t kind;          /* enum mpm_object_kind */
    int32_t bc;            /* enum mpm_bc          */
    double friction;
    double param;          /* GROUND: y_ground; GATE: height; SILO: cut; otherwise unused */
    double center[3];      /* BOX_ROTATED */
    double half_extents[3];/* BOX_ROTATED */
    double rotation[9];    /* BOX_ROTATED, dim x dim column-major (first dim*dim entries used) */
} mpm_object;

/* Level-set boundary object sampled on a regular lattice: the data of ObjectVdb (src/objects/object_vdb.hpp:17-94),
 * whose OpenVDB grid a caller copies to a dense array (e.g. openvdb::tools::copyToDense over its active bounding
 * box). Lattice point (i, j, k) sits at origin + spacing * (i, j, k) and holds phi[(i*n[1] + j)*n[2] + k], the
 * signed distance (negative inside); n[2] = 1 in 2D. */
typedef struct mpm_sdf {
    int32_t bc;            /* enum mpm_bc */
    int32_t reserved;
    double friction;
    double origin[3];
    double spacing;
    uint32_t n[3];         /* lattice points per axis, at least 2 (n[2] = 1 in 2D) */
    uint32_t reserved1;
    const float* phi;      /* n[0]*n[1]*n[2] values, copied by the call */
} mpm_sdf;

/* Host views of particle state (Particles, src/data_structures.hpp:8-47). Any pointer may be NULL
 * to skip that field. Sizes: x, v: n*dim; F, Bmat: n*dim*dim; scalars: n. */
typedef struct mpm_particle_arrays {
    double* x;
    double* v;
    double* F;
    double* Bmat;
    double* eps_pl_dev;
    double* eps_pl_vol;
    double* eps_pl_vol_pradhana;
    double* delta_gamma;
    double* viscosity;
    double* muI;
} mpm_particle_arrays;

/* Result of mpm_step: the state simulate() keeps between steps (simulation.cpp:175-200). */
typedef struct mpm_step_info {
    double dt;              /* dt of the last step taken                          */
    double time;            /* simulation time after the last step                */
    uint64_t step;          /* current_time_step after the last step              */
    uint32_t frame;         /* frame counter after the last step                  */
    uint32_t plastic_count; /* "Proj: n / Np" of the last step (deformation_update.cpp:50-51) */
    uint32_t steps_done;    /* steps executed by this call                        */
    uint32_t status;        /* MPM_ERR_* / MPM_WARN_* bits raised by this call    */
    int32_t frame_ended;    /* the last step completed a frame                    */
    int32_t finished;       /* time reached end_frame/fps (simulation.cpp:191)    */
} mpm_step_info;

/* Geometry of the dense node grid of the last step (remesh.cpp:282-301). */
typedef struct mpm_grid_info {
    uint32_t Nx, Ny, Nz;    /* Nz = 1 in 2D */
    uint32_t grid_nodes;
    double xc, yc, zc;      /* origin = first node coordinate per axis */
} mpm_grid_info;

typedef struct mpm_context mpm_context;

/* Fills the reference's default member values (simulation.hpp:28-106). Replaces the
 * default member initialisers of class Simulation. */
void mpm_config_defaults(mpm_config* cfg);

/* Creates a context on the current HIP device and derives the constants that simulate()
 * computes before its loop (simulation.cpp:94-145: apicDinverse, lambda, mu, K, wave speed,
 * dt_max, q/d/e/f/rma prefactors, fac_Q). Replaces `Simulation sim; sim.<fields> = ...;`
 * plus the set-up part of Simulation::simulate(). */
int mpm_create(const mpm_config* cfg, mpm_context** out);

/* Releases all device memory of the context. Replaces Simulation::~Simulation (simulation.hpp:26). */
void mpm_destroy(mpm_context* ctx);

/* Appends a plate. Replaces `plates.push_back(std::make_unique<ObjectPlate>(...))`
 * (simulation.hpp:109, object_plate.hpp:13-28). Plates are applied in insertion order. */
int mpm_add_plate(mpm_context* ctx, const mpm_plate* plate);

/* Appends an analytic object. Replaces `objects.push_back(std::make_unique<Object...>(...))`
 * (simulation.hpp:110, object_general.hpp:8-23). Objects are applied in insertion order, before plates. */
int mpm_add_object(mpm_context* ctx, const mpm_object* object);

/* Appends a level-set object. Replaces `objects.push_back(std::make_unique<ObjectVdb>(file, bc, friction))`
 * (simulation.hpp:110, object_vdb.hpp:26-42); applied in insertion order with the other objects. As ObjectVdb:
 * inside(X) is phi(X) <= 0 with phi interpolated trilinearly (bilinearly in 2D) between lattice points (the
 * BoxSampler of object_vdb.hpp:44-56); normal(X) is the likewise interpolated gradient of phi, taken by central
 * differences at the lattice points (one-sided at the border; openvdb::tools::Gradient), normalised, or zero
 * where it vanishes (:58-77). Outside the lattice's box the object is absent (inside false). The lattice is
 * copied to the device and held by the context. MPM_ERR_INVALID_ARGUMENT for a bad bc, spacing or size. */
int mpm_add_sdf_object(mpm_context* ctx, const mpm_sdf* sdf);

/* Uploads particle state (Np = n). Fields passed as NULL take the values of the reference's
 * Particles constructor (data_structures.hpp:10-27: zeros, F = I). Replaces writing
 * `sim.Np`, `sim.particles` before simulate(). If use_mibf, muI is then filled as simulate()
 * does (simulation.cpp:138-145). Resets time, frame and step to zero. */
int mpm_set_particles(mpm_context* ctx, uint32_t n, const mpm_particle_arrays* in);

/* Advances up to nsteps steps. Each step is one iteration of simulate()'s loop
 * (simulation.cpp:175-200): advanceStep() (simulation.cpp:219-273), then time += dt,
 * current_time_step++, and the frame bookkeeping. Stops early at the end of the simulation
 * (final_time), after an exit-class error (the failing step completes, time does not
 * advance, as simulation.cpp:175-177), or, with MPM_STEP_STOP_AT_FRAME_END, after the step
 * that completes a frame. Returns the OR of MPM_ERR_* bits raised (warnings excluded). */
int mpm_step(mpm_context* ctx, uint32_t nsteps, uint32_t flags, mpm_step_info* info);

/* mpm_step without device allocation. No counterpart in the reference: for callers that cannot allocate
 * device memory while they step (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same steps, results and return value as mpm_step, but the grid, block arrays and staging
 * tiles of the steps live in a static device array of MPM_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends) instead of the context's own grid, and no hipMalloc, hipFree or allocating library call is
 * made. The particle state is the context's (allocated by mpm_set_particles). A step whose grid (grown by a
 * margin of 8 steps) does not fit returns MPM_ERR_SCRATCH_LIMIT before it runs; in 3D with the quadratic
 * spline the array holds a grid of about 60 nodes per axis (FP64) or 75 (split), margin included. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race (one host thread per GPU), and mpm_get_grid after this call reads the array, whose
 * contents the next mpm_step_noalloc on that GPU replaces. A later mpm_step returns to the context's own grid. */
#define MPM_NOALLOC_SCRATCH_BYTES (64u << 20)
int mpm_step_noalloc(mpm_context* ctx, uint32_t nsteps, uint32_t flags, mpm_step_info* info);

/* Downloads particle state in the order of mpm_set_particles. Replaces reading
 * `sim.particles` (tests/tests.cpp:49,141-143,359). */
int mpm_get_particles(mpm_context* ctx, mpm_particle_arrays* out);

/* Number of particles of the context. Replaces reading `sim.Np`. */
uint32_t mpm_num_particles(const mpm_context* ctx);

/* Downloads the dense grid of the last step in the reference's row-major node order
 * ind(i,j,k) = (i*Ny + j)*Nz + k (simulation.hpp:197-203): node coordinates
 * (x[Nx], y[Ny], z[Nz]), mass[grid_nodes], v and flip [grid_nodes*dim], friction[grid_nodes]
 * (use_mibf only). flip is formed only when something reads it (flip_ratio >= -1, or a step run with
 * MPM_STEP_CHECK_CONSERVATION); otherwise it is returned as zeros. Any output pointer may be NULL.
 * Call first with all NULL to get the sizes. */
int mpm_get_grid(mpm_context* ctx, mpm_grid_info* info, double* x, double* y, double* z,
                 double* mass, double* v, double* flip, double* friction);

/* Downloads plate state after moveObjects (simulation.cpp:277-281): pos_object, pos_lower,
 * pos_upper and v_object per plate, in insertion order. */
int mpm_get_plates(mpm_context* ctx, mpm_plate* plates, uint32_t capacity, uint32_t* count);

/* Cell binning of the current particle positions in the grid frame of the last step, in the order of
 * mpm_set_particles: cell_base[n*dim] = floor((x - origin) / dx) - 1 per axis, the first node of the
 * 4-node window of p2g.cpp:21-24 / g2p.cpp:29-32 (before the P2G clamp at 0); block_key[n] = the
 * engine's sort key, (bx*nby + by)*nbz + bz with b = w / 4 per axis (3D), or bx*nby + by with
 * b = w / 8 (2D), where w = max(cell_base, 0) is moved up by one node for the quadratic spline when
 * that node's weight is exactly zero (w is then the first node of the non-zero window).
 * Either pointer may be NULL. The reference has no binning; the cell index replaces the
 * i_base/j_base/k_base expressions of the lines cited above. */
int mpm_get_binning(mpm_context* ctx, int32_t* cell_base, uint32_t* block_key);

/* Per-particle output quantities of Simulation::saveParticleData (save_data.cpp:9-38), computed on the
 * device in the order of mpm_set_particles: p = -tr(tau)/dim, q = sqrt(3/2 tau_dev:tau_dev) if use_mises_q
 * else sqrt(1/2 tau_dev:tau_dev), Je = det(F), with tau = HenckyPiola(F) F^T (NeoHookeanPiola for the
 * Neo-Hookean model). Any pointer may be NULL. */
int mpm_get_stress_measures(mpm_context* ctx, double* p, double* q, double* Je);

/* Writes the particle file of Simulation::saveParticleData (save_data.cpp:7-171): binary little-endian
 * PLY, element "vertex" with float64 properties x y [z] vx vy [vz] p q Je, then, when a plastic model is
 * set, eps_pl_vol eps_pl_dev delta_gamma, then muI (DPVisc or MCCVisc with use_mibf, DPMui, MCCMui) and
 * viscosity (DPMui, MCCMui), one row per particle in the order of mpm_set_particles. */
int mpm_write_particles_ply(mpm_context* ctx, const char* path);

/* Writes the grid file of Simulation::saveGridData (save_data.cpp:173-259): binary little-endian PLY,
 * float64 properties x y [z] vx vy [vz] mass [friction (use_mibf)], one row per node of the last step's
 * grid in the order ind(i,j,k) (node coordinates from grid.x/y/z). */
int mpm_write_grid_ply(mpm_context* ctx, const char* path);

/* Accumulated device time of one kernel of the step pipeline. */
typedef struct mpm_kernel_timing {
    char name[32];       /* kernel name, e.g. "k_p2g" */
    double total_ms;     /* sum of HIP-event durations of its launches */
    uint64_t launches;   /* launches that were timed */
} mpm_kernel_timing;

/* Enables (1) or disables (0) per-kernel HIP-event timing of mpm_step and resets the accumulators.
 * Replaces the phase timers of advanceStep (`timer t_p2g ...`, simulation.cpp:240-267, accumulated in
 * runtime_p2g/euler/g2p/defgrad, simulation.hpp:154-158) with one timer per kernel on the context's stream. */
int mpm_set_profiling(mpm_context* ctx, int enable);

/* Enables HIP-event timing for the one kernel called `name` (e.g. "k_defgrad") and resets the accumulators:
 * a timed run then carries two events per step instead of two per launch. No counterpart in the reference
 * (its phase timers are always on); for callers that measure one kernel inside a throughput run.
 * MPM_ERR_INVALID_ARGUMENT for an unknown name. */
int mpm_set_profiling_kernel(mpm_context* ctx, const char* name);

/* Enables (1, the default) or disables (0) HIP-graph replay of mpm_step batches. With it, a scene of at most
 * 262,144 particles stepped at least 16 steps per host synchronisation has each batch's launches captured once
 * and replayed (the results are those of ordinary launches, bit for bit); other calls launch as usual. No
 * counterpart in the reference (a CPU loop); for callers that compare the two launch paths or step while
 * another thread captures on the same device. */
int mpm_set_step_graphs(mpm_context* ctx, int enable);

/* Copies the accumulated timings (one entry per kernel that ran) into out[0..capacity). */
int mpm_get_kernel_times(mpm_context* ctx, mpm_kernel_timing* out, uint32_t capacity, uint32_t* count);

/* ---- Batched form: many independent scenes advanced together ------------------------------------------
 * The reference advances one Simulation at a time; independent instances share nothing
 * (tests/tests.cpp:313-465 runs three of them one after another). A batch advances many contexts
 * together: each kernel of the step is launched once for all member scenes that share dimension, spline
 * degree, precision form, transfer variant and material, and a chunk of steps is replayed as one HIP
 * graph. Every member keeps its own dt, grid, time, frame and stop conditions, so its results are those
 * of mpm_step on the same context (bit for bit: no step sum depends on another scene). None of these
 * calls has a counterpart in the reference. */
typedef struct mpm_batch mpm_batch;

typedef struct mpm_batch_options {
    uint32_t chunk_steps;   /* steps enqueued per host synchronisation (default 16)              */
    int32_t use_graph;      /* capture a chunk as a HIP graph and replay it (default 1)          */
} mpm_batch_options;

typedef struct mpm_batch_info {
    uint32_t n_scenes;       /* members of the batch                                          */
    uint32_t n_active;       /* members with steps left in their budget after the call        */
    uint32_t n_classes;      /* launch classes (distinct compile-time variants) of the members */
    uint32_t chunks;         /* chunks of steps launched by the call                          */
    uint32_t status;         /* batch-level MPM_ERR_* bits (HIP errors, out of memory)        */
    uint32_t reserved;
    uint64_t particle_steps; /* sum over members of particles x steps taken by the call       */
} mpm_batch_info;

/* Creates an empty batch on the current HIP device, with its own stream. */
int mpm_batch_create(const mpm_batch_options* opt, mpm_batch** out);

/* Destroys the batch; its members stay valid contexts (no longer members). */
void mpm_batch_destroy(mpm_batch* b);

/* Adds a context whose particles are set (mpm_set_particles) with a budget of nsteps steps. While
 * it is a member, mpm_step on it returns MPM_ERR_INVALID_ARGUMENT; mpm_get_particles/grid and
 * mpm_scene_info read its state after any mpm_batch_step. The context must outlive its membership. */
int mpm_batch_add(mpm_batch* b, mpm_context* scene, uint32_t nsteps);

/* Removes a member (e.g. a finished request); its state is kept. */
int mpm_batch_remove(mpm_batch* b, mpm_context* scene);

/* Sets the remaining step budget of a member. */
int mpm_batch_set_budget(mpm_batch* b, mpm_context* scene, uint32_t nsteps);

/* Advances every member by up to min(max_steps, its budget) steps; a member stops early on the
 * conditions of mpm_step (final time, an exit-class error, with MPM_STEP_STOP_AT_FRAME_END a frame
 * end), and a member that finished or failed gets budget 0. Returns the batch-level error bits;
 * per-member results come from mpm_scene_info. */
int mpm_batch_step(mpm_batch* b, uint32_t max_steps, uint32_t flags, mpm_batch_info* info);

/* Picks the chunk length (steps per host synchronisation and per graph) of the batch: for each candidate
 * in {4, 8, 16, 32} it advances the members by one warm-up chunk (graph capture) and two timed chunks,
 * within their budgets and at most max_steps steps in all, and keeps the length with the highest
 * particle-updates per second. The steps taken are ordinary steps of the members (results do not depend on
 * the chunk length). *chosen receives the length; *rate (optional) the measured particle-updates/s. */
int mpm_batch_autotune(mpm_batch* b, uint32_t max_steps, uint32_t* chosen, double* rate);

/* Result of the member's part of the last mpm_batch_step (as mpm_step's info) and its remaining budget. */
int mpm_scene_info(const mpm_context* scene, mpm_step_info* info, uint32_t* budget_left);

/* Device bytes held by a context (particles, sort arrays, grid, staging): the footprint that an
 * admission policy charges for it. */
uint64_t mpm_device_bytes(const mpm_context* ctx);

/* ---- Served form: a request loop with continuous batching ---------------------------------------------
 * Requests (a scene and a number of steps) arrive over time. The server admits a request when its
 * device footprint fits the HBM budget beside the resident scenes, prepares it (context, upload of the
 * particles) on a host thread while the resident scenes keep stepping on the device, adds it to the
 * running batch at the next chunk boundary, and retires it when its steps are done. No counterpart in
 * the reference (one Simulation per process run). */
typedef struct mpm_server mpm_server;

typedef struct mpm_server_options {
    uint64_t hbm_budget_bytes; /* admission limit on the summed footprint of resident scenes (0: 80% of free memory) */
    uint32_t max_scenes;       /* resident scenes at most (0: 256)                                           */
    uint32_t chunk_steps;      /* steps between admission points (0: 8)                                       */
    int32_t use_graph;         /* replay chunks as HIP graphs                                                 */
    int32_t keep_results;      /* keep each request's final particle state for mpm_server_result              */
} mpm_server_options;

typedef struct mpm_request_desc {
    const mpm_config* cfg;
    const mpm_plate* plates;
    uint32_t n_plates;
    uint32_t n_objects;
    const mpm_object* objects;
    const mpm_particle_arrays* particles; /* copied by mpm_server_submit */
    uint32_t n;                           /* particles */
    uint32_t nsteps;                      /* steps the request asks for */
    double arrival_s;                     /* arrival, seconds after the start of mpm_server_run */
} mpm_request_desc;

typedef struct mpm_server_stats {
    uint32_t submitted, completed, failed, peak_resident;
    uint64_t peak_resident_bytes;
    uint64_t particle_steps;  /* particle-updates done by the run                                */
    double wall_s;            /* duration of the run                                             */
    double throughput;        /* particle_steps / wall_s                                         */
    double step_ms_p50, step_ms_p99;        /* per step of a resident request: chunk time / chunk steps */
    double request_ms_p50, request_ms_p99;  /* completion - arrival                              */
    double queue_ms_p50, queue_ms_p99;      /* admission (resident on the device) - arrival      */
} mpm_server_stats;

int mpm_server_create(const mpm_server_options* opt, mpm_server** out);
void mpm_server_destroy(mpm_server* s);

/* Queues a request (its arrays are copied); *id numbers requests from 0 in submission order. */
int mpm_server_submit(mpm_server* s, const mpm_request_desc* req, uint64_t* id);

/* Runs the request loop until every submitted request has completed or failed, or time_limit_s has
 * passed; requests arrive at their arrival_s after the start of the call. */
int mpm_server_run(mpm_server* s, double time_limit_s, mpm_server_stats* stats);

/* Final state of a completed request (keep_results), its last step info and its latency
 * (completion - arrival, ms). Any output may be NULL. */
int mpm_server_result(mpm_server* s, uint64_t id, mpm_particle_arrays* out, mpm_step_info* info, double* latency_ms);

/* Human-readable name of the lowest set status bit. */
const char* mpm_status_string(uint32_t status);

/* MPM_ABI_VERSION of the library. */
int mpm_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MATTER_MPM_H */
