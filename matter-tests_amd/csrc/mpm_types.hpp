// mpm_types.hpp — plain-old-data records shared by the host side of the engine (context, C ABI) and the
// gfx950 kernels. No HIP types appear here, so the header compiles in both passes.
#pragma once
#include <stdint.h>

namespace mpm {

constexpr int kMaxPlates = 16;
constexpr int kMaxObjects = 16;

// Plate state (ObjectPlate, src/objects/object_plate.hpp:164-188). Mutable: moveObjects() advances it.
struct DevPlate {
    double pos_object, pos_lower, pos_upper;
    double v_object[3], v_object_original[3];
    double friction, vmin_factor, load_factor;
    int32_t plate_type, bc;
};

// Sampled level set of an MPM_OBJECT_SDF object: per lattice point (i, j, k), at index (i*n[1] + j)*n[2] + k,
// four floats: phi, dphi/dx, dphi/dy, dphi/dz (gradient by central differences, one-sided at the border).
struct DevSdf {
    const float* field;
    uint32_t n[3];  // points per axis (n[2] = 1 in 2D)
    uint32_t pad;
};

// Boundary object (ObjectGeneral subclasses, src/objects/object_*.hpp, or a sampled level set). Immutable.
// The kind-specific data share storage so that Params (passed to every kernel by value) stays small.
struct DevObject {
    int32_t kind, bc;
    double friction, param;  // param: GROUND y, GATE height, SILO cut, SDF lattice spacing
    double center[3];        // BOX_ROTATED centre; SDF position of lattice point (0, 0, 0)
    union {
        struct {
            double half_extents[3];
            double rot[9];   // dim x dim column-major
        };
        DevSdf sdf;
    };
};

// Constants of a context: the simulate() set-up values (src/simulation/simulation.cpp:94-145) and the
// material/boundary parameters (src/simulation/simulation.hpp:28-110). Passed to kernels by value.
struct Params {
    int32_t dim, spline_degree, elastic_model, plastic_model, hardening_law;
    int32_t use_pradhana, use_mises_q, use_mibf, use_musl, gravity_special, change_particle_positions;
    int32_t pbc;     // periodic in x (pbc.cpp, remesh.cpp:14-36, position_update.cpp:26-33)
    int32_t n_objects, n_plates;
    uint32_t end_frame;
    double dx, one_over_dx, apicDinverse, particle_mass, particle_volume;
    double cfl, dt_max, frame_dt, final_time, wave_speed, flip_ratio, gravity_time;
    double gravity_final[3], grid_reference_point[3];
    double L[3];     // Lx, Ly, Lz: the periodic interval [0, Lx) and the fixed grid's extent
    double E, nu, mu, lambda, K, fac_Q, stress_tolerance, xi;
    double q_prefac, d_prefac, e_mu_prefac, f_mu_prefac, rma_prefac;
    double q_max, q_min, p_min, M, q_cohesion, perzyna_exp, perzyna_visc, beta, p0;
    double rho_s, grain_diameter, I_ref, mu_1, mu_2;
    double svd_tol;  // Jacobi SVD: rotate while an off-diagonal entry exceeds svd_tol * max |diagonal|
    DevObject objects[kMaxObjects];
};

// Partial reductions of the particle kernels. Each workgroup folds its result into slot
// (blockIdx % kRedSlots) so that no single address takes one atomic per wavefront of the launch;
// k_setup and k_finish fold the slots into StepState and reset them.
constexpr int kRedSlots = 128;
struct RedSlot {
    unsigned long long v2;  // max |v|^2 bits
    long long mn[3], mx[3]; // order-preserving keys of the coordinate bounds
    uint32_t plastic, pad;
};

// Device-resident loop state of simulate() plus the per-step grid geometry and reductions.
// Written by the single-workgroup set-up/finish kernels, read by every other kernel.
struct StepState {
    // loop state (simulation.hpp:151-153, simulation.cpp:154-156)
    double time, dt;
    uint64_t step;
    uint32_t frame, pad0;
    double gravity[3];
    // grid geometry of the current step (remesh.cpp:282-301)
    double origin[3];
    uint32_t n[3], grid_nodes;
    // remeshFixedInit bookkeeping (remesh.cpp:106-131)
    double min_init[3], max_init[3], low_init[3], high_init[3];
    // reductions over the particles of the previous step's end (update_dt.cpp:7-12, remesh.cpp:186-223)
    unsigned long long vmax2_bits;   // max |v|^2 as the bit pattern of a non-negative double
    long long xmin_key[3], xmax_key[3]; // order-preserving int64 keys of min/max coordinate
    uint32_t plastic_count, err, warn;
    uint32_t stop;          // nonzero: remaining kernels of this batch are no-ops
    uint32_t need_realloc;  // the grid of the next step exceeds capacity
    uint32_t steps_done, frame_ended, finished;
    uint32_t step_limit;    // stop once steps_done reaches it (a scene's step budget in a batch)
    uint32_t n_ghosts;      // periodic copies binned this step (k_bin_scan)
    uint32_t stop_at_frame_end, check_conservation;
    uint32_t cap_axis, cap_nodes;
    uint32_t xsel;          // which particle-state buffer is current (toggles per executed step)
    uint32_t nblocks;       // blocks of the counting sort (BlockEdge^dim nodes each)
    uint32_t nb[3];         // blocks per axis
    uint32_t nb_m[3], nb_l[3];  // division by nb[d] as a multiply (block_div_magic)
    uint32_t cap_blocks;
    // conservation sums (simulation.cpp:292-333), accumulated after P2G when requested
    double cons_grid_mass, cons_grid_mom[3], cons_part_mom[3];
    DevPlate plates[kMaxPlates];
    RedSlot red[kRedSlots];
};

}  // namespace mpm
