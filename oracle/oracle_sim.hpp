// oracle_sim.hpp — TEST INFRASTRUCTURE ONLY (the CPU oracle). A plain FP64 restatement of the reference's
// Simulation step, used by tests/ as the checker and by bench.py as the timed CPU baseline. The product
// never includes, links or calls this file.
//
// Follows, line by line (reference paths relative to the repository root of twmeehan/matter-tests):
//   src/simulation/simulation.hpp:23-238, src/simulation/simulation.cpp:73-333,
//   src/simulation/{update_dt,remesh,p2g,explicit_euler_update,boundary_collision,g2p,musl,
//                   deformation_update,plasticity,position_update}.cpp,
//   src/plasticity_helpers/{mcc_rma_explicit,mcc_rma_implicit_exponential_onevar,
//                   mcc_rma_implicit_sinh_onevar,limited_search_exponential,limited_search_sinh}.hpp,
//   src/objects/{object_plate,object_general,object_box,object_bump,object_ground,object_ramp,
//                object_gate,object_silo,object_curve}.hpp,
//   src/simulation/save_data.cpp:264-292 (computeAvgData).
// The OpenMP structure of the reference (per-thread dense grid copies merged in a critical section,
// per-thread Np-sized G2P buffers) is kept, so the timed baseline does the same work as the reference.
// With n_threads = 1 every result is deterministic.
//
// Deliberate differences from the reference, each forced by a defect that must not be inherited:
//   * debug() messages become error bits (MPM_ERR_* of include/mpm.h); `exit` is their OR.
//   * ObjectGround::normal sets n(2)=0 in 3D (the reference leaves it uninitialised, object_ground.hpp:29-34).
//   * Stencil nodes outside the grid raise MPM_ERR_GRID_BOUNDS instead of reading out of bounds.
//   * Unqualified abs() on doubles (mcc_rma_explicit.hpp:71, mcc_rma_implicit_exponential_onevar.hpp:96)
//     is std::abs(double), the overload the reference binds to through Eigen's SSE headers.
#pragma once
#include "oracle_math.hpp"
#include "../include/mpm.h"
#include <chrono>
#include <memory>
#include <string>
#include <cstdio>
#include <omp.h>

namespace oracle {

enum class PlateType { top, bottom, left, right, front, back };                 // tools.hpp:53
enum class ElasticModel { Hencky, NeoHookean };                                  // tools.hpp:54
enum class PlasticModel { NoPlasticity, VM, DP, DPSoft, MCC, VMVisc, DPVisc, MCCVisc, DPMui, MCCMui, Snow };  // tools.hpp:55
enum class HardeningLaw { NoHard, ExpoExpl, ExpoImpl, SinhExpl, SinhImpl };      // tools.hpp:56
enum class BC { NoSlip, SlipStick, SlipFree };                                   // tools.hpp:57

// ---------------------------------------------------------------------------------------------
// Boundary objects
// ---------------------------------------------------------------------------------------------
// ObjectGeneral, src/objects/object_general.hpp:8-23
template <int D> struct ObjectGeneral {
    BC bc; T friction; std::string name;
    ObjectGeneral(BC b, T f, std::string n = "") : bc(b), friction(f), name(n) {}
    virtual ~ObjectGeneral() {}
    virtual bool inside(const Vec<D>& X) const = 0;
    virtual Vec<D> normal(const Vec<D>& X) const = 0;
};

// ObjectBoxRotated, src/objects/object_box.hpp:6-44
template <int D> struct ObjectBoxRotated : ObjectGeneral<D> {
    Vec<D> center, half_extents; Mat<D> rotation;
    ObjectBoxRotated(BC b, T f, const Vec<D>& c, const Vec<D>& h, const Mat<D>& R)
        : ObjectGeneral<D>(b, f), center(c), half_extents(h), rotation(R) {}
    bool inside(const Vec<D>& X) const override {
        Vec<D> local = matvec(transpose(rotation), X - center);
        for (int i = 0; i < D; ++i) if (std::abs(local(i)) > half_extents(i)) return false;
        return true;
    }
    Vec<D> normal(const Vec<D>& X) const override {
        Vec<D> local = matvec(transpose(rotation), X - center);
        int closest_axis = 0; T max_dist = 0.0;
        for (int i = 0; i < D; ++i) {
            T dist = std::abs(half_extents(i) - std::abs(local(i)));
            if (dist > max_dist) { max_dist = dist; closest_axis = i; }
        }
        Vec<D> ln = Vec<D>::Zero();
        ln(closest_axis) = (local(closest_axis) > 0) ? 1 : -1;
        return matvec(rotation, ln);
    }
};
// ObjectBump, src/objects/object_bump.hpp:8-42
template <int D> struct ObjectBump : ObjectGeneral<D> {
    ObjectBump(BC b, T f) : ObjectGeneral<D>(b, f) {}
    bool inside(const Vec<D>& X) const override { T x = X(0), y = X(1); T y_limit = 0.0475 / std::cosh(25 * (x - 0.43)); return y < y_limit; }
    Vec<D> normal(const Vec<D>& X) const override {
        T x = X(0); Vec<D> n = Vec<D>::Zero();
        T arg = 25 * (x - 0.43);
        T b_der = -25 * 0.0475 * std::tanh(arg) / std::cosh(arg);
        n(0) = -b_der; n(1) = 1;
        return normalized(n);
    }
};
// ObjectGround, src/objects/object_ground.hpp:8-38
template <int D> struct ObjectGround : ObjectGeneral<D> {
    T y_ground;
    ObjectGround(BC b, T f, T yg = 0) : ObjectGeneral<D>(b, f), y_ground(yg) {}
    bool inside(const Vec<D>& X) const override { return !(X(1) > y_ground); }
    Vec<D> normal(const Vec<D>&) const override { Vec<D> n = Vec<D>::Zero(); n(0) = 0; n(1) = 1; return n; }
};
// ObjectRamp, src/objects/object_ramp.hpp:8-41
template <int D> struct ObjectRamp : ObjectGeneral<D> {
    ObjectRamp(BC b, T f) : ObjectGeneral<D>(b, f) {}
    bool inside(const Vec<D>& X) const override { T x = X(0), y = X(1); T y_limit = 0.1 * std::tanh(-100 * x); return y < y_limit; }
    Vec<D> normal(const Vec<D>& X) const override {
        T x = X(0); Vec<D> n = Vec<D>::Zero();
        T tmp = std::cosh(-100 * x);
        T b_der = 0.1 * (-100) / (tmp * tmp);
        n(0) = -b_der; n(1) = 1;
        return normalized(n);
    }
};
// ObjectGate, src/objects/object_gate.hpp:8-43
template <int D> struct ObjectGate : ObjectGeneral<D> {
    T height;
    ObjectGate(BC b, T f, T h = 0.016) : ObjectGeneral<D>(b, f), height(h) {}
    bool inside(const Vec<D>& X) const override { T x = X(0), y = X(1); T y_limit = height + 100 * x * x; return y > y_limit; }
    Vec<D> normal(const Vec<D>& X) const override {
        T x = X(0); Vec<D> n = Vec<D>::Zero();
        T b_der = 2.0 * 100 * x;
        n(0) = -b_der; n(1) = 1;
        for (int i = 0; i < D; ++i) n(i) *= -1;
        return normalized(n);
    }
};
// ObjectSilo (3D only), src/objects/object_silo.hpp:8-56
template <int D> struct ObjectSilo : ObjectGeneral<D> {
    T cut;
    ObjectSilo(BC b, T f, T c = -1) : ObjectGeneral<D>(b, f), cut(c) {}
    bool inside(const Vec<D>& X) const override {
        T x = X(0), y = X(1), z = (D == 3) ? X(D - 1) : 0.0;
        if (y < cut) return false;
        T r_surface = std::tanh(y) + 1;
        T r_surface_sq = r_surface * r_surface;
        T r_point_sq = x * x + z * z;
        return !(r_point_sq < r_surface_sq);
    }
    Vec<D> normal(const Vec<D>& X) const override {
        Vec<D> n = Vec<D>::Zero();
        T x = X(0), y = X(1), z = (D == 3) ? X(D - 1) : 0.0;
        T theta = std::atan2(z, x);
        T sech_y = 1.0 / std::cosh(y);
        n(0) = -std::cos(theta); n(1) = sech_y * sech_y;
        if (D == 3) n(D - 1) = -std::sin(theta);
        return normalized(n);
    }
};
// ObjectCurve, src/objects/object_curve.hpp:8-39
template <int D> struct ObjectCurve : ObjectGeneral<D> {
    ObjectCurve(BC b, T f) : ObjectGeneral<D>(b, f) {}
    bool inside(const Vec<D>& X) const override { T x = X(0), y = X(1); return !(y > x * x); }
    Vec<D> normal(const Vec<D>& X) const override {
        T x = X(0); Vec<D> n = Vec<D>::Zero();
        n(0) = -2 * x; n(1) = 1;
        return normalized(n);
    }
};

// ObjectVdb, src/objects/object_vdb.hpp:17-94, on a dense lattice instead of an OpenVDB grid (OpenVDB is not
// available): point (i, j, k) at origin + h (i, j, k) holds phi; the gradient grid of openvdb::tools::Gradient is
// formed by central differences (one-sided at the lattice border); both are sampled by trilinear (2D: bilinear)
// interpolation, the BoxSampler of :44-77. Outside the lattice box the object is absent.
template <int D> struct ObjectSdf : ObjectGeneral<D> {
    Vec<D> origin; T h; unsigned n[3];
    std::vector<float> field;  // per point: phi, dphi/dx, dphi/dy, dphi/dz
    ObjectSdf(BC b, T f, const double* org, T spacing, const unsigned* nn, const float* phi) : ObjectGeneral<D>(b, f), h(spacing) {
        for (int d = 0; d < D; ++d) origin(d) = org[d];
        for (int d = 0; d < 3; ++d) n[d] = d < D ? nn[d] : 1;
        const size_t npts = (size_t)n[0] * n[1] * n[2];
        field.assign(4 * npts, 0.0f);
        for (unsigned i = 0; i < n[0]; ++i)
            for (unsigned j = 0; j < n[1]; ++j)
                for (unsigned k = 0; k < n[2]; ++k) {
                    const size_t q = ((size_t)i * n[1] + j) * n[2] + k;
                    field[4 * q] = phi[q];
                    const unsigned c[3] = {i, j, k};
                    for (int d = 0; d < D; ++d) {
                        unsigned a[3] = {i, j, k}, b[3] = {i, j, k};
                        a[d] = c[d] > 0 ? c[d] - 1 : c[d];
                        b[d] = c[d] + 1 < n[d] ? c[d] + 1 : c[d];
                        const double pa = phi[((size_t)a[0] * n[1] + a[1]) * n[2] + a[2]];
                        const double pb = phi[((size_t)b[0] * n[1] + b[1]) * n[2] + b[2]];
                        field[4 * q + 1 + d] = (float)((pb - pa) / ((double)(b[d] - a[d]) * h));
                    }
                }
    }
    bool sample(const Vec<D>& X, T& phi, T* g) const {
        int i0[3] = {0, 0, 0}; T fr[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
            const T u = (X(d) - origin(d)) / h;
            if (!(u >= 0.0 && u <= (T)(n[d] - 1))) return false;
            const int fi = (int)std::floor(u), last = (int)n[d] - 2;
            i0[d] = fi < last ? fi : last;
            fr[d] = u - (T)i0[d];
        }
        phi = 0; for (int d = 0; d < 3; ++d) g[d] = 0;
        for (int c = 0; c < (1 << D); ++c) {
            T w = 1.0; size_t idx = 0;
            for (int d = 0; d < D; ++d) {
                const int b = (c >> (D - 1 - d)) & 1;
                w = w * (b ? fr[d] : 1.0 - fr[d]);
                idx = idx * n[d] + (size_t)(i0[d] + b);
            }
            phi += w * (T)field[4 * idx];
            for (int d = 0; d < D; ++d) g[d] += w * (T)field[4 * idx + 1 + d];
        }
        return true;
    }
    bool inside(const Vec<D>& X) const override { T phi, g[3]; return sample(X, phi, g) && phi <= 0; }
    Vec<D> normal(const Vec<D>& X) const override {
        T phi, g[3]; Vec<D> r = Vec<D>::Zero();
        if (sample(X, phi, g)) for (int d = 0; d < D; ++d) r(d) = g[d];
        return normalized(r);
    }
};

// ObjectPlate, src/objects/object_plate.hpp:8-190
template <int D> struct ObjectPlate {
    T pos_object; PlateType plate_type; BC bc; T friction; T pos_upper; T pos_lower;
    T v_object[3]; T v_object_original[3]; T vmin_factor; T load_factor;
    ObjectPlate(T pos, PlateType pt, BC b = BC::NoSlip, T fr = 0, T lower = -1e15, T upper = 1e15,
                T vx = 0, T vy = 0, T vz = 0, T vmin = 1.0, T load = 0.0)
        : pos_object(pos), plate_type(pt), bc(b), friction(fr), pos_upper(upper), pos_lower(lower),
          vmin_factor(vmin), load_factor(load) {
        v_object[0] = vx; v_object[1] = vy; v_object[2] = vz;
        for (int i = 0; i < 3; ++i) v_object_original[i] = v_object[i];
    }
    // inside: 3D object_plate.hpp:30-78, 2D :98-131
    bool inside(const Vec<D>& X) const {
        if (plate_type == PlateType::left) return X(1) < pos_upper && X(1) > pos_lower && (X(0) - pos_object) <= 0;
        if (plate_type == PlateType::right) return X(1) < pos_upper && X(1) > pos_lower && (pos_object - X(0)) <= 0;
        if (plate_type == PlateType::bottom) return X(0) < pos_upper && X(0) > pos_lower && (X(1) - pos_object) <= 0;
        if (plate_type == PlateType::top) return X(0) < pos_upper && X(0) > pos_lower && (pos_object - X(1)) <= 0;
        if (D == 3 && plate_type == PlateType::back) return X(D - 1) - pos_object <= 0;
        if (D == 3 && plate_type == PlateType::front) return pos_object - X(D - 1) <= 0;
        return false;  // "FATAL: invalid plate type"
    }
    // move, object_plate.hpp:136-162
    void move(T dt, T frame_dt, T time) {
        T factor = (time < load_factor * frame_dt) ? (1.0 / vmin_factor) : 1.0;
        for (int i = 0; i < D; ++i) v_object[i] = v_object_original[i] * factor;
        if (plate_type == PlateType::left || plate_type == PlateType::right) {
            pos_object += dt * v_object[0]; pos_upper += dt * v_object[1]; pos_lower += dt * v_object[1];
        } else if (plate_type == PlateType::bottom || plate_type == PlateType::top) {
            pos_object += dt * v_object[1]; pos_upper += dt * v_object[0]; pos_lower += dt * v_object[0];
        } else if (D == 3 && (plate_type == PlateType::back || plate_type == PlateType::front)) {
            pos_object += dt * v_object[2];
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Particles, src/data_structures.hpp:8-47 ; Grid, src/data_structures.hpp:49-66
// ---------------------------------------------------------------------------------------------
template <int D> struct Particles {
    std::vector<Vec<D>> x, v, pic, flip;
    std::vector<T> eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana, delta_gamma, viscosity, muI;
    std::vector<Mat<D>> F, Bmat;
    explicit Particles(unsigned Np = 1) {
        x.assign(Np, Vec<D>::Zero()); v.assign(Np, Vec<D>::Zero());
        pic.assign(Np, Vec<D>::Zero()); flip.assign(Np, Vec<D>::Zero());
        eps_pl_dev.assign(Np, 0.0); eps_pl_vol.assign(Np, 0.0); eps_pl_vol_pradhana.assign(Np, 0.0);
        delta_gamma.assign(Np, 0.0); viscosity.assign(Np, 0.0); muI.assign(Np, 0.0);
        F.assign(Np, Mat<D>::Identity()); Bmat.assign(Np, Mat<D>::Zero());
    }
};
template <int D> struct Grid {
    std::vector<T> x, y, z;
    std::vector<Vec<D>> v, flip;
    std::vector<T> mass, friction;
    T xc = 0, yc = 0, zc = 0;
};

// ---------------------------------------------------------------------------------------------
// Simulation, src/simulation/simulation.hpp:23-225
// ---------------------------------------------------------------------------------------------
template <int D, int SPL> class Simulation {
public:
    typedef Vec<D> TV; typedef Mat<D> TM; typedef Spline<SPL> S;
    uint32_t exit = 0;  // OR of MPM_ERR_* bits; nonzero == the reference's exit = 1
    uint32_t warnings = 0;
    const unsigned dim = D;
    unsigned n_threads = 1;
    unsigned end_frame = 1;
    bool is_initialized = true;
    bool reduce_verbose = true;
    bool pbc = false, change_particle_positions = false, gravity_special = false;
    bool use_mibf = false, use_musl = false;
    TV grid_reference_point = TV::Ones() * 2e10;
    TV gravity = TV::Zero();
    T fps = 1, cfl = 0.5, cfl_elastic = 0.5, flip_ratio = -0.95, rho = 1000, gravity_time = 0;
    T Lx = 1, Ly = 1, Lz = 1;
    Particles<D> particles;
    unsigned Np = 0;
    T particle_mass = 0, particle_volume = 0, dx = 0;
    ElasticModel elastic_model = ElasticModel::Hencky;
    PlasticModel plastic_model = PlasticModel::NoPlasticity;
    HardeningLaw hardening_law = HardeningLaw::ExpoImpl;
    bool use_pradhana = true, use_mises_q = false;
    T E = 1e6, nu = 0.3, stress_tolerance = 1e-5, xi = 0;
    T q_max = 100, q_min = 100, p_min = -1e20;
    T M = 1, q_cohesion = 0;
    T perzyna_exp = 1, perzyna_visc = 0;
    T beta = 0, p0 = 1000;
    T rho_s = 2500, grain_diameter = 1e-3, I_ref = 0.279;
    T mu_1 = std::tan(20.9 * M_PI / 180.0), mu_2 = std::tan(32.8 * M_PI / 180.0);
    std::vector<std::unique_ptr<ObjectPlate<D>>> plates;
    std::vector<std::unique_ptr<ObjectGeneral<D>>> objects;

    // loop state of simulate() (simulation.hpp:151-153)
    unsigned long long current_time_step = 0;
    unsigned frame = 0;
    T time = 0;
    unsigned last_plastic_count = 0;
    bool check_conservation = false;

    T final_time = 0, frame_dt = 0, dt = 0, dt_max = 5e-3, wave_speed = 0;
    T mu = 0, lambda = 0, K = 0, fac_Q = 0;
    T q_prefac = 0, d_prefac = 0, e_mu_prefac = 0, f_mu_prefac = 0, rma_prefac = 0;
    T one_over_dx = 0, one_over_dx_square = 0, apicDinverse = 0;
    TV gravity_final = TV::Zero();
    Grid<D> grid;
    unsigned Nx = 0, Ny = 0, Nz = 1, grid_nodes = 0;
    T min_init[3] = {0, 0, 0}, max_init[3] = {0, 0, 0}, low_init[3] = {0, 0, 0}, high_init[3] = {0, 0, 0};
    unsigned num_add_pbc_particles = 0;  // simulation.hpp:68
    // phase timers of advanceStep (simulation.hpp:154-157, timer.hpp), in seconds
    double runtime_p2g = 0, runtime_euler = 0, runtime_g2p = 0, runtime_defgrad = 0;
    static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

    inline unsigned ind(unsigned i, unsigned j, unsigned k) const { return (i * Ny + j) * Nz + k; }  // simulation.hpp:197-199
    inline unsigned ind(unsigned i, unsigned j) const { return (i * Ny + j); }                         // simulation.hpp:201-203

    T calculateBulkModulus() const { return lambda + 2.0 * mu / dim; }  // simulation.cpp:288-290

    // HenckyPiola, simulation.hpp:231-238
    TM HenckyPiola(const TM& Fe) const {
        TM U, V; TV sigma; svd(Fe, U, sigma, V);
        TV logS, invS;
        for (int i = 0; i < D; ++i) { logS(i) = std::log(std::abs(sigma(i))); invS(i) = 1.0 / sigma(i); }
        T trlog = vsum(logS);
        TV diag; for (int i = 0; i < D; ++i) diag(i) = 2 * mu * invS(i) * logS(i) + lambda * trlog * invS(i);
        return udvt(U, diag, V);
    }
    // NeoHookeanPiola, simulation.hpp:227-229
    TM NeoHookeanPiola(const TM& Fe) const {
        TM FinvT = transpose(inverse(Fe));
        return mu * (Fe - FinvT) + (lambda * std::log(determinant(Fe))) * FinvT;
    }

    // Set-up part of simulate(), simulation.cpp:80-157
    bool prepare() {
        if (elastic_model != ElasticModel::Hencky && plastic_model != PlasticModel::NoPlasticity) return false;
        apicDinverse = S::apic_D_factor / (dx * dx);
        lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
        mu = E / (2.0 * (1.0 + nu));
        K = calculateBulkModulus();
        wave_speed = std::sqrt(E / rho);
        dt_max = cfl_elastic * dx / wave_speed;
        frame_dt = 1.0 / fps;
        gravity_final = gravity;
        one_over_dx = 1.0 / dx;
        one_over_dx_square = one_over_dx * one_over_dx;
        if (use_mises_q) { q_prefac = std::sqrt(3.0) / std::sqrt(2.0); d_prefac = std::sqrt(2.0) / std::sqrt(3.0); }
        else { q_prefac = 1.0 / std::sqrt(2.0); d_prefac = std::sqrt(2.0); }
        e_mu_prefac = 2 * q_prefac * mu;
        f_mu_prefac = 2 * q_prefac / d_prefac * mu;
        rma_prefac = 2 * q_prefac * q_prefac;
        fac_Q = I_ref / (grain_diameter * std::sqrt(rho_s));
        if (use_mibf) {
            T fill = (plastic_model == PlasticModel::DPMui || plastic_model == PlasticModel::MCCMui) ? mu_1 : M;
            std::fill(particles.muI.begin(), particles.muI.end(), fill);
        }
        time = 0; frame = 0;
        final_time = end_frame * frame_dt;
        return true;
    }

    // One iteration of the frame loop of simulate(), simulation.cpp:175-200 (without I/O).
    // Returns false when the loop would stop (exit set, or final time reached).
    bool loopIteration(bool& frame_ended) {
        const T time_tolerance = 1e-15;  // simulation.cpp:158
        frame_ended = false;
        advanceStep();
        if (exit) return false;
        time += dt;
        current_time_step++;
        if (frame_dt * (frame + 1) - time < time_tolerance) { frame++; frame_ended = true; }
        if (std::abs(final_time - time) < time_tolerance || final_time < time) return false;
        return true;
    }

    // simulate(), simulation.cpp:73-213 (frame loop; no file output)
    void simulate() {
        if (!prepare()) return;
        while (frame < end_frame) { bool fe; if (!loopIteration(fe)) break; }
    }

    // advanceStep(), simulation.cpp:219-273
    void advanceStep() {
        updateDt();
        if (pbc) {
            if (current_time_step == 0) remeshFixed(4);
        } else {
            if (current_time_step == 0) remeshFixedInit(2, 2, 2); else remeshFixedCont();
        }
        if (pbc) PBCAddParticles(4);
        resizeGrid();
        double t = now_s();
        P2G();
        runtime_p2g += now_s() - t;
        if (check_conservation) { checkMassConservation(); checkMomentumConservation(); }
        t = now_s();
        explicitEulerUpdate();
        runtime_euler += now_s() - t;
        if (pbc) PBCDelParticles();
        t = now_s();
        G2P();
        runtime_g2p += now_s() - t;
        if (use_musl) MUSL();
        t = now_s();
        deformationUpdate();
        runtime_defgrad += now_s() - t;
        positionUpdate();
        moveObjects();
    }

    // updateDt, update_dt.cpp:5-66
    void updateDt() {
        unsigned imax = 0; T best = 0;
        for (unsigned p = 0; p < Np; ++p) { T s = sqnorm(particles.v[p]); if (p == 0 || best < s) { best = s; imax = p; } }
        T max_speed = Np ? vnorm(particles.v[imax]) : 0.0;
        if (max_speed >= wave_speed) warnings |= MPM_WARN_SUPERSONIC;
        if (std::abs(max_speed) > 1e-10) { T dt_cfl = cfl * dx / max_speed; dt = std::min(dt_cfl, dt_max); }
        else dt = dt_max;
        dt = std::min(dt, frame_dt * (frame + 1) - time);
        dt = std::min(dt, final_time - time);
        if (gravity_special) {
            if (time < gravity_time) gravity = gravity_final * time / gravity_time;
            else gravity = gravity_final;
        }
    }

    void particleBounds(T mn[3], T mx[3]) const {
        for (int d = 0; d < D; ++d) { mn[d] = particles.x[0](d); mx[d] = particles.x[0](d); }
        for (unsigned p = 1; p < Np; ++p)
            for (int d = 0; d < D; ++d) { T v = particles.x[p](d); if (v < mn[d]) mn[d] = v; if (mx[d] < v) mx[d] = v; }
    }
    std::vector<T>& axis(int d) { return d == 0 ? grid.x : (d == 1 ? grid.y : grid.z); }
    void setGridSize() {
        Nx = grid.x.size(); Ny = grid.y.size(); Nz = (D == 3) ? grid.z.size() : 1;
        grid_nodes = Nx * Ny * Nz;
        grid.xc = grid.x[0]; grid.yc = grid.y[0]; if (D == 3) grid.zc = grid.z[0];
    }
    // remeshFixed, remesh.cpp:14-36: the fixed grid of a periodic run, set once at the first step.
    void remeshFixed(unsigned extra_nodes) {
        grid.x = arange(-dx * (T)(1 + extra_nodes), Lx + (T)(2 + extra_nodes) * dx, dx);
        grid.y = arange(-dx, Ly + (T)(2 + extra_nodes) * dx, dx);
        if (D == 3) grid.z = arange(-dx * (T)(1 + extra_nodes), Lz + (T)(2 + extra_nodes) * dx, dx);
        setGridSize();
    }
    // PBCAddParticles, pbc.cpp:37-83: a particle within safety_factor*dx of x = Lx (strictly inside) gets a
    // copy at x - Lx, written -(Lx - x); otherwise one within that distance of x = 0 gets a copy at Lx + x.
    // The copies are appended in particle order (the reference's OpenMP loop appends in thread order).
    // The copy carries x, v, F and Bmat as in the reference, and also muI: the reference leaves muI
    // unextended and its MIBF P2G reads muI[p] past the end for the copies (p2g.cpp:46,61), so the copy of
    // the source value is this restatement's definition of that read.
    void PBCAddParticles(unsigned safety_factor) {
        num_add_pbc_particles = 0;
        const unsigned n0 = Np;
        const T band = (T)safety_factor * dx;
        for (unsigned p = 0; p < n0; ++p) {
            TV part_x = particles.x[p];
            T diff = Lx - part_x(0);
            bool add = false;
            if (diff < band && diff > 0) { part_x(0) = -diff; add = true; }
            else {
                diff = part_x(0);
                if (diff < band && diff >= 0) { part_x(0) = Lx + diff; add = true; }
            }
            if (!add) continue;
            particles.x.push_back(part_x);
            particles.v.push_back(particles.v[p]);
            particles.F.push_back(particles.F[p]);
            particles.Bmat.push_back(particles.Bmat[p]);
            particles.muI.push_back(particles.muI[p]);
            ++num_add_pbc_particles;
        }
        Np = (unsigned)particles.x.size();
    }
    // PBCDelParticles, pbc.cpp:85-101
    void PBCDelParticles() {
        const size_t k = num_add_pbc_particles;
        particles.x.resize(particles.x.size() - k);
        particles.v.resize(particles.v.size() - k);
        particles.F.resize(particles.F.size() - k);
        particles.Bmat.resize(particles.Bmat.size() - k);
        particles.muI.resize(particles.muI.size() - k);
        Np = (unsigned)particles.x.size();
    }
    // remeshFixedInit, remesh.cpp:38-182
    void remeshFixedInit(unsigned sfx, unsigned sfy, unsigned sfz) {
        T mn[3], mx[3]; particleBounds(mn, mx);
        if (grid_reference_point(0) < 1e10) {
            for (int d = 0; d < D; ++d) {
                if (grid_reference_point(d) < mn[d]) mn[d] = grid_reference_point(d);
                else if (grid_reference_point(d) > mx[d]) mx[d] = grid_reference_point(d);
            }
        }
        unsigned sf[3] = {sfx, sfy, sfz};
        for (int d = 0; d < D; ++d) {
            max_init[d] = mx[d]; min_init[d] = mn[d];
            low_init[d] = mn[d] - dx * sf[d];
            high_init[d] = mx[d] + dx * sf[d];
            axis(d) = arange(low_init[d], high_init[d] + dx, dx);
        }
        setGridSize();
    }
    // remeshFixedCont, remesh.cpp:184-311
    void remeshFixedCont() {
        T mn[3], mx[3]; particleBounds(mn, mx);
        for (int d = 0; d < D; ++d) {
            T high, low;
            if (mx[d] < max_init[d]) { unsigned rf = std::floor(std::max(0.0, (max_init[d] - mx[d]) / dx - 1e-8 * dx)); high = high_init[d] - rf * dx; }
            else { unsigned ef = std::floor(std::max(0.0, (mx[d] - max_init[d]) / dx - 1e-8 * dx)); high = high_init[d] + ef * dx; }
            if (mn[d] > min_init[d]) { unsigned rf = std::floor(std::max(0.0, (mn[d] - min_init[d]) / dx - 1e-8 * dx)); low = low_init[d] + rf * dx; }
            else { unsigned ef = std::floor(std::max(0.0, (min_init[d] - mn[d]) / dx - 1e-8 * dx)); low = low_init[d] - ef * dx; }
            axis(d) = arange(low, high + dx, dx);
        }
        setGridSize();
    }
    // resizeGrid, remesh.cpp:5-11
    void resizeGrid() {
        grid.v.assign(grid_nodes, TV::Zero());
        grid.flip.assign(grid_nodes, TV::Zero());
        grid.mass.assign(grid_nodes, 0.0);
        if (use_mibf) grid.friction.assign(grid_nodes, 0.0);
    }

    // Base index of the stencil: floor((x - origin)/dx) - 1 (p2g.cpp:21-24 clamps at 0, g2p.cpp:29-32 does not).
    // A NaN or out-of-range coordinate maps to an index far outside every grid, which stencilInside rejects
    // (the reference converts it unchecked and reads out of bounds).
    inline long baseIndex(T xp, T origin) const {
        const T f = std::floor((xp - origin) * one_over_dx);
        return f < 2.0e9 ? (f > -2.0e9 ? (long)f - 1 : -(1L << 40)) : (1L << 40);
    }
    inline bool stencilInside(long b[3]) const {
        long n[3] = {(long)Nx, (long)Ny, (long)Nz};
        for (int d = 0; d < D; ++d) if (b[d] < 0 || b[d] + 3 >= n[d]) return false;
        return true;
    }
    void stencilBase(const TV& xp, bool clamp, long b[3]) {
        T org[3] = {grid.xc, grid.yc, grid.zc};
        b[2] = 0;
        for (int d = 0; d < D; ++d) { b[d] = baseIndex(xp(d), org[d]); if (clamp && b[d] < 0) b[d] = 0; }
    }

    // P2G, p2g.cpp:6-100
    void P2G() {
        uint32_t err = 0;
        #pragma omp parallel num_threads(n_threads) reduction(|:err)
        {
            std::vector<TV> gv(grid_nodes, TV::Zero());
            std::vector<T> gm(grid_nodes, 0.0), gf(grid_nodes, 0.0);
            #pragma omp for
            for (long p = 0; p < (long)Np; ++p) {
                const TV xp = particles.x[p];
                long b[3]; stencilBase(xp, true, b);
                if (!stencilInside(b)) { err |= MPM_ERR_GRID_BOUNDS; continue; }
                for (long i = b[0]; i < b[0] + 4; ++i) {
                    T xi = grid.x[i];
                    for (long j = b[1]; j < b[1] + 4; ++j) {
                        T yi = grid.y[j];
                        if constexpr (D == 3) {
                            for (long k = b[2]; k < b[2] + 4; ++k) {
                                T zi = grid.z[k];
                                T weight = wip3<SPL>(xp(0), xp(1), xp(D - 1), xi, yi, zi, one_over_dx);
                                if (weight > 1e-25) {
                                    unsigned id = ind(i, j, k);
                                    gm[id] += weight;
                                    gv[id] += particles.v[p] * weight;
                                    if (flip_ratio < 0) {
                                        TV pd; pd(0) = xi - xp(0); pd(1) = yi - xp(1); pd(D - 1) = zi - xp(D - 1);
                                        gv[id] += matvec(particles.Bmat[p], pd) * apicDinverse * weight;
                                    }
                                    if (use_mibf) gf[id] += particles.muI[p] * weight;
                                }
                            }
                        } else {
                            T weight = wip2<SPL>(xp(0), xp(1), xi, yi, one_over_dx);
                            if (weight > 1e-25) {
                                unsigned id = ind(i, j);
                                gm[id] += weight;
                                gv[id] += particles.v[p] * weight;
                                if (flip_ratio < 0) {
                                    TV pd; pd(0) = xi - xp(0); pd(1) = yi - xp(1);
                                    gv[id] += matvec(particles.Bmat[p], pd) * apicDinverse * weight;
                                }
                                if (use_mibf) gf[id] += particles.muI[p] * weight;
                            }
                        }
                    }
                }
            }
            #pragma omp critical
            {
                for (unsigned l = 0; l < grid_nodes; ++l) {
                    grid.mass[l] += gm[l];
                    grid.v[l] += gv[l];
                    if (use_mibf) grid.friction[l] += gf[l];
                }
            }
        }
        exit |= err;
        // grid.mass holds m_i / m_p here (p2g.cpp:80-82)
        #pragma omp parallel for num_threads(n_threads)
        for (long l = 0; l < (long)grid_nodes; ++l) {
            T mi = grid.mass[l];
            if (mi > 0) grid.v[l] /= mi; else grid.v[l] = TV::Zero();
            if (use_mibf) grid.friction[l] /= mi;
        }
        for (auto& m : grid.mass) m *= particle_mass;
    }

    // checkMomentumConservation / checkMassConservation, simulation.cpp:292-333
    void checkMomentumConservation() {
        TV mg = TV::Zero();
        for (unsigned l = 0; l < grid_nodes; ++l) mg += grid.v[l] * grid.mass[l];
        TV mp = TV::Zero();
        for (unsigned p = 0; p < Np; ++p) mp += particle_mass * particles.v[p];
        if (vnorm(mg - mp) > 1e-10 * vnorm(mp)) exit |= MPM_ERR_CONSERVATION;
    }
    void checkMassConservation() {
        T pm = particle_mass * Np, gm = 0;
        for (auto m : grid.mass) gm += m;
        if (std::abs(gm - pm) > 1e-10 * pm) exit |= MPM_ERR_CONSERVATION;
    }

    TM kirchhoff(const TM& Fe) const {
        TM P = (elastic_model == ElasticModel::NeoHookean) ? NeoHookeanPiola(Fe) : HenckyPiola(Fe);
        return matmul(P, transpose(Fe));
    }

    // explicitEulerUpdate, explicit_euler_update.cpp:7-134
    void explicitEulerUpdate() {
        std::vector<TV> grid_force(grid_nodes, TV::Zero());
        #pragma omp parallel num_threads(n_threads)
        {
            std::vector<TV> gfl(grid_nodes, TV::Zero());
            #pragma omp for
            for (long p = 0; p < (long)Np; ++p) {
                TM tau = kirchhoff(particles.F[p]);
                const TV xp = particles.x[p];
                long b[3]; stencilBase(xp, true, b);
                if (!stencilInside(b)) continue;
                for (long i = b[0]; i < b[0] + 4; ++i) {
                    T xi = grid.x[i];
                    for (long j = b[1]; j < b[1] + 4; ++j) {
                        T yi = grid.y[j];
                        if constexpr (D == 3) {
                            for (long k = b[2]; k < b[2] + 4; ++k) {
                                T zi = grid.z[k];
                                unsigned id = ind(i, j, k);
                                if (grid.mass[id] > 0) {
                                    TV g = vecFrom3(grad_wip3<SPL>(xp(0), xp(1), xp(D - 1), xi, yi, zi, one_over_dx));
                                    gfl[id] += matvec(tau, g);
                                }
                            }
                        } else {
                            unsigned id = ind(i, j);
                            if (grid.mass[id] > 0) {
                                TV g = vecFrom2(grad_wip2<SPL>(xp(0), xp(1), xi, yi, one_over_dx));
                                gfl[id] += matvec(tau, g);
                            }
                        }
                    }
                }
            }
            #pragma omp critical
            { for (unsigned l = 0; l < grid_nodes; ++l) grid_force[l] += gfl[l]; }
        }
        T dt_particle_volume = dt * particle_volume;
        TV dt_gravity = dt * gravity;
        #pragma omp parallel for num_threads(n_threads)
        for (long l = 0; l < (long)grid_nodes; ++l) {
            unsigned index = (unsigned)l;
            T mi = grid.mass[index];
            if (mi > 0) {
                TV velocity_increment = ((-dt_particle_volume) * grid_force[index]) / mi + dt_gravity;
                TV old_vi = grid.v[index];
                TV new_vi = old_vi + velocity_increment;
                unsigned i = index / (Ny * Nz), j = (index / Nz) % Ny, k = index % Nz;
                TV Xi; Xi(0) = grid.x[i]; Xi(1) = grid.y[j]; if (D == 3) Xi(D - 1) = grid.z[k];
                uint32_t e = boundaryCollision(index, Xi, new_vi);
                if (e) {
                    #pragma omp atomic
                    exit |= e;
                }
                grid.v[index] = new_vi;
                grid.flip[index] = new_vi - old_vi;
            }
        }
    }
    static TV vecFrom3(const Vec<3>& g) { TV r; for (int i = 0; i < D; ++i) r(i) = g(i); return r; }
    static TV vecFrom2(const Vec<2>& g) { TV r = TV::Zero(); for (int i = 0; i < 2 && i < D; ++i) r(i) = g(i); return r; }

    // Plate friction in 3D (boundary_collision.cpp:103-119 and siblings): reduce the two tangential
    // components sequentially; the second line uses the already-updated first component.
    static void plateTangent3(T& a, T& b, T n_comp, T friction) {
        T vel_t = std::sqrt(a * a + b * b);
        T fric_vel_n = friction * std::abs(n_comp);
        if (friction == 0) {}
        else if (vel_t <= fric_vel_n) { a = 0; b = 0; }
        else {
            a = sgn(a) * (vel_t - fric_vel_n) / std::sqrt(1 + b * b / (a * a));
            b = sgn(b) * (vel_t - fric_vel_n) / std::sqrt(1 + a * a / (b * b));
        }
    }
    // Plate friction in 2D (boundary_collision.cpp:284-301): strict '<' stick test.
    static void plateTangent2(T& a, T n_comp, T friction) {
        if (friction == 0) {}
        else if (std::abs(a) < friction * std::abs(n_comp)) a = 0;
        else { if (a > 0) a -= friction * std::abs(n_comp); else a += friction * std::abs(n_comp); }
    }

    // boundaryCollision, boundary_collision.cpp:5-394
    uint32_t boundaryCollision(unsigned index, TV Xi, TV& vi) {
        TV vi_orig = vi;
        Xi += dt * vi;
        for (auto& obj : objects) {
            if (obj->inside(Xi)) {
                TV v_rel = vi_orig;
                if (obj->bc == BC::NoSlip) v_rel = TV::Zero();
                else if (obj->bc == BC::SlipFree) {
                    TV n = obj->normal(Xi);
                    T dt_ = dot(v_rel, n);
                    if (dt_ < 0) {
                        T friction = obj->friction;
                        if (use_mibf) friction = grid.friction[index];
                        TV v_tang = v_rel - dt_ * n;
                        if (friction > 0) {
                            if ((-dt_) * friction < vnorm(v_tang)) v_rel = v_tang + normalized(v_tang) * dt_ * friction;
                            else v_rel = TV::Zero();
                        } else v_rel = v_tang;
                    }
                } else if (obj->bc == BC::SlipStick) {
                    TV n = obj->normal(Xi);
                    T dt_ = dot(v_rel, n);
                    v_rel = v_rel - dt_ * n;
                    if (dt_ < 0) {
                        T friction = obj->friction;
                        if (use_mibf) friction = grid.friction[index];
                        if (friction > 0) {
                            if ((-dt_) * friction < vnorm(v_rel)) v_rel = v_rel + normalized(v_rel) * dt_ * friction;
                            else v_rel = TV::Zero();
                        }
                    }
                } else return MPM_ERR_INVALID_ARGUMENT;
                vi = v_rel;
                vi_orig = vi;
            }
        }
        for (auto& obj : plates) {
            if (!obj->inside(Xi)) continue;
            T vr[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) vr[d] = vi_orig(d) - obj->v_object[d];
            T friction = obj->friction;
            if (obj->bc == BC::NoSlip) { vr[0] = vr[1] = vr[2] = 0; }
            else if (obj->bc == BC::SlipStick || obj->bc == BC::SlipFree) {
                if (use_mibf) friction = grid.friction[index];
                bool slipfree = obj->bc == BC::SlipFree;
                PlateType pt = obj->plate_type;
                if (D == 3) {
                    if ((pt == PlateType::top || pt == PlateType::bottom) &&
                        (!slipfree || (pt == PlateType::top && vr[1] > 0) || (pt == PlateType::bottom && vr[1] < 0))) {
                        plateTangent3(vr[0], vr[2], vr[1], friction); vr[1] = 0;
                    } else if ((pt == PlateType::left || pt == PlateType::right) &&
                        (!slipfree || (pt == PlateType::left && vr[0] < 0) || (pt == PlateType::right && vr[0] > 0))) {
                        plateTangent3(vr[1], vr[2], vr[0], friction); vr[0] = 0;
                    } else if ((pt == PlateType::front || pt == PlateType::back) &&
                        (!slipfree || (pt == PlateType::front && vr[2] > 0) || (pt == PlateType::back && vr[2] < 0))) {
                        plateTangent3(vr[0], vr[1], vr[2], friction); vr[2] = 0;
                    }
                } else {
                    if ((pt == PlateType::top || pt == PlateType::bottom) &&
                        (!slipfree || (pt == PlateType::top && vr[1] > 0) || (pt == PlateType::bottom && vr[1] < 0))) {
                        plateTangent2(vr[0], vr[1], friction); vr[1] = 0;
                    } else if ((pt == PlateType::left || pt == PlateType::right) &&
                        (!slipfree || (pt == PlateType::left && vr[0] < 0) || (pt == PlateType::right && vr[0] > 0))) {
                        plateTangent2(vr[1], vr[0], friction); vr[0] = 0;
                    }
                }
            } else return MPM_ERR_INVALID_ARGUMENT;
            for (int d = 0; d < D; ++d) vi(d) = vr[d] + obj->v_object[d];
            vi_orig = vi;
        }
        return 0;
    }

    // G2P, g2p.cpp:7-94
    void G2P() {
        std::fill(particles.pic.begin(), particles.pic.end(), TV::Zero());
        std::fill(particles.flip.begin(), particles.flip.end(), TV::Zero());
        std::fill(particles.Bmat.begin(), particles.Bmat.end(), TM::Zero());
        uint32_t err = 0;
        #pragma omp parallel num_threads(n_threads) reduction(|:err)
        {
            std::vector<TV> pic_l(Np, TV::Zero()), flip_l(Np, TV::Zero());
            std::vector<TM> B_l(Np, TM::Zero());
            #pragma omp for
            for (long p = 0; p < (long)Np; ++p) {
                const TV xp = particles.x[p];
                TV vp = TV::Zero(), flipp = TV::Zero(); TM Bp = TM::Zero();
                long b[3]; stencilBase(xp, false, b);
                if (!stencilInside(b)) { err |= MPM_ERR_GRID_BOUNDS; continue; }
                for (long i = b[0]; i < b[0] + 4; ++i) {
                    T xi = grid.x[i];
                    for (long j = b[1]; j < b[1] + 4; ++j) {
                        T yi = grid.y[j];
                        for (long k = b[2]; k < b[2] + (D == 3 ? 4 : 1); ++k) {
                            unsigned id; T weight; TV pd;
                            pd(0) = xi - xp(0); pd(1) = yi - xp(1);
                            if constexpr (D == 3) { T zi = grid.z[k]; weight = wip3<SPL>(xp(0), xp(1), xp(D - 1), xi, yi, zi, one_over_dx); id = ind(i, j, k); pd(D - 1) = zi - xp(D - 1); }
                            else { weight = wip2<SPL>(xp(0), xp(1), xi, yi, one_over_dx); id = ind(i, j); }
                            vp += grid.v[id] * weight;
                            if (flip_ratio < 0) Bp += outer(grid.v[id], pd) * weight;
                            if (flip_ratio >= -1) flipp += grid.flip[id] * weight;
                        }
                    }
                }
                pic_l[p] = vp;
                if (flip_ratio < 0) B_l[p] = Bp;
                if (flip_ratio >= -1) flip_l[p] = flipp;
            }
            #pragma omp critical
            {
                for (unsigned p = 0; p < Np; ++p) {
                    particles.pic[p] += pic_l[p];
                    if (flip_ratio < 0) particles.Bmat[p] += B_l[p];
                    if (flip_ratio >= -1) particles.flip[p] += flip_l[p];
                }
            }
        }
        exit |= err;
    }

    // Particle velocity update shared by MUSL (musl.cpp:13-25) and positionUpdate (position_update.cpp:14-23)
    void velocityUpdate(unsigned p) {
        if (flip_ratio < -1) particles.v[p] = particles.pic[p];
        else if (flip_ratio < 0) particles.v[p] = (-flip_ratio) * (particles.v[p] + particles.flip[p]) + (1 - (-flip_ratio)) * particles.pic[p];
        else particles.v[p] = flip_ratio * (particles.v[p] + particles.flip[p]) + (1 - flip_ratio) * particles.pic[p];
    }

    // MUSL, musl.cpp:6-96
    void MUSL() {
        #pragma omp parallel for num_threads(n_threads)
        for (long p = 0; p < (long)Np; ++p) velocityUpdate(p);
        std::fill(grid.v.begin(), grid.v.end(), TV::Zero());
        #pragma omp parallel num_threads(n_threads)
        {
            std::vector<TV> gv(grid_nodes, TV::Zero());
            #pragma omp for
            for (long p = 0; p < (long)Np; ++p) {
                const TV xp = particles.x[p];
                long b[3]; stencilBase(xp, true, b);
                if (!stencilInside(b)) continue;
                for (long i = b[0]; i < b[0] + 4; ++i) {
                    T xi = grid.x[i];
                    for (long j = b[1]; j < b[1] + 4; ++j) {
                        T yi = grid.y[j];
                        for (long k = b[2]; k < b[2] + (D == 3 ? 4 : 1); ++k) {
                            unsigned id; T weight; TV pd;
                            pd(0) = xi - xp(0); pd(1) = yi - xp(1);
                            if constexpr (D == 3) { T zi = grid.z[k]; weight = wip3<SPL>(xp(0), xp(1), xp(D - 1), xi, yi, zi, one_over_dx); id = ind(i, j, k); pd(D - 1) = zi - xp(D - 1); }
                            else { weight = wip2<SPL>(xp(0), xp(1), xi, yi, one_over_dx); id = ind(i, j); }
                            if (weight > 1e-25) {
                                gv[id] += particles.v[p] * weight;
                                if (flip_ratio < 0) gv[id] += matvec(particles.Bmat[p], pd) * apicDinverse * weight;
                            }
                        }
                    }
                }
            }
            #pragma omp critical
            { for (unsigned l = 0; l < grid_nodes; ++l) grid.v[l] += gv[l]; }
        }
        for (unsigned l = 0; l < grid_nodes; ++l) {
            T mi = grid.mass[l];
            if (mi > 0) grid.v[l] /= (mi / particle_mass); else grid.v[l] = TV::Zero();
        }
    }

    // deformationUpdate, deformation_update.cpp:7-53
    void deformationUpdate() {
        std::fill(particles.delta_gamma.begin(), particles.delta_gamma.end(), 0.0);
        unsigned plastic_count = 0;
        uint32_t err = 0;
        #pragma omp parallel for reduction(+:plastic_count) reduction(|:err) num_threads(n_threads)
        for (long p = 0; p < (long)Np; ++p) {
            TM sum = TM::Zero();
            const TV xp = particles.x[p];
            long b[3]; stencilBase(xp, false, b);
            if (!stencilInside(b)) { err |= MPM_ERR_GRID_BOUNDS; continue; }
            for (long i = b[0]; i < b[0] + 4; ++i) {
                T xi = grid.x[i];
                for (long j = b[1]; j < b[1] + 4; ++j) {
                    T yi = grid.y[j];
                    for (long k = b[2]; k < b[2] + (D == 3 ? 4 : 1); ++k) {
                        TV g; unsigned id;
                        if constexpr (D == 3) { T zi = grid.z[k]; g = vecFrom3(grad_wip3<SPL>(xp(0), xp(1), xp(D - 1), xi, yi, zi, one_over_dx)); id = ind(i, j, k); }
                        else { g = vecFrom2(grad_wip2<SPL>(xp(0), xp(1), xi, yi, one_over_dx)); id = ind(i, j); }
                        sum += outer(grid.v[id], g);
                    }
                }
            }
            TM Fe_trial = particles.F[p];
            Fe_trial = Fe_trial + matmul(dt * sum, Fe_trial);
            particles.F[p] = Fe_trial;
            err |= plasticity(p, plastic_count, Fe_trial);
        }
        exit |= err;
        last_plastic_count = plastic_count;
    }

    // ---- MCC return-mapping helpers (src/plasticity_helpers/) ----
    // MCCRMAExplicit, mcc_rma_explicit.hpp:8-96
    static bool MCCRMAExplicit(T& p, T& q, uint32_t&, T M_, T p0_, T beta_, T mu_, T K_, T rma) {
        T y = M_ * M_ * (p - p0_) * (p + beta_ * p0_) + (q * q);
        if (y > 0) {
            T delta_gamma = 0, pt = p, qt = q;
            T dkdp = 2 * M_ * M_, dldq = 2;
            T max_iter = 40;
            for (int iter = 0; iter < max_iter; iter++) {
                y = M_ * M_ * (p - p0_) * (p + beta_ * p0_) + (q * q);
                T k = M_ * M_ * (beta_ * p0_ + 2 * p - p0_);
                T l = 2 * q;
                T r1 = pt - p - K_ * delta_gamma * k;
                T r2 = qt - q - rma * mu_ * delta_gamma * l;
                if (iter > 4 && std::abs(y) < 1e-3 && std::abs(r1) < 1e-3 && std::abs(r2) < 1e-3) break;
                if (iter == max_iter - 1) {
                    if (p0_ > 0.1) break;          // "RMA: FATAL did not exit loop" (exit not set)
                    else { p = 1e-15; q = 1e-15; break; }
                }
                T J11 = -1 - K_ * delta_gamma * dkdp;
                T J13 = -K_ * k;
                T J22 = -1 - rma * mu_ * delta_gamma * dldq;
                T J23 = -rma * mu_ * l;
                T J31 = k, J32 = l;
                T det = J11 * (-J32 * J23) + J13 * (-J31 * J22);
                if (std::abs(det) < T(1e-6)) {
                    p -= 0.001 * r1; q -= 0.001 * r2; delta_gamma -= 0.001 * y / (p0_ * p0_);
                } else {
                    p -= (-J32 * J23 * r1 + J32 * J13 * r2 - J22 * J13 * y) / det;
                    q -= (J31 * J23 * r1 - J31 * J13 * r2 - J11 * J23 * y) / det;
                    delta_gamma -= (-J31 * J22 * r1 - J11 * J32 * r2 + J11 * J22 * y) / det;
                }
                if (q < 1e-15) q = 1e-15;
            }
            p = std::max(p, -beta_ * p0_);
            p = std::min(p, p0_);
            q = M_ * std::sqrt((p0_ - p) * (beta_ * p0_ + p));
            return true;
        }
        return false;
    }
    // limited_search_exponential, limited_search_exponential.hpp:8-71
    static T limited_search_exponential(T epv, T epv_t, T epv_delta, T xi_, T K_, T p_t, T p00, T beta_, int& right) {
        T scale = 1.; int tries = 0; int max_tries = 32; int f = 0; T test_epv = epv;
        while (tries < max_tries) {
            test_epv = epv - scale * epv_delta;
            T pp0 = p00 * std::exp(-xi_ * test_epv);
            T pp = K_ * (test_epv - epv_t) + p_t;
            T val = 2. * pp + (beta_ - 1.) * pp0;
            T midpoint = (pp0 - beta_ * pp0) * 0.5;
            if (std::abs(pp - midpoint) <= T(1.e-5)) break;
            if (((right == 1 && (pp < midpoint || p_t < midpoint)) || (right == 0 && (pp > midpoint || p_t > midpoint))) && tries > -1) {
                test_epv = epv; right = -1; break;
            }
            if (sgn(epv_t - test_epv) == sgn(val) || sgn(epv_t - test_epv) == 0) break;
            scale *= 0.5;
            tries++;
            if (tries != max_tries) continue;
            tries = 0; scale = 1.; epv_delta *= -1.; f++;
            if (f > 1) return std::numeric_limits<double>::infinity();
        }
        return test_epv;
    }
    // MCCRMAImplicitExponentialOnevar, mcc_rma_implicit_exponential_onevar.hpp:9-166
    static bool MCCRMAImplicitExponentialOnevar(T& p, T& q, uint32_t& ex, T M_, T p00, T beta_, T mu_, T K_, T xi_, T, T epv) {
        T p0_t = p00 * std::exp(-xi_ * epv);
        T Msq = M_ * M_;
        T damping_factor = 1;
        T y = Msq * (p * p + (beta_ - 1.) * p0_t * p - beta_ * p0_t * p0_t) + damping_factor * q * q;
        if (y <= 0) return false;
        T scale_factor = 1. / p0_t;
        T p0_ = 1.;
        p00 *= scale_factor; p *= scale_factor; q *= scale_factor; mu_ *= scale_factor; K_ *= scale_factor;
        T p_t = p, q_t = q, epv_t = epv;
        int right = 0;
        if (p_t > 0.5 * (p0_ - beta_ * p0_)) right = 1;
        int max_iter = 30;
        for (int iter = 0; iter < max_iter; iter++) {
            T midpoint = (p0_ - beta_ * p0_) * 0.5;
            if (std::abs(p - midpoint) < 1e-4) {
                q = M_ * std::sqrt((p0_ - p) * (beta_ * p0_ + p) / damping_factor);
                q *= p0_t; p *= p0_t;
                return true;
            }
            p = (epv - epv_t) * K_ + p_t;
            p0_ = p00 * std::exp(-xi_ * epv);
            T denom = 2. * K_ * Msq * p + Msq * (beta_ - 1.) * K_ * p0_;
            T dg_step = (p_t - p) / denom;
            T q_denom = 1 + 2. * mu_ * damping_factor * dg_step;
            q = q_t / q_denom;
            y = Msq * (p - p0_) * (p + beta_ * p0_) + damping_factor * q * q;
            if (iter > 5 && std::abs(y) < 1.e-4) break;
            // iter == max_iter-1: "RMA: FATAL did not exit loop" (exit not set)
            T dp_depv = K_;
            T dp0_depv = -xi_ * p0_;
            T ddelta_gamma_depv = (-dp_depv / denom) - (p_t - p) * (2. * K_ * Msq * dp_depv + Msq * (beta_ - 1.) * K_ * dp0_depv) / (denom * denom);
            T dq_depv = -q_t * 2. * mu_ * damping_factor * ddelta_gamma_depv / (q_denom * q_denom);
            T dy_depv = Msq * (dp_depv - dp0_depv) * (p + beta_ * p0_) + Msq * (p - p0_) * (beta_ * dp0_depv + dp_depv) + 2. * damping_factor * q * dq_depv;
            epv = limited_search_exponential(epv, epv_t, y / dy_depv, xi_, K_, p_t, p00, beta_, right);
            if (epv == std::numeric_limits<double>::infinity()) { ex |= MPM_ERR_MCC_NO_STEP; break; }
        }
        p = std::max(p, -beta_ * p0_);
        p = std::min(p, p0_);
        q = M_ * std::sqrt((p0_ - p) * (beta_ * p0_ + p) / damping_factor);
        p *= p0_t; q *= p0_t;
        return true;
    }
    // limited_search_sinh, limited_search_sinh.hpp:8-72
    static T limited_search_sinh(T epv, T epv_t, T epv_delta, T xi_, T K_, T p_t, T p00, T p0_min, T beta_, int& right) {
        T scale = 1.; int tries = 0; int max_tries = 32; int f = 0; T test_epv = epv;
        while (tries < max_tries) {
            test_epv = epv - scale * epv_delta;
            T pp0 = std::max(p0_min, K_ * std::sinh(-xi_ * test_epv + std::asinh(p00 / K_)));
            T pp = K_ * (test_epv - epv_t) + p_t;
            T val = 2. * pp + (beta_ - 1.) * pp0;
            T midpoint = (pp0 - beta_ * pp0) * 0.5;
            if (std::abs(pp - midpoint) <= T(1.e-5)) break;
            if (((right == 1 && (pp < midpoint || p_t < midpoint)) || (right == 0 && (pp > midpoint || p_t > midpoint))) && tries > -1) {
                test_epv = epv; right = -1; break;
            }
            if (sgn(epv_t - test_epv) == sgn(val) || sgn(epv_t - test_epv) == 0) break;
            scale *= 0.5;
            tries++;
            if (tries != max_tries) continue;
            tries = 0; scale = 1.; epv_delta *= -1.; f++;
            if (f > 1) return std::numeric_limits<double>::infinity();
        }
        return test_epv;
    }
    // MCCRMAImplicitSinhOnevar, mcc_rma_implicit_sinh_onevar.hpp:9-171
    static bool MCCRMAImplicitSinhOnevar(T& p, T& q, uint32_t& ex, T M_, T p00, T beta_, T mu_, T K_, T xi_, T, T epv) {
        T p0_ = std::max(T(0.01), K_ * std::sinh(-xi_ * epv + std::asinh(p00 / K_)));
        T Msq = M_ * M_;
        T damping_factor = 1;
        T y = Msq * (p * p + (beta_ - 1.) * p0_ * p - beta_ * p0_ * p0_) + damping_factor * q * q;
        if (y <= 0) return false;
        T p0_t = p0_;
        T scale_factor = 1. / p0_t;
        T p0_min = 0.01 * scale_factor; p00 *= scale_factor; p0_ = 1.; p *= scale_factor; q *= scale_factor; mu_ *= scale_factor; K_ *= scale_factor;
        T p_t = p, q_t = q, epv_t = epv;
        int right = 0;
        if (p_t > 0.5 * (p0_ - beta_ * p0_)) right = 1;
        int max_iter = 40;
        for (int iter = 0; iter < max_iter; iter++) {
            T midpoint = (p0_ - beta_ * p0_) * 0.5;
            if (std::abs(p - midpoint) < T(1.e-4)) {
                q = M_ * std::sqrt((p0_ - p) * (beta_ * p0_ + p) / damping_factor);
                p *= p0_t; q *= p0_t;
                return true;
            }
            p = (epv - epv_t) * K_ + p_t;
            T p0_helper = K_ * std::sinh(-xi_ * epv + std::asinh(p00 / K_));
            p0_ = std::max(p0_min, p0_helper);
            T denom = 2. * K_ * Msq * p + Msq * (beta_ - 1.) * K_ * p0_;
            T dg_step = (p_t - p) / denom;
            T q_denom = 1. + 2. * mu_ * damping_factor * dg_step;
            q = q_t / q_denom;
            y = Msq * (p - p0_) * (p + beta_ * p0_) + damping_factor * q * q;
            if (iter > 5 && std::abs(y) < 1e-4) break;
            if (iter == max_iter - 1) { if (p0_ > p0_min) ex |= MPM_ERR_MCC_NO_STEP; }
            T dp_depv = K_;
            T dp0_depv;
            T dp0_depv_helper = -xi_ * K_ * std::cosh(-xi_ * epv + std::asinh(p00 / K_));
            if (p0_helper > p0_min) dp0_depv = 0.5 * (dp0_depv_helper + dp0_depv_helper - p0_min);
            else dp0_depv = 0.5 * (dp0_depv_helper + p0_min - dp0_depv_helper);
            T ddelta_gamma_depv = (-dp_depv / denom) - (p_t - p) * (2. * K_ * Msq * dp_depv + Msq * (beta_ - 1.) * K_ * dp0_depv) / (denom * denom);
            T dq_depv = -q_t * 2. * mu_ * damping_factor * ddelta_gamma_depv / (q_denom * q_denom);
            T dy_depv = Msq * (dp_depv - dp0_depv) * (p + beta_ * p0_) + Msq * (p - p0_) * (beta_ * dp0_depv + dp_depv) + 2. * damping_factor * q * dq_depv;
            epv = limited_search_sinh(epv, epv_t, y / dy_depv, xi_, K_, p_t, p00, p0_min, beta_, right);
            if (epv == std::numeric_limits<double>::infinity()) { ex |= MPM_ERR_MCC_NO_STEP; break; }
        }
        p = std::max(p, -beta_ * p0_);
        p = std::min(p, p0_);
        q = M_ * std::sqrt((p0_ - p) * (beta_ * p0_ + p) / damping_factor);
        p *= p0_t; q *= p0_t;
        return true;
    }

    // Perzyna Newton iteration for VMVisc (plasticity.cpp:197-229) when xi_vm is true, DPVisc (:293-323)
    // and MCCVisc (:558-588) otherwise.
    T perzynaNewton(T q_trial, T q_yield, unsigned p, bool vm_hardening, uint32_t& err) const {
        T delta_gamma = 0.01 * (q_trial - q_yield) / f_mu_prefac;
        int max_iter = 60;
        for (int iter = 0; iter < max_iter; iter++) {
            if (iter == max_iter - 1) err |= MPM_ERR_NEWTON_MAXITER;
            T tm = perzyna_visc * delta_gamma + dt;
            T tmp = dt / tm;
            T tmp1 = std::pow(tmp, perzyna_exp);
            T q_yield_diff = 0;
            if (vm_hardening) q_yield = q_min + (q_max - q_min) * std::exp(-xi * (particles.eps_pl_dev[p] + (1.0 / d_prefac) * delta_gamma));
            T residual = (q_trial - f_mu_prefac * delta_gamma) * tmp1 - q_yield;
            if (std::abs(residual) < 1e-2) break;
            if (vm_hardening) q_yield_diff = -xi / d_prefac * (q_max - q_min) * std::exp(-xi * (particles.eps_pl_dev[p] + (1.0 / d_prefac) * delta_gamma));
            T residual_diff = -f_mu_prefac * tmp1 + (q_trial - f_mu_prefac * delta_gamma) * perzyna_exp * std::pow(tmp, perzyna_exp - 1) * (-perzyna_visc * dt) / (tm * tm);
            if (vm_hardening) residual_diff -= q_yield_diff;
            if (std::abs(residual_diff) < 1e-14) err |= MPM_ERR_NEWTON_SMALL_DERIV;
            delta_gamma -= residual / residual_diff;
            if (delta_gamma < 0) delta_gamma = 1e-10;
        }
        return delta_gamma;
    }

    TM recompose(const TM& U, const TV& h, const TM& V) const {
        TV e; for (int i = 0; i < D; ++i) e(i) = std::exp(h(i));
        return udvt(U, e, V);
    }

    // plasticity, plasticity.cpp:11-665 (Snow unsupported)
    uint32_t plasticity(unsigned p, unsigned& plastic_count, TM& Fe_trial) {
        uint32_t err = 0;
        if (plastic_model == PlasticModel::NoPlasticity) return 0;
        Particles<D>& P = particles;
        TM U, V; TV sv; svd(Fe_trial, U, sv, V);
        TV hencky; for (int i = 0; i < D; ++i) hencky(i) = std::log(std::max(std::abs(sv(i)), 1e-4));
        T hencky_trace = vsum(hencky);
        TV hencky_deviatoric = hencky - (hencky_trace / dim) * TV::Ones();
        T hencky_deviatoric_norm = vnorm(hencky_deviatoric);
        if (hencky_deviatoric_norm > 0) hencky_deviatoric /= hencky_deviatoric_norm;

        if (plastic_model == PlasticModel::VM) {  // :30-49
            T q_yield = q_max;
            T delta_gamma = hencky_deviatoric_norm - q_yield / e_mu_prefac;
            if (delta_gamma > 0) {
                plastic_count++;
                P.delta_gamma[p] = d_prefac * delta_gamma / dt;
                hencky -= delta_gamma * hencky_deviatoric;
                P.F[p] = recompose(U, hencky, V);
                P.eps_pl_dev[p] += delta_gamma;
            }
        } else if (plastic_model == PlasticModel::DP) {  // :51-85
            T p_trial = -K * hencky_trace;
            T q_yield = M * p_trial + q_cohesion;
            if (q_yield < 1e-10) {
                plastic_count++;
                T p_proj = -q_cohesion / M;
                hencky = (-p_proj / (K * dim)) * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
                T delta_gamma = d_prefac * hencky_deviatoric_norm;
                P.delta_gamma[p] = delta_gamma / dt;
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.eps_pl_vol[p] += (p_proj - p_trial) / K;
            } else {
                T delta_gamma = d_prefac * (hencky_deviatoric_norm - q_yield / e_mu_prefac);
                if (delta_gamma > 0) {
                    plastic_count++;
                    P.delta_gamma[p] = delta_gamma / dt;
                    hencky -= ((1.0 / d_prefac) * delta_gamma) * hencky_deviatoric;
                    P.F[p] = recompose(U, hencky, V);
                    P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                }
            }
        } else if (plastic_model == PlasticModel::DPSoft) {  // :87-162
            T p_trial = -K * hencky_trace;
            T q_trial = e_mu_prefac * hencky_deviatoric_norm;
            T p_tip_orig = -q_cohesion / M;
            T p_tip = p_tip_orig * std::exp(-xi * P.eps_pl_dev[p]);
            T p_shift = 0;
            if (use_pradhana) p_shift = -K * P.eps_pl_vol_pradhana[p];
            T q_yield = M * (p_trial + p_shift) + (-p_tip * M);
            if ((p_trial + p_shift) <= p_tip) {
                plastic_count++;
                T p_proj = p_tip_orig * std::exp(-xi * P.eps_pl_dev[p]);
                hencky = (-p_proj / (K * dim)) * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
                T delta_gamma = d_prefac * hencky_deviatoric_norm;
                P.delta_gamma[p] = delta_gamma / dt;
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                T eps_pl_vol_inst = (p_proj - p_trial) / K;
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                if (use_pradhana) P.eps_pl_vol_pradhana[p] += eps_pl_vol_inst;
            }
            if ((p_trial + p_shift) > p_tip && q_trial <= q_yield) {
                if (use_pradhana) P.eps_pl_vol_pradhana[p] = 0;
                P.delta_gamma[p] = 0;
            }
            if ((p_trial + p_shift) > p_tip && q_trial > q_yield) {
                plastic_count++;
                T temp_eps_pl_dev = P.eps_pl_dev[p] + (hencky_deviatoric_norm - q_yield / e_mu_prefac);
                T p_proj = p_tip_orig * std::exp(-xi * temp_eps_pl_dev);
                if ((p_trial + p_shift) < p_proj) {
                    T delta_gamma = d_prefac * hencky_deviatoric_norm;
                    P.delta_gamma[p] = delta_gamma / dt;
                    P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                    T eps_pl_vol_inst = (p_proj - p_trial) / K;
                    P.eps_pl_vol[p] += eps_pl_vol_inst;
                    if (use_pradhana) P.eps_pl_vol_pradhana[p] += eps_pl_vol_inst;
                    hencky = (-p_proj / (K * dim)) * TV::Ones();
                    P.F[p] = recompose(U, hencky, V);
                } else {
                    T q_yield_new = M * (p_trial + p_shift) + (-p_proj * M);
                    T delta_gamma = d_prefac * (hencky_deviatoric_norm - q_yield_new / e_mu_prefac);
                    hencky -= ((1.0 / d_prefac) * delta_gamma) * hencky_deviatoric;
                    P.F[p] = recompose(U, hencky, V);
                    P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                    P.delta_gamma[p] = delta_gamma / dt;
                    if (use_pradhana) P.eps_pl_vol_pradhana[p] = 0;
                }
            }
        } else if (plastic_model == PlasticModel::VMVisc) {  // :164-238
            T q_trial = e_mu_prefac * hencky_deviatoric_norm;
            T q_yield = q_min + (q_max - q_min) * std::exp(-xi * P.eps_pl_dev[p]);
            T p_trial = -K * hencky_trace;
            if (p_trial < p_min * std::exp(-xi * P.eps_pl_vol[p])) {
                T delta_gamma = q_trial / f_mu_prefac;
                T eps_pl_vol_inst = -p_trial / K;
                P.F[p] = matmul(U, transpose(V));
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.delta_gamma[p] = delta_gamma / dt;
            } else if (q_trial > q_yield) {
                plastic_count++;
                T delta_gamma;
                if (perzyna_exp == 1 && xi == 0) {
                    delta_gamma = (q_trial - q_yield) / (f_mu_prefac + q_yield * perzyna_visc / dt);
                    if (delta_gamma < 0) err |= MPM_ERR_PERZYNA_NEGATIVE;
                } else {
                    delta_gamma = perzynaNewton(q_trial, q_yield, p, true, err);
                }
                hencky -= ((1.0 / d_prefac) * delta_gamma) * hencky_deviatoric;
                P.F[p] = recompose(U, hencky, V);
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.delta_gamma[p] = delta_gamma / dt;
            }
        } else if (plastic_model == PlasticModel::DPVisc) {  // :240-339
            T p_trial = -K * hencky_trace;
            T q_trial = e_mu_prefac * hencky_deviatoric_norm;
            T p_tip = -q_cohesion / M;
            T p_shift = 0;
            if (use_pradhana) p_shift = -K * P.eps_pl_vol_pradhana[p];
            T q_yield = M * (p_trial + p_shift) + q_cohesion;
            if (use_mibf) P.muI[p] = M;
            if ((p_trial + p_shift) < p_tip) {
                T delta_gamma = d_prefac * hencky_deviatoric_norm;
                T p_proj = p_tip;
                T eps_pl_vol_inst = (p_proj - p_trial) / K;
                plastic_count++;
                hencky = (-p_proj / (K * dim)) * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
                P.delta_gamma[p] = delta_gamma / dt;
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                if (use_pradhana) P.eps_pl_vol_pradhana[p] += eps_pl_vol_inst;
            } else {
                if (use_pradhana) P.eps_pl_vol_pradhana[p] = 0;
                P.delta_gamma[p] = 0;
            }
            if ((p_trial + p_shift) > p_tip && q_trial > q_yield) {
                plastic_count++;
                T delta_gamma;
                if (perzyna_exp == 1) {
                    delta_gamma = (q_trial - q_yield) / (f_mu_prefac + q_yield * perzyna_visc / dt);
                    if (delta_gamma < 0) err |= MPM_ERR_PERZYNA_NEGATIVE;
                } else {
                    delta_gamma = perzynaNewton(q_trial, q_yield, p, false, err);
                }
                if (use_mibf) {
                    if (std::abs(p_trial + p_shift) < 1e-10) P.muI[p] = 1e10;
                    else P.muI[p] = ((q_trial - f_mu_prefac * delta_gamma) - q_cohesion) / (p_trial + p_shift);
                }
                hencky -= ((1.0 / d_prefac) * delta_gamma) * hencky_deviatoric;
                P.F[p] = recompose(U, hencky, V);
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.delta_gamma[p] = delta_gamma / dt;
            }
        } else if (plastic_model == PlasticModel::DPMui) {  // :342-403
            T p_trial = -K * hencky_trace;
            T q_trial = e_mu_prefac * hencky_deviatoric_norm;
            T p_tip = -q_cohesion / mu_1;
            T p_shift = 0;
            if (use_pradhana) p_shift = -K * P.eps_pl_vol_pradhana[p];
            P.muI[p] = mu_1;
            P.viscosity[p] = 0;
            if ((p_trial + p_shift) < p_tip) {
                T delta_gamma = d_prefac * hencky_deviatoric_norm;
                T eps_pl_vol_inst = (p_tip - p_trial) / K;
                plastic_count++;
                hencky = (-p_tip / (K * dim)) * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
                P.delta_gamma[p] = delta_gamma;  // NB: not divided by dt (plasticity.cpp:364)
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                P.eps_pl_vol_pradhana[p] += eps_pl_vol_inst;
            } else {
                P.eps_pl_vol_pradhana[p] = 0;
            }
            T q_yield = mu_1 * (p_trial + p_shift) + q_cohesion;
            if ((p_trial + p_shift) > p_tip && q_trial > (q_yield + stress_tolerance)) {
                plastic_count++;
                T p_special = p_trial + p_shift - p_tip + stress_tolerance;
                T fac_a = f_mu_prefac * dt;
                T fac_b = p_trial * (mu_2 - mu_1) + f_mu_prefac * dt * fac_Q * std::sqrt(p_special) - (q_trial - q_yield);
                T fac_c = -(q_trial - q_yield) * fac_Q * std::sqrt(p_special);
                T delta_gamma = (-fac_b + std::sqrt(fac_b * fac_b - 4 * fac_a * fac_c)) / (2 * fac_a);
                T mu_i = mu_1 + (mu_2 - mu_1) / (fac_Q * std::sqrt(p_special) / delta_gamma + 1.0);
                P.muI[p] = mu_i;
                P.viscosity[p] = (mu_i - mu_1) * p_special / delta_gamma;
                delta_gamma *= dt;
                hencky -= ((1.0 / d_prefac) * delta_gamma) * hencky_deviatoric;
                P.F[p] = recompose(U, hencky, V);
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.delta_gamma[p] = delta_gamma / dt;
            }
        } else if (plastic_model == PlasticModel::MCCMui) {  // :405-484
            T p_stress = -K * hencky_trace;
            T q_stress = e_mu_prefac * hencky_deviatoric_norm;
            T q_trial = q_stress;
            bool perform_rma = false;
            T p_c = 0;
            if (hardening_law == HardeningLaw::ExpoExpl) {
                p_c = std::max(stress_tolerance, p0 * std::exp(-xi * P.eps_pl_vol[p]));
                perform_rma = MCCRMAExplicit(p_stress, q_stress, err, mu_1, p_c, beta, mu, K, rma_prefac);
            } else if (hardening_law == HardeningLaw::SinhExpl) {
                p_c = std::max(stress_tolerance, K * std::sinh(-xi * P.eps_pl_vol[p] + std::asinh(p0 / K)));
                perform_rma = MCCRMAExplicit(p_stress, q_stress, err, mu_1, p_c, beta, mu, K, rma_prefac);
            } else if (hardening_law == HardeningLaw::ExpoImpl) {
                perform_rma = MCCRMAImplicitExponentialOnevar(p_stress, q_stress, err, mu_1, p0, beta, mu, K, xi, rma_prefac, P.eps_pl_vol[p]);
            } else if (hardening_law == HardeningLaw::SinhImpl) {
                perform_rma = MCCRMAImplicitSinhOnevar(p_stress, q_stress, err, mu_1, p0, beta, mu, K, xi, rma_prefac, P.eps_pl_vol[p]);
            } else {
                err |= MPM_ERR_INVALID_ARGUMENT;
            }
            P.muI[p] = mu_1;
            P.viscosity[p] = 0;
            if (perform_rma) {
                plastic_count++;
                T ep = p_stress / (K * dim);
                T eps_pl_vol_inst = hencky_trace + dim * ep;
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                if (q_trial > (q_stress + stress_tolerance)) {
                    if (hardening_law == HardeningLaw::ExpoImpl) p_c = std::max(p0 * std::exp(-xi * P.eps_pl_vol[p]), p_stress + stress_tolerance);
                    else if (hardening_law == HardeningLaw::SinhImpl) p_c = std::max(K * std::sinh(-xi * P.eps_pl_vol[p] + std::asinh(p0 / K)), p_stress + stress_tolerance);
                    T p_special = p_stress + beta * p_c + stress_tolerance;
                    T fac_a = f_mu_prefac * dt;
                    T fac_b = std::sqrt((p_c - p_stress) * p_special) * (mu_2 - mu_1) + f_mu_prefac * dt * fac_Q * std::sqrt(p_special) - (q_trial - q_stress);
                    T fac_c = -(q_trial - q_stress) * fac_Q * std::sqrt(p_special);
                    T gamma_dot_S = (-fac_b + std::sqrt(fac_b * fac_b - 4 * fac_a * fac_c)) / (2 * fac_a);
                    q_stress = std::max(q_stress, q_trial - f_mu_prefac * dt * gamma_dot_S);
                    T mu_i = mu_1 + (mu_2 - mu_1) / (fac_Q * std::sqrt(p_special) / gamma_dot_S + 1.0);
                    P.muI[p] = mu_i;
                    P.viscosity[p] = (mu_i - mu_1) * std::sqrt((p_c - p_stress) * p_special) / gamma_dot_S;
                    T delta_gamma = (q_trial - q_stress) / f_mu_prefac;
                    P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                    P.delta_gamma[p] = delta_gamma / dt;
                }
                hencky = (q_stress / e_mu_prefac) * hencky_deviatoric - ep * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
            }
        } else if (plastic_model == PlasticModel::MCC || plastic_model == PlasticModel::MCCVisc) {  // :487-621
            T p_stress = -K * hencky_trace;
            T q_stress = e_mu_prefac * hencky_deviatoric_norm;
            T q_trial = q_stress;
            if (plastic_model == PlasticModel::MCCVisc && use_mibf) P.muI[p] = M;
            bool perform_rma = false;
            T p_c = 0;
            if (hardening_law == HardeningLaw::NoHard) {
                perform_rma = MCCRMAExplicit(p_stress, q_stress, err, M, p0, beta, mu, K, rma_prefac);
            } else if (hardening_law == HardeningLaw::ExpoExpl) {
                p_c = std::max(stress_tolerance, p0 * std::exp(-xi * P.eps_pl_vol[p]));
                perform_rma = MCCRMAExplicit(p_stress, q_stress, err, M, p_c, beta, mu, K, rma_prefac);
            } else if (hardening_law == HardeningLaw::SinhExpl) {
                p_c = std::max(stress_tolerance, K * std::sinh(-xi * P.eps_pl_vol[p] + std::asinh(p0 / K)));
                perform_rma = MCCRMAExplicit(p_stress, q_stress, err, M, p_c, beta, mu, K, rma_prefac);
            } else if (hardening_law == HardeningLaw::ExpoImpl) {
                perform_rma = MCCRMAImplicitExponentialOnevar(p_stress, q_stress, err, M, p0, beta, mu, K, xi, rma_prefac, P.eps_pl_vol[p]);
            } else if (hardening_law == HardeningLaw::SinhImpl) {
                perform_rma = MCCRMAImplicitSinhOnevar(p_stress, q_stress, err, M, p0, beta, mu, K, xi, rma_prefac, P.eps_pl_vol[p]);
            } else {
                err |= MPM_ERR_INVALID_ARGUMENT;
            }
            if (perform_rma) {
                plastic_count++;
                T ep = p_stress / (K * dim);
                T eps_pl_vol_inst = hencky_trace + dim * ep;
                P.eps_pl_vol[p] += eps_pl_vol_inst;
                T delta_gamma;
                if (plastic_model == PlasticModel::MCCVisc) {
                    T q_yield = q_stress;
                    if (perzyna_exp == 1) {
                        delta_gamma = (q_trial - q_yield) / (f_mu_prefac + q_yield * perzyna_visc / dt);
                        if (delta_gamma < 0) {
                            if (delta_gamma > -1e-14) delta_gamma = 0;
                            else err |= MPM_ERR_PERZYNA_NEGATIVE;
                        }
                    } else {
                        delta_gamma = perzynaNewton(q_trial, q_yield, p, false, err);
                    }
                    q_stress = std::max(q_yield, q_trial - f_mu_prefac * delta_gamma);
                    if (use_mibf) {
                        if (hardening_law == HardeningLaw::ExpoImpl) p_c = std::max(T(0), p0 * std::exp(-xi * P.eps_pl_vol[p]));
                        else if (hardening_law == HardeningLaw::SinhImpl) p_c = std::max(T(0), K * std::sinh(-xi * P.eps_pl_vol[p] + std::asinh(p0 / K)));
                        if ((p_stress < -beta * p_c + stress_tolerance) || (p_stress > p_c - stress_tolerance)) P.muI[p] = M;
                        else P.muI[p] = q_stress / std::sqrt((p_stress + beta * p_c) * (p_c - p_stress));
                    }
                }
                delta_gamma = (q_trial - q_stress) / f_mu_prefac;
                P.eps_pl_dev[p] += (1.0 / d_prefac) * delta_gamma;
                P.delta_gamma[p] = delta_gamma / dt;
                hencky = (q_stress / e_mu_prefac) * hencky_deviatoric - ep * TV::Ones();
                P.F[p] = recompose(U, hencky, V);
            }
        } else {
            err |= MPM_ERR_UNSUPPORTED;  // Snow
        }
        return err;
    }

    // positionUpdate, position_update.cpp:6-46
    void positionUpdate() {
        #pragma omp parallel for num_threads(n_threads)
        for (long p = 0; p < (long)Np; ++p) {
            particles.x[p] = particles.x[p] + dt * particles.pic[p];
            if (!use_musl) velocityUpdate(p);
            if (pbc) {  // position_update.cpp:26-33
                if (particles.x[p](0) > Lx) particles.x[p](0) = particles.x[p](0) - Lx;
                else if (particles.x[p](0) < 0) particles.x[p](0) = Lx + particles.x[p](0);
            }
            if (change_particle_positions) {
                if (particles.x[p](0) > 0.8) { particles.x[p](0) = -0.1; particles.x[p](1) += 0.27; }
            }
        }
    }

    // moveObjects, simulation.cpp:277-281
    void moveObjects() { for (auto& obj : plates) obj->move(dt, frame_dt, time); }

    // computeAvgData, save_data.cpp:264-292
    void computeAvgData(TM& volavg_cauchy, TM& volavg_kirchh, T& Javg) const {
        volavg_cauchy = TM::Zero(); volavg_kirchh = TM::Zero();
        T Jsum = 0;
        for (unsigned p = 0; p < Np; ++p) {
            const TM& Fe = particles.F[p];
            TM tau = kirchhoff(Fe);
            T Je = determinant(Fe);
            T J = Je * std::exp(particles.eps_pl_vol[p]);
            Jsum += J;
            volavg_cauchy += tau;
            volavg_kirchh += tau * J;
        }
        volavg_cauchy /= Jsum; volavg_kirchh /= Jsum;
        Javg = Jsum / Np;
    }
};

}  // namespace oracle
