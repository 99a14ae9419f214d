// matter.hpp — C++ host API of the MI355X engine with the names and meanings of the reference's
// Simulation interface (src/simulation/simulation.hpp:23-225, src/tools.hpp:53-57,
// src/data_structures.hpp:8-47, src/objects/*.hpp), so that set-up code written for the reference
// compiles against it. Every step runs on the GPU through the C ABI of include/mpm.h; this header only
// keeps host copies of the particles and translates the objects into plain descriptors.
//
// Compile-time switches as in the reference's tools.hpp:19-21: define MATTER_TWODIM for 2D (default 3D) and
// MATTER_SPLINEDEG as 2 (quadratic, default) or 3 (cubic). Eigen is not needed: TV and TM are small
// fixed-size types with the element access the reference's user code uses (operator(), operator[],
// Zero/Ones/Identity).
#pragma once
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <fstream>
#include <sys/stat.h>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpm.h"

namespace matter {

typedef double T;
#ifdef MATTER_TWODIM
constexpr int kDim = 2;
#else
constexpr int kDim = 3;
#endif
#ifndef MATTER_SPLINEDEG
#define MATTER_SPLINEDEG 2
#endif

// enums, src/tools.hpp:53-57 (same order, so the integer values match include/mpm.h)
enum class PlateType { top, bottom, left, right, front, back };
enum class ElasticModel { Hencky, NeoHookean };
enum class PlasticModel { NoPlasticity, VM, DP, DPSoft, MCC, VMVisc, DPVisc, MCCVisc, DPMui, MCCMui, Snow };
enum class HardeningLaw { NoHard, ExpoExpl, ExpoImpl, SinhExpl, SinhImpl };
enum class BC { NoSlip, SlipStick, SlipFree };

struct TV {
    T a[kDim];
    T& operator()(int i) { return a[i]; }
    const T& operator()(int i) const { return a[i]; }
    T& operator[](int i) { return a[i]; }
    const T& operator[](int i) const { return a[i]; }
    static TV Zero() { TV v; for (int i = 0; i < kDim; ++i) v.a[i] = 0; return v; }
    static TV Ones() { TV v; for (int i = 0; i < kDim; ++i) v.a[i] = 1; return v; }
    T squaredNorm() const { T s = 0; for (int i = 0; i < kDim; ++i) s += a[i] * a[i]; return s; }
    T norm() const { return std::sqrt(squaredNorm()); }
};
inline TV operator*(T s, const TV& v) { TV r; for (int i = 0; i < kDim; ++i) r.a[i] = s * v.a[i]; return r; }
inline TV operator*(const TV& v, T s) { return s * v; }

// column-major like Eigen's default storage
struct TM {
    T a[kDim * kDim];
    T& operator()(int r, int c) { return a[r + c * kDim]; }
    const T& operator()(int r, int c) const { return a[r + c * kDim]; }
    static TM Zero() { TM m; for (int i = 0; i < kDim * kDim; ++i) m.a[i] = 0; return m; }
    static TM Identity() { TM m = Zero(); for (int i = 0; i < kDim; ++i) m(i, i) = 1; return m; }
    T determinant() const {
        const TM& m = *this;
        if (kDim == 2) return m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1);
        return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) - m(1, 0) * (m(0, 1) * m(2, 2) - m(0, 2) * m(2, 1)) +
               m(2, 0) * (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1));
    }
};

// Particles, src/data_structures.hpp:8-47 (the fields the engine evolves)
class Particles {
public:
    explicit Particles(unsigned Np = 1)
        : x(Np, TV::Zero()), v(Np, TV::Zero()), eps_pl_dev(Np, 0.0), eps_pl_vol(Np, 0.0), eps_pl_vol_pradhana(Np, 0.0),
          delta_gamma(Np, 0.0), viscosity(Np, 0.0), muI(Np, 0.0), F(Np, TM::Identity()), Bmat(Np, TM::Zero()) {}
    std::vector<TV> x, v;
    std::vector<T> eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana, delta_gamma, viscosity, muI;
    std::vector<TM> F, Bmat;
};

// ObjectGeneral, src/objects/object_general.hpp:8-23: the analytic subclasses become descriptors.
class ObjectGeneral {
public:
    BC bc; T friction; std::string name;
    ObjectGeneral(BC bc_, T friction_, std::string name_) : bc(bc_), friction(friction_), name(std::move(name_)) {}
    virtual ~ObjectGeneral() {}
    virtual mpm_object descriptor() const = 0;
    // Appends the object to a context (mpm_add_object; the level-set object uses mpm_add_sdf_object).
    virtual int attach(mpm_context* ctx) const { const mpm_object d = descriptor(); return mpm_add_object(ctx, &d); }
protected:
    mpm_object base(int kind) const {
        mpm_object o; std::memset(&o, 0, sizeof(o));
        o.kind = kind; o.bc = (int32_t)bc; o.friction = friction;
        return o;
    }
};
class ObjectBoxRotated : public ObjectGeneral {  // object_box.hpp:6-16
public:
    TV center, half_extents; TM rotation;
    ObjectBoxRotated(BC b, T f, const TV& c, const TV& h, const TM& R, std::string n = "")
        : ObjectGeneral(b, f, n), center(c), half_extents(h), rotation(R) {}
    mpm_object descriptor() const override {
        mpm_object o = base(MPM_OBJECT_BOX_ROTATED);
        for (int i = 0; i < kDim; ++i) { o.center[i] = center(i); o.half_extents[i] = half_extents(i); }
        for (int i = 0; i < kDim * kDim; ++i) o.rotation[i] = rotation.a[i];
        return o;
    }
};
class ObjectBump : public ObjectGeneral {  // object_bump.hpp:13
public:
    ObjectBump(BC b, T f, std::string n = "") : ObjectGeneral(b, f, n) {}
    mpm_object descriptor() const override { return base(MPM_OBJECT_BUMP); }
};
class ObjectGround : public ObjectGeneral {  // object_ground.hpp:14
public:
    T y_ground;
    ObjectGround(BC b = BC::NoSlip, T f = 0.0, std::string n = "", T yg = 0) : ObjectGeneral(b, f, n), y_ground(yg) {}
    mpm_object descriptor() const override { mpm_object o = base(MPM_OBJECT_GROUND); o.param = y_ground; return o; }
};
class ObjectRamp : public ObjectGeneral {  // object_ramp.hpp:13
public:
    ObjectRamp(BC b, T f, std::string n = "") : ObjectGeneral(b, f, n) {}
    mpm_object descriptor() const override { return base(MPM_OBJECT_RAMP); }
};
class ObjectGate : public ObjectGeneral {  // object_gate.hpp:14
public:
    T height;
    ObjectGate(BC b, T f, std::string n = "", T h = 0.016) : ObjectGeneral(b, f, n), height(h) {}
    mpm_object descriptor() const override { mpm_object o = base(MPM_OBJECT_GATE); o.param = height; return o; }
};
class ObjectSilo : public ObjectGeneral {  // object_silo.hpp:14
public:
    T cut;
    ObjectSilo(BC b, T f, std::string n = "", T c = -1) : ObjectGeneral(b, f, n), cut(c) {}
    mpm_object descriptor() const override { mpm_object o = base(MPM_OBJECT_SILO); o.param = cut; return o; }
};
class ObjectCurve : public ObjectGeneral {  // object_curve.hpp:13
public:
    ObjectCurve(BC b = BC::NoSlip, T f = 0.0, std::string n = "") : ObjectGeneral(b, f, n) {}
    mpm_object descriptor() const override { return base(MPM_OBJECT_CURVE); }
};

// ObjectVdb, src/objects/object_vdb.hpp:17-94: the level set as a dense lattice of signed distances (a caller
// with OpenVDB copies its grid's active bounding box to dense, openvdb::tools::copyToDense), point (i, j, k) at
// origin + spacing * (i, j, k), index (i*n[1] + j)*n[2] + k, n[2] = 1 in 2D.
class ObjectSdf : public ObjectGeneral {
public:
    std::vector<float> phi; unsigned n[3]; TV origin; T spacing;
    ObjectSdf(std::vector<float> phi_, const unsigned* n_, const TV& origin_, T spacing_, BC b = BC::NoSlip, T f = 0.0,
              std::string nm = "")
        : ObjectGeneral(b, f, nm), phi(std::move(phi_)), origin(origin_), spacing(spacing_) { for (int d = 0; d < 3; ++d) n[d] = n_[d]; }
    mpm_object descriptor() const override { return base(MPM_OBJECT_SDF); }
    int attach(mpm_context* ctx) const override {
        mpm_sdf s; std::memset(&s, 0, sizeof(s));
        s.bc = (int32_t)bc; s.friction = friction; s.spacing = spacing; s.phi = phi.data();
        for (int d = 0; d < 3; ++d) { s.n[d] = n[d]; s.origin[d] = d < kDim ? origin(d) : 0.0; }
        return mpm_add_sdf_object(ctx, &s);
    }
};

// ObjectPlate, src/objects/object_plate.hpp:13-28 (3D) / :83-96 (2D): same constructor arguments.
class ObjectPlate {
public:
#ifdef MATTER_TWODIM
    ObjectPlate(T pos, PlateType pt, BC b = BC::NoSlip, T fr = 0, T lower = -1e15, T upper = 1e15, T vx = 0.0, T vy = 0.0,
                T vmin = 1.0, T load = 0.0, std::string n = "")
        : pos_object(pos), plate_type(pt), bc(b), friction(fr), pos_upper(upper), pos_lower(lower), vx_object(vx), vy_object(vy),
          vz_object(0), vmin_factor(vmin), load_factor(load), name(n) {}
#else
    ObjectPlate(T pos, PlateType pt, BC b = BC::NoSlip, T fr = 0, T lower = -1e15, T upper = 1e15, T vx = 0.0, T vy = 0.0,
                T vz = 0.0, T vmin = 1.0, T load = 0.0, std::string n = "")
        : pos_object(pos), plate_type(pt), bc(b), friction(fr), pos_upper(upper), pos_lower(lower), vx_object(vx), vy_object(vy),
          vz_object(vz), vmin_factor(vmin), load_factor(load), name(n) {}
#endif
    T pos_object; PlateType plate_type; BC bc; T friction; T pos_upper, pos_lower;
    T vx_object, vy_object, vz_object; T vmin_factor, load_factor; std::string name;
    mpm_plate descriptor() const {
        mpm_plate p; std::memset(&p, 0, sizeof(p));
        p.pos_object = pos_object; p.plate_type = (int32_t)plate_type; p.bc = (int32_t)bc; p.friction = friction;
        p.pos_lower = pos_lower; p.pos_upper = pos_upper;
        p.v_object[0] = vx_object; p.v_object[1] = vy_object; p.v_object[2] = vz_object;
        p.vmin_factor = vmin_factor; p.load_factor = load_factor;
        return p;
    }
};

// Simulation, src/simulation/simulation.hpp:23-225: public members with the reference's names and defaults.
class Simulation {
public:
    Simulation() {}
    ~Simulation() { if (ctx) mpm_destroy(ctx); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    int exit = 0;
    const unsigned int dim = kDim;
    unsigned int n_threads = 1;   // OpenMP threads of the reference; the GPU ignores it
    unsigned int end_frame = 1;
    bool is_initialized = false, save_sim = true, reduce_verbose = false, pbc = false;
    bool change_particle_positions = false, gravity_special = false, save_grid = false, use_mibf = false, use_musl = false;
    TV grid_reference_point = 2e10 * TV::Ones();
    TV gravity = TV::Zero();
    T fps = 1, cfl = 0.5, cfl_elastic = 0.5, flip_ratio = -0.95, rho = 1000, gravity_time = 0;
    T Lx = 1, Ly = 1, Lz = 1;
    Particles particles;
    unsigned int Np = 0;
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
    std::vector<std::unique_ptr<ObjectPlate>> plates;
    std::vector<std::unique_ptr<ObjectGeneral>> objects;

    // simulation.cpp:24-46 and createDirectory (simulation.cpp:49-71): the output folders are created when
    // saving; the copy of the set-up source that the reference keeps as provenance is not made.
    void initialize(bool save = true, std::string dir = "output/", std::string name = "dummy") {
        save_sim = save; directory = dir; sim_name = name; is_initialized = true;
        if (save_sim) {
            if (mkdir(directory.c_str(), 0777) == -1) std::cerr << "Directory " << directory << " has already been created" << std::endl;
            if (mkdir((directory + sim_name).c_str(), 0777) == -1) std::cerr << "Simulation " << sim_name << " has already been created" << std::endl;
        }
    }

    // simulation.cpp:73-213: the frame loop, one device batch per frame; particles are copied back to
    // `particles` at the end (and at every frame end when save_sim would have written them).
    void simulate() {
        if (!is_initialized) { std::cout << "Simulation not initialized. Call the initialize(...) function in your input file." << std::endl; return; }
        if (elastic_model != ElasticModel::Hencky && plastic_model != PlasticModel::NoPlasticity) {
            std::cout << "This plastic model is only compatible with Hencky's elasticity model" << std::endl; return;
        }
        if (!start()) return;
        frame = 0;
        if (save_sim) {  // simulation.cpp:160-163
            saveInfo();
            saveParticleData();
            mpm_set_profiling(ctx, 1);
        }
        const auto begin = std::chrono::steady_clock::now();
        while (true) {
            mpm_step_info info;
            const int rc = mpm_step(ctx, 1u << 30, MPM_STEP_STOP_AT_FRAME_END, &info);
            absorb(info);
            if (rc) { exit = 1; report(rc); break; }
            if (info.frame_ended) {  // simulation.cpp:180-190
                if (!reduce_verbose) std::cout << "End of frame " << info.frame << std::endl;
                if (save_sim) { saveParticleData(); if (save_grid) saveGridData(); }
            }
            if (info.finished) {  // simulation.cpp:191-199: the last frame is written again
                if (save_sim) { saveParticleData(); if (save_grid) saveGridData(); }
                break;
            }
            if (info.steps_done == 0) break;
        }
        runtime_total = (T)std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - begin).count();
        download();
        if (save_sim) saveTiming();
    }

    // Output files of save_data.cpp, written by the engine (binary little-endian PLY, float64 properties).
    // saveParticleData, save_data.cpp:7-171: particles_f<extra><frame>.ply and last_written.txt.
    void saveParticleData(std::string extra = "") {
        const std::string f = directory + sim_name + "/particles_f" + extra + std::to_string(frame) + ".ply";
        if (mpm_write_particles_ply(ctx, f.c_str())) { std::cerr << "Could not write " << f << std::endl; exit = 1; return; }
        std::ofstream out(directory + sim_name + "/last_written.txt");
        out << std::to_string(frame) << "\n";
    }
    // saveGridData, save_data.cpp:173-259: grid_f<extra><frame>.ply.
    void saveGridData(std::string extra = "") {
        const std::string f = directory + sim_name + "/grid_f" + extra + std::to_string(frame) + ".ply";
        if (mpm_write_grid_ply(ctx, f.c_str())) { std::cerr << "Could not write " << f << std::endl; exit = 1; }
    }
    // saveInfo, save_data.cpp:352-361.
    void saveInfo() const {
        std::ofstream info(directory + sim_name + "/info.txt");
        info << end_frame << "\n" << fps << "\n" << dx << "\n" << Np << "\n" << particle_volume << "\n";
    }
    // saveTiming, save_data.cpp:363-373. The phase timers are the engine's per-kernel device times, grouped
    // as the reference's phases: p2g (sort + P2G with the stress force), euler (grid update), g2p (G2P, MUSL
    // re-scatter), defgrad (deformation update and return map; in the fused USL kernel it counts with g2p).
    void saveTiming() {
        mpm_kernel_timing kt[16];
        uint32_t cnt = 0;
        T p2g = 0, euler = 0, g2p = 0, defgrad = 0;
        if (ctx && mpm_get_kernel_times(ctx, kt, 16, &cnt) == 0) {
            for (uint32_t i = 0; i < cnt && i < 16; ++i) {
                const std::string k = kt[i].name;
                const T sec = kt[i].total_ms * 1e-3;
                if (k == "k_p2g" || k.rfind("k_bin", 0) == 0 || k == "k_zero_grid") p2g += sec;
                else if (k == "k_grid_update") euler += sec;
                else if (k == "k_defgrad") defgrad += sec;
                else if (k.rfind("k_g2p", 0) == 0 || k.rfind("k_musl", 0) == 0) g2p += sec;
            }
        }
        std::ofstream out(directory + sim_name + "/info_timing.txt");
        out << step_ << "\n" << runtime_total / ((T)1000) << "\n" << p2g << "\n" << g2p << "\n" << euler << "\n" << defgrad << "\n";
    }

    // simulation.cpp:219-273 plus the loop bookkeeping of simulate() (time += dt, step++, frame);
    // the particles are copied back after the step.
    void advanceStep() {
        if (!ctx && !start()) return;
        mpm_step_info info;
        const int rc = mpm_step(ctx, 1, 0, &info);
        absorb(info);
        if (rc) { exit = 1; report(rc); }
        download();
    }

    T calculateBulkModulus() const {  // simulation.cpp:288-290
        const T lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
        return lambda + 2.0 * mu / dim;
    }

    // save_data.cpp:264-292 on the host copy (Hencky only): tau = 2 mu ln V + lambda tr(ln V) I.
    void computeAvgData(TM& volavg_cauchy, TM& volavg_kirchh, T& Javg);

    T time() const { return time_; }
    unsigned long long currentTimeStep() const { return step_; }
    T lastDt() const { return dt_; }
    mpm_context* context() { return ctx; }

private:
    mpm_context* ctx = nullptr;
    std::string directory, sim_name;
    unsigned int frame = 0;
    T runtime_total = 0;
    T time_ = 0, dt_ = 0;
    unsigned long long step_ = 0;

    bool start() {
        mpm_config c;
        mpm_config_defaults(&c);
        c.dim = kDim; c.spline_degree = MATTER_SPLINEDEG; c.precision = MPM_PRECISION_FP64; c.n_threads = (int32_t)n_threads;
        c.elastic_model = (int32_t)elastic_model; c.plastic_model = (int32_t)plastic_model; c.hardening_law = (int32_t)hardening_law;
        c.use_pradhana = use_pradhana; c.use_mises_q = use_mises_q; c.use_mibf = use_mibf; c.use_musl = use_musl;
        c.gravity_special = gravity_special; c.pbc = pbc; c.change_particle_positions = change_particle_positions;
        c.end_frame = end_frame;
        for (int d = 0; d < kDim; ++d) { c.grid_reference_point[d] = grid_reference_point(d); c.gravity[d] = gravity(d); }
        c.fps = fps; c.cfl = cfl; c.cfl_elastic = cfl_elastic; c.flip_ratio = flip_ratio; c.rho = rho; c.gravity_time = gravity_time;
        c.Lx = Lx; c.Ly = Ly; c.Lz = Lz;
        c.particle_mass = particle_mass; c.particle_volume = particle_volume; c.dx = dx;
        c.E = E; c.nu = nu; c.stress_tolerance = stress_tolerance; c.xi = xi;
        c.q_max = q_max; c.q_min = q_min; c.p_min = p_min; c.M = M; c.q_cohesion = q_cohesion;
        c.perzyna_exp = perzyna_exp; c.perzyna_visc = perzyna_visc; c.beta = beta; c.p0 = p0;
        c.rho_s = rho_s; c.grain_diameter = grain_diameter; c.I_ref = I_ref; c.mu_1 = mu_1; c.mu_2 = mu_2;
        if (ctx) { mpm_destroy(ctx); ctx = nullptr; }
        int rc = mpm_create(&c, &ctx);
        for (auto& p : plates) { if (rc) break; mpm_plate d = p->descriptor(); rc |= mpm_add_plate(ctx, &d); }
        for (auto& o : objects) { if (rc) break; rc |= o->attach(ctx); }
        if (!rc) {
            mpm_particle_arrays a = arrays();
            rc = mpm_set_particles(ctx, Np, &a);
        }
        if (rc) { exit = 1; report(rc); return false; }
        time_ = 0; step_ = 0;
        return true;
    }
    mpm_particle_arrays arrays() {
        mpm_particle_arrays a;
        a.x = particles.x.empty() ? nullptr : particles.x[0].a;
        a.v = particles.v.empty() ? nullptr : particles.v[0].a;
        a.F = particles.F.empty() ? nullptr : particles.F[0].a;
        a.Bmat = particles.Bmat.empty() ? nullptr : particles.Bmat[0].a;
        a.eps_pl_dev = particles.eps_pl_dev.data(); a.eps_pl_vol = particles.eps_pl_vol.data();
        a.eps_pl_vol_pradhana = particles.eps_pl_vol_pradhana.data(); a.delta_gamma = particles.delta_gamma.data();
        a.viscosity = particles.viscosity.data(); a.muI = particles.muI.data();
        return a;
    }
    void download() {
        if (!ctx) return;
        mpm_particle_arrays a = arrays();
        if (mpm_get_particles(ctx, &a)) exit = 1;
    }
    void absorb(const mpm_step_info& info) { time_ = info.time; step_ = info.step; dt_ = info.dt; frame = info.frame; }
    void report(int rc) const { std::cout << "Simulation stopped: " << mpm_status_string((uint32_t)rc) << std::endl; }
};

inline void Simulation::computeAvgData(TM& volavg_cauchy, TM& volavg_kirchh, T& Javg) {
    // V^2 = F F^T; ln V by Jacobi eigen-decomposition of the symmetric F F^T
    const T lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    volavg_cauchy = TM::Zero(); volavg_kirchh = TM::Zero();
    T Jsum = 0;
    for (unsigned p = 0; p < Np; ++p) {
        const TM& F = particles.F[p];
        T b[3][3] = {{0}};
        for (int r = 0; r < kDim; ++r) for (int c = 0; c < kDim; ++c) { T s = 0; for (int k = 0; k < kDim; ++k) s += F(r, k) * F(c, k); b[r][c] = s; }
        T Q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int sweep = 0; sweep < 50; ++sweep) {
            T off = 0; for (int r = 0; r < kDim; ++r) for (int c = r + 1; c < kDim; ++c) off += b[r][c] * b[r][c];
            if (off < 1e-30) break;
            for (int pp = 0; pp < kDim; ++pp) for (int qq = pp + 1; qq < kDim; ++qq) {
                if (b[pp][qq] == 0) continue;
                T th = (b[qq][qq] - b[pp][pp]) / (2 * b[pp][qq]);
                T t = (th >= 0 ? 1 : -1) / (std::fabs(th) + std::sqrt(th * th + 1));
                T c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < kDim; ++k) { T x = b[k][pp], y = b[k][qq]; b[k][pp] = c * x - s * y; b[k][qq] = s * x + c * y; }
                for (int k = 0; k < kDim; ++k) { T x = b[pp][k], y = b[qq][k]; b[pp][k] = c * x - s * y; b[qq][k] = s * x + c * y; }
                for (int k = 0; k < kDim; ++k) { T x = Q[k][pp], y = Q[k][qq]; Q[k][pp] = c * x - s * y; Q[k][qq] = s * x + c * y; }
            }
        }
        T l[3], tr = 0; for (int i = 0; i < kDim; ++i) { l[i] = 0.5 * std::log(b[i][i]); tr += l[i]; }
        TM tau;
        for (int r = 0; r < kDim; ++r) for (int c = 0; c < kDim; ++c) {
            T s = 0; for (int k = 0; k < kDim; ++k) s += Q[r][k] * (2 * mu * l[k]) * Q[c][k];
            tau(r, c) = s + (r == c ? lambda * tr : 0.0);
        }
        const T J = F.determinant() * std::exp(particles.eps_pl_vol[p]);
        Jsum += J;
        for (int i = 0; i < kDim * kDim; ++i) { volavg_cauchy.a[i] += tau.a[i]; volavg_kirchh.a[i] += tau.a[i] * J; }
    }
    for (int i = 0; i < kDim * kDim; ++i) { volavg_cauchy.a[i] /= Jsum; volavg_kirchh.a[i] /= Jsum; }
    Javg = Jsum / Np;
}

// Poisson-disk sampling of the box [lo, hi] with minimum distance r: Bridson's algorithm (R. Bridson, "Fast
// Poisson disk sampling in arbitrary dimensions", SIGGRAPH 2007 sketches). A background grid of cell size
// r/sqrt(dim) holds at most one sample per cell; a random active sample spawns up to `attempts` candidates
// drawn uniformly from the shell [r, 2r) around it and retires when none of them is at least r from every
// sample. The first sample is uniform in the box. Same role and density as tph_poisson's sampler
// (sampling_particles.hpp:23), with this file's own random stream (std::mt19937_64 seeded with `seed`).
inline std::vector<std::array<T, 3>> poissonDiskSampling(T r, const T* lo, const T* hi, std::uint32_t attempts, std::uint32_t seed) {
    const T cell = r / std::sqrt(T(kDim));
    long long n[3] = {1, 1, 1};
    for (int d = 0; d < kDim; ++d) n[d] = std::max(1LL, (long long)std::ceil((hi[d] - lo[d]) / cell));
    std::vector<int> grid((size_t)(n[0] * n[1] * n[2]), -1);
    std::vector<std::array<T, 3>> pts;
    std::vector<int> active;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<T> u01(0.0, 1.0);
    auto cell_of = [&](const std::array<T, 3>& x, long long* c) {
        for (int d = 0; d < 3; ++d) c[d] = d < kDim ? std::min(n[d] - 1, std::max(0LL, (long long)std::floor((x[d] - lo[d]) / cell))) : 0;
    };
    auto flat = [&](const long long* c) { return (size_t)((c[0] * n[1] + c[1]) * n[2] + c[2]); };
    auto add = [&](const std::array<T, 3>& x) {
        long long c[3]; cell_of(x, c);
        grid[flat(c)] = (int)pts.size();
        active.push_back((int)pts.size());
        pts.push_back(x);
    };
    auto far_enough = [&](const std::array<T, 3>& x) {
        long long c[3]; cell_of(x, c);
        const int reach = kDim == 3 ? 2 : 2;  // cells within distance r of a point, cell = r/sqrt(dim)
        long long a[3], b[3];
        for (int d = 0; d < 3; ++d) { a[d] = d < kDim ? std::max(0LL, c[d] - reach) : 0; b[d] = d < kDim ? std::min(n[d] - 1, c[d] + reach) : 0; }
        for (long long i = a[0]; i <= b[0]; ++i) for (long long j = a[1]; j <= b[1]; ++j) for (long long k = a[2]; k <= b[2]; ++k) {
            const long long cc[3] = {i, j, k};
            const int q = grid[flat(cc)];
            if (q < 0) continue;
            T s = 0; for (int d = 0; d < kDim; ++d) { const T e = pts[q][d] - x[d]; s += e * e; }
            if (s < r * r) return false;
        }
        return true;
    };
    std::array<T, 3> x0 = {0, 0, 0};
    for (int d = 0; d < kDim; ++d) x0[d] = lo[d] + u01(rng) * (hi[d] - lo[d]);
    add(x0);
    while (!active.empty()) {
        const size_t ai = (size_t)(u01(rng) * active.size()) % active.size();
        const std::array<T, 3> base = pts[active[ai]];
        bool placed = false;
        for (std::uint32_t t = 0; t < attempts && !placed; ++t) {
            T dir[3] = {0, 0, 0}, nn = 0;
            do {  // uniform direction by rejection from the unit ball
                nn = 0;
                for (int d = 0; d < kDim; ++d) { dir[d] = 2 * u01(rng) - 1; nn += dir[d] * dir[d]; }
            } while (nn > 1 || nn < 1e-12);
            nn = std::sqrt(nn);
            const T rr = kDim == 3 ? std::cbrt(r * r * r * (1 + 7 * u01(rng))) : std::sqrt(r * r * (1 + 3 * u01(rng)));
            std::array<T, 3> x = {0, 0, 0};
            bool in = true;
            for (int d = 0; d < kDim; ++d) { x[d] = base[d] + rr * dir[d] / nn; in = in && x[d] >= lo[d] && x[d] <= hi[d]; }
            if (in && far_enough(x)) { add(x); placed = true; }
        }
        if (!placed) { active[ai] = active.back(); active.pop_back(); }
    }
    return pts;
}

// sampleParticles (src/sampling/sampling_particles.hpp:12-71; 2D :75-123): Poisson-disk samples of
// [0,Lx] x [0,Ly] (x [0,Lz]) with radius kRadius, then dx^dim = ppc * Lx*Ly*Lz / N, particle_volume = dx^dim/ppc,
// particle_mass = rho * particle_volume. crop_to_shape 1 and 2 (cylinder, silo / 2D gate) are not offered.
template <class S>
void sampleParticles(S& sim, T kRadius, T ppc = (kDim == 3 ? 8 : 6), unsigned crop_to_shape = 0, std::uint32_t attempts = (kDim == 3 ? 30 : 200),
                     std::uint32_t seed = 42) {
    if (crop_to_shape != 0) throw std::invalid_argument("sampleParticles: crop_to_shape is not supported");
    const T lo[3] = {0, 0, 0};
    const T hi[3] = {sim.Lx, sim.Ly, kDim == 3 ? sim.Lz : 0.0};
    const std::vector<std::array<T, 3>> s = poissonDiskSampling(kRadius, lo, hi, attempts, seed);
    const T vol = sim.Lx * sim.Ly * (kDim == 3 ? sim.Lz : 1.0);
    sim.dx = std::pow(ppc / T(s.size()) * vol, 1.0 / kDim);
    sim.particle_volume = std::pow(sim.dx, kDim) / ppc;
    sim.particle_mass = sim.rho * sim.particle_volume;
    sim.Np = (unsigned)s.size();
    sim.particles = Particles(sim.Np);
    for (unsigned p = 0; p < sim.Np; ++p) for (int d = 0; d < kDim; ++d) sim.particles.x[p](d) = s[p][d];
}

}  // namespace matter
