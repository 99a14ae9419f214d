// oracle_abi.cpp — TEST INFRASTRUCTURE ONLY. C entry points into the CPU oracle (oracle_sim.hpp) so that
// pytest can drive it through ctypes with the same mpm_config / mpm_particle_arrays structs as the product's
// C ABI (include/mpm.h). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this.
#include "oracle_sim.hpp"
#include <cstring>
#include <chrono>

using namespace oracle;

namespace {

struct OracleBase {
    virtual ~OracleBase() {}
    virtual int add_plate(const mpm_plate*) = 0;
    virtual int add_object(const mpm_object*) = 0;
    virtual int set_particles(uint32_t n, const mpm_particle_arrays*) = 0;
    virtual int step(uint32_t nsteps, uint32_t flags, mpm_step_info*) = 0;
    virtual int get_particles(mpm_particle_arrays*) = 0;
    virtual int get_grid(mpm_grid_info*, double*, double*, double*, double*, double*, double*, double*) = 0;
    virtual int get_plates(mpm_plate*, uint32_t, uint32_t*) = 0;
    virtual uint32_t num_particles() const = 0;
    virtual void set_threads(int) = 0;
    virtual int avg_data(double*, double*, double*) = 0;
    virtual double bulk_modulus() const = 0;
    virtual void phase_times(double*) const = 0;
    virtual int object_eval(uint32_t, const double*, int*, double*) const = 0;
    virtual int add_sdf(const mpm_sdf*) = 0;
};

template <int D, int SPL> struct OracleImpl : OracleBase {
    typedef Simulation<D, SPL> Sim;
    Sim sim;
    bool prepared = false;
    uint32_t sticky = 0;

    explicit OracleImpl(const mpm_config* c) {
        sim.n_threads = c->n_threads > 0 ? c->n_threads : 1;
        sim.end_frame = c->end_frame;
        sim.elastic_model = (ElasticModel)c->elastic_model;
        sim.plastic_model = (PlasticModel)c->plastic_model;
        sim.hardening_law = (HardeningLaw)c->hardening_law;
        sim.use_pradhana = c->use_pradhana; sim.use_mises_q = c->use_mises_q; sim.use_mibf = c->use_mibf;
        sim.use_musl = c->use_musl; sim.gravity_special = c->gravity_special;
        sim.change_particle_positions = c->change_particle_positions;
        sim.pbc = c->pbc != 0;
        for (int d = 0; d < D; ++d) { sim.grid_reference_point(d) = c->grid_reference_point[d]; sim.gravity(d) = c->gravity[d]; }
        sim.fps = c->fps; sim.cfl = c->cfl; sim.cfl_elastic = c->cfl_elastic; sim.flip_ratio = c->flip_ratio;
        sim.rho = c->rho; sim.gravity_time = c->gravity_time;
        sim.Lx = c->Lx; sim.Ly = c->Ly; sim.Lz = c->Lz;
        sim.particle_mass = c->particle_mass; sim.particle_volume = c->particle_volume; sim.dx = c->dx;
        sim.E = c->E; sim.nu = c->nu; sim.stress_tolerance = c->stress_tolerance; sim.xi = c->xi;
        sim.q_max = c->q_max; sim.q_min = c->q_min; sim.p_min = c->p_min;
        sim.M = c->M; sim.q_cohesion = c->q_cohesion;
        sim.perzyna_exp = c->perzyna_exp; sim.perzyna_visc = c->perzyna_visc;
        sim.beta = c->beta; sim.p0 = c->p0;
        sim.rho_s = c->rho_s; sim.grain_diameter = c->grain_diameter; sim.I_ref = c->I_ref;
        sim.mu_1 = c->mu_1; sim.mu_2 = c->mu_2;
    }
    int add_plate(const mpm_plate* p) override {
        sim.plates.push_back(std::make_unique<ObjectPlate<D>>(p->pos_object, (PlateType)p->plate_type, (BC)p->bc, p->friction,
            p->pos_lower, p->pos_upper, p->v_object[0], p->v_object[1], D == 3 ? p->v_object[2] : 0.0, p->vmin_factor, p->load_factor));
        return 0;
    }
    int add_object(const mpm_object* o) override {
        BC bc = (BC)o->bc; T f = o->friction;
        switch (o->kind) {
            case MPM_OBJECT_BOX_ROTATED: {
                Vec<D> c, h; Mat<D> R;
                for (int i = 0; i < D; ++i) { c(i) = o->center[i]; h(i) = o->half_extents[i]; }
                for (int i = 0; i < D * D; ++i) R.a[i] = o->rotation[i];
                sim.objects.push_back(std::make_unique<ObjectBoxRotated<D>>(bc, f, c, h, R)); break;
            }
            case MPM_OBJECT_BUMP: sim.objects.push_back(std::make_unique<ObjectBump<D>>(bc, f)); break;
            case MPM_OBJECT_GROUND: sim.objects.push_back(std::make_unique<ObjectGround<D>>(bc, f, o->param)); break;
            case MPM_OBJECT_RAMP: sim.objects.push_back(std::make_unique<ObjectRamp<D>>(bc, f)); break;
            case MPM_OBJECT_GATE: sim.objects.push_back(std::make_unique<ObjectGate<D>>(bc, f, o->param)); break;
            case MPM_OBJECT_SILO: if (D != 3) return MPM_ERR_UNSUPPORTED; sim.objects.push_back(std::make_unique<ObjectSilo<D>>(bc, f, o->param)); break;
            case MPM_OBJECT_CURVE: sim.objects.push_back(std::make_unique<ObjectCurve<D>>(bc, f)); break;
            default: return MPM_ERR_INVALID_ARGUMENT;
        }
        return 0;
    }
    int set_particles(uint32_t n, const mpm_particle_arrays* in) override {
        sim.Np = n;
        sim.particles = Particles<D>(n);
        for (uint32_t p = 0; p < n; ++p) {
            if (in->x) for (int d = 0; d < D; ++d) sim.particles.x[p](d) = in->x[p * D + d];
            if (in->v) for (int d = 0; d < D; ++d) sim.particles.v[p](d) = in->v[p * D + d];
            if (in->F) for (int i = 0; i < D * D; ++i) sim.particles.F[p].a[i] = in->F[p * D * D + i];
            if (in->Bmat) for (int i = 0; i < D * D; ++i) sim.particles.Bmat[p].a[i] = in->Bmat[p * D * D + i];
            if (in->eps_pl_dev) sim.particles.eps_pl_dev[p] = in->eps_pl_dev[p];
            if (in->eps_pl_vol) sim.particles.eps_pl_vol[p] = in->eps_pl_vol[p];
            if (in->eps_pl_vol_pradhana) sim.particles.eps_pl_vol_pradhana[p] = in->eps_pl_vol_pradhana[p];
            if (in->delta_gamma) sim.particles.delta_gamma[p] = in->delta_gamma[p];
            if (in->viscosity) sim.particles.viscosity[p] = in->viscosity[p];
            if (in->muI) sim.particles.muI[p] = in->muI[p];
        }
        sim.current_time_step = 0;
        sim.exit = 0;
        if (!sim.prepare()) return MPM_ERR_UNSUPPORTED;
        prepared = true;
        return 0;
    }
    int step(uint32_t nsteps, uint32_t flags, mpm_step_info* info) override {
        if (!prepared) return MPM_ERR_NOT_READY;
        sim.check_conservation = (flags & MPM_STEP_CHECK_CONSERVATION) != 0;
        uint32_t done = 0; bool fe = false, finished = false;
        sim.warnings = 0;
        uint32_t before = sim.exit;
        for (uint32_t s = 0; s < nsteps; ++s) {
            if (sim.frame >= sim.end_frame) { finished = true; break; }
            bool cont = sim.loopIteration(fe);
            ++done;
            if (!cont) { finished = !sim.exit; break; }
            if (fe && (flags & MPM_STEP_STOP_AT_FRAME_END)) break;
        }
        if (sim.frame >= sim.end_frame) finished = true;
        uint32_t raised = sim.exit & ~before;
        if (info) {
            info->dt = sim.dt; info->time = sim.time; info->step = sim.current_time_step; info->frame = sim.frame;
            info->plastic_count = sim.last_plastic_count; info->steps_done = done;
            info->status = raised | sim.warnings; info->frame_ended = fe; info->finished = finished;
        }
        return (int)(raised & MPM_ERR_EXIT_MASK);
    }
    int get_particles(mpm_particle_arrays* out) override {
        for (uint32_t p = 0; p < sim.Np; ++p) {
            if (out->x) for (int d = 0; d < D; ++d) out->x[p * D + d] = sim.particles.x[p](d);
            if (out->v) for (int d = 0; d < D; ++d) out->v[p * D + d] = sim.particles.v[p](d);
            if (out->F) for (int i = 0; i < D * D; ++i) out->F[p * D * D + i] = sim.particles.F[p].a[i];
            if (out->Bmat) for (int i = 0; i < D * D; ++i) out->Bmat[p * D * D + i] = sim.particles.Bmat[p].a[i];
            if (out->eps_pl_dev) out->eps_pl_dev[p] = sim.particles.eps_pl_dev[p];
            if (out->eps_pl_vol) out->eps_pl_vol[p] = sim.particles.eps_pl_vol[p];
            if (out->eps_pl_vol_pradhana) out->eps_pl_vol_pradhana[p] = sim.particles.eps_pl_vol_pradhana[p];
            if (out->delta_gamma) out->delta_gamma[p] = sim.particles.delta_gamma[p];
            if (out->viscosity) out->viscosity[p] = sim.particles.viscosity[p];
            if (out->muI) out->muI[p] = sim.particles.muI[p];
        }
        return 0;
    }
    int get_grid(mpm_grid_info* info, double* x, double* y, double* z, double* mass, double* v, double* flip, double* fric) override {
        if (info) {
            info->Nx = sim.Nx; info->Ny = sim.Ny; info->Nz = (D == 3) ? sim.Nz : 1; info->grid_nodes = sim.grid_nodes;
            info->xc = sim.grid.xc; info->yc = sim.grid.yc; info->zc = (D == 3) ? sim.grid.zc : 0.0;
        }
        if (x) std::memcpy(x, sim.grid.x.data(), sizeof(double) * sim.grid.x.size());
        if (y) std::memcpy(y, sim.grid.y.data(), sizeof(double) * sim.grid.y.size());
        if (z && D == 3) std::memcpy(z, sim.grid.z.data(), sizeof(double) * sim.grid.z.size());
        for (unsigned l = 0; l < sim.grid_nodes && l < sim.grid.mass.size(); ++l) {
            if (mass) mass[l] = sim.grid.mass[l];
            if (v) for (int d = 0; d < D; ++d) v[l * D + d] = sim.grid.v[l](d);
            if (flip) for (int d = 0; d < D; ++d) flip[l * D + d] = sim.grid.flip[l](d);
            if (fric && sim.use_mibf) fric[l] = sim.grid.friction[l];
        }
        return 0;
    }
    int get_plates(mpm_plate* out, uint32_t cap, uint32_t* count) override {
        if (count) *count = sim.plates.size();
        for (uint32_t i = 0; i < sim.plates.size() && i < cap; ++i) {
            const ObjectPlate<D>& p = *sim.plates[i];
            out[i].pos_object = p.pos_object; out[i].plate_type = (int32_t)p.plate_type; out[i].bc = (int32_t)p.bc;
            out[i].friction = p.friction; out[i].pos_lower = p.pos_lower; out[i].pos_upper = p.pos_upper;
            for (int d = 0; d < 3; ++d) out[i].v_object[d] = p.v_object[d];
            out[i].vmin_factor = p.vmin_factor; out[i].load_factor = p.load_factor;
        }
        return 0;
    }
    uint32_t num_particles() const override { return sim.Np; }
    void set_threads(int t) override { sim.n_threads = t > 0 ? t : 1; }
    int avg_data(double* cauchy, double* kirchh, double* Javg) override {
        Mat<D> c, k; T J; sim.computeAvgData(c, k, J);
        for (int i = 0; i < D * D; ++i) { cauchy[i] = c.a[i]; kirchh[i] = k.a[i]; }
        *Javg = J;
        return 0;
    }
    double bulk_modulus() const override { return sim.calculateBulkModulus(); }
    void phase_times(double* t) const override {
        t[0] = sim.runtime_p2g; t[1] = sim.runtime_euler; t[2] = sim.runtime_g2p; t[3] = sim.runtime_defgrad;
    }
    int add_sdf(const mpm_sdf* sd) override {
        unsigned nn[3] = {sd->n[0], sd->n[1], sd->n[2]};
        sim.objects.push_back(std::make_unique<ObjectSdf<D>>((BC)sd->bc, sd->friction, sd->origin, sd->spacing, nn, sd->phi));
        return 0;
    }
    int object_eval(uint32_t k, const double* X, int* inside, double* normal) const override {
        if (k >= sim.objects.size()) return MPM_ERR_INVALID_ARGUMENT;
        Vec<D> x; for (int d = 0; d < D; ++d) x(d) = X[d];
        *inside = sim.objects[k]->inside(x) ? 1 : 0;
        const Vec<D> n = sim.objects[k]->normal(x);
        for (int d = 0; d < D; ++d) normal[d] = n(d);
        return 0;
    }
};

OracleBase* make(const mpm_config* c) {
    if (c->dim == 3 && c->spline_degree == 2) return new OracleImpl<3, 2>(c);
    if (c->dim == 3 && c->spline_degree == 3) return new OracleImpl<3, 3>(c);
    if (c->dim == 2 && c->spline_degree == 2) return new OracleImpl<2, 2>(c);
    if (c->dim == 2 && c->spline_degree == 3) return new OracleImpl<2, 3>(c);
    return nullptr;
}

template <int D> void svd_impl(const double* A, double* U, double* S, double* V) {
    Mat<D> a, u, v; Vec<D> s;
    for (int i = 0; i < D * D; ++i) a.a[i] = A[i];
    svd(a, u, s, v);
    for (int i = 0; i < D * D; ++i) { U[i] = u.a[i]; V[i] = v.a[i]; }
    for (int i = 0; i < D; ++i) S[i] = s(i);
}

}  // namespace

extern "C" {

void* oracle_create(const mpm_config* cfg) { return cfg ? make(cfg) : nullptr; }
void oracle_destroy(void* h) { delete (OracleBase*)h; }
int oracle_add_plate(void* h, const mpm_plate* p) { return ((OracleBase*)h)->add_plate(p); }
int oracle_add_object(void* h, const mpm_object* o) { return ((OracleBase*)h)->add_object(o); }
int oracle_add_sdf_object(void* h, const mpm_sdf* s) { return ((OracleBase*)h)->add_sdf(s); }
int oracle_set_particles(void* h, uint32_t n, const mpm_particle_arrays* in) { return ((OracleBase*)h)->set_particles(n, in); }
int oracle_step(void* h, uint32_t nsteps, uint32_t flags, mpm_step_info* info) { return ((OracleBase*)h)->step(nsteps, flags, info); }
int oracle_get_particles(void* h, mpm_particle_arrays* out) { return ((OracleBase*)h)->get_particles(out); }
uint32_t oracle_num_particles(void* h) { return ((OracleBase*)h)->num_particles(); }
int oracle_get_grid(void* h, mpm_grid_info* info, double* x, double* y, double* z, double* mass, double* v, double* flip, double* fric) {
    return ((OracleBase*)h)->get_grid(info, x, y, z, mass, v, flip, fric);
}
int oracle_get_plates(void* h, mpm_plate* out, uint32_t cap, uint32_t* count) { return ((OracleBase*)h)->get_plates(out, cap, count); }
void oracle_set_threads(void* h, int t) { ((OracleBase*)h)->set_threads(t); }
int oracle_avg_data(void* h, double* cauchy, double* kirchh, double* Javg) { return ((OracleBase*)h)->avg_data(cauchy, kirchh, Javg); }
double oracle_bulk_modulus(void* h) { return ((OracleBase*)h)->bulk_modulus(); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

// ---- primitives for known-answer tests ----
double oracle_spline_N(int degree, double u) { return degree == 3 ? Spline<3>::N(u) : Spline<2>::N(u); }
double oracle_spline_dN(int degree, double u) { return degree == 3 ? Spline<3>::dNdu(u) : Spline<2>::dNdu(u); }
int oracle_arange(double start, double stop, double step, double* out, int cap) {
    std::vector<double> v = arange(start, stop, step);
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
    return (int)v.size();
}
void oracle_svd(int dim, const double* A, double* U, double* S, double* V) {
    if (dim == 3) svd_impl<3>(A, U, S, V); else svd_impl<2>(A, U, S, V);
}

// One call of plasticity() on a single particle with trial deformation Fe_trial (column-major dim x dim).
// scalars (in/out): eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana, delta_gamma, viscosity, muI.
// Returns the error bits; *plastic receives the plastic_count increment.
int oracle_plasticity_single(const mpm_config* cfg, double dt, const double* Fe_trial, double* scalars, double* F_out, uint32_t* plastic) {
    OracleBase* b = make(cfg);
    if (!b) return MPM_ERR_UNSUPPORTED;
    int err = 0;
    auto body = [&](auto* impl) {
        auto& s = impl->sim;
        const int D = s.dim;
        s.Np = 1;
        s.particles = decltype(s.particles)(1);
        s.prepare();
        s.dt = dt;
        s.particles.eps_pl_dev[0] = scalars[0]; s.particles.eps_pl_vol[0] = scalars[1];
        s.particles.eps_pl_vol_pradhana[0] = scalars[2]; s.particles.delta_gamma[0] = scalars[3];
        s.particles.viscosity[0] = scalars[4]; s.particles.muI[0] = scalars[5];
        auto Fe = s.particles.F[0];
        for (int i = 0; i < D * D; ++i) Fe.a[i] = Fe_trial[i];
        s.particles.F[0] = Fe;
        unsigned pc = 0;
        err = s.plasticity(0, pc, Fe);
        for (int i = 0; i < D * D; ++i) F_out[i] = s.particles.F[0].a[i];
        scalars[0] = s.particles.eps_pl_dev[0]; scalars[1] = s.particles.eps_pl_vol[0];
        scalars[2] = s.particles.eps_pl_vol_pradhana[0]; scalars[3] = s.particles.delta_gamma[0];
        scalars[4] = s.particles.viscosity[0]; scalars[5] = s.particles.muI[0];
        if (plastic) *plastic = pc;
    };
    if (auto* i32 = dynamic_cast<OracleImpl<3, 2>*>(b)) body(i32);
    else if (auto* i33 = dynamic_cast<OracleImpl<3, 3>*>(b)) body(i33);
    else if (auto* i22 = dynamic_cast<OracleImpl<2, 2>*>(b)) body(i22);
    else if (auto* i23 = dynamic_cast<OracleImpl<2, 3>*>(b)) body(i23);
    delete b;
    return err;
}

// Kirchhoff stress tau = P(F) F^T of the configured elastic model (HenckyPiola, simulation.hpp:231-238).
int oracle_kirchhoff(const mpm_config* cfg, const double* F, double* tau) {
    OracleBase* b = make(cfg);
    if (!b) return MPM_ERR_UNSUPPORTED;
    auto body = [&](auto* impl) {
        auto& s = impl->sim;
        const int D = s.dim;
        s.prepare();
        auto Fm = s.particles.F[0];
        for (int i = 0; i < D * D; ++i) Fm.a[i] = F[i];
        auto t = s.kirchhoff(Fm);
        for (int i = 0; i < D * D; ++i) tau[i] = t.a[i];
    };
    if (auto* i32 = dynamic_cast<OracleImpl<3, 2>*>(b)) body(i32);
    else if (auto* i33 = dynamic_cast<OracleImpl<3, 3>*>(b)) body(i33);
    else if (auto* i22 = dynamic_cast<OracleImpl<2, 2>*>(b)) body(i22);
    else if (auto* i23 = dynamic_cast<OracleImpl<2, 3>*>(b)) body(i23);
    delete b;
    return 0;
}

// boundaryCollision on one node (boundary_collision.cpp:5-394): plates and objects of the handle, with the
// handle's dt set to `dt` and grid friction `node_friction` when use_mibf.
int oracle_boundary_single(void* h, double dt, const double* Xi, double* vi, double node_friction) {
    OracleBase* b = (OracleBase*)h;
    int err = 0;
    auto body = [&](auto* impl) {
        auto& s = impl->sim;
        const int D = s.dim;
        s.dt = dt;
        s.grid.friction.assign(1, node_friction);
        auto X = s.particles.x[0]; auto v = s.particles.v[0];
        for (int d = 0; d < D; ++d) { X(d) = Xi[d]; v(d) = vi[d]; }
        err = s.boundaryCollision(0, X, v);
        for (int d = 0; d < D; ++d) vi[d] = v(d);
    };
    if (auto* i32 = dynamic_cast<OracleImpl<3, 2>*>(b)) body(i32);
    else if (auto* i33 = dynamic_cast<OracleImpl<3, 3>*>(b)) body(i33);
    else if (auto* i22 = dynamic_cast<OracleImpl<2, 2>*>(b)) body(i22);
    else if (auto* i23 = dynamic_cast<OracleImpl<2, 3>*>(b)) body(i23);
    return err;
}

// Wall-clock seconds of `nsteps` loop iterations (steady clock), for the CPU baseline.
double oracle_time_steps(void* h, uint32_t nsteps, mpm_step_info* info) {
    auto t0 = std::chrono::steady_clock::now();
    ((OracleBase*)h)->step(nsteps, 0, info);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// Accumulated phase timers of advanceStep in seconds: P2G, explicitEulerUpdate, G2P, deformationUpdate
// (simulation.cpp:237-266; the reference prints them at the end of simulate(), :205-209).
void oracle_phase_times(void* h, double* out) { ((OracleBase*)h)->phase_times(out); }

// inside() and normal() of analytic object k of the handle (ObjectGeneral, object_general.hpp:19-21) at X.
int oracle_object_eval(void* h, uint32_t k, const double* X, int* inside, double* normal) {
    return ((OracleBase*)h)->object_eval(k, X, inside, normal);
}

}  // extern "C"
