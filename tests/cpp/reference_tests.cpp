// The reference's physics tests (tests/tests.cpp:13-736) written against the C++ host API of the engine
// (matter-tests_amd/host/matter.hpp). Each test keeps the reference's parameters and tolerance; particle
// set-up uses the Poisson-disk sampleParticles of matter.hpp (Bridson's algorithm) in place of tph_poisson.
//
//   reference_tests [Suite.Name ...]     runs the named tests (all when none is given); exit code 0 = pass
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <string>

#include "../../matter-tests_amd/host/matter.hpp"

using namespace matter;

namespace {
std::map<std::string, std::function<bool()>>& registry() { static std::map<std::string, std::function<bool()>> r; return r; }
struct Reg { Reg(const char* n, std::function<bool()> f) { registry()[n] = std::move(f); } };
bool check_near(const char* what, T val, T ref, T tol) {
    const bool ok = std::fabs(val - ref) <= tol;
    std::printf("  %-44s %.6e (ref %.6e, tol %.1e) %s\n", what, (double)val, (double)ref, (double)tol, ok ? "ok" : "FAIL");
    return ok;
}
T max_x(const Simulation& s) {
    return (*std::max_element(s.particles.x.begin(), s.particles.x.end(), [](const TV& a, const TV& b) { return a(0) < b(0); }))(0);
}
}  // namespace

#define MATTER_TEST(suite, name) \
    static bool suite##_##name(); \
    static Reg reg_##suite##_##name(#suite "." #name, suite##_##name); \
    static bool suite##_##name()

// BoundaryTest.AnalyticSlipFree / AnalyticSlipStick (tests/tests.cpp:13-98): one particle released at
// (-1, 1) on the frictionless parabola y = x^2 reaches v_x = sqrt(2 g) at the bottom after half a swing.
static bool slide_on_curve(BC bc) {
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 100;
    sim.fps = 100 / 0.5949238427;
    sim.cfl = 0.5;
    sim.flip_ratio = -0.95;
    sim.E = 1e6; sim.nu = 0.3; sim.rho = 1000;
    sim.gravity[1] = -9.81;
    sim.dx = 0.001;
    sim.particle_volume = std::pow(sim.dx, sim.dim);
    sim.particle_mass = sim.rho * sim.particle_volume;
    sim.Np = 1;
    sim.particles = Particles(sim.Np);
    sim.particles.x[0](0) = -1;
    sim.particles.x[0](1) = 1;
    sim.objects.push_back(std::make_unique<ObjectCurve>(bc, 0.0));
    sim.simulate();
    const T v_true = std::sqrt(2 * 9.81);
    return sim.exit == 0 && check_near("relative error of v_x", std::fabs(sim.particles.v[0](0) - v_true) / v_true, 0.0, 1e-3);
}
MATTER_TEST(BoundaryTest, AnalyticSlipFree) { return slide_on_curve(BC::SlipFree); }
MATTER_TEST(BoundaryTest, AnalyticSlipStick) { return slide_on_curve(BC::SlipStick); }

// CoulombFrictionTest.* (tests/tests.cpp:100-306): block sliding down a floor with Coulomb friction
// tan 15 deg under gravity tilted 24 deg; mean x after 5 s is g t^2 (sin - mu cos)/2.
static bool coulomb(BC bc, bool general) {
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 5;
    sim.fps = 1;
    sim.cfl = 0.5;
    sim.flip_ratio = -0.95;
    const T theta = 24 * M_PI / 180;
    sim.gravity = TV::Zero();
    sim.gravity[0] = +9.81 * std::sin(theta);
    sim.gravity[1] = -9.81 * std::cos(theta);
    sim.E = 1e6; sim.nu = 0.3; sim.rho = 1000;
    sim.Lx = 0.1; sim.Ly = 0.05;
    sampleParticles(sim, 0.005);
    for (unsigned p = 0; p < sim.Np; ++p) { sim.particles.x[p](0) -= 0.5 * sim.Lx; sim.particles.x[p](1) += 0.5 * sim.dx; }
    sim.grid_reference_point = TV::Zero();
    const T friction = std::tan(15.0 * M_PI / 180.0);
    if (general) sim.objects.push_back(std::make_unique<ObjectGround>(bc, friction));
    else sim.plates.push_back(std::make_unique<ObjectPlate>(0, PlateType::bottom, bc, friction));
    sim.simulate();
    T mean_x = 0;
    for (unsigned p = 0; p < sim.Np; ++p) mean_x += sim.particles.x[p](0);
    mean_x /= (T)sim.Np;
    const T t = (T)sim.end_frame / sim.fps;
    const T ref = 0.5 * 9.81 * t * t * (std::sin(theta) - friction * std::cos(theta));
    return sim.exit == 0 && check_near("relative error of mean x", std::fabs(mean_x - ref) / ref, 0.0, 1e-3);
}
MATTER_TEST(CoulombFrictionTest, PlateSlipFree) { return coulomb(BC::SlipFree, false); }
MATTER_TEST(CoulombFrictionTest, PlateSlipStick) { return coulomb(BC::SlipStick, false); }
MATTER_TEST(CoulombFrictionTest, GeneralSlipFree) { return coulomb(BC::SlipFree, true); }
MATTER_TEST(CoulombFrictionTest, GeneralSlipStick) { return coulomb(BC::SlipStick, true); }

// BoundaryTest.MIBF (tests/tests.cpp:308-474): DPVisc with zero viscosity on a 32 deg slope; plate friction,
// MIBF friction and the von Mises q definition with M*sqrt(3) give the same maximum x.
static T mibf_run(bool mibf, bool mises) {
    const T friction = std::tan(30.0 * M_PI / 180.0);
    const T theta = 32 * M_PI / 180;
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 1;
    sim.fps = 2;
    sim.cfl = 0.5;
    sim.flip_ratio = -0.95;
    sim.gravity = TV::Zero();
    sim.gravity[0] = +9.81 * std::sin(theta);
    sim.gravity[1] = -9.81 * std::cos(theta);
    sim.E = 1e5; sim.nu = 0.3; sim.rho = 1000;
    sim.Lx = 0.1; sim.Ly = 0.05; sim.Lz = 0.05;
    sampleParticles(sim, 0.001);
    for (unsigned p = 0; p < sim.Np; ++p) { sim.particles.x[p](0) -= 0.5 * sim.Lx; sim.particles.x[p](1) += 0.5 * sim.dx; }
    sim.grid_reference_point = TV::Zero();
    sim.plates.push_back(std::make_unique<ObjectPlate>(0, PlateType::bottom, BC::SlipFree, friction));
    sim.plastic_model = PlasticModel::DPVisc;
    sim.use_pradhana = false;
    sim.use_mises_q = mises;
    sim.use_mibf = mibf;
    sim.M = mises ? std::sqrt(3.0) * friction : friction;
    sim.q_cohesion = 0; sim.perzyna_exp = 1; sim.perzyna_visc = 0;
    sim.simulate();
    return sim.exit ? NAN : max_x(sim);
}
MATTER_TEST(BoundaryTest, MIBF) {
    const T a = mibf_run(false, false), b = mibf_run(true, false), c = mibf_run(false, true);
    const char* tol_env = std::getenv("MATTER_MIBF_TOL");
    const T tol = tol_env ? std::atof(tol_env) : 1e-13;
    return check_near("|max x (plate mu) - max x (MIBF)|", std::fabs(a - b), 0.0, tol) &
           check_near("|max x (plate mu) - max x (mises q)|", std::fabs(a - c), 0.0, tol);
}

// ElasticityTest.BulkModulus (tests/tests.cpp:476-541): plates compress a slab; the bulk modulus measured
// from the volume-averaged pressure and log J matches lambda + 2 mu / dim within 3 %.
MATTER_TEST(ElasticityTest, BulkModulus) {
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 100;
    sim.fps = 1;
    sim.cfl = 0.6;
    sim.flip_ratio = -0.95;
    sim.gravity = TV::Zero();
    sim.E = 1e6; sim.nu = 0.3; sim.rho = 10000;
    sim.Lx = 1; sim.Ly = 1; sim.Lz = 0.2;
    sampleParticles(sim, 0.02, 8);
    const T vel = 0.001, vmin_factor = 10, load_factor = 1;
    sim.plates.push_back(std::make_unique<ObjectPlate>(0 - 0.5 * sim.dx, PlateType::bottom, BC::NoSlip, 0, -1e15, 1e15, 0, vel, 0, vmin_factor, load_factor));
    sim.plates.push_back(std::make_unique<ObjectPlate>(sim.Ly + 0.5 * sim.dx, PlateType::top, BC::NoSlip, 0, -1e15, 1e15, 0, -vel, 0, vmin_factor, load_factor));
    sim.simulate();
    TM cauchy = TM::Zero(), kirchh = TM::Zero();
    T Javg;
    sim.computeAvgData(cauchy, kirchh, Javg);
    const T p = -1.0 * (kirchh(0, 0) + kirchh(1, 1) + kirchh(2, 2)) / 3;
    const T K = p / (-std::log(Javg));
    const T Kt = sim.calculateBulkModulus();
    T jmin = 1e300, jmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (unsigned q = 0; q < sim.Np; ++q) {
        const T J = sim.particles.F[q].determinant();
        jmin = std::min(jmin, J); jmax = std::max(jmax, J);
        ymin = std::min(ymin, sim.particles.x[q](1)); ymax = std::max(ymax, sim.particles.x[q](1));
    }
    std::printf("  particles %u, dx %.5f, J avg %.6f min %.6f max %.6f, y in [%.5f, %.5f], p %.6e, K %.6e, K true %.6e\n",
                sim.Np, (double)sim.dx, (double)Javg, (double)jmin, (double)jmax, (double)ymin, (double)ymax, (double)p, (double)K, (double)Kt);
    return sim.exit == 0 && check_near("relative error of K", std::fabs(K - Kt) / Kt, 0.0, 0.03);
}

// EnergyTest.Rotation (tests/tests.cpp:543-603): rigid rotation with pure AFLIP loses at most 1e-3 of its
// kinetic energy and never gains.
MATTER_TEST(EnergyTest, Rotation) {
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 20;
    sim.fps = 1;
    sim.gravity = TV::Zero();
    sim.cfl = 0.5;
    sim.flip_ratio = -1;
    sim.E = 1e6; sim.nu = 0.3; sim.rho = 1550;
    sim.Lx = 1; sim.Ly = 1; sim.Lz = 0.05;
    sampleParticles(sim, 0.01);
    T e0 = 0;
    for (unsigned p = 0; p < sim.Np; ++p) {
        sim.particles.x[p](0) -= 0.5 * sim.Lx;
        sim.particles.x[p](1) -= 0.5 * sim.Ly;
        const T vx = -1.0 * sim.particles.x[p](1) + 0.5, vy = 1.0 * sim.particles.x[p](0) + 0.5;
        sim.particles.v[p](0) = vx; sim.particles.v[p](1) = vy;
        e0 += 0.5 * (vx * vx + vy * vy);
    }
    sim.simulate();
    T e1 = 0;
    for (unsigned p = 0; p < sim.Np; ++p) e1 += 0.5 * sim.particles.v[p].squaredNorm();
    const T rel = (e0 - e1) / e0;
    std::printf("  relative energy loss %.6e\n", (double)rel);
    return sim.exit == 0 && rel >= 0 && rel <= 1e-3;
}

// CollapseTest.DruckerPragerOne / Two (tests/tests.cpp:605-736): a 0.2 x 0.15 x 0.1 column collapses under
// gravity tilted 10 deg against SlipFree back/front/left walls on a NoSlip floor; runout max x = 0.56 +- 0.011.
static bool collapse(PlasticModel model) {
    Simulation sim;
    sim.initialize(false);
    sim.reduce_verbose = true;
    sim.end_frame = 70;
    sim.fps = 50;
    sim.cfl = 0.5;
    sim.flip_ratio = -0.95;
    const T theta = 10 * M_PI / 180;
    sim.gravity = TV::Zero();
    sim.gravity[0] = +9.81 * std::sin(theta);
    sim.gravity[1] = -9.81 * std::cos(theta);
    sim.E = 1e6; sim.nu = 0.3; sim.rho = 1000;
    sim.Lx = 0.20; sim.Ly = 0.15; sim.Lz = 0.10;
    const char* kr = std::getenv("MATTER_COLLAPSE_KRAD");
    sampleParticles(sim, kr ? std::atof(kr) : 0.0015, 8);
    for (unsigned p = 0; p < sim.Np; ++p) sim.particles.x[p](1) += 0.5 * sim.dx;
    sim.grid_reference_point = TV::Zero();
    sim.plastic_model = model;
    sim.use_mises_q = false;
    sim.use_pradhana = true;
    sim.xi = 0;
    sim.perzyna_visc = 0; sim.perzyna_exp = 1;
    sim.q_cohesion = 0;
    sim.M = std::tan(30.0 * M_PI / 180.0);
    sim.plates.push_back(std::make_unique<ObjectPlate>(0, PlateType::back, BC::SlipFree));
    sim.plates.push_back(std::make_unique<ObjectPlate>(sim.Lz, PlateType::front, BC::SlipFree));
    sim.plates.push_back(std::make_unique<ObjectPlate>(0, PlateType::left, BC::SlipFree));
    sim.plates.push_back(std::make_unique<ObjectPlate>(0, PlateType::bottom, BC::NoSlip));
    sim.simulate();
    std::printf("  particles %u, dx %.5f\n", sim.Np, (double)sim.dx);
    return sim.exit == 0 && check_near("max x", max_x(sim), 0.56, 0.011);
}
MATTER_TEST(CollapseTest, DruckerPragerOne) { return collapse(PlasticModel::DPSoft); }
MATTER_TEST(CollapseTest, DruckerPragerTwo) { return collapse(PlasticModel::DPVisc); }

int main(int argc, char** argv) {
    int failed = 0, ran = 0;
    for (auto& kv : registry()) {
        bool want = argc < 2;
        for (int i = 1; i < argc; ++i) want = want || kv.first == argv[i];
        if (!want) continue;
        std::printf("[ RUN  ] %s\n", kv.first.c_str());
        std::fflush(stdout);
        const bool ok = kv.second();
        std::printf("[ %s ] %s\n", ok ? " OK " : "FAIL", kv.first.c_str());
        std::fflush(stdout);
        ++ran;
        failed += ok ? 0 : 1;
    }
    if (ran == 0) { std::printf("no test matched\n"); return 2; }
    return failed ? 1 : 0;
}
