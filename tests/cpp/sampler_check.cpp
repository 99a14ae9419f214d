// Runs matter.hpp's sampleParticles (the host sampler that stands in for tph_poisson,
// src/sampling/sampling_particles.hpp:12-123) and prints its results for tests/test_sampler_cpu.py:
//   line 1: N dx particle_volume particle_mass
//   then N lines of coordinates.
// Arguments: kRadius Lx Ly Lz ppc rho seed. Built twice, with and without MATTER_TWODIM.
#include <cstdio>
#include <cstdlib>
#include "../../matter-tests_amd/host/matter.hpp"

struct Box {
    matter::T Lx = 1, Ly = 1, Lz = 1, rho = 1000, dx = 0, particle_volume = 0, particle_mass = 0;
    unsigned Np = 0;
    matter::Particles particles;
};

int main(int argc, char** argv) {
    if (argc != 8) { std::fprintf(stderr, "usage: sampler_check r Lx Ly Lz ppc rho seed\n"); return 2; }
    Box b;
    const double r = std::atof(argv[1]);
    b.Lx = std::atof(argv[2]); b.Ly = std::atof(argv[3]); b.Lz = std::atof(argv[4]);
    const double ppc = std::atof(argv[5]);
    b.rho = std::atof(argv[6]);
    matter::sampleParticles(b, r, ppc, 0, matter::kDim == 3 ? 30 : 200, (std::uint32_t)std::atoi(argv[7]));
    std::printf("%u %.17g %.17g %.17g\n", b.Np, b.dx, b.particle_volume, b.particle_mass);
    for (unsigned p = 0; p < b.Np; ++p) {
        for (int d = 0; d < matter::kDim; ++d) std::printf(d ? " %.17g" : "%.17g", b.particles.x[p](d));
        std::printf("\n");
    }
    return 0;
}
