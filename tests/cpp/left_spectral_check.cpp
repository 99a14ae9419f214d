// Host build of the engine's left spectral decomposition (matter-tests_amd/csrc/mpm_math.hpp:
// left_spectral, recompose_left, hencky_tau), the path the deformation update takes instead of the
// two-sided SVD. Reads lines "dim, dim*dim column-major values of F" on stdin and writes per line: accepted
// (0/1), lam, U (column-major), the recomposed F' = U diag(exp(h)/sqrt lam) U^T F with h = ln sqrt(lam) + shift,
// shift = -0.3 times the deviatoric part of ln sqrt(lam) (the shape of a Drucker-Prager projection; it does not
// depend on the order of the eigenpairs), and the Hencky stress of h with mu = 1, lambda = 1.5 (packed upper
// triangle).
#include <cstdio>
#include "../../matter-tests_amd/csrc/mpm_types.hpp"
#include "../../matter-tests_amd/csrc/mpm_math.hpp"
using namespace mpm;

template <int D> void run(const double* a) {
    M<D> F; for (int i = 0; i < D * D; ++i) F.m[i] = a[i];
    M<D> U; V<D> lam;
    const bool ok = left_spectral<D>(F, U, lam, kSvdTolEigen);
    V<D> h, e;
    double mean = 0.0;
    for (int k = 0; k < D; ++k) { h[k] = 0.5 * log_d(lam[k]); mean += h[k] / D; }
    for (int k = 0; k < D; ++k) {
        h[k] = h[k] - 0.3 * (h[k] - mean);
        e[k] = exp(h[k]) * rsqrt_d(lam[k]);
    }
    const M<D> Fn = recompose_left<D>(U, e, F);
    double tau[6];
    hencky_tau<D>(U, h, 1.0, 1.5, tau);
    std::printf("%d ", ok ? 1 : 0);
    for (int i = 0; i < D; ++i) std::printf("%.17g ", lam[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", U.m[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", Fn.m[i]);
    for (int i = 0; i < SymPack<D>::n; ++i) std::printf("%.17g ", tau[i]);
    std::printf("\n");
}

int main() {
    int d;
    double a[9];
    while (std::scanf("%d", &d) == 1) {
        for (int i = 0; i < d * d; ++i) if (std::scanf("%lf", &a[i]) != 1) return 1;
        if (d == 3) run<3>(a); else run<2>(a);
    }
    return 0;
}
