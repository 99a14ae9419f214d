// Host build of the engine's 3x3 / 2x2 SVD (matter-tests_amd/csrc/mpm_math.hpp compiles for the host as
// well as for gfx950) at the tolerances of both forms: reads matrices (dim, then dim*dim column-major values
// per line) on stdin and writes, per matrix and tolerance, the singular values and U, W (column-major).
#include <cstdio>
#include "../../matter-tests_amd/csrc/mpm_types.hpp"
#include "../../matter-tests_amd/csrc/mpm_math.hpp"
using namespace mpm;

template <int D> void run(const double* a) {
    M<D> F; for (int i = 0; i < D * D; ++i) F.m[i] = a[i];
    M<D> U, W; V<D> s;
    svd<D>(F, U, s, W, kSvdTolEigen);
    for (int i = 0; i < D; ++i) std::printf("%.17g ", s[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", U.m[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", W.m[i]);
    svd<D>(F, U, s, W, 1e-12);  // the split form's tolerance
    for (int i = 0; i < D; ++i) std::printf("%.17g ", s[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", U.m[i]);
    for (int i = 0; i < D * D; ++i) std::printf("%.17g ", W.m[i]);
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
