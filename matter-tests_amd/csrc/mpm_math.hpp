// mpm_math.hpp — per-particle and per-node arithmetic of the MPM step for gfx950 kernels.
//
// Everything here is FP64 and branch-for-branch equivalent to the reference's formulas (cited per
// function); the organisation is the GPU's: the return map works on principal Hencky strains only and
// reports whether F must be recomposed, errors become bits instead of console messages, and boundary
// objects are a closed tagged union instead of virtual classes. MPM_HD functions also compile for the
// host so they can be exercised without a GPU.
#pragma once
#include <stdint.h>
#include <math.h>
#include "mpm_types.hpp"
#include "../../include/mpm.h"

#if defined(__HIPCC__)
#define MPM_HD __host__ __device__ inline
#else
#define MPM_HD inline
#endif

namespace mpm {

// ------------------------------------------------------------------------------------------------
// B-spline kernels (src/tools.hpp:110-175): N(u) and dN/du for the quadratic and cubic B-spline.
// ------------------------------------------------------------------------------------------------
MPM_HD int sgn_d(double v) { return (0.0 < v) - (v < 0.0); }

template <int SPL> struct BSpline;
// The piecewise formulas are evaluated unconditionally and selected (the selects are marked unpredictable
// so that the compiler keeps them as selects rather than divergent branches); each piece is the same
// expression, with the same operations in the same order, as in the reference.
#if defined(__HIP_DEVICE_COMPILE__)
#define MPM_SEL(c) __builtin_unpredictable(c)
#else
#define MPM_SEL(c) (c)
#endif

// Natural logarithm for positive normal x (other inputs go to the library log): x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1), ln m = 2 atanh(f) = 2 sum_k f^(2k+1)/(2k+1) (|f| <= 0.1716, ten
// terms leave a remainder below 1e-18 relative), ln x = e ln 2 + ln m with ln 2 split in two parts so that
// e ln 2 carries no rounding error of its own. Accurate to about 1 ulp, in a third of the library's
// instructions.
MPM_HD double log_d(double x) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log(x);
    int e;
    double m = frexp(x, &e);                       // m in [0.5, 1)
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = (m - 1.0) / (m + 1.0);
    const double s = f * f;
    double q = 1.0 / 21.0;
    q = fma(q, s, 1.0 / 19.0); q = fma(q, s, 1.0 / 17.0); q = fma(q, s, 1.0 / 15.0); q = fma(q, s, 1.0 / 13.0);
    q = fma(q, s, 1.0 / 11.0); q = fma(q, s, 1.0 / 9.0); q = fma(q, s, 1.0 / 7.0); q = fma(q, s, 1.0 / 5.0);
    q = fma(q, s, 1.0 / 3.0);
    const double f2 = f + f;
    const double lm = fma(f2 * s, q, f2);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

// Multiplier and shift for the quotient n / d of n < 2^31 by a block count d >= 1 (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994): with l = ceil(log2 d) and m = ceil(2^(31+l) / d),
// m d - 2^(31+l) < d <= 2^l, so floor(n / d) = floor(n m / 2^(31+l)) for every n < 2^31; m < 2^32 because
// d > 2^(l-1). The quotient is taken as mulhi(2n, m) >> l (block_div). m starts from the binary64 quotient (within
// one of the exact one) and is corrected with integer products.
MPM_HD void block_div_magic(uint32_t d, uint32_t& m, uint32_t& l) {
    l = 0;
    while (l < 31 && (1ull << l) < (unsigned long long)d) ++l;
    const unsigned long long p = 1ull << (31 + l);
    unsigned long long q = (unsigned long long)((double)p / (double)d);
    while (q * d < p) ++q;
    while ((q - 1) * d >= p) --q;
    m = (uint32_t)q;
}
MPM_HD uint32_t block_div(uint32_t n, uint32_t m, uint32_t l) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(n + n, m) >> l;
#else
    return (uint32_t)(((unsigned long long)(n + n) * m) >> 32) >> l;
#endif
}

// 1/x for the split form's arithmetic: the hardware reciprocal estimate refined by two Newton steps (about 1 ulp,
// against the correctly rounded IEEE division's 10 instructions; each step at least doubles the correct bits of
// an estimate that starts near single precision). For positive normal x. The plain form keeps IEEE division, so
// that its results round as the reference's.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double rcp_d(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
#else
MPM_HD double rcp_d(double x) { return 1.0 / x; }
#endif

// log_d with the reciprocal of rcp_d in place of the division (m - 1)/(m + 1): for the split form.
MPM_HD double log_d_split(double x) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log(x);
    int e;
    double m = frexp(x, &e);
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = (m - 1.0) * rcp_d(m + 1.0);
    const double s = f * f;
    double q = 1.0 / 21.0;
    q = fma(q, s, 1.0 / 19.0); q = fma(q, s, 1.0 / 17.0); q = fma(q, s, 1.0 / 15.0); q = fma(q, s, 1.0 / 13.0);
    q = fma(q, s, 1.0 / 11.0); q = fma(q, s, 1.0 / 9.0); q = fma(q, s, 1.0 / 7.0); q = fma(q, s, 1.0 / 5.0);
    q = fma(q, s, 1.0 / 3.0);
    const double f2 = f + f;
    const double lm = fma(f2 * s, q, f2);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

// 1/sqrt(x) for the split form's Jacobi rotations (positive normal x): the hardware estimate (measured on gfx950:
// within 2^-24 relative) and one Newton step, within 2^-47 (4.2e-15, measured over 2^20 arguments on [1e-6, 1e6])
// against the library's 1 ulp in two steps and special-case handling.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double rsqrt_d_split(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return y * fma(-0.5 * x * y, y, 1.5);
}
#else
MPM_HD double rsqrt_d_split(double x) { return 1.0 / sqrt(x); }
#endif

// 1/sqrt(x): the device library's reciprocal square root (a hardware estimate refined to about 1 ulp).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double rsqrt_d(double x) { return rsqrt(x); }
#else
MPM_HD double rsqrt_d(double x) { return 1.0 / sqrt(x); }
#endif
// BSpline<SPL>::node<K>(u, w, dw): weight and derivative of node K of a particle's non-zero window
// (K = 0 .. SPL), u = (x_p - x_i)/dx. Within the window the node fixes the polynomial piece: u lies in
// [0.5,1.5], [-0.5,0.5], [-1.5,-0.5] for K = 0,1,2 of the quadratic spline and in [1,2], [0,1], [-1,0],
// [-2,-1] for K = 0..3 of the cubic one, so the piece is chosen at compile time. The pieces meet with equal
// value and slope, so a u that rounds across a boundary changes w and dw by a few ulp at most.
template <> struct BSpline<2> {
    static constexpr double apic_factor = 4.0;  // apicDinverse = 4/dx^2 (simulation.cpp:98)
    template <int K> MPM_HD static void node(double u, double& w, double& dw) {
        if (K == 1) { w = 0.75 - u * u; dw = -2 * u; }
        else if (K == 0) { const double r = 1.5 - u; w = 0.5 * r * r; dw = u - 1.5; }
        else { const double r = 1.5 + u; w = 0.5 * r * r; dw = u + 1.5; }
    }
    MPM_HD static double N(double u) {
        const double a = fabs(u);
        const double inner = 0.75 - a * a;
        const double outer = 0.5 * (1.5 - a) * (1.5 - a);
        const double r = MPM_SEL(a < 1.5) ? outer : 0.0;
        return MPM_SEL(a < 0.5) ? inner : r;
    }
    MPM_HD static double dN(double u) {
        const double a = fabs(u);
        const double inner = -2 * u;
        const double sg = MPM_SEL(u > 0.0) ? 1.0 : (MPM_SEL(u < 0.0) ? -1.0 : 0.0);
        const double outer = u - 1.5 * sg;
        const double r = MPM_SEL(a < 1.5) ? outer : 0.0;
        return MPM_SEL(a < 0.5) ? inner : r;
    }
};
template <> struct BSpline<3> {
    static constexpr double apic_factor = 3.0;  // apicDinverse = 3/dx^2 (simulation.cpp:95)
    template <int K> MPM_HD static void node(double u, double& w, double& dw) {
        constexpr double c23 = 0.6666666666666666666666666666666666666666666666666;
        constexpr double c16 = 0.1666666666666666666666666666666666666666666;
        if (K == 0) { const double r = 2.0 - u; w = c16 * r * r * r; dw = -0.5 * r * r; }
        else if (K == 1) { w = 0.5 * u * u * u - u * u + c23; dw = u * (1.5 * u - 2.0); }
        else if (K == 2) { const double a = -u; w = 0.5 * a * a * a - a * a + c23; dw = u * (1.5 * a - 2.0); }
        else { const double r = 2.0 + u; w = c16 * r * r * r; dw = 0.5 * r * r; }
    }
    MPM_HD static double N(double u) {
        const double a = fabs(u);
        const double inner = 0.5 * a * a * a - a * a + 0.6666666666666666666666666666666666666666666666666;
        const double outer = 0.1666666666666666666666666666666666666666666 * (2.0 - a) * (2.0 - a) * (2.0 - a);
        const double r = MPM_SEL(a < 2.0) ? outer : 0.0;
        return MPM_SEL(a < 1.0) ? inner : r;
    }
    MPM_HD static double dN(double u) {
        const double a = fabs(u);
        const double inner = u * (1.5 * a - 2.0);
        const double sg = MPM_SEL(u > 0.0) ? 1.0 : (MPM_SEL(u < 0.0) ? -1.0 : 0.0);
        const double outer = -0.5 * sg * (2.0 - a) * (2.0 - a);
        const double r = MPM_SEL(a < 2.0) ? outer : 0.0;
        return MPM_SEL(a < 1.0) ? inner : r;
    }
};

// ------------------------------------------------------------------------------------------------
// Small dense matrices, column-major m[r + c*D].
// ------------------------------------------------------------------------------------------------
template <int D> struct M { double m[D * D]; MPM_HD double& operator()(int r, int c) { return m[r + c * D]; } MPM_HD double operator()(int r, int c) const { return m[r + c * D]; } };
template <int D> struct V { double v[D]; MPM_HD double& operator[](int i) { return v[i]; } MPM_HD double operator[](int i) const { return v[i]; } };

template <int D> MPM_HD void set_identity(M<D>& a) { for (int i = 0; i < D * D; ++i) a.m[i] = 0.0; for (int i = 0; i < D; ++i) a(i, i) = 1.0; }
// (A*B)(r,c) = sum_k A(r,k) B(k,c), summed in k order.
template <int D> MPM_HD M<D> mul(const M<D>& A, const M<D>& B) {
    #pragma clang fp contract(on)  // a product and the sum it feeds in one statement fuse into an FMA, decided per statement, so every kernel that inlines this rounds alike
    M<D> R;
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) { double s = A(r, 0) * B(0, c); for (int k = 1; k < D; ++k) s += A(r, k) * B(k, c); R(r, c) = s; }
    return R;
}
// U * diag(d) * V^T evaluated as (U*diag(d)) * V^T.
template <int D> MPM_HD M<D> udvt(const M<D>& U, const V<D>& d, const M<D>& W) {
    #pragma clang fp contract(on)
    M<D> UD; for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) UD(r, c) = U(r, c) * d[c];
    M<D> R;
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) { double s = UD(r, 0) * W(c, 0); for (int k = 1; k < D; ++k) s += UD(r, k) * W(c, k); R(r, c) = s; }
    return R;
}
MPM_HD double det(const M<2>& a) { return a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1); }
MPM_HD double det(const M<3>& a) {
    return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(1, 0) * (a(0, 1) * a(2, 2) - a(0, 2) * a(2, 1))
         + a(2, 0) * (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1));
}

// ------------------------------------------------------------------------------------------------
// Two-sided Jacobi SVD, A = U diag(s) W^T with s >= 0 sorted descending (the convention of the
// reference's Eigen::JacobiSVD, simulation.hpp:232 and plasticity.cpp:19). Each sweep visits the
// pairs (1,0),(2,0),(2,1); a pair is first made symmetric by a plane rotation on the left, then
// diagonalised by a symmetric Jacobi rotation applied on both sides. Iterates until every off-diagonal
// entry is below 2*eps times the largest diagonal entry (at most 24 sweeps).
// ------------------------------------------------------------------------------------------------
struct PlaneRot { double c, s; };

MPM_HD PlaneRot sym_jacobi(double x, double y, double z) {  // annihilates y in [[x y][y z]]
    PlaneRot r{1.0, 0.0};
    double ay = fabs(y);
    if (2.0 * ay < 2.2250738585072014e-308) return r;
    double tau = (x - z) / (2.0 * ay);
    double w = sqrt(tau * tau + 1.0);
    double t = 1.0 / (tau > 0 ? tau + w : tau - w);
    double n = 1.0 / sqrt(t * t + 1.0);
    double st = t > 0 ? 1.0 : -1.0;
    r.s = -st * copysign(1.0, y) * fabs(t) * n;  // y / |y| is exactly the sign of y
    r.c = n;
    return r;
}

// One two-sided Jacobi step on the pair (p, q); compile-time indices keep X, U, W in registers.
template <int D, int p, int q>
MPM_HD void jacobi_pair(M<D>& X, M<D>& U, M<D>& W, double& dmax, bool& rotated, double tol) {
    #pragma clang fp contract(on)
    const double thr = fmax(2.2250738585072014e-308, tol * dmax);
    if (!(fabs(X(p, q)) > thr || fabs(X(q, p)) > thr)) return;
    rotated = true;
    // symmetrising rotation r1 acting on rows (p,q)
    const double a = X(p, p), b = X(p, q), c = X(q, p), d = X(q, q);
    const double tsum = a + d, diff = c - b;
    PlaneRot r1{1.0, 0.0}, jr;
    if (fmax(fabs(tsum), fabs(diff)) > 1e-140) {
        // The same rotations in closed form, without divisions: with u = tsum/diff, 1/sqrt(1+u^2) =
        // |diff|/hypot(tsum,diff) and u/sqrt(1+u^2) = sgn(diff) tsum/hypot; for the symmetric step with
        // alpha = x - z, beta = 2|y|, r = hypot(alpha, beta): c = (|alpha|+r)/hypot(|alpha|+r, beta) and
        // |s| = beta/hypot(|alpha|+r, beta). The squares cannot underflow to zero while the larger operand
        // exceeds 1e-140.
        if (fabs(diff) >= 2.2250738585072014e-308) {
            const double ih = rsqrt_d(fma(tsum, tsum, diff * diff));
            r1.s = fabs(diff) * ih;
            r1.c = (diff < 0 ? -tsum : tsum) * ih;
        }
        const double s00 = r1.c * a + r1.s * c, s01 = r1.c * b + r1.s * d, s11 = -r1.s * b + r1.c * d;
        const double al = s00 - s11, be = 2.0 * fabs(s01);
        jr = PlaneRot{1.0, 0.0};
        if (be >= 2.2250738585072014e-308 && fmax(fabs(al), be) > 1e-140) {
            const double r2 = fma(al, al, be * be);
            const double g = fabs(al) + r2 * rsqrt_d(r2);
            const double ig = rsqrt_d(fma(g, g, be * be));
            jr.c = g * ig;
            const double sy = copysign(be * ig, s01);  // sgn(y) |t| n
            jr.s = al > 0 ? -sy : sy;                  // times -sgn(t), where t > 0 exactly when alpha > 0
        } else {
            jr = sym_jacobi(s00, s01, s11);
        }
    } else {
        if (fabs(diff) >= 2.2250738585072014e-308) { const double u = tsum / diff; const double hi = 1.0 / sqrt(1.0 + u * u); r1.s = hi; r1.c = u * hi; }
        const double s00 = r1.c * a + r1.s * c, s01 = r1.c * b + r1.s * d, s11 = -r1.s * b + r1.c * d;
        jr = sym_jacobi(s00, s01, s11);
    }
    const PlaneRot jl{r1.c * jr.c + r1.s * jr.s, -r1.c * jr.s + r1.s * jr.c};  // jl = r1 * jr^T
#pragma unroll
    for (int k = 0; k < D; ++k) {  // rows p,q of X <- jl applied on the left; U <- U * jl^T
        const double xp = X(p, k), xq = X(q, k);
        X(p, k) = jl.c * xp + jl.s * xq; X(q, k) = -jl.s * xp + jl.c * xq;
        const double up = U(k, p), uq = U(k, q);
        U(k, p) = jl.c * up + jl.s * uq; U(k, q) = -jl.s * up + jl.c * uq;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {  // X <- X * jr ; W <- W * jr  (columns p,q)
        const double xp = X(k, p), xq = X(k, q);
        X(k, p) = jr.c * xp - jr.s * xq; X(k, q) = jr.s * xp + jr.c * xq;
        const double wp = W(k, p), wq = W(k, q);
        W(k, p) = jr.c * wp - jr.s * wq; W(k, q) = jr.s * wp + jr.c * wq;
    }
    dmax = fmax(dmax, fmax(fabs(X(p, p)), fabs(X(q, q))));
}

// Orders singular values i < j descending, swapping the columns of U and W along (branch-free selects).
template <int D, int i, int j>
MPM_HD void sv_order(V<D>& s, M<D>& U, M<D>& W) {
    const bool sw = s[j] > s[i];
    const double a = s[i], b = s[j];
    s[i] = sw ? b : a; s[j] = sw ? a : b;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double ui = U(k, i), uj = U(k, j), wi = W(k, i), wj = W(k, j);
        U(k, i) = sw ? uj : ui; U(k, j) = sw ? ui : uj;
        W(k, i) = sw ? wj : wi; W(k, j) = sw ? wi : wj;
    }
}

// Eigen's JacobiSVD threshold: twice the binary64 epsilon.
constexpr double kSvdTolEigen = 2.0 * 2.220446049250313e-16;

template <int D> MPM_HD void svd(const M<D>& A, M<D>& U, V<D>& s, M<D>& W, double tol = kSvdTolEigen) {
    double scale = 0.0;
#pragma unroll
    for (int i = 0; i < D * D; ++i) scale = fmax(scale, fabs(A.m[i]));
    if (scale == 0.0) scale = 1.0;
    const double inv = 1.0 / scale;
    M<D> X;
#pragma unroll
    for (int i = 0; i < D * D; ++i) X.m[i] = A.m[i] * inv;
    set_identity(U); set_identity(W);
    double dmax = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) dmax = fmax(dmax, fabs(X(i, i)));
    for (int sweep = 0; sweep < 24; ++sweep) {  // sweeps visit the pairs (1,0), (2,0), (2,1)
        bool rotated = false;
        jacobi_pair<D, 1, 0>(X, U, W, dmax, rotated, tol);
        if constexpr (D == 3) {
            jacobi_pair<D, 2, 0>(X, U, W, dmax, rotated, tol);
            jacobi_pair<D, 2, 1>(X, U, W, dmax, rotated, tol);
        }
        if (!rotated) break;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const double a = X(i, i);
        s[i] = fabs(a) * scale;
        const double sg = a < 0 ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < D; ++k) U(k, i) = sg * U(k, i);
    }
    sv_order<D, 0, 1>(s, U, W);  // descending order: compare-swap network
    if constexpr (D == 3) {
        sv_order<D, 1, 2>(s, U, W);
        sv_order<D, 0, 1>(s, U, W);
    }
}

// ------------------------------------------------------------------------------------------------
// Left spectral decomposition: b = F F^T = U diag(lam) U^T by cyclic Jacobi rotations of the symmetric b,
// lam descending. U holds the left singular vectors of F and lam its squared singular values, so
// F = U diag(sqrt lam) W^T with W = F^T U diag(1/sqrt lam). The Hencky stress and the return map need only U
// and lam, and the plastic update F' = U diag(exp h) W^T = U diag(exp(h)/sqrt lam) U^T F, so this replaces
// the two-sided SVD with one rotation per pair instead of two and no W. Squaring F squares its condition
// number: the smallest lam carries an absolute error of about eps * lam_max, so the decomposition is
// accepted (returns true) only when lam_min >= 1e-4 lam_max (singular values within a factor 100, their
// logarithms then within about 1e-12 of the two-sided SVD's), lam_min >= 1e-8 (no singular value under the
// return map's 1e-4 clamp), every lam is finite and the sweeps converged; otherwise the caller takes svd().
// ------------------------------------------------------------------------------------------------
template <int D, int p, int q>
MPM_HD void sym_jacobi_rot(M<D>& A, M<D>& U, bool& rotated, double thr) {
    // every fused operation is written out, so that all kernels that inline this round alike
    const double apq = A(p, q);
    if (!(fabs(apq) > thr)) return;
    rotated = true;
    // theta = (a_qq - a_pp) / (2 a_pq), t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)) (the smaller rotation),
    // c = 1 / sqrt(1 + t^2), s = t c; with al = a_qq - a_pp, be = 2 a_pq and g = |al| + hypot(al, be):
    // t = sgn(al) be / g, c = g / hypot(g, be), s = sgn(al) be / hypot(g, be)
    const double al = A(q, q) - A(p, p), be = 2.0 * apq;
    const double r2 = fma(al, al, be * be);
    const double r = r2 * rsqrt_d_split(r2);
    const double g = fabs(al) + r;
    const double ig = rsqrt_d_split(fma(g, g, be * be));
    const double sbe = al < 0 ? -be : be;
    const double c = g * ig, s = sbe * ig;
    // t a_pq = sbe a_pq / g, and g^2 + be^2 = 2 r g (r^2 = al^2 + be^2), so 1/g = 2 r ig^2: no reciprocal (g > 0;
    // the split form only: the plain form takes svd())
    const double tapq = sbe * apq * ((r + r) * (ig * ig));
    A(p, p) = A(p, p) - tapq;
    A(q, q) = A(q, q) + tapq;
    A(p, q) = 0.0; A(q, p) = 0.0;
    if constexpr (D == 3) {
        constexpr int r = 3 - p - q;
        const double arp = A(r, p), arq = A(r, q);
        const double np = fma(c, arp, -(s * arq)), nq = fma(s, arp, c * arq);
        A(r, p) = np; A(p, r) = np; A(r, q) = nq; A(q, r) = nq;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double up = U(k, p), uq = U(k, q);
        U(k, p) = fma(c, up, -(s * uq)); U(k, q) = fma(s, up, c * uq);
    }
}

template <int D> MPM_HD bool left_spectral(const M<D>& F, M<D>& U, V<D>& lam, double tol = kSvdTolEigen) {
    M<D> A;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = r; c < D; ++c) {
            double v = F(r, 0) * F(c, 0);
            for (int k = 1; k < D; ++k) v = fma(F(r, k), F(c, k), v);
            A(r, c) = v; A(c, r) = v;
        }
    set_identity(U);
    bool converged = false;
    for (int sweep = 0; sweep < 24; ++sweep) {
        double dmax = fabs(A(0, 0));
#pragma unroll
        for (int k = 1; k < D; ++k) dmax = fmax(dmax, fabs(A(k, k)));
        const double thr = fmax(2.2250738585072014e-308, tol * dmax);
        bool rotated = false;
        sym_jacobi_rot<D, 0, 1>(A, U, rotated, thr);
        if constexpr (D == 3) {
            sym_jacobi_rot<D, 0, 2>(A, U, rotated, thr);
            sym_jacobi_rot<D, 1, 2>(A, U, rotated, thr);
        }
        if (!rotated) { converged = true; break; }
    }
    // The eigenpairs are left in the order the rotations leave them: everything the caller forms from them
    // (the trace of the Hencky strain, the return maps' deviatoric parts, U diag(.) U^T) is symmetric in it.
    double lmin = A(0, 0), lmax = A(0, 0);
#pragma unroll
    for (int k = 0; k < D; ++k) { lam[k] = A(k, k); lmin = fmin(lmin, lam[k]); lmax = fmax(lmax, lam[k]); }
    return converged && lmax < 1.0e300 && lmin >= 1.0e-8 && lmin >= 1.0e-4 * lmax;
}

// F' = U diag(e) U^T F: the recomposition U diag(exp h) W^T of the plastic update with e = exp(h) / sqrt(lam).
template <int D> MPM_HD M<D> recompose_left(const M<D>& U, const V<D>& e, const M<D>& F) {
    M<D> Y;  // diag(e) U^T F
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double v = U(0, k) * F(0, c);
            for (int r = 1; r < D; ++r) v = fma(U(r, k), F(r, c), v);
            Y(k, c) = e[k] * v;
        }
    M<D> R;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double v = U(r, 0) * Y(0, c);
            for (int k = 1; k < D; ++k) v = fma(U(r, k), Y(k, c), v);
            R(r, c) = v;
        }
    return R;
}

// Symmetric Kirchhoff stress tau = P(F) F^T stored as packed upper triangle:
// 3D (xx, yy, zz, xy, xz, yz); 2D (xx, yy, xy).
template <int D> struct SymPack { static constexpr int n = D == 3 ? 6 : 3; };

// Hencky: tau = U diag(2 mu ln s + lambda tr(ln s)) U^T, the product HenckyPiola(F) * F^T of
// simulation.hpp:231-238 (P = U diag((2 mu ln s + lambda tr ln s)/s) W^T, F^T = W diag(s) U^T).
template <int D> MPM_HD void hencky_tau(const M<D>& U, const V<D>& lnS, double mu, double lambda, double* tau) {
    #pragma clang fp contract(on)
    double tr = lnS[0]; for (int i = 1; i < D; ++i) tr += lnS[i];
    V<D> d; for (int i = 0; i < D; ++i) d[i] = 2.0 * mu * lnS[i] + lambda * tr;
    auto T = [&](int r, int c) { double s = U(r, 0) * d[0] * U(c, 0); for (int k = 1; k < D; ++k) s += U(r, k) * d[k] * U(c, k); return s; };
    if (D == 3) { tau[0] = T(0, 0); tau[1] = T(1, 1); tau[2] = T(2, 2); tau[3] = T(0, 1); tau[4] = T(0, 2); tau[5] = T(1, 2); }
    else { tau[0] = T(0, 0); tau[1] = T(1, 1); tau[2] = T(0, 1); }
}
// Neo-Hookean: tau = mu (F F^T - I) + lambda ln(J) I  (NeoHookeanPiola * F^T, simulation.hpp:227-229).
template <int D> MPM_HD void neohookean_tau(const M<D>& F, double mu, double lambda, double* tau) {
    double lj = lambda * log(det(F));
    auto B = [&](int r, int c) { double s = F(r, 0) * F(c, 0); for (int k = 1; k < D; ++k) s += F(r, k) * F(c, k); return s; };
    if (D == 3) {
        tau[0] = mu * (B(0, 0) - 1.0) + lj; tau[1] = mu * (B(1, 1) - 1.0) + lj; tau[2] = mu * (B(2, 2) - 1.0) + lj;
        tau[3] = mu * B(0, 1); tau[4] = mu * B(0, 2); tau[5] = mu * B(1, 2);
    } else {
        tau[0] = mu * (B(0, 0) - 1.0) + lj; tau[1] = mu * (B(1, 1) - 1.0) + lj; tau[2] = mu * B(0, 1);
    }
}

// ------------------------------------------------------------------------------------------------
// Return mapping on principal Hencky strains (plasticity.cpp:11-665).
// ------------------------------------------------------------------------------------------------
struct PlasticState { double eps_dev, eps_vol, eps_prad, dgamma, visc, muI; };

struct RMResult {
    bool recompose;   // F must be rebuilt as U exp(h) W^T
    bool plastic;     // counts towards plastic_count (may be counted twice, as the reference does)
    int plastic_incr; // number of plastic_count++ executed
    uint32_t err;
};

// MCCRMAExplicit, mcc_rma_explicit.hpp:8-96
MPM_HD bool mcc_explicit(double& p, double& q, double Mm, double p0, double beta, double mu, double K, double rma) {
    double y = Mm * Mm * (p - p0) * (p + beta * p0) + (q * q);
    if (!(y > 0)) return false;
    double dg = 0.0, pt = p, qt = q;
    const double dkdp = 2 * Mm * Mm, dldq = 2;
    const int max_iter = 40;
    for (int iter = 0; iter < max_iter; ++iter) {
        y = Mm * Mm * (p - p0) * (p + beta * p0) + (q * q);
        double k = Mm * Mm * (beta * p0 + 2 * p - p0);
        double l = 2 * q;
        double r1 = pt - p - K * dg * k;
        double r2 = qt - q - rma * mu * dg * l;
        if (iter > 4 && fabs(y) < 1e-3 && fabs(r1) < 1e-3 && fabs(r2) < 1e-3) break;
        if (iter == max_iter - 1) { if (!(p0 > 0.1)) { p = 1e-15; q = 1e-15; } break; }
        double J11 = -1 - K * dg * dkdp, J13 = -K * k;
        double J22 = -1 - rma * mu * dg * dldq, J23 = -rma * mu * l;
        double J31 = k, J32 = l;
        double dt_ = J11 * (-J32 * J23) + J13 * (-J31 * J22);
        if (fabs(dt_) < 1e-6) {
            p -= 0.001 * r1; q -= 0.001 * r2; dg -= 0.001 * y / (p0 * p0);
        } else {
            p -= (-J32 * J23 * r1 + J32 * J13 * r2 - J22 * J13 * y) / dt_;
            q -= (J31 * J23 * r1 - J31 * J13 * r2 - J11 * J23 * y) / dt_;
            dg -= (-J31 * J22 * r1 - J11 * J32 * r2 + J11 * J22 * y) / dt_;
        }
        if (q < 1e-15) q = 1e-15;
    }
    p = fmax(p, -beta * p0);
    p = fmin(p, p0);
    q = Mm * sqrt((p0 - p) * (beta * p0 + p));
    return true;
}

// Consolidation pressure laws of the implicit MCC maps: exponential p00*exp(-xi*epv)
// (mcc_rma_implicit_exponential_onevar.hpp:136,181) and sinh K*sinh(-xi*epv + asinh(p00/K)) floored at
// p0_min (mcc_rma_implicit_sinh_onevar.hpp:37,78-79).
// Admissible scaled Newton step: limited_search_exponential.hpp:8-71 / limited_search_sinh.hpp:8-72.
template <bool SINH>
MPM_HD double mcc_limited_search(double epv, double epv_t, double delta, double xi, double K, double p_t, double p00,
                                 double p0_min, double beta, int& right) {
    double scale = 1.0; int tries = 0; const int max_tries = 32; int f = 0; double test = epv;
    while (tries < max_tries) {
        test = epv - scale * delta;
        double pp0 = SINH ? fmax(p0_min, K * sinh(-xi * test + asinh(p00 / K))) : p00 * exp(-xi * test);
        double pp = K * (test - epv_t) + p_t;
        double val = 2. * pp + (beta - 1.) * pp0;
        double mid = (pp0 - beta * pp0) * 0.5;
        if (fabs(pp - mid) <= 1.e-5) break;
        if ((right == 1 && (pp < mid || p_t < mid)) || (right == 0 && (pp > mid || p_t > mid))) { test = epv; right = -1; break; }
        if (sgn_d(epv_t - test) == sgn_d(val) || sgn_d(epv_t - test) == 0) break;
        scale *= 0.5;
        tries++;
        if (tries != max_tries) continue;
        tries = 0; scale = 1.; delta *= -1.; f++;
        if (f > 1) return INFINITY;
    }
    return test;
}

// MCCRMAImplicitExponentialOnevar (mcc_rma_implicit_exponential_onevar.hpp:9-166) and
// MCCRMAImplicitSinhOnevar (mcc_rma_implicit_sinh_onevar.hpp:9-171): 1-variable Newton on epv with
// values normalised by the trial consolidation pressure.
template <bool SINH>
MPM_HD bool mcc_implicit(double& p, double& q, uint32_t& err, double Mm, double p00, double beta, double mu, double K,
                         double xi, double epv) {
    double p0 = SINH ? fmax(0.01, K * sinh(-xi * epv + asinh(p00 / K))) : p00 * exp(-xi * epv);
    const double Msq = Mm * Mm, damp = 1.0;
    double y = Msq * (p * p + (beta - 1.) * p0 * p - beta * p0 * p0) + damp * q * q;
    if (y <= 0) return false;
    const double p0_t = p0;
    const double sf = 1. / p0_t;
    double p0_min = SINH ? 0.01 * sf : 0.0;
    p00 *= sf; p0 = 1.; p *= sf; q *= sf; mu *= sf; K *= sf;
    const double p_t = p, q_t = q, epv_t = epv;
    int right = (p_t > 0.5 * (p0 - beta * p0)) ? 1 : 0;
    const int max_iter = SINH ? 40 : 30;
    for (int iter = 0; iter < max_iter; ++iter) {
        double mid = (p0 - beta * p0) * 0.5;
        if (fabs(p - mid) < 1e-4) {
            q = Mm * sqrt((p0 - p) * (beta * p0 + p) / damp);
            p *= p0_t; q *= p0_t;
            return true;
        }
        p = (epv - epv_t) * K + p_t;
        double p0h = 0.0;
        if (SINH) { p0h = K * sinh(-xi * epv + asinh(p00 / K)); p0 = fmax(p0_min, p0h); }
        else p0 = p00 * exp(-xi * epv);
        double denom = 2. * K * Msq * p + Msq * (beta - 1.) * K * p0;
        double dg = (p_t - p) / denom;
        double qd = 1 + 2. * mu * damp * dg;
        q = q_t / qd;
        y = Msq * (p - p0) * (p + beta * p0) + damp * q * q;
        if (iter > 5 && fabs(y) < 1e-4) break;
        if (SINH && iter == max_iter - 1 && p0 > p0_min) err |= MPM_ERR_MCC_NO_STEP;
        double dp = K, dp0;
        if (SINH) {
            double hh = -xi * K * cosh(-xi * epv + asinh(p00 / K));
            dp0 = (p0h > p0_min) ? 0.5 * (hh + hh - p0_min) : 0.5 * (hh + p0_min - hh);
        } else dp0 = -xi * p0;
        double ddg = (-dp / denom) - (p_t - p) * (2. * K * Msq * dp + Msq * (beta - 1.) * K * dp0) / (denom * denom);
        double dq = -q_t * 2. * mu * damp * ddg / (qd * qd);
        double dy = Msq * (dp - dp0) * (p + beta * p0) + Msq * (p - p0) * (beta * dp0 + dp) + 2. * damp * q * dq;
        epv = mcc_limited_search<SINH>(epv, epv_t, y / dy, xi, K, p_t, p00, p0_min, beta, right);
        if (epv == INFINITY) { err |= MPM_ERR_MCC_NO_STEP; break; }
    }
    p = fmax(p, -beta * p0);
    p = fmin(p, p0);
    q = Mm * sqrt((p0 - p) * (beta * p0 + p) / damp);
    p *= p0_t; q *= p0_t;
    return true;
}

// Perzyna Newton iteration on delta_gamma (VMVisc plasticity.cpp:197-229 with hardening of q_yield;
// DPVisc :293-323 and MCCVisc :558-588 without).
MPM_HD double perzyna_newton(const Params& P, double dt, double q_trial, double q_yield, double eps_dev, bool vm_hard, uint32_t& err) {
    double dg = 0.01 * (q_trial - q_yield) / P.f_mu_prefac;
    const int max_iter = 60;
    for (int iter = 0; iter < max_iter; ++iter) {
        if (iter == max_iter - 1) err |= MPM_ERR_NEWTON_MAXITER;
        double tm = P.perzyna_visc * dg + dt;
        double tmp = dt / tm;
        double tmp1 = pow(tmp, P.perzyna_exp);
        if (vm_hard) q_yield = P.q_min + (P.q_max - P.q_min) * exp(-P.xi * (eps_dev + (1.0 / P.d_prefac) * dg));
        double res = (q_trial - P.f_mu_prefac * dg) * tmp1 - q_yield;
        if (fabs(res) < 1e-2) break;
        double rd = -P.f_mu_prefac * tmp1 + (q_trial - P.f_mu_prefac * dg) * P.perzyna_exp * pow(tmp, P.perzyna_exp - 1) * (-P.perzyna_visc * dt) / (tm * tm);
        if (vm_hard) rd -= -P.xi / P.d_prefac * (P.q_max - P.q_min) * exp(-P.xi * (eps_dev + (1.0 / P.d_prefac) * dg));
        if (fabs(rd) < 1e-14) err |= MPM_ERR_NEWTON_SMALL_DERIV;
        dg -= res / rd;
        if (dg < 0) dg = 1e-10;
    }
    return dg;
}

// Compile-time material variants beyond the plastic-model values: the Perzyna laws whose parameters select
// the closed-form rate update (plasticity.cpp:188-194 for VMVisc: exponent 1 and no softening;
// :284-290 for DPVisc: exponent 1) carry only that branch, not the Newton iteration and its registers.
// The arithmetic is the same as the generic variant's closed form.
constexpr int kMatVMViscLinear = 17;
constexpr int kMatDPViscLinear = 18;

// plasticity(), plasticity.cpp:17-621. `h` holds the clamped principal Hencky strains
// ln(max(|s|,1e-4)) of F_trial on entry and the projected strains on exit when recompose is set.
// For VMVisc's capped branch the result is F = U W^T, i.e. h = 0. PM >= 0 fixes the plastic model at
// compile time (the kernel then carries only that model's code and registers); PM < 0 reads it from P.
// EXACT = false (the split form): the divisions of the shared head and of the von Mises and Drucker-Prager branches
// by dim, |dev|, d_prefac, dt and K become products with rcp_d reciprocals (about 1 ulp each).
template <int D, int PM = -1, bool EXACT = true>
MPM_HD RMResult return_map(const Params& P, double dt, V<D>& h, PlasticState& st) {
    RMResult r{false, false, 0, 0u};
    const double dim = D;
    double tr = h[0]; for (int i = 1; i < D; ++i) tr += h[i];
    V<D> dev; for (int i = 0; i < D; ++i) dev[i] = h[i] - (EXACT ? tr / dim : tr * (1.0 / dim)) * 1.0;
    double nrm2 = dev[0] * dev[0]; for (int i = 1; i < D; ++i) nrm2 += dev[i] * dev[i];
    const double dn = sqrt(nrm2);
    if (dn > 0) {
        if (EXACT) { for (int i = 0; i < D; ++i) dev[i] /= dn; }
        else { const double idn = rcp_d(dn); for (int i = 0; i < D; ++i) dev[i] *= idn; }
    }
    const double K = P.K, inv_d = EXACT ? 1.0 / P.d_prefac : rcp_d(P.d_prefac);
    auto div = [&](double a, double b) { return EXACT ? a / b : a * rcp_d(b); };
    auto set_iso = [&](double pp) { double e = EXACT ? -pp / (K * dim) : -pp * rcp_d(K * dim); for (int i = 0; i < D; ++i) h[i] = e * 1.0; r.recompose = true; };
    auto shift_dev = [&](double amount) { for (int i = 0; i < D; ++i) h[i] -= amount * dev[i]; r.recompose = true; };

    constexpr bool kLin = PM == kMatVMViscLinear || PM == kMatDPViscLinear;
    constexpr int PMB = PM == kMatVMViscLinear ? MPM_PLASTIC_VMVISC : (PM == kMatDPViscLinear ? MPM_PLASTIC_DPVISC : PM);
    switch (PMB >= 0 ? PMB : P.plastic_model) {
    case MPM_PLASTIC_VM: {  // :30-49
        double dg = dn - div(P.q_max, P.e_mu_prefac);
        if (dg > 0) {
            r.plastic_incr++;
            st.dgamma = div(P.d_prefac * dg, dt);
            shift_dev(dg);
            st.eps_dev += dg;
        }
    } break;
    case MPM_PLASTIC_DP: {  // :51-85
        double p_trial = -K * tr;
        double q_yield = P.M * p_trial + P.q_cohesion;
        if (q_yield < 1e-10) {
            r.plastic_incr++;
            double p_proj = div(-P.q_cohesion, P.M);
            set_iso(p_proj);
            double dg = P.d_prefac * dn;
            st.dgamma = div(dg, dt);
            st.eps_dev += inv_d * dg;
            st.eps_vol += div(p_proj - p_trial, K);
        } else {
            double dg = P.d_prefac * (dn - div(q_yield, P.e_mu_prefac));
            if (dg > 0) {
                r.plastic_incr++;
                st.dgamma = div(dg, dt);
                shift_dev(inv_d * dg);
                st.eps_dev += inv_d * dg;
            }
        }
    } break;
    case MPM_PLASTIC_DPSOFT: {  // :87-162
        double p_trial = -K * tr, q_trial = P.e_mu_prefac * dn;
        double p_tip_orig = -P.q_cohesion / P.M;
        double p_tip = p_tip_orig * exp(-P.xi * st.eps_dev);
        double p_shift = P.use_pradhana ? -K * st.eps_prad : 0.0;
        double q_yield = P.M * (p_trial + p_shift) + (-p_tip * P.M);
        double ps = p_trial + p_shift;
        if (ps <= p_tip) {
            r.plastic_incr++;
            double p_proj = p_tip_orig * exp(-P.xi * st.eps_dev);
            set_iso(p_proj);
            double dg = P.d_prefac * dn;
            st.dgamma = dg / dt;
            st.eps_dev += inv_d * dg;
            double inst = (p_proj - p_trial) / K;
            st.eps_vol += inst;
            if (P.use_pradhana) st.eps_prad += inst;
        }
        if (ps > p_tip && q_trial <= q_yield) {
            if (P.use_pradhana) st.eps_prad = 0;
            st.dgamma = 0;
        }
        if (ps > p_tip && q_trial > q_yield) {
            r.plastic_incr++;
            double tmp_dev = st.eps_dev + (dn - q_yield / P.e_mu_prefac);
            double p_proj = p_tip_orig * exp(-P.xi * tmp_dev);
            if (ps < p_proj) {
                double dg = P.d_prefac * dn;
                st.dgamma = dg / dt;
                st.eps_dev += inv_d * dg;
                double inst = (p_proj - p_trial) / K;
                st.eps_vol += inst;
                if (P.use_pradhana) st.eps_prad += inst;
                set_iso(p_proj);
            } else {
                double q_yield_new = P.M * ps + (-p_proj * P.M);
                double dg = P.d_prefac * (dn - q_yield_new / P.e_mu_prefac);
                shift_dev(inv_d * dg);
                st.eps_dev += inv_d * dg;
                st.dgamma = dg / dt;
                if (P.use_pradhana) st.eps_prad = 0;
            }
        }
    } break;
    case MPM_PLASTIC_VMVISC: {  // :164-238
        double q_trial = P.e_mu_prefac * dn;
        double q_yield = P.q_min + (P.q_max - P.q_min) * exp(-P.xi * st.eps_dev);
        double p_trial = -K * tr;
        if (p_trial < P.p_min * exp(-P.xi * st.eps_vol)) {
            double dg = q_trial / P.f_mu_prefac;
            double inst = -p_trial / K;
            for (int i = 0; i < D; ++i) h[i] = 0.0;
            r.recompose = true;
            st.eps_vol += inst;
            st.eps_dev += inv_d * dg;
            st.dgamma = dg / dt;
        } else if (q_trial > q_yield) {
            r.plastic_incr++;
            double dg;
            if (kLin || (P.perzyna_exp == 1 && P.xi == 0)) {
                dg = (q_trial - q_yield) / (P.f_mu_prefac + q_yield * P.perzyna_visc / dt);
                if (dg < 0) r.err |= MPM_ERR_PERZYNA_NEGATIVE;
            } else {
                if constexpr (!kLin) dg = perzyna_newton(P, dt, q_trial, q_yield, st.eps_dev, true, r.err);
            }
            shift_dev(inv_d * dg);
            st.eps_dev += inv_d * dg;
            st.dgamma = dg / dt;
        }
    } break;
    case MPM_PLASTIC_DPVISC: {  // :240-339
        double p_trial = -K * tr, q_trial = P.e_mu_prefac * dn;
        double p_tip = -P.q_cohesion / P.M;
        double p_shift = P.use_pradhana ? -K * st.eps_prad : 0.0;
        double ps = p_trial + p_shift;
        double q_yield = P.M * ps + P.q_cohesion;
        if (P.use_mibf) st.muI = P.M;
        if (ps < p_tip) {
            double dg = P.d_prefac * dn;
            double p_proj = p_tip;
            double inst = (p_proj - p_trial) / K;
            r.plastic_incr++;
            set_iso(p_proj);
            st.dgamma = dg / dt;
            st.eps_dev += inv_d * dg;
            st.eps_vol += inst;
            if (P.use_pradhana) st.eps_prad += inst;
        } else {
            if (P.use_pradhana) st.eps_prad = 0;
            st.dgamma = 0;
        }
        if (ps > p_tip && q_trial > q_yield) {
            r.plastic_incr++;
            double dg;
            if (kLin || P.perzyna_exp == 1) {
                dg = (q_trial - q_yield) / (P.f_mu_prefac + q_yield * P.perzyna_visc / dt);
                if (dg < 0) r.err |= MPM_ERR_PERZYNA_NEGATIVE;
            } else {
                if constexpr (!kLin) dg = perzyna_newton(P, dt, q_trial, q_yield, st.eps_dev, false, r.err);
            }
            if (P.use_mibf) {
                if (fabs(ps) < 1e-10) st.muI = 1e10;
                else st.muI = ((q_trial - P.f_mu_prefac * dg) - P.q_cohesion) / ps;
            }
            shift_dev(inv_d * dg);
            st.eps_dev += inv_d * dg;
            st.dgamma = dg / dt;
        }
    } break;
    case MPM_PLASTIC_DPMUI: {  // :342-403
        double p_trial = -K * tr, q_trial = P.e_mu_prefac * dn;
        double p_tip = -P.q_cohesion / P.mu_1;
        double p_shift = P.use_pradhana ? -K * st.eps_prad : 0.0;
        double ps = p_trial + p_shift;
        st.muI = P.mu_1;
        st.visc = 0;
        if (ps < p_tip) {
            double dg = P.d_prefac * dn;
            double inst = (p_tip - p_trial) / K;
            r.plastic_incr++;
            set_iso(p_tip);
            st.dgamma = dg;  // not divided by dt (plasticity.cpp:364)
            st.eps_dev += inv_d * dg;
            st.eps_vol += inst;
            st.eps_prad += inst;
        } else {
            st.eps_prad = 0;
        }
        double q_yield = P.mu_1 * ps + P.q_cohesion;
        if (ps > p_tip && q_trial > (q_yield + P.stress_tolerance)) {
            r.plastic_incr++;
            double p_special = ps - p_tip + P.stress_tolerance;
            double fa = P.f_mu_prefac * dt;
            double fb = p_trial * (P.mu_2 - P.mu_1) + P.f_mu_prefac * dt * P.fac_Q * sqrt(p_special) - (q_trial - q_yield);
            double fc = -(q_trial - q_yield) * P.fac_Q * sqrt(p_special);
            double dg = (-fb + sqrt(fb * fb - 4 * fa * fc)) / (2 * fa);
            double mu_i = P.mu_1 + (P.mu_2 - P.mu_1) / (P.fac_Q * sqrt(p_special) / dg + 1.0);
            st.muI = mu_i;
            st.visc = (mu_i - P.mu_1) * p_special / dg;
            dg *= dt;
            shift_dev(inv_d * dg);
            st.eps_dev += inv_d * dg;
            st.dgamma = dg / dt;
        }
    } break;
    case MPM_PLASTIC_MCCMUI:
    case MPM_PLASTIC_MCC:
    case MPM_PLASTIC_MCCVISC: {  // MCCMui :405-484, MCC/MCCVisc :487-621
        const bool mui = P.plastic_model == MPM_PLASTIC_MCCMUI;
        const double Mm = mui ? P.mu_1 : P.M;
        double p_stress = -K * tr, q_stress = P.e_mu_prefac * dn;
        const double q_trial = q_stress;
        if (P.plastic_model == MPM_PLASTIC_MCCVISC && P.use_mibf) st.muI = P.M;
        bool rma = false;
        double p_c = 0;
        switch (P.hardening_law) {
        case MPM_HARD_NONE:
            if (mui) r.err |= MPM_ERR_INVALID_ARGUMENT;
            else rma = mcc_explicit(p_stress, q_stress, Mm, P.p0, P.beta, P.mu, K, P.rma_prefac);
            break;
        case MPM_HARD_EXPO_EXPL:
            p_c = fmax(P.stress_tolerance, P.p0 * exp(-P.xi * st.eps_vol));
            rma = mcc_explicit(p_stress, q_stress, Mm, p_c, P.beta, P.mu, K, P.rma_prefac);
            break;
        case MPM_HARD_SINH_EXPL:
            p_c = fmax(P.stress_tolerance, K * sinh(-P.xi * st.eps_vol + asinh(P.p0 / K)));
            rma = mcc_explicit(p_stress, q_stress, Mm, p_c, P.beta, P.mu, K, P.rma_prefac);
            break;
        case MPM_HARD_EXPO_IMPL:
            rma = mcc_implicit<false>(p_stress, q_stress, r.err, Mm, P.p0, P.beta, P.mu, K, P.xi, st.eps_vol);
            break;
        case MPM_HARD_SINH_IMPL:
            rma = mcc_implicit<true>(p_stress, q_stress, r.err, Mm, P.p0, P.beta, P.mu, K, P.xi, st.eps_vol);
            break;
        default: r.err |= MPM_ERR_INVALID_ARGUMENT; break;
        }
        if (mui) { st.muI = P.mu_1; st.visc = 0; }
        if (rma) {
            r.plastic_incr++;
            double ep = p_stress / (K * dim);
            double inst = tr + dim * ep;
            st.eps_vol += inst;
            if (mui) {
                if (q_trial > (q_stress + P.stress_tolerance)) {
                    if (P.hardening_law == MPM_HARD_EXPO_IMPL) p_c = fmax(P.p0 * exp(-P.xi * st.eps_vol), p_stress + P.stress_tolerance);
                    else if (P.hardening_law == MPM_HARD_SINH_IMPL) p_c = fmax(K * sinh(-P.xi * st.eps_vol + asinh(P.p0 / K)), p_stress + P.stress_tolerance);
                    double p_special = p_stress + P.beta * p_c + P.stress_tolerance;
                    double fa = P.f_mu_prefac * dt;
                    double fb = sqrt((p_c - p_stress) * p_special) * (P.mu_2 - P.mu_1) + P.f_mu_prefac * dt * P.fac_Q * sqrt(p_special) - (q_trial - q_stress);
                    double fc = -(q_trial - q_stress) * P.fac_Q * sqrt(p_special);
                    double gds = (-fb + sqrt(fb * fb - 4 * fa * fc)) / (2 * fa);
                    q_stress = fmax(q_stress, q_trial - P.f_mu_prefac * dt * gds);
                    double mu_i = P.mu_1 + (P.mu_2 - P.mu_1) / (P.fac_Q * sqrt(p_special) / gds + 1.0);
                    st.muI = mu_i;
                    st.visc = (mu_i - P.mu_1) * sqrt((p_c - p_stress) * p_special) / gds;
                    double dg = (q_trial - q_stress) / P.f_mu_prefac;
                    st.eps_dev += inv_d * dg;
                    st.dgamma = dg / dt;
                }
            } else {
                if (P.plastic_model == MPM_PLASTIC_MCCVISC) {
                    double q_yield = q_stress, dg;
                    if (P.perzyna_exp == 1) {
                        dg = (q_trial - q_yield) / (P.f_mu_prefac + q_yield * P.perzyna_visc / dt);
                        if (dg < 0) { if (dg > -1e-14) dg = 0; else r.err |= MPM_ERR_PERZYNA_NEGATIVE; }
                    } else {
                        dg = perzyna_newton(P, dt, q_trial, q_yield, st.eps_dev, false, r.err);
                    }
                    q_stress = fmax(q_yield, q_trial - P.f_mu_prefac * dg);
                    if (P.use_mibf) {
                        if (P.hardening_law == MPM_HARD_EXPO_IMPL) p_c = fmax(0.0, P.p0 * exp(-P.xi * st.eps_vol));
                        else if (P.hardening_law == MPM_HARD_SINH_IMPL) p_c = fmax(0.0, K * sinh(-P.xi * st.eps_vol + asinh(P.p0 / K)));
                        if ((p_stress < -P.beta * p_c + P.stress_tolerance) || (p_stress > p_c - P.stress_tolerance)) st.muI = P.M;
                        else st.muI = q_stress / sqrt((p_stress + P.beta * p_c) * (p_c - p_stress));
                    }
                }
                double dg = (q_trial - q_stress) / P.f_mu_prefac;
                st.eps_dev += inv_d * dg;
                st.dgamma = dg / dt;
            }
            for (int i = 0; i < D; ++i) h[i] = (q_stress / P.e_mu_prefac) * dev[i] - ep * 1.0;
            r.recompose = true;
        }
    } break;
    default: break;  // NoPlasticity
    }
    r.plastic = r.plastic_incr > 0;
    return r;
}

// ------------------------------------------------------------------------------------------------
// Boundary conditions on one grid node (boundaryCollision, boundary_collision.cpp:5-394).
// ------------------------------------------------------------------------------------------------
template <int D> MPM_HD V<D> normalized_v(V<D> n) {
    double z = n[0] * n[0]; for (int i = 1; i < D; ++i) z += n[i] * n[i];
    if (z > 0) { double s = sqrt(z); for (int i = 0; i < D; ++i) n[i] = n[i] / s; }
    return n;
}

// Level set of an MPM_OBJECT_SDF object at X: phi and its gradient interpolated (trilinear in 3D, bilinear in
// 2D, the BoxSampler of object_vdb.hpp:44-77) from the lattice's 2^dim surrounding points. False outside the
// lattice's box.
template <int D> MPM_HD bool sdf_sample(const DevObject& o, const V<D>& X, double& phi, double* g) {
    const DevSdf& L = o.sdf;
    int i0[3] = {0, 0, 0};
    double f[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        const double u = (X[d] - o.center[d]) / o.param;
        if (!(u >= 0.0 && u <= (double)(L.n[d] - 1))) return false;
        const int fi = (int)floor(u), last = (int)L.n[d] - 2;
        i0[d] = fi < last ? fi : last;
        f[d] = u - (double)i0[d];
    }
    phi = 0.0;
    for (int d = 0; d < 3; ++d) g[d] = 0.0;
    for (int c = 0; c < (1 << D); ++c) {
        double w = 1.0;
        uint32_t idx = 0;
        for (int d = 0; d < D; ++d) {
            const int b = (c >> (D - 1 - d)) & 1;
            w = w * (b ? f[d] : 1.0 - f[d]);
            idx = idx * L.n[d] + (uint32_t)(i0[d] + b);
        }
        const float* s = L.field + 4 * (size_t)idx;
        phi += w * (double)s[0];
        for (int d = 0; d < D; ++d) g[d] += w * (double)s[1 + d];
    }
    return true;
}

// inside()/normal() of the analytic objects (src/objects/object_*.hpp) and of sampled level sets.
template <int D> MPM_HD bool object_inside(const DevObject& o, const V<D>& X) {
    const double x = X[0], y = X[1];
    switch (o.kind) {
    case MPM_OBJECT_BOX_ROTATED: {  // object_box.hpp:18-24
        for (int i = 0; i < D; ++i) {
            double l = o.rot[0 + i * D] * (X[0] - o.center[0]);
            for (int k = 1; k < D; ++k) l += o.rot[k + i * D] * (X[k] - o.center[k]);
            if (fabs(l) > o.half_extents[i]) return false;
        }
        return true;
    }
    case MPM_OBJECT_BUMP: return y < 0.0475 / cosh(25 * (x - 0.43));          // object_bump.hpp:15-27
    case MPM_OBJECT_GROUND: return !(y > o.param);                              // object_ground.hpp:16-25
    case MPM_OBJECT_RAMP: return y < 0.1 * tanh(-100 * x);                      // object_ramp.hpp:15-26
    case MPM_OBJECT_GATE: return y > o.param + 100 * x * x;                     // object_gate.hpp:16-28
    case MPM_OBJECT_SILO: {                                                     // object_silo.hpp:16-35
        double z = X[D - 1];
        if (y < o.param) return false;
        double rs = tanh(y) + 1;
        return !(x * x + z * z < rs * rs);
    }
    case MPM_OBJECT_CURVE: return !(y > x * x);                                 // object_curve.hpp:15-25
    case MPM_OBJECT_SDF: {                                                      // object_vdb.hpp:44-56
        double phi, g[3];
        return sdf_sample<D>(o, X, phi, g) && phi <= 0.0;
    }
    default: return false;
    }
}
template <int D> MPM_HD V<D> object_normal(const DevObject& o, const V<D>& X) {
    V<D> n; for (int i = 0; i < D; ++i) n[i] = 0.0;
    const double x = X[0];
    switch (o.kind) {
    case MPM_OBJECT_BOX_ROTATED: {  // object_box.hpp:26-43
        double loc[3] = {0, 0, 0};
        for (int i = 0; i < D; ++i) { double l = o.rot[0 + i * D] * (X[0] - o.center[0]); for (int k = 1; k < D; ++k) l += o.rot[k + i * D] * (X[k] - o.center[k]); loc[i] = l; }
        int ax = 0; double md = 0.0;
        for (int i = 0; i < D; ++i) { double dd = fabs(o.half_extents[i] - fabs(loc[i])); if (dd > md) { md = dd; ax = i; } }
        double sgn = (loc[ax] > 0) ? 1.0 : -1.0;
        double ln[3] = {0, 0, 0}; ln[ax] = sgn;
        for (int r = 0; r < D; ++r) { double s = o.rot[r + 0 * D] * ln[0]; for (int k = 1; k < D; ++k) s += o.rot[r + k * D] * ln[k]; n[r] = s; }
        return n;
    }
    case MPM_OBJECT_BUMP: { double arg = 25 * (x - 0.43); double b = -25 * 0.0475 * tanh(arg) / cosh(arg); n[0] = -b; n[1] = 1; return normalized_v(n); }
    case MPM_OBJECT_GROUND: n[1] = 1; return n;
    case MPM_OBJECT_RAMP: { double t = cosh(-100 * x); double b = 0.1 * (-100) / (t * t); n[0] = -b; n[1] = 1; return normalized_v(n); }
    case MPM_OBJECT_GATE: { double b = 2.0 * 100 * x; n[0] = -b; n[1] = 1; for (int i = 0; i < D; ++i) n[i] *= -1; return normalized_v(n); }
    case MPM_OBJECT_SILO: { double z = X[D - 1], y = X[1]; double th = atan2(z, x); double sy = 1.0 / cosh(y); n[0] = -cos(th); n[1] = sy * sy; n[D - 1] = -sin(th); return normalized_v(n); }
    case MPM_OBJECT_CURVE: n[0] = -2 * x; n[1] = 1; return normalized_v(n);
    case MPM_OBJECT_SDF: {                                                      // object_vdb.hpp:58-77
        double phi, g[3];
        if (sdf_sample<D>(o, X, phi, g)) for (int i = 0; i < D; ++i) n[i] = g[i];
        return normalized_v(n);
    }
    default: return n;
    }
}

// ObjectPlate::inside, object_plate.hpp:30-78 (3D) / :98-131 (2D)
template <int D> MPM_HD bool plate_inside(const DevPlate& p, const V<D>& X) {
    switch (p.plate_type) {
    case MPM_PLATE_LEFT: return X[1] < p.pos_upper && X[1] > p.pos_lower && (X[0] - p.pos_object) <= 0;
    case MPM_PLATE_RIGHT: return X[1] < p.pos_upper && X[1] > p.pos_lower && (p.pos_object - X[0]) <= 0;
    case MPM_PLATE_BOTTOM: return X[0] < p.pos_upper && X[0] > p.pos_lower && (X[1] - p.pos_object) <= 0;
    case MPM_PLATE_TOP: return X[0] < p.pos_upper && X[0] > p.pos_lower && (p.pos_object - X[1]) <= 0;
    case MPM_PLATE_BACK: return D == 3 && X[D - 1] - p.pos_object <= 0;
    case MPM_PLATE_FRONT: return D == 3 && p.pos_object - X[D - 1] <= 0;
    default: return false;
    }
}

// 3D Coulomb reduction of the two tangential components; the second uses the updated first
// (boundary_collision.cpp:106-119 and its five siblings).
MPM_HD void plate_tangent3(double& a, double& b, double n_comp, double fr) {
    double vt = sqrt(a * a + b * b);
    double fn = fr * fabs(n_comp);
    if (fr == 0) return;
    if (vt <= fn) { a = 0; b = 0; return; }
    a = sgn_d(a) * (vt - fn) / sqrt(1 + b * b / (a * a));
    b = sgn_d(b) * (vt - fn) / sqrt(1 + a * a / (b * b));
}
// 2D Coulomb reduction with the strict stick test (boundary_collision.cpp:286-297).
MPM_HD void plate_tangent2(double& a, double n_comp, double fr) {
    if (fr == 0) return;
    if (fabs(a) < fr * fabs(n_comp)) { a = 0; return; }
    if (a > 0) a -= fr * fabs(n_comp); else a += fr * fabs(n_comp);
}

// boundaryCollision(index, Xi, vi): objects first (chained), then plates (chained).
template <int D>
MPM_HD void boundary_collision(const Params& P, const DevPlate* plates, double dt, double node_friction, V<D> X, V<D>& vi) {
    V<D> vo = vi;
    for (int i = 0; i < D; ++i) X[i] += dt * vi[i];
    for (int k = 0; k < P.n_objects; ++k) {
        const DevObject& o = P.objects[k];
        if (!object_inside<D>(o, X)) continue;
        V<D> vr = vo;
        if (o.bc == MPM_BC_NOSLIP) { for (int i = 0; i < D; ++i) vr[i] = 0.0; }
        else {
            V<D> n = object_normal<D>(o, X);
            double dn = vr[0] * n[0]; for (int i = 1; i < D; ++i) dn += vr[i] * n[i];
            double fr = P.use_mibf ? node_friction : o.friction;
            if (o.bc == MPM_BC_SLIPFREE) {
                if (dn < 0) {
                    V<D> vt; for (int i = 0; i < D; ++i) vt[i] = vr[i] - dn * n[i];
                    if (fr > 0) {
                        double z = vt[0] * vt[0]; for (int i = 1; i < D; ++i) z += vt[i] * vt[i];
                        if ((-dn) * fr < sqrt(z)) { V<D> u = normalized_v(vt); for (int i = 0; i < D; ++i) vr[i] = vt[i] + u[i] * dn * fr; }
                        else for (int i = 0; i < D; ++i) vr[i] = 0.0;
                    } else vr = vt;
                }
            } else {  // SlipStick
                for (int i = 0; i < D; ++i) vr[i] = vr[i] - dn * n[i];
                if (dn < 0 && fr > 0) {
                    double z = vr[0] * vr[0]; for (int i = 1; i < D; ++i) z += vr[i] * vr[i];
                    if ((-dn) * fr < sqrt(z)) { V<D> u = normalized_v(vr); for (int i = 0; i < D; ++i) vr[i] = vr[i] + u[i] * dn * fr; }
                    else for (int i = 0; i < D; ++i) vr[i] = 0.0;
                }
            }
        }
        vi = vr;
        vo = vi;
    }
    for (int k = 0; k < P.n_plates; ++k) {
        const DevPlate& pl = plates[k];
        if (!plate_inside<D>(pl, X)) continue;
        double v[3] = {0, 0, 0};
        for (int i = 0; i < D; ++i) v[i] = vo[i] - pl.v_object[i];
        if (pl.bc == MPM_BC_NOSLIP) { v[0] = v[1] = v[2] = 0; }
        else {
            const double fr = P.use_mibf ? node_friction : pl.friction;
            const bool sf = pl.bc == MPM_BC_SLIPFREE;
            const int t = pl.plate_type;
            const bool tb = t == MPM_PLATE_TOP || t == MPM_PLATE_BOTTOM;
            const bool lr = t == MPM_PLATE_LEFT || t == MPM_PLATE_RIGHT;
            const bool fbk = t == MPM_PLATE_FRONT || t == MPM_PLATE_BACK;
            if (tb && (!sf || (t == MPM_PLATE_TOP && v[1] > 0) || (t == MPM_PLATE_BOTTOM && v[1] < 0))) {
                if (D == 3) plate_tangent3(v[0], v[2], v[1], fr); else plate_tangent2(v[0], v[1], fr);
                v[1] = 0;
            } else if (lr && (!sf || (t == MPM_PLATE_LEFT && v[0] < 0) || (t == MPM_PLATE_RIGHT && v[0] > 0))) {
                if (D == 3) plate_tangent3(v[1], v[2], v[0], fr); else plate_tangent2(v[1], v[0], fr);
                v[0] = 0;
            } else if (D == 3 && fbk && (!sf || (t == MPM_PLATE_FRONT && v[2] > 0) || (t == MPM_PLATE_BACK && v[2] < 0))) {
                plate_tangent3(v[0], v[1], v[2], fr);
                v[2] = 0;
            }
        }
        for (int i = 0; i < D; ++i) vi[i] = v[i] + pl.v_object[i];
        vo = vi;
    }
}

// ObjectPlate::move, object_plate.hpp:136-162
template <int D> MPM_HD void plate_move(DevPlate& p, double dt, double frame_dt, double time) {
    double factor = (time < p.load_factor * frame_dt) ? (1.0 / p.vmin_factor) : 1.0;
    for (int i = 0; i < D; ++i) p.v_object[i] = p.v_object_original[i] * factor;
    if (p.plate_type == MPM_PLATE_LEFT || p.plate_type == MPM_PLATE_RIGHT) {
        p.pos_object += dt * p.v_object[0]; p.pos_upper += dt * p.v_object[1]; p.pos_lower += dt * p.v_object[1];
    } else if (p.plate_type == MPM_PLATE_BOTTOM || p.plate_type == MPM_PLATE_TOP) {
        p.pos_object += dt * p.v_object[1]; p.pos_upper += dt * p.v_object[0]; p.pos_lower += dt * p.v_object[0];
    } else if (D == 3 && (p.plate_type == MPM_PLATE_BACK || p.plate_type == MPM_PLATE_FRONT)) {
        p.pos_object += dt * p.v_object[2];
    }
}

}  // namespace mpm
