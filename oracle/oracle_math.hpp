// oracle_math.hpp — TEST INFRASTRUCTURE ONLY. Part of the CPU oracle: a plain FP64 restatement of the
// reference's arithmetic used to check the GPU engine. Nothing in the product links, includes or calls it.
//
// Small fixed-size vector/matrix kit that replaces the Eigen 3 types of the reference
// (TV/TM/TA, src/tools.hpp:38-50), the B-spline primitives (src/tools.hpp:108-252), arange
// (src/tools.hpp:291-296) and Eigen::JacobiSVD (called at src/simulation/simulation.hpp:232 and
// src/simulation/plasticity.cpp:19).
//
// Eigen is a third-party dependency that is absent from this machine (required version >= 3.3,
// src/CMakeLists.txt:1; exact version unpinned). Its JacobiSVD is restated here from Eigen's published
// two-sided Jacobi algorithm (square case: no QR preconditioner; 2x2 real SVD by a symmetrising rotation
// followed by a Jacobi rotation; stop at 2*epsilon relative to the largest diagonal entry; make the
// diagonal non-negative by negating U columns; sort descending). Evaluation order inside Eigen's
// expression templates is taken as plain left-to-right; Eigen's exact rounding is therefore not pinned.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>
#include <algorithm>

namespace oracle {

typedef double T;

template <int D> struct Vec {
    T a[D];
    T& operator()(int i) { return a[i]; }
    const T& operator()(int i) const { return a[i]; }
    T& operator[](int i) { return a[i]; }
    const T& operator[](int i) const { return a[i]; }
    static Vec Zero() { Vec v; for (int i = 0; i < D; ++i) v.a[i] = 0.0; return v; }
    static Vec Ones() { Vec v; for (int i = 0; i < D; ++i) v.a[i] = 1.0; return v; }
};

// Column-major storage like Eigen's default: element (r,c) at a[r + c*D].
template <int D> struct Mat {
    T a[D * D];
    T& operator()(int r, int c) { return a[r + c * D]; }
    const T& operator()(int r, int c) const { return a[r + c * D]; }
    static Mat Zero() { Mat m; for (int i = 0; i < D * D; ++i) m.a[i] = 0.0; return m; }
    static Mat Identity() { Mat m = Zero(); for (int i = 0; i < D; ++i) m(i, i) = 1.0; return m; }
};

template <int D> inline Vec<D> operator+(const Vec<D>& x, const Vec<D>& y) { Vec<D> r; for (int i = 0; i < D; ++i) r.a[i] = x.a[i] + y.a[i]; return r; }
template <int D> inline Vec<D> operator-(const Vec<D>& x, const Vec<D>& y) { Vec<D> r; for (int i = 0; i < D; ++i) r.a[i] = x.a[i] - y.a[i]; return r; }
template <int D> inline Vec<D> operator*(const Vec<D>& x, T s) { Vec<D> r; for (int i = 0; i < D; ++i) r.a[i] = x.a[i] * s; return r; }
template <int D> inline Vec<D> operator*(T s, const Vec<D>& x) { Vec<D> r; for (int i = 0; i < D; ++i) r.a[i] = s * x.a[i]; return r; }
template <int D> inline Vec<D> operator/(const Vec<D>& x, T s) { Vec<D> r; for (int i = 0; i < D; ++i) r.a[i] = x.a[i] / s; return r; }
template <int D> inline Vec<D>& operator+=(Vec<D>& x, const Vec<D>& y) { for (int i = 0; i < D; ++i) x.a[i] += y.a[i]; return x; }
template <int D> inline Vec<D>& operator-=(Vec<D>& x, const Vec<D>& y) { for (int i = 0; i < D; ++i) x.a[i] -= y.a[i]; return x; }
template <int D> inline Vec<D>& operator/=(Vec<D>& x, T s) { for (int i = 0; i < D; ++i) x.a[i] /= s; return x; }

// Sequential left-to-right reductions: sum() and squaredNorm() as ((a0 op a1) op a2).
template <int D> inline T vsum(const Vec<D>& x) { T s = x.a[0]; for (int i = 1; i < D; ++i) s += x.a[i]; return s; }
template <int D> inline T sqnorm(const Vec<D>& x) { T s = x.a[0] * x.a[0]; for (int i = 1; i < D; ++i) s += x.a[i] * x.a[i]; return s; }
template <int D> inline T vnorm(const Vec<D>& x) { return std::sqrt(sqnorm(x)); }
template <int D> inline T dot(const Vec<D>& x, const Vec<D>& y) { T s = x.a[0] * y.a[0]; for (int i = 1; i < D; ++i) s += x.a[i] * y.a[i]; return s; }
// Eigen >= 3.3 normalized(): returns the input unchanged when its squared norm is not positive.
template <int D> inline Vec<D> normalized(const Vec<D>& x) { T z = sqnorm(x); if (z > 0) return x / std::sqrt(z); return x; }

template <int D> inline Vec<D> matvec(const Mat<D>& m, const Vec<D>& x) {
    Vec<D> r;
    for (int i = 0; i < D; ++i) { T s = m(i, 0) * x(0); for (int k = 1; k < D; ++k) s += m(i, k) * x(k); r(i) = s; }
    return r;
}
template <int D> inline Mat<D> matmul(const Mat<D>& A, const Mat<D>& B) {
    Mat<D> r;
    for (int c = 0; c < D; ++c)
        for (int i = 0; i < D; ++i) { T s = A(i, 0) * B(0, c); for (int k = 1; k < D; ++k) s += A(i, k) * B(k, c); r(i, c) = s; }
    return r;
}
template <int D> inline Mat<D> transpose(const Mat<D>& A) { Mat<D> r; for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) r(i, j) = A(j, i); return r; }
template <int D> inline Mat<D> operator+(const Mat<D>& A, const Mat<D>& B) { Mat<D> r; for (int i = 0; i < D * D; ++i) r.a[i] = A.a[i] + B.a[i]; return r; }
template <int D> inline Mat<D> operator-(const Mat<D>& A, const Mat<D>& B) { Mat<D> r; for (int i = 0; i < D * D; ++i) r.a[i] = A.a[i] - B.a[i]; return r; }
template <int D> inline Mat<D> operator*(T s, const Mat<D>& A) { Mat<D> r; for (int i = 0; i < D * D; ++i) r.a[i] = s * A.a[i]; return r; }
template <int D> inline Mat<D> operator*(const Mat<D>& A, T s) { Mat<D> r; for (int i = 0; i < D * D; ++i) r.a[i] = A.a[i] * s; return r; }
template <int D> inline Mat<D>& operator+=(Mat<D>& A, const Mat<D>& B) { for (int i = 0; i < D * D; ++i) A.a[i] += B.a[i]; return A; }
template <int D> inline Mat<D>& operator/=(Mat<D>& A, T s) { for (int i = 0; i < D * D; ++i) A.a[i] /= s; return A; }
// Outer product a * b^T, coefficient a(r)*b(c).
template <int D> inline Mat<D> outer(const Vec<D>& a, const Vec<D>& b) { Mat<D> r; for (int c = 0; c < D; ++c) for (int i = 0; i < D; ++i) r(i, c) = a(i) * b(c); return r; }
// U * diag(d) * V^T evaluated as (U*diag(d)) * V^T.
template <int D> inline Mat<D> udvt(const Mat<D>& U, const Vec<D>& d, const Mat<D>& V) {
    Mat<D> UD; for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) UD(r, c) = U(r, c) * d(c);
    return matmul(UD, transpose(V));
}
template <int D> inline T trace(const Mat<D>& A) { T s = A(0, 0); for (int i = 1; i < D; ++i) s += A(i, i); return s; }

inline T determinant(const Mat<2>& m) { return m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1); }
inline T determinant(const Mat<3>& m) {
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1))
         - m(1, 0) * (m(0, 1) * m(2, 2) - m(0, 2) * m(2, 1))
         + m(2, 0) * (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1));
}
inline Mat<2> inverse(const Mat<2>& m) {
    T det = determinant(m); Mat<2> r;
    r(0, 0) = m(1, 1) / det; r(1, 1) = m(0, 0) / det; r(0, 1) = -m(0, 1) / det; r(1, 0) = -m(1, 0) / det;
    return r;
}
inline Mat<3> inverse(const Mat<3>& m) {
    Mat<3> c;  // cofactor-transposed (adjugate)
    c(0, 0) = m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1);
    c(1, 0) = m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2);
    c(2, 0) = m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0);
    c(0, 1) = m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2);
    c(1, 1) = m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0);
    c(2, 1) = m(0, 1) * m(2, 0) - m(0, 0) * m(2, 1);
    c(0, 2) = m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1);
    c(1, 2) = m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2);
    c(2, 2) = m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0);
    T det = m(0, 0) * c(0, 0) + m(0, 1) * c(1, 0) + m(0, 2) * c(2, 0);
    for (int i = 0; i < 9; ++i) c.a[i] /= det;
    return c;
}

// sgn, src/tools.hpp:90-92
inline int sgn(T val) { return (T(0) < val) - (val < T(0)); }

// ---------------------------------------------------------------------------------------------
// B-splines. Cubic: src/tools.hpp:110-134. Quadratic: src/tools.hpp:151-175.
// ---------------------------------------------------------------------------------------------
template <int SPL> struct Spline;
template <> struct Spline<3> {
    static T N(T u) {
        T uabs = std::abs(u);
        if (uabs < 1.0) return 0.5 * uabs * uabs * uabs - uabs * uabs + 0.6666666666666666666666666666666666666666666666666;
        else if (uabs < 2.0) return 0.1666666666666666666666666666666666666666666 * (2.0 - uabs) * (2.0 - uabs) * (2.0 - uabs);
        else return 0;
    }
    static T dNdu(T u) {
        T uabs = std::abs(u);
        if (uabs < 1.0) return u * (1.5 * uabs - 2.0);
        else if (uabs < 2.0) return -0.5 * sgn(u) * (2.0 - uabs) * (2.0 - uabs);
        else return 0;
    }
    static constexpr T apic_D_factor = 3.0;  // simulation.cpp:95
};
template <> struct Spline<2> {
    static T N(T u) {
        T uabs = std::abs(u);
        if (uabs < 0.5) return 0.75 - uabs * uabs;
        else if (uabs < 1.5) return 0.5 * (1.5 - uabs) * (1.5 - uabs);
        else return 0;
    }
    static T dNdu(T u) {
        T uabs = std::abs(u);
        if (uabs < 0.5) return (-2 * u);
        else if (uabs < 1.5) return (u - 1.5 * sgn(u));
        else return 0;
    }
    static constexpr T apic_D_factor = 4.0;  // simulation.cpp:98
};

// wip / grad_wip, 3D src/tools.hpp:222-232, 2D src/tools.hpp:243-252.
template <int SPL> inline T wip3(T xp, T yp, T zp, T xi, T yi, T zi, T h1) {
    typedef Spline<SPL> S;
    return S::N((xp - xi) * h1) * S::N((yp - yi) * h1) * S::N((zp - zi) * h1);
}
template <int SPL> inline Vec<3> grad_wip3(T xp, T yp, T zp, T xi, T yi, T zi, T h1) {
    typedef Spline<SPL> S;
    Vec<3> o;
    o(0) = S::dNdu((xp - xi) * h1) * S::N((yp - yi) * h1) * S::N((zp - zi) * h1) * h1;
    o(1) = S::dNdu((yp - yi) * h1) * S::N((xp - xi) * h1) * S::N((zp - zi) * h1) * h1;
    o(2) = S::dNdu((zp - zi) * h1) * S::N((xp - xi) * h1) * S::N((yp - yi) * h1) * h1;
    return o;
}
template <int SPL> inline T wip2(T xp, T yp, T xi, T yi, T h1) {
    typedef Spline<SPL> S;
    return S::N((xp - xi) * h1) * S::N((yp - yi) * h1);
}
template <int SPL> inline Vec<2> grad_wip2(T xp, T yp, T xi, T yi, T h1) {
    typedef Spline<SPL> S;
    Vec<2> o;
    o(0) = S::dNdu((xp - xi) * h1) * S::N((yp - yi) * h1) * h1;
    o(1) = S::dNdu((yp - yi) * h1) * S::N((xp - xi) * h1) * h1;
    return o;
}

// arange, src/tools.hpp:291-296 (accumulating, excludes stop).
inline std::vector<T> arange(T start, T stop, T step) {
    std::vector<T> values;
    for (T value = start; value < stop; value += step) values.push_back(value);
    return values;
}

// ---------------------------------------------------------------------------------------------
// Two-sided Jacobi SVD (restatement of Eigen::JacobiSVD for square real matrices, see header note).
// ---------------------------------------------------------------------------------------------
struct Rot { T c, s; };

// Rows p,q of W <- J * (rows p,q):  x' = c x + s y ; y' = -s x + c y
template <int D> inline void rot_rows(Mat<D>& W, int p, int q, Rot j) {
    if (j.c == 1 && j.s == 0) return;
    for (int i = 0; i < D; ++i) { T x = W(p, i), y = W(q, i); W(p, i) = j.c * x + j.s * y; W(q, i) = -j.s * x + j.c * y; }
}
// Columns p,q of W <- (cols p,q) * J, applied as the plane rotation by J^T on the column pair.
template <int D> inline void rot_cols(Mat<D>& W, int p, int q, Rot j) {
    Rot jt = {j.c, -j.s};
    if (jt.c == 1 && jt.s == 0) return;
    for (int i = 0; i < D; ++i) { T x = W(i, p), y = W(i, q); W(i, p) = jt.c * x + jt.s * y; W(i, q) = -jt.s * x + jt.c * y; }
}
// Symmetric Jacobi rotation annihilating y in [[x y][y z]].
inline Rot make_jacobi(T x, T y, T z) {
    Rot r;
    T deno = 2.0 * std::abs(y);
    if (deno < std::numeric_limits<T>::min()) { r.c = 1; r.s = 0; return r; }
    T tau = (x - z) / deno;
    T w = std::sqrt(tau * tau + 1.0);
    T t = (tau > 0) ? 1.0 / (tau + w) : 1.0 / (tau - w);
    T sign_t = t > 0 ? 1 : -1;
    T n = 1.0 / std::sqrt(t * t + 1.0);
    r.s = -sign_t * (y / std::abs(y)) * std::abs(t) * n;
    r.c = n;
    return r;
}
template <int D> inline void real_2x2_jacobi_svd(const Mat<D>& W, int p, int q, Rot& jl, Rot& jr) {
    T m00 = W(p, p), m01 = W(p, q), m10 = W(q, p), m11 = W(q, q);
    Rot r1;
    T t = m00 + m11;
    T d = m10 - m01;
    if (std::abs(d) < std::numeric_limits<T>::min()) { r1.s = 0; r1.c = 1; }
    else { T u = t / d; T tmp = std::sqrt(1.0 + u * u); r1.s = 1.0 / tmp; r1.c = u / tmp; }
    // m <- r1 * m (rows 0,1 of the 2x2)
    T n00 = r1.c * m00 + r1.s * m10, n01 = r1.c * m01 + r1.s * m11;
    T n11 = -r1.s * m01 + r1.c * m11;
    jr = make_jacobi(n00, n01, n11);
    // jl = r1 * jr^T ; product of rotations (c1 c2 - s1 s2, c1 s2 + s1 c2)
    Rot jrt = {jr.c, -jr.s};
    jl.c = r1.c * jrt.c - r1.s * jrt.s;
    jl.s = r1.c * jrt.s + r1.s * jrt.c;
}

template <int D> inline void svd(const Mat<D>& A, Mat<D>& U, Vec<D>& S, Mat<D>& V) {
    const T precision = 2.0 * std::numeric_limits<T>::epsilon();
    const T considerAsZero = std::numeric_limits<T>::min();
    T scale = 0; for (int i = 0; i < D * D; ++i) scale = std::max(scale, std::abs(A.a[i]));
    if (scale == 0) scale = 1;
    Mat<D> W; for (int i = 0; i < D * D; ++i) W.a[i] = A.a[i] / scale;
    U = Mat<D>::Identity(); V = Mat<D>::Identity();
    T maxDiag = 0; for (int i = 0; i < D; ++i) maxDiag = std::max(maxDiag, std::abs(W(i, i)));
    bool finished = false;
    int sweeps = 0;
    while (!finished && sweeps < 100) {
        finished = true; ++sweeps;
        for (int p = 1; p < D; ++p) for (int q = 0; q < p; ++q) {
            T threshold = std::max(considerAsZero, precision * maxDiag);
            if (std::abs(W(p, q)) > threshold || std::abs(W(q, p)) > threshold) {
                finished = false;
                Rot jl, jr;
                real_2x2_jacobi_svd(W, p, q, jl, jr);
                rot_rows(W, p, q, jl);
                rot_cols(U, p, q, Rot{jl.c, -jl.s});   // U <- U * jl^T
                rot_cols(W, p, q, jr);
                rot_cols(V, p, q, jr);
                maxDiag = std::max(maxDiag, std::max(std::abs(W(p, p)), std::abs(W(q, q))));
            }
        }
    }
    for (int i = 0; i < D; ++i) {
        T a = W(i, i);
        S(i) = std::abs(a);
        if (a < 0) for (int r = 0; r < D; ++r) U(r, i) = -U(r, i);
    }
    for (int i = 0; i < D; ++i) S(i) *= scale;
    for (int i = 0; i < D; ++i) {
        int pos = i; T mx = S(i);
        for (int k = i + 1; k < D; ++k) if (S(k) > mx) { mx = S(k); pos = k; }
        if (mx == 0) break;
        if (pos != i) {
            std::swap(S(i), S(pos));
            for (int r = 0; r < D; ++r) { std::swap(U(r, i), U(r, pos)); std::swap(V(r, i), V(r, pos)); }
        }
    }
}

}  // namespace oracle
