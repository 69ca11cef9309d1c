// Scalar building blocks of the hot path, compiled for gfx950 and for the host set-up code.
//
// Every function reproduces the Julia operation order of the reference so that the kernels give the
// same IEEE-754 results (the library is built with -ffp-contract=off; FP64 division and sqrt are
// correctly rounded on gfx950, as on the host).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

#define HCC_HD __host__ __device__ __forceinline__

namespace hcc {

constexpr double kEps = 2.220446049250313e-16;          // eps(Float64)
constexpr double kFloatMin = 2.2250738585072014e-308;   // floatmin(Float64)
constexpr int kMaxStages = 5;

// Julia Base.max / Base.min on IEEE floats (julia/base/math.jl: diff = x - y, pick by signbit(diff), diff if either
// is NaN): NaN-propagating with -0 < +0 -- the IEEE 754-2019 maximum / minimum operations, whose values these
// builtins return (a NaN result is the canonical quiet NaN rather than x - y; no NaN payload is observable here).
// On gfx950 each is one v_max_f64 / v_min_f64 (which order -0 below +0) plus a NaN select.
HCC_HD double jmax(double x, double y) { return __builtin_elementwise_maximum(x, y); }
HCC_HD double jmin(double x, double y) { return __builtin_elementwise_minimum(x, y); }
// Julia Base.mod for floats: r = rem(x, y) (exact), copysign at zero, r + y on a sign mismatch.
HCC_HD double jmod(double x, double y) {
    double r = fmod(x, y);
    if (r == 0.0) return copysign(r, y);
    if ((r > 0) != (y > 0)) return r + y;
    return r;
}

// ExponentialWeightFunction: arg = (-alpha * d2) * (1/R^2), 1 / P4(-arg)
// (src/Grid/MLSWeightFunctions.jl:68-82)
HCC_HD double mls_weight(double alpha, double inv_range_sq, double d2) {
    double arg = -alpha * d2 * inv_range_sq;
    double y = -arg;
    double den = 1.0 + y * (1.0 + y * (0.5 + y * (0.16666666666666666 + y * 0.041666666666666664)));
    return 1.0 / den;
}

// Venkatakrishnan / Barth-Jespersen phi (src/Interpolation/Limiter.jl:35-45)
HCC_HD double vk_phi(double r) {
    // both branches' numerator and denominator are formed and one quotient taken: the same operations, no branch
    const double r_pos = jmax(0.0, r);
    const double s = 1.0 / jmax(1.0, r_pos);
    const bool large = r_pos > 1.0;
    const double num = large ? 1.0 + 2.0 * s : r_pos * r_pos + 2.0 * r_pos;
    const double den = large ? 1.0 + s + 2.0 * (s * s) : r_pos * r_pos + r_pos + 2.0;
    return num / den;
}
HCC_HD double bj_phi(double r) { return jmin(1.0, r); }

// The limiter's phi of one component without a quotient per interface. _bounding_phi (Limiter.jl:93-107) folds
// phi = min(phi, limiter_phi(r_k)) from phi = 1 over the limited slots, r_k = (bound - U_i) / rec_k with the bound on
// rec_k's side: dup = u_max - U_i >= 0 for rec_k > 0, ddn = u_min - U_i <= 0 for rec_k < 0. Within one side the
// rounded quotient is monotone in rec_k (the numerator is fixed and division rounds monotonically), so the smallest
// ratio is that of the largest increment p1 (resp. the most negative n1), and every other slot's ratio is at least that
// of the second largest distinct increment p2 (n2). A slot scan keeps these four numbers (0: none; a limited
// increment exceeds its tolerance, so it is non-zero and not NaN) and phi_scan_finish forms phi from them:
// - Barth-Jespersen: min(1, r) is monotone, so phi = min(1, r*), r* = min(r(p1), r(n1)) (signed zeros included:
//   a zero bound difference gives the same zero ratio for its whole side).
// - Venkatakrishnan: phi = min(1, vk(r*)) when every other ratio r is r*, >= r* + kPhiGap, or >= kPhiFlat:
//   * vk(r) >= 1 exactly for r >= 2, and the rounded large branch keeps that for r >= kPhiFlat: there
//     s = fl(1/r) <= 0.5 - 2e-10, the margin (1 + 2s) - (1 + s + 2s^2) = s (1 - 2s) >= 1e-10 exceeds the rounding of
//     numerator and denominator (< 1e-15), so num >= den and the quotient rounds to >= 1.
//   * on [0, kPhiFlat] vk rises with slope >= 0.124 (vk' = (4 + 4r - r^2) / (r^2 + r + 2)^2, smallest at the right
//     end), so ratios kPhiGap apart differ by >= 1.2e-13 in vk, against a rounding error of the computed vk below
//     1e-15 (a handful of correctly rounded operations on values <= 1.21): the farther ratio's rounded phi is not
//     smaller.
//   Otherwise (a near tie) it returns false and the caller folds every slot as the reference does.
constexpr double kPhiFlat = 2.000000001, kPhiGap = 1e-12;
struct PhiScan {
    double p1, p2, n1, n2;
};
HCC_HD void phi_scan_init(PhiScan& t) { t.p1 = t.p2 = t.n1 = t.n2 = 0.0; }
// v: a limited increment (|v| above its side's tolerance)
HCC_HD void phi_scan_add(PhiScan& t, double v) {
    const bool pos = v > 0.0;
    const bool ptop = pos && v > t.p1, pmid = pos && v < t.p1 && v > t.p2;
    const bool ntop = !pos && v < t.n1, nmid = !pos && v > t.n1 && v < t.n2;
    t.p2 = ptop ? t.p1 : (pmid ? v : t.p2);
    t.p1 = ptop ? v : t.p1;
    t.n2 = ntop ? t.n1 : (nmid ? v : t.n2);
    t.n1 = ntop ? v : t.n1;
}
HCC_HD bool phi_scan_finish(const PhiScan& t, double dup, double ddn, bool bj, double& phi) {
    const double inf = __builtin_inf();
    const double rp = t.p1 > 0.0 ? dup / t.p1 : inf, rn = t.n1 < 0.0 ? ddn / t.n1 : inf;
    const double rs = jmin(rp, rn);
    if (bj) { phi = jmin(1.0, rs); return true; }
    if (rs != rs) return false;
    if (!(rs < kPhiFlat)) { phi = 1.0; return true; }
    const double lo = rs + kPhiGap;
    auto safe = [&](double x) { return x >= lo || x >= kPhiFlat; };
    const double rp2 = t.p2 > 0.0 ? dup / t.p2 : inf, rn2 = t.n2 < 0.0 ? ddn / t.n2 : inf;
    // a zero bound difference gives its whole side the same zero ratio
    if (!((rp == rs || safe(rp)) && (rn == rs || safe(rn)) && (dup == 0.0 || safe(rp2)) && (ddn == 0.0 || safe(rn2))))
        return false;
    phi = jmin(1.0, vk_phi(rs));
    return true;
}

// Compressible Euler (build-defined; the reference ships no Euler PDE): state (rho, m_1..m_D, E),
// u = m * (1/rho), p = (gamma - 1) (E - 0.5 sum m_d u_d), c = sqrt((gamma p) * (1/rho)).
template <int D>
struct EulerPrim {
    double inv, u[D], p;
};
template <int D>
HCC_HD EulerPrim<D> euler_prim(double gamma, const double* U) {
    EulerPrim<D> q;
    q.inv = 1.0 / U[0];
    for (int d = 0; d < D; ++d) q.u[d] = U[1 + d] * q.inv;
    double ke = U[1] * q.u[0];
    for (int d = 1; d < D; ++d) ke = ke + U[1 + d] * q.u[d];
    ke = 0.5 * ke;
    q.p = (gamma - 1.0) * (U[D + 1] - ke);
    return q;
}

// Cholesky of a small SPD moment matrix, column by column with a multiply-add per update and a
// reciprocal of each pivot, and the conditional Tikhonov retry of _cholesky_moment
// (src/Interpolation/Interpolators.jl:46-64): the plain factorisation is kept if every pivot U[k,k]^2
// exceeds 10 eps N[k,k]; otherwise sqrt(eps) tr(N)/B is added to the diagonal and the factorisation
// repeated; false if that fails too.
template <int B>
HCC_HD bool chol_plain(const double (&A)[B][B], double (&U)[B][B]) {
    double L[B][B];
    for (int r = 0; r < B; ++r) for (int c = 0; c < B; ++c) L[r][c] = 0.0;
    for (int n = 0; n < B; ++n) {
        for (int m = n; m < B; ++m) L[m][n] = A[n][m];
        for (int k = 0; k < n; ++k)
            for (int m = n; m < B; ++m) L[m][n] = fma(-L[m][k], L[n][k], L[m][n]);
        if (!(L[n][n] > 0.0)) return false;
        L[n][n] = sqrt(L[n][n]);
        double linv = 1.0 / L[n][n];
        for (int m = n + 1; m < B; ++m) L[m][n] *= linv;
    }
    for (int r = 0; r < B; ++r) for (int c = 0; c < B; ++c) U[r][c] = (c >= r) ? L[c][r] : 0.0;
    return true;
}
template <int B>
HCC_HD bool chol_moment(const double (&A)[B][B], double (&U)[B][B]) {
    if (chol_plain<B>(A, U)) {
        bool ok = true;
        for (int k = 0; k < B; ++k) if (!(U[k][k] * U[k][k] > (10 * kEps) * A[k][k])) ok = false;
        if (ok) return true;
    }
    double tr = A[0][0];
    for (int k = 1; k < B; ++k) tr = tr + A[k][k];
    double reg = 1.4901161193847656e-08 * tr / (double)B;     // sqrt(eps(Float64)) tr / B
    double A2[B][B];
    for (int r = 0; r < B; ++r) for (int c = 0; c < B; ++c) A2[r][c] = A[r][c] + ((r == c) ? reg : 0.0);
    return chol_plain<B>(A2, U);
}
// inv(C) = U \ (U' \ I), one column at a time: forward then back substitution.
template <int B>
HCC_HD void chol_inverse(const double (&U)[B][B], double (&X)[B][B]) {
    for (int j = 0; j < B; ++j) {
        double y[B];
        for (int i = 0; i < B; ++i) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s = s - U[k][i] * y[k];
            y[i] = s / U[i][i];
        }
        for (int i = B - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < B; ++k) s = s - U[i][k] * X[k][j];
            X[i][j] = s / U[i][i];
        }
    }
}

}  // namespace hcc
