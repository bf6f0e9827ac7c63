// Small dense FP64 solvers used in the epilogue of the scan kernels.
// Sizes are compile-time (NB <= RR_MAX_NBASIS) and fully unrolled so the
// matrices stay in registers.
#pragma once
#include <hip/hip_runtime.h>

namespace rr {

// Solve M c = y by LU with partial pivoting (first maximum |a_ik| as the
// pivot, like LAPACK getrf).  M is given as a full NB x NB array and is
// destroyed.  Returns false when a pivot is exactly zero — the case in which
// np.linalg.solve raises LinAlgError (zscan.py:1150, :769-771).
template <int NB>
__device__ __forceinline__ bool lu_solve(double (&a)[NB][NB], double (&c)[NB]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        double best = fabs(a[k][k]);
        int piv = k;
#pragma unroll
        for (int i = k + 1; i < NB; ++i) {
            const double v = fabs(a[i][k]);
            if (v > best) { best = v; piv = i; }
        }
        // row interchange as predicated moves (piv is a runtime value)
#pragma unroll
        for (int i = k + 1; i < NB; ++i) {
            const bool sw = (i == piv);
#pragma unroll
            for (int j = k; j < NB; ++j) {
                const double t0 = a[k][j], t1 = a[i][j];
                a[k][j] = sw ? t1 : t0;
                a[i][j] = sw ? t0 : t1;
            }
            const double r0 = c[k], r1 = c[i];
            c[k] = sw ? r1 : r0;
            c[i] = sw ? r0 : r1;
        }
        const double p = a[k][k];
        if (p == 0.0) ok = false;
        const double inv = (p == 0.0) ? 0.0 : 1.0 / p;
#pragma unroll
        for (int i = k + 1; i < NB; ++i) {
            const double l = a[i][k] * inv;
#pragma unroll
            for (int j = k + 1; j < NB; ++j) a[i][j] = fma(-l, a[k][j], a[i][j]);
            c[i] = fma(-l, c[k], c[i]);
        }
    }
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
        double s = c[k];
#pragma unroll
        for (int j = k + 1; j < NB; ++j) s = fma(-a[k][j], c[j], s);
        c[k] = (a[k][k] == 0.0) ? 0.0 : s / a[k][k];
    }
    return ok;
}

// Cholesky solve of a symmetric positive definite system, lower triangle of
// `a` used.  Returns false if a pivot is not strictly positive.
template <int NB>
__device__ __forceinline__ bool chol_solve(double (&a)[NB][NB], double (&c)[NB]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double d = a[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) d = fma(-a[j][k], a[j][k], d);
        if (!(d > 0.0)) ok = false;
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        const double inv = 1.0 / ljj;
        a[j][j] = ljj;
#pragma unroll
        for (int i = j + 1; i < NB; ++i) {
            double s = a[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) s = fma(-a[i][k], a[j][k], s);
            a[i][j] = s * inv;
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        double s = c[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = fma(-a[i][k], c[k], s);
        c[i] = s / a[i][i];
    }
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
        double s = c[i];
#pragma unroll
        for (int k = i + 1; k < NB; ++k) s = fma(-a[k][i], c[k], s);
        c[i] = s / a[i][i];
    }
    return ok;
}

// Error-free transformations for the compensated chi2 = f'Wf - c'y.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// chi2(c) = fwf - 2 c'y + c'Mc for the unregularised M (no prior): the
// reference's explicit residual sum(w (f - A c)^2) expanded, stationary in c
// so that solve errors enter only at second order.  The final sum
// fwf + sum_b c_b (t_b - 2 y_b), t = M c, is carried in double-double
// (Ogita-Rump-Oishi Dot2) because fwf and c'y cancel at high S/N.
template <int NB>
__device__ __forceinline__ double compensated_chi2(double fwf, const double (&c)[NB],
                                                   const double (&y)[NB],
                                                   const double (&t)[NB]) {
    double s = fwf, comp = 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const double g = t[b] - 2.0 * y[b];
        const double p = c[b] * g;
        const double pe = fma(c[b], g, -p);
        double u, e;
        two_sum(s, p, u, e);
        s = u;
        comp += e + pe;
    }
    return s + comp;
}

}  // namespace rr
