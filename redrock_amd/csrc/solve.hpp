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

// min |M[:, F] x - y|^2 (x_j = 0 outside the free set F) by Householder QR of
// the free columns taken in index order.  A column with nothing left after
// the reflections of the earlier ones (zero, or exactly dependent on them)
// gets x_j = 0, the basic solution.
template <int NB>
__device__ inline void ls_free(const double (&M)[NB][NB], const double (&y)[NB], unsigned F, double (&x)[NB]) {
    double A[NB][NB], b[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        b[i] = y[i];
#pragma unroll
        for (int j = 0; j < NB; ++j) A[i][j] = M[i][j];
    }
    int piv[NB];
    int r = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        piv[j] = -1;
        if (!((F >> j) & 1u)) continue;
        // reflect rows r.. of column j onto row r
        double nrm2 = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (i >= r) nrm2 = fma(A[i][j], A[i][j], nrm2);
        if (!(nrm2 > 0.0)) continue;
        double arr = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (i == r) arr = A[i][j];
        const double alpha = arr >= 0.0 ? -sqrt(nrm2) : sqrt(nrm2);
        double v[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) v[i] = (i > r) ? A[i][j] : (i == r ? arr - alpha : 0.0);
        const double vv = (arr - alpha) * (arr - alpha) + (nrm2 - arr * arr);
        if (vv > 0.0) {
            const double s2 = 2.0 / vv;
#pragma unroll
            for (int j2 = 0; j2 < NB; ++j2) {
                if (j2 <= j || !((F >> j2) & 1u)) continue;
                double d = 0.0;
#pragma unroll
                for (int i = 0; i < NB; ++i) d = fma(v[i], A[i][j2], d);
                d *= s2;
#pragma unroll
                for (int i = 0; i < NB; ++i) A[i][j2] = fma(-d, v[i], A[i][j2]);
            }
            double d = 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) d = fma(v[i], b[i], d);
            d *= s2;
#pragma unroll
            for (int i = 0; i < NB; ++i) b[i] = fma(-d, v[i], b[i]);
#pragma unroll
            for (int i = 0; i < NB; ++i)
                if (i == r) A[i][j] = alpha;
        }
        piv[j] = r;
        ++r;
    }
    // back substitution, last free column first
#pragma unroll
    for (int j = NB - 1; j >= 0; --j) {
        x[j] = 0.0;
        if (piv[j] < 0) continue;
        const int pr = piv[j];
        double s = 0.0, diag = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (i == pr) {
                s = b[i];
                diag = A[i][j];
#pragma unroll
                for (int j2 = 0; j2 < NB; ++j2)
                    if (j2 > j) s = fma(-A[i][j2], x[j2], s);
            }
        x[j] = s / diag;
    }
}


// Non-negative least squares min |M x - y|^2 over x_j >= 0 for the
// coefficients j in `bounded` (the others free) by the active-set method of
// Lawson and Hanson (Solving Least Squares Problems, 1974, ch. 23, with the
// free coefficients kept in the passive set throughout).  All coefficients
// bounded: the problem scipy.optimize.nnls(M, y) solves for the reference's
// NMF templates (zscan.py:1153-1172); some bounded below by 0 and the rest
// free: scipy lsq_linear(M, y, bounds, 'bvls') with the archetype bounds of
// zscan.py:289-294.  Each passive-set subproblem is ls_free (Householder QR
// of the passive columns).  The solution is unique when M has full rank.
// Returns false when 3 NB inner iterations did not settle (scipy raises
// there; the reference then gives HUGE_CHI2, zscan.py:769-771).
template <int NB>
__device__ inline bool nnls_solve(const double (&M)[NB][NB], const double (&y)[NB], double (&x)[NB],
                                  unsigned bounded) {
    double scale = 0.0, ymax = 0.0;          // |M|_1 |y|_inf: the scale of the gradient
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double cs = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i) cs += fabs(M[i][j]);
        scale = fmax(scale, cs);
        ymax = fmax(ymax, fabs(y[j]));
    }
    const double tol = 10.0 * NB * 2.220446049250313e-16 * scale * ymax;
    double w[NB];                            // M^T (y - M x), the negative gradient
    auto gradient = [&]() {
        double r[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double v = y[i];
#pragma unroll
            for (int j = 0; j < NB; ++j) v = fma(-M[i][j], x[j], v);
            r[i] = v;
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) v = fma(M[i][j], r[i], v);
            w[j] = v;
        }
    };
    constexpr unsigned all = NB >= 32 ? 0xffffffffu : ((1u << NB) - 1u);
    bounded &= all;
    unsigned P = all & ~bounded;             // passive set: the free coefficients always
    if (P)
        ls_free<NB>(M, y, P, x);
    else {
#pragma unroll
        for (int j = 0; j < NB; ++j) x[j] = 0.0;
    }
    gradient();
    unsigned rejected = 0u;                  // indices that cannot enter at this x
    int iter = 0;
    for (int outer = 0; outer < 4 * NB; ++outer) {
        int t = -1;
        double wt = tol;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (((bounded & ~(P | rejected)) >> j & 1u) && w[j] > wt) {
                wt = w[j];
                t = j;
            }
        if (t < 0) return true;              // Kuhn-Tucker conditions hold
        const unsigned P0 = P;
        P |= 1u << t;
        double s[NB];
        ls_free<NB>(M, y, P, s);
        bool tpos = false;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j == t) tpos = s[j] > 0.0;
        if (!tpos) {                         // t would enter at a non-positive value
            P = P0;
            rejected |= 1u << t;
            continue;
        }
        for (;;) {
            bool feasible = true;
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if ((((P & bounded) >> j) & 1u) && !(s[j] > 0.0)) feasible = false;
            if (feasible) {
#pragma unroll
                for (int j = 0; j < NB; ++j) x[j] = s[j];
                break;
            }
            if (++iter > 3 * NB) return false;
            // move from x towards s until the first passive coefficient reaches 0
            double alpha = 2.0;
            int jmin = -1;
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if ((((P & bounded) >> j) & 1u) && !(s[j] > 0.0)) {
                    const double d = x[j] - s[j];
                    const double a = d > 0.0 ? x[j] / d : 0.0;
                    if (a < alpha) {
                        alpha = a;
                        jmin = j;
                    }
                }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                x[j] = fma(alpha, s[j] - x[j], x[j]);
                if (((bounded >> j) & 1u) &&
                    (j == jmin || (((P >> j) & 1u) && !(x[j] > 0.0)))) {
                    P &= ~(1u << j);
                    x[j] = 0.0;
                }
            }
            if (P == 0u) break;
            ls_free<NB>(M, y, P, s);
        }
        rejected = 0u;
        gradient();
    }
    return false;
}

// Bounded-variable least squares min |M x - y|^2 over lo <= x <= hi (either
// bound may be infinite): scipy lsq_linear(M, y, bounds, 'bvls'), the
// reference's BVLS branch (zscan.py:1175-1201) for any bounds.  The active-set
// method of Stark and Parker (Bounded-variable least squares, Computational
// Statistics 10, 1995), in the form of nnls_solve above with two-sided
// bounds: a coefficient with a finite bound starts at it (the lower one when
// both are finite), the unbounded ones stay free throughout; the bound
// coefficient whose gradient points most into the box enters the free set;
// an infeasible free-set solution is cut back to the first bound crossed.
// Each free-set subproblem is ls_free on y - M_B x_B.  lo[j] < hi[j] for every
// j (the host checks).  Returns false when 3 NB inner iterations did not
// settle (the reference's LinAlgError, zscan.py:769-771).
template <int NB>
__device__ inline bool bvls_solve(const double (&M)[NB][NB], const double (&y)[NB], double (&x)[NB],
                                  const double (&lo)[NB], const double (&hi)[NB]) {
    double scale = 0.0, ymax = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double cs = 0.0;
#pragma unroll
        for (int i = 0; i < NB; ++i) cs += fabs(M[i][j]);
        scale = fmax(scale, cs);
        ymax = fmax(ymax, fabs(y[j]));
    }
    const double tol = 10.0 * NB * 2.220446049250313e-16 * scale * ymax;
    unsigned F = 0u, U = 0u;                 // free set; at the upper bound (others: lower)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const bool hl = lo[j] > -INFINITY, hh = hi[j] < INFINITY;
        if (!hl && !hh) {
            F |= 1u << j;
            x[j] = 0.0;
        } else if (hl) {
            x[j] = lo[j];
        } else {
            x[j] = hi[j];
            U |= 1u << j;
        }
    }
    // s = argmin over the free set F with the others held at x
    auto solve_free = [&](unsigned P, double (&s)[NB]) {
        double rb[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double v = y[i];
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (!((P >> j) & 1u)) v = fma(-M[i][j], x[j], v);
            rb[i] = v;
        }
        ls_free<NB>(M, rb, P, s);
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (!((P >> j) & 1u)) s[j] = x[j];
    };
    double w[NB];                            // M^T (y - M x), the negative gradient
    auto gradient = [&]() {
        double r[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double v = y[i];
#pragma unroll
            for (int j = 0; j < NB; ++j) v = fma(-M[i][j], x[j], v);
            r[i] = v;
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) v = fma(M[i][j], r[i], v);
            w[j] = v;
        }
    };
    if (F) {
        double s[NB];
        solve_free(F, s);
#pragma unroll
        for (int j = 0; j < NB; ++j) x[j] = s[j];
    }
    gradient();
    unsigned rejected = 0u;
    int iter = 0;
    for (int outer = 0; outer < 4 * NB; ++outer) {
        int t = -1;
        double wt = tol;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (((F | rejected) >> j) & 1u) continue;
            const double cand = ((U >> j) & 1u) ? -w[j] : w[j];
            if (cand > wt) {
                wt = cand;
                t = j;
            }
        }
        if (t < 0) return true;              // Kuhn-Tucker conditions hold
        const unsigned F0 = F, U0 = U;
        const bool from_upper = (U >> t) & 1u;
        F |= 1u << t;
        U &= ~(1u << t);
        double s[NB];
        solve_free(F, s);
        bool moves = false;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j == t) moves = from_upper ? s[j] < x[j] : s[j] > x[j];
        if (!moves) {                        // t would not move into the box
            F = F0;
            U = U0;
            rejected |= 1u << t;
            continue;
        }
        for (;;) {
            bool feasible = true;
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (((F >> j) & 1u) && !(s[j] > lo[j] && s[j] < hi[j])) feasible = false;
            if (feasible) {
#pragma unroll
                for (int j = 0; j < NB; ++j) x[j] = s[j];
                break;
            }
            if (++iter > 3 * NB) return false;
            // move from x towards s until the first free coefficient reaches a bound
            double alpha = 2.0;
            int jmin = -1;
            bool jup = false;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (!((F >> j) & 1u)) continue;
                if (!(s[j] > lo[j])) {
                    const double d = x[j] - s[j];
                    const double a = d > 0.0 ? (x[j] - lo[j]) / d : 0.0;
                    if (a < alpha) {
                        alpha = a;
                        jmin = j;
                        jup = false;
                    }
                } else if (!(s[j] < hi[j])) {
                    const double d = s[j] - x[j];
                    const double a = d > 0.0 ? (hi[j] - x[j]) / d : 0.0;
                    if (a < alpha) {
                        alpha = a;
                        jmin = j;
                        jup = true;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (!((F >> j) & 1u)) continue;
                x[j] = fma(alpha, s[j] - x[j], x[j]);
                const bool at_lo = (j == jmin && !jup) || !(x[j] > lo[j]);
                const bool at_hi = (j == jmin && jup) || !(x[j] < hi[j]);
                if (at_lo) {
                    F &= ~(1u << j);
                    x[j] = lo[j];
                } else if (at_hi) {
                    F &= ~(1u << j);
                    U |= 1u << j;
                    x[j] = hi[j];
                }
            }
            if (F == 0u) break;
            solve_free(F, s);
        }
        rejected = 0u;
        gradient();
    }
    return false;
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
