// Per-item statistics and the FP64 epilogue of rr_zscan.
//
// zscan_stats:  f'Wf and sum(w) of each item's target (one wave per item),
//               exact FP64 sums; f'Wf is the leading term of chi2.
// zscan_finish: one thread per (item, z): reads that redshift's normal
//               equations [M | y] from the workspace, adds the prior, solves
//               in FP64 and writes chi2, coefficients, [OII] penalty and
//               status (zscan.py:711-780, :783-800).
//   F64 mode : LU with partial pivoting, as np.linalg.solve (gesv); an exact
//              zero pivot -> HUGE_CHI2 and zero coefficients (:769-771).
//   mixed    : Cholesky (M is a Gram matrix), LU when Cholesky meets a
//              non-positive pivot.
//   NNLS     : (the NMF / NNLS templates, either precision) min |Mc - y|^2
//              over c >= 0 (scipy.optimize.nnls(M, y), zscan.py:1153-1172),
//              or over c_j >= 0 for the coefficients of nonneg_mask only
//              (the reference's BVLS archetype bounds, zscan.py:289-294).
//   BVLS     : min |Mc - y|^2 over lo <= c <= hi for any bounds (scipy
//              lsq_linear(M, y, bounds, 'bvls'), zscan.py:1175-1201).
// chi2 = f'Wf - 2 c'y + c'Mc with the unregularised M (compensated sum): the
// reference's explicit residual sum(w (f - A c)^2) (:777-780) expanded, which
// stays exact with a prior and is stationary in c.
#include "common.hpp"
#include "solve.hpp"
#include "zscan_args.hpp"

namespace {

constexpr double HUGE_CHI2 = 9999999.0;   // zscan.py:18

__global__ __launch_bounds__(64) void stats_kernel(rr::ZArgs a) {
    const int item = blockIdx.x;
    const int lane = threadIdx.x;
    const int t = a.item_target[item];
    const long long npp = a.npix_pad;
    const double *W = a.w + t * npp;
    const double *WF = a.wf + t * npp;
    const double *F = a.f + t * npp;
    double s1 = 0.0, s2 = 0.0;
    for (int p = lane; p < a.npix_pad; p += 64) {
        s1 = fma(WF[p], F[p], s1);
        s2 += W[p];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        double *st = static_cast<double *>(a.work);
        st[2 * item] = s1;
        st[2 * item + 1] = s2;
    }
}

template <int NB, bool MIXED, int SOLVE>
__global__ __launch_bounds__(64) void finish_kernel(rr::ZArgs a, long long ne_off) {
    const int item = blockIdx.y;
    const int iz = blockIdx.x * 64 + threadIdx.x;
    if (item >= a.nitems) return;
    const int nzi = a.item_nz[item];
    if (iz >= nzi) return;
    const long long row = a.item_out[item] + iz;
    const long long R = a.total_rows;
    const double *st = static_cast<const double *>(a.work);
    const double fwf = st[2 * item];
    const double sumw = st[2 * item + 1];
    const char *wbase = static_cast<const char *>(a.work) + ne_off;

    // element (i, j), i <= j, of the unregularised [M | y] in FP64
    double ms = 1.0, ys = 1.0;
    if constexpr (MIXED) {
        const int t = a.item_target[item];
        const double rscale = a.scales[2 * t], fscale = a.scales[2 * t + 1];
        // cache tiles hold tscale*T, R tiles sqrt(w)*rscale*R, f columns
        // sqrt(w)*f*fscale:  G_bb = (tscale rscale)^2 M, G_bf = tscale rscale fscale y
        ms = 1.0 / (a.tscale * a.tscale * rscale * rscale);
        ys = 1.0 / (a.tscale * rscale * fscale);
    }
    auto m_at = [&](int i, int j) -> double {
        if constexpr (MIXED)
            return reinterpret_cast<const double *>(wbase)[rr::ne_index(i, j, NB, true) * R + row] * ms;
        else
            return reinterpret_cast<const double *>(wbase)[rr::ne_index(i, j, NB, false) * R + row];
    };
    double M[NB][NB], y[NB], c[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
        for (int j = i; j < NB; ++j) {
            const double v = m_at(i, j);
            M[i][j] = v;
            M[j][i] = v;
        }
        if constexpr (MIXED) {
            const double *ne = reinterpret_cast<const double *>(wbase);
            y[i] = (ne[rr::ne_index(i, NB, NB, true) * R + row] +
                    ne[rr::ne_index(i, NB + 1, NB, true) * R + row]) * ys;
        } else {
            y[i] = reinterpret_cast<const double *>(wbase)[rr::ne_index(i, NB, NB, false) * R + row];
        }
    }
    if (a.prior) {
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) M[i][j] += a.prior[i * NB + j];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) c[i] = y[i];
    bool ok;
    if constexpr (SOLVE == RR_SOLVE_NNLS) {
        ok = rr::nnls_solve<NB>(M, y, c, (unsigned)a.nonneg_mask);
    } else if constexpr (SOLVE == RR_SOLVE_BVLS) {
        double lo[NB], hi[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            lo[j] = a.bounds[j];
            hi[j] = a.bounds[NB + j];
        }
        ok = rr::bvls_solve<NB>(M, y, c, lo, hi);
    } else if constexpr (MIXED) {
        double L[NB][NB];
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) L[i][j] = M[i][j];
        ok = rr::chol_solve<NB>(L, c);
        if (!ok) {
#pragma unroll
            for (int i = 0; i < NB; ++i) c[i] = y[i];
            ok = rr::lu_solve<NB>(M, c);
        }
    } else {
        ok = rr::lu_solve<NB>(M, c);
    }
    unsigned char status = RR_STATUS_OK;
    double chi2;
    if (sumw == 0.0) {
        chi2 = HUGE_CHI2;
#pragma unroll
        for (int b = 0; b < NB; ++b) c[b] = 0.0;
    } else if (!ok) {
        chi2 = HUGE_CHI2;
        status = RR_STATUS_SINGULAR;
#pragma unroll
        for (int b = 0; b < NB; ++b) c[b] = 0.0;
    } else {
        // t = M c with the unregularised M re-read from the workspace
        double tv[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NB; ++j) s = fma(i <= j ? m_at(i, j) : m_at(j, i), c[j], s);
            tv[i] = s;
        }
        chi2 = rr::compensated_chi2<NB>(fwf, c, y, tv);
    }
    a.zchi2[row] = chi2;
#pragma unroll
    for (int b = 0; b < NB; ++b) a.zcoeff[row * NB + b] = c[b];
    if (a.penalty) {
        double oii = 0.0;
        if (a.oiisum) {
#pragma unroll
            for (int b = 0; b < NB; ++b) oii = fma(c[b], a.oiisum[b], oii);
        }
        a.penalty[row] = (oii < 0.0) ? -oii : 0.0;
    }
    if (a.status) a.status[row] = status;
}

template <int NB>
int launch_finish(const rr::ZArgs &a, hipStream_t s) {
    const bool mixed = a.precision != RR_PREC_F64;
    const long long ne_off = rr::align256((long long)a.nitems * 2 * 8);
    dim3 grid((a.max_item_nz + 63) / 64, a.nitems);
    // the bounded solves are their own instantiations: their active-set loops
    // would raise the register use (and spill) of the plain solves
    if (a.solve == RR_SOLVE_NNLS) {
        if (mixed)
            hipLaunchKernelGGL((finish_kernel<NB, true, RR_SOLVE_NNLS>), grid, dim3(64), 0, s, a, ne_off);
        else
            hipLaunchKernelGGL((finish_kernel<NB, false, RR_SOLVE_NNLS>), grid, dim3(64), 0, s, a, ne_off);
    } else if (a.solve == RR_SOLVE_BVLS) {
        if (mixed)
            hipLaunchKernelGGL((finish_kernel<NB, true, RR_SOLVE_BVLS>), grid, dim3(64), 0, s, a, ne_off);
        else
            hipLaunchKernelGGL((finish_kernel<NB, false, RR_SOLVE_BVLS>), grid, dim3(64), 0, s, a, ne_off);
    } else if (mixed) {
        hipLaunchKernelGGL((finish_kernel<NB, true, RR_SOLVE_PCA>), grid, dim3(64), 0, s, a, ne_off);
    } else {
        hipLaunchKernelGGL((finish_kernel<NB, false, RR_SOLVE_PCA>), grid, dim3(64), 0, s, a, ne_off);
    }
    return rr::check_launch("rr_zscan(finish)");
}

}  // namespace

namespace rr {

long long workspace_bytes(int nitems, int nb, bool mixed, long long total_rows) {
    return align256((long long)nitems * 2 * 8) +
           align256((long long)ne_count(nb, mixed) * total_rows * 8);
}

int zscan_stats(const ZArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(stats_kernel, dim3(a.nitems), dim3(64), 0, s, a);
    return check_launch("rr_zscan(stats)");
}

int zscan_finish(const ZArgs &a, hipStream_t s) {
    switch (a.nb) {
        case 1: return launch_finish<1>(a, s);
        case 2: return launch_finish<2>(a, s);
        case 3: return launch_finish<3>(a, s);
        case 4: return launch_finish<4>(a, s);
        case 5: return launch_finish<5>(a, s);
        case 6: return launch_finish<6>(a, s);
        case 7: return launch_finish<7>(a, s);
        case 8: return launch_finish<8>(a, s);
        case 9: return launch_finish<9>(a, s);
        case 10: return launch_finish<10>(a, s);
        case 11: return launch_finish<11>(a, s);
        case 12: return launch_finish<12>(a, s);
        default: return fail(RR_EUNSUPPORTED, "rr_zscan(finish): nb=%d unsupported", a.nb);
    }
}

}  // namespace rr
