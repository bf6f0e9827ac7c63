/*
 * CPU oracle for the trapezoidal rebin (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of redrock's serial two-pointer sweep:
 *   - _trapz_rebin_1d       py/redrock/rebin.py:126-176
 *   - _trapz_rebin_batch    py/redrock/rebin.py:178-209
 * The arithmetic is written expression-for-expression in the same order as
 * the reference so that, compiled with -ffp-contract=off, every output is the
 * IEEE-754 double the reference's sweep produces.  It is loaded by the tests
 * and by the CPU-baseline leg of the benchmark through ctypes; the product
 * path never calls it.
 */
#include <stddef.h>
#include <string.h>

/* One basis, one redshift: results has nbin entries and is overwritten. */
static void trapz_rebin_1d(const double *x, const double *y, int nx,
                           const double *edges, int nbin, double *results,
                           ptrdiff_t rstride)
{
    int i = 0, j = 0;
    double yedge, area;
    (void)nx;
    for (int k = 0; k < nbin; ++k) results[k * rstride] = 0.0;
    while (i < nbin) {
        /* seek next sample beyond the bin edge (rebin.py:142) */
        while (x[j] <= edges[i]) j += 1;
        /* interpolated y where the sample grid crosses the lower edge (:146) */
        yedge = y[j - 1] + (edges[i] - x[j - 1]) * (y[j] - y[j - 1]) / (x[j] - x[j - 1]);
        if (x[j] < edges[i + 1]) {
            area = 0.5 * (y[j] + yedge) * (x[j] - edges[i]);
            results[i * rstride] += area;
            /* interior samples (:154-157) */
            while (x[j + 1] < edges[i + 1]) {
                j += 1;
                area = 0.5 * (y[j] + y[j - 1]) * (x[j] - x[j - 1]);
                results[i * rstride] += area;
            }
            /* upper edge (:160-162) */
            yedge = y[j] + (edges[i + 1] - x[j]) * (y[j + 1] - y[j]) / (x[j + 1] - x[j]);
            area = 0.5 * (yedge + y[j]) * (edges[i + 1] - x[j]);
            results[i * rstride] += area;
        } else {
            /* one sample interval spans the whole bin (:165-169) */
            double ylo = y[j] + (edges[i] - x[j]) * (y[j] - y[j - 1]) / (x[j] - x[j - 1]);
            double yhi = y[j] + (edges[i + 1] - x[j]) * (y[j] - y[j - 1]) / (x[j] - x[j - 1]);
            area = 0.5 * (ylo + yhi) * (edges[i + 1] - edges[i]);
            results[i * rstride] += area;
        }
        i += 1;
    }
    for (i = 0; i < nbin; ++i) results[i * rstride] /= edges[i + 1] - edges[i];
}

/*
 * Batch form (rebin.py:179-209): out has shape (nz, nbin, nb), C order.
 * work must hold nx doubles (the redshifted x grid).
 * The caller performs the edge-range check (rebin.py:332-341) first; the
 * sweep assumes edges lie inside [x[0](1+zmax), x[-1](1+zmin)].
 */
void oracle_trapz_rebin_batch(const double *x, int nx, const double *y, int nb,
                              const double *edges, int nbin, const double *z,
                              int nz, double *out, double *work)
{
    for (int iz = 0; iz < nz; ++iz) {
        for (int i = 0; i < nx; ++i) work[i] = x[i] * (1. + z[iz]);
        for (int b = 0; b < nb; ++b)
            trapz_rebin_1d(work, y + (size_t)b * nx, nx, edges, nbin,
                           out + (size_t)iz * nbin * nb + b, nb);
    }
}
