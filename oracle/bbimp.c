/* bbimp.c -- CPU oracle (test infrastructure only) of the broadband
 * impedance in the frequency domain, longitudinal, dipolar and quadrupolar
 * planes: MPBunch::BBImpBunchInteraction (src/MPBunch.cpp:2143-2312) with
 * the grid of BoardBandImp::ReadInImp (src/BoardBandImp.cpp:71-90) and the
 * smoothing of GetSmoothedBunchProfileGassionFilter
 * (src/MPBunch.cpp:2348-2410).
 * FFTW's r2c and c2r transforms are written as direct sums (c2r of odd
 * length n: y_j = Re X_0 + 2 sum_{k=1}^{(n-1)/2} Re(X_k e^{2 pi i j k/n})),
 * and the GSL 2.8 Gaussian filter (gsl_filter_gaussian, the dependency the
 * reference links, not part of it) from its published definition: window
 * K = 51, kernel exp(-alpha^2 x^2/2) on x = (i - 25)/25 scaled to unit sum,
 * the input padded with its end values.  Pinned in tests/test_bbimp.py by a
 * constant resistive impedance (V = -R I), a linear reactive one (the
 * spectral derivative of a cosine), a constant profile (smoothing changes
 * nothing), a single spike (it returns the normalised kernel) and a
 * quadrupolar Z_Q = -i R (V_Q = +R I, and the px kick), a dipolar
 * Z_D = -i R (V_D = +R I_x from the smoothed x-weighted profile, and the
 * px kick), and a bunch at one common offset x0, whose dipolar kick equals
 * its quadrupolar kick for the same impedance table.
 *
 * The time-domain method, MPBunch::BBImpBunchInteractionTD
 * (src/MPBunch.cpp:2029-2142), is orc_bbi_td_kick, and the wake table's
 * preparation, MPBeam::GetQuasiWakePoten (src/MPBeam.cpp:748-812), is
 * orc_bbi_td_wake.  Pinned in tests/test_bbimp.py by a wake of one row
 * (V = -w_0 rho, the smoothed profile itself), a wake of one row at lag d
 * (V_i = -w_d rho_{i+d}: the sum runs over the bins ahead, j >= i), a
 * constant wake (V_0 = -w Q, the bunch's total charge, and V decreasing
 * to the tail), a single particle (rho is the normalised filter kernel
 * times its charge, centred on bin (int)((z + L/2)/dz)), a bunch at one
 * common offset x0 (the dipolar kick equals the quadrupolar kick of the
 * same table, and the mean x of every occupied bin is x0), the grid size
 * (int)(L/dz) and the fold of a causal wake (a shift) and of a symmetric
 * one (doubled behind z = 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cts_oracle.h"

#define PI 3.14159265358979323846
#define TAPS 51

void orc_bbi_smooth(double *p, int n)
{
    const int half = (TAPS - 1) / 2;
    const double alpha = 10.0;
    double k[TAPS], sum = 0;
    for (int i = 0; i < TAPS; i++) {
        const double xi = ((double)i - 0.5 * (TAPS - 1.0)) / (0.5 * (TAPS - 1.0));
        k[i] = exp(-0.5 * alpha * alpha * xi * xi);
        sum += k[i];
    }
    for (int i = 0; i < TAPS; i++) k[i] *= 1.0 / sum;
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < TAPS; j++) {
            int q = i + j - half;
            q = q < 0 ? 0 : (q >= n ? n - 1 : q);
            s += p[q] * k[j];
        }
        y[i] = s;
    }
    double s0 = 0, s1 = 0;
    for (int i = 0; i < n; i++) {
        s0 += p[i];
        s1 += y[i];
        p[i] = y[i];
    }
    s0 = s0 / s1;
    for (int i = 0; i < n; i++) p[i] *= s0;
    free(y);
}

/* X_k = sum_j p_j e^{-2 pi i j k / n}, k < nimp (FFTW r2c) */
static void spectrum(const double *p, int nimp, double *xr, double *xi)
{
    const int n = 2 * nimp - 1;
    for (int k = 0; k < nimp; k++) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            const double ph = -2 * PI * (double)(((int64_t)j * k) % n) / n;
            sr += p[j] * cos(ph);
            si += p[j] * sin(ph);
        }
        xr[k] = sr;
        xi[k] = si;
    }
}

/* out_j = s c2r(X Z)_j / n, with X Z times i when rot (the quadrupolar planes) */
static void product_c2r(const double *xr, const double *xi, int nimp, const double *zre, const double *zim, int rot,
                        double s, double *out)
{
    const int n = 2 * nimp - 1;
    double *yr = (double *)malloc(sizeof(double) * nimp), *yi = (double *)malloc(sizeof(double) * nimp);
    for (int k = 0; k < nimp; k++) {
        const double ar = xr[k] * zre[k] - xi[k] * zim[k], ai = xr[k] * zim[k] + xi[k] * zre[k];
        yr[k] = rot ? -ai : ar;
        yi[k] = rot ? ar : ai;
    }
    for (int j = 0; j < n; j++) {
        double t = 0;
        for (int k = 1; k < nimp; k++) {
            const double ph = 2 * PI * (double)(((int64_t)j * k) % n) / n;
            t += yr[k] * cos(ph) - yi[k] * sin(ph);
        }
        out[j] = s * (yr[0] + 2 * t) / n;
    }
    free(yr);
    free(yi);
}

/* wake[j] = -c2r(r2c(p) Z)[j] / n, n = 2 nimp - 1 */
void orc_bbi_potential(const double *p, int nimp, const double *zre, const double *zim, double *wake)
{
    double *xr = (double *)malloc(sizeof(double) * nimp), *xi = (double *)malloc(sizeof(double) * nimp);
    spectrum(p, nimp, xr, xi);
    product_c2r(xr, xi, nimp, zre, zim, 0, -1.0, wake);
    free(xr);
    free(xi);
}

/* One bunch's interaction: survivors binned on the grid
 * (int)((z + z_max + dz/2)/dz) of n = 2 nimp - 1 bins; three profiles, each
 * a running sum per particle as the reference forms them: the charge
 * profile (src/MPBunch.cpp:2166-2171) and, for the dipolar planes, the
 * x- and y-weighted ones (2256-2261, 2290-2295); each smoothed over the
 * bucket on its own.  The induced voltages, in the reference's flag order:
 * longitudinal V = -c2r(X Z)/n (planes[0]); dipolar V_Dx = c2r(X_x Z_Dx i)/n
 * (planes[1]), V_Dy likewise (planes[2]); quadrupolar V_Q = c2r(X Z_Q i)/n
 * (planes[3], planes[4]).  Kicks in the reference's order: delta +=
 * V/(E beta^2); px += V_Qx/(E beta^2) x/beta_x, py likewise; px +=
 * V_Dx/(E beta^2)/beta_x, py likewise.  zre, zim [5][nimp] and wake [5][n]
 * in flag order; prof [3][n] (charge, x-weighted, y-weighted). */
void orc_bbi_kick(orc_bunch *b, int nimp, double f_max, const int *planes, const double *zre, const double *zim,
                  double bx, double by, double macro_charge, double rbeta, double t0, int harmonic, double energy,
                  double *wake, double *prof)
{
    const int n = 2 * nimp - 1;
    const double dt = 1.0 / (2 * f_max), dz = dt * ORC_CLIGHT * rbeta;
    const double zmin = -dz * (nimp - 1);
    const double qc = macro_charge * ORC_ECHARGE / dz * ORC_CLIGHT;
    int *idx = (int *)malloc(sizeof(int) * (size_t)(b->n > 0 ? b->n : 1));
    for (int i = 0; i < 3 * n; i++) prof[i] = 0;
    for (int64_t i = 0; i < b->n; i++) {
        idx[i] = -1;
        if (b->lost[i] != 0) continue;
        const double q = (b->z[i] - zmin + dz / 2) / dz;
        if (!(q > -1.0 && q < (double)n)) continue;
        idx[i] = (int)q;
        prof[idx[i]] += qc;
        prof[n + idx[i]] += qc * b->x[i];
        prof[2 * n + idx[i]] += qc * b->y[i];
    }
    const double rf_len = t0 / harmonic * ORC_CLIGHT;
    const int is = (int)((-rf_len / 2. - zmin + dz / 2) / dz);
    const int ie = (int)((rf_len / 2. - zmin + dz / 2) / dz);
    for (int r = 0; r < 3; r++) orc_bbi_smooth(prof + (size_t)r * n + is, ie - is);
    double *xr = (double *)malloc(sizeof(double) * nimp), *xi = (double *)malloc(sizeof(double) * nimp);
    for (int q = 0; q < 5; q++)
        for (int j = 0; j < n; j++) wake[(size_t)q * n + j] = 0;
    /* profile of each plane: charge for Z and the quadrupolar planes */
    const int src[5] = {0, 1, 2, 0, 0};
    for (int r = 0; r < 3; r++) {
        int used = 0;
        for (int q = 0; q < 5; q++) used |= planes[q] && src[q] == r;
        if (!used) continue;
        spectrum(prof + (size_t)r * n, nimp, xr, xi);
        for (int q = 0; q < 5; q++)
            if (planes[q] && src[q] == r)
                product_c2r(xr, xi, nimp, zre + (size_t)q * nimp, zim + (size_t)q * nimp, q > 0, q == 0 ? -1.0 : 1.0,
                            wake + (size_t)q * n);
    }
    const double eb = energy * pow(rbeta, 2);
    for (int64_t i = 0; i < b->n; i++) {
        const int k = idx[i];
        if (k < 0) continue;
        if (planes[0]) b->dp[i] += wake[k] / eb;
        if (planes[3]) b->px[i] += wake[3 * n + k] / eb * b->x[i] / bx;
        if (planes[4]) b->py[i] += wake[4 * n + k] / eb * b->y[i] / by;
        if (planes[1]) b->px[i] += wake[n + k] / eb / bx;
        if (planes[2]) b->py[i] += wake[2 * n + k] / eb / by;
    }
    free(xr);
    free(xi);
    free(idx);
}

/* MPBeam::GetQuasiWakePoten (src/MPBeam.cpp:748-812) after the file is
 * read: rows [n][6] = the file's columns z, wz, wDx, wDy, wQx, wQy; out
 * [6][n] = binPosZ (z moved to start at 0) and the five wakes times -1, the
 * longitudinal one then folded about the row whose z is exactly 0 (cen; 0
 * when no row is): w[0] = t[cen], w[i] = t[cen + i] + t[cen - i] for
 * 0 < i < cen, w[i] = t[cen + i] up to n - cen, zeros after.  Returns cen,
 * or -1 when 2 cen > n (the reference's fold reads past the table). */
int orc_bbi_td_wake(int n, const double *rows, double *out)
{
    if (n < 1) return -1;
    int cen = 0;
    for (int i = 0; i < n; i++) {
        out[i] = rows[6 * (size_t)i];
        for (int c = 1; c < 6; c++) out[(size_t)c * n + i] = -1 * rows[6 * (size_t)i + c];
        if (rows[6 * (size_t)i] == 0) cen = i;
    }
    if (2 * cen > n) return -1; /* the fold would read past the table's end */
    const double z0 = out[0];
    for (int i = 0; i < n; i++) out[i] -= z0;
    double *wz = out + n;
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    if (!t) return -1;
    for (int i = 0; i < n; i++) t[i] = wz[i];
    wz[0] = t[cen];
    for (int i = 1; i < cen; i++) wz[i] = t[cen + i] + t[cen - i];
    for (int i = cen; i < n - cen; i++) wz[i] = t[cen + i];
    for (int i = n - cen; i < n; i++) wz[i] = 0.0;
    free(t);
    return cen;
}

/* bins of the time-domain grid: (int)(L/dz), L = T0/h c */
int orc_bbi_td_bins(double t0, int harmonic, double dz)
{
    const double rf_len = t0 / harmonic * ORC_CLIGHT;
    return (int)(1 * rf_len / dz);
}

/* One bunch's time-domain interaction.  The survivors are binned over one
 * RF bucket, (int)((z + L/2)/dz) with dz = |pos[1] - pos[0]| (particles
 * off the grid are skipped), the charge profile rho (a running sum of
 * macro_charge e per particle) is smoothed over the whole grid, <x>_j and
 * <y>_j are the bin's sums over its particles divided by their count; then
 * for each bin i, over j = i, i + 1, ... while j - i < nw:
 * V_z -= w_z[j - i] rho_j, V_Dx -= w_Dx[j - i] rho_j <x>_j, V_Dy likewise,
 * V_Qx -= w_Qx[j - i] rho_j, V_Qy likewise, and the particles of bin i get
 * delta += V_z/E/beta^2, px += V_Dx/E/beta^2/beta_x,
 * py += V_Dy/E/beta^2/beta_y, px += V_Qx/E/beta^2/beta_x x,
 * py += V_Qy/E/beta^2/beta_y y, in that order.  w [5][nw] in flag order;
 * V [5][n] and prof [3][n] (rho, <x>, <y>) for n = orc_bbi_td_bins(). */
void orc_bbi_td_kick(orc_bunch *b, int nw, const double *pos, const double *w, const int *planes, double bx,
                     double by, double macro_charge, double t0, int harmonic, double rbeta, double energy, double *V,
                     double *prof)
{
    const double rf_len = t0 / harmonic * ORC_CLIGHT;
    const double dz = fabs(pos[1] - pos[0]);
    const double zmin = -1 * rf_len / 2;
    const int n = (int)(1 * rf_len / dz);
    int *idx = (int *)malloc(sizeof(int) * (size_t)(b->n > 0 ? b->n : 1));
    int *cnt = (int *)calloc((size_t)n, sizeof(int));
    double *rho = prof, *ax = prof + n, *ay = prof + 2 * (size_t)n;
    for (int i = 0; i < 3 * n; i++) prof[i] = 0;
    for (int64_t i = 0; i < b->n; i++) {
        idx[i] = -1;
        if (b->lost[i] != 0) continue;
        const double q = (b->z[i] - zmin) / dz;
        if (!(q > -1.0 && q < (double)n)) continue;
        idx[i] = (int)q;
        rho[idx[i]] += macro_charge * ORC_ECHARGE;
        cnt[idx[i]]++;
    }
    orc_bbi_smooth(rho, n);
    for (int64_t i = 0; i < b->n; i++) {
        if (idx[i] < 0) continue;
        ax[idx[i]] += b->x[i];
        ay[idx[i]] += b->y[i];
    }
    for (int j = 0; j < n; j++) {
        if (cnt[j] == 0) continue;
        ax[j] /= cnt[j];
        ay[j] /= cnt[j];
    }
    const double *wz = w, *wdx = w + nw, *wdy = w + 2 * (size_t)nw, *wqx = w + 3 * (size_t)nw,
                 *wqy = w + 4 * (size_t)nw;
    for (int i = 0; i < 5 * n; i++) V[i] = 0;
    for (int i = 0; i < n; i++) {
        double t[5] = {0, 0, 0, 0, 0};
        for (int j = i; j < n && j - i < nw; j++) {
            const int tij = j - i;
            if (planes[0]) t[0] -= wz[tij] * rho[j];
            if (planes[1]) t[1] -= wdx[tij] * rho[j] * ax[j];
            if (planes[2]) t[2] -= wdy[tij] * rho[j] * ay[j];
            if (planes[3]) t[3] -= wqx[tij] * rho[j];
            if (planes[4]) t[4] -= wqy[tij] * rho[j];
        }
        for (int q = 0; q < 5; q++) V[(size_t)q * n + i] = t[q];
    }
    for (int64_t i = 0; i < b->n; i++) {
        const int k = idx[i];
        if (k < 0) continue;
        if (planes[0]) b->dp[i] += V[k] / energy / pow(rbeta, 2);
        if (planes[1]) b->px[i] += V[n + k] / energy / pow(rbeta, 2) / bx;
        if (planes[2]) b->py[i] += V[2 * (size_t)n + k] / energy / pow(rbeta, 2) / by;
        if (planes[3]) b->px[i] += V[3 * (size_t)n + k] / energy / pow(rbeta, 2) / bx * b->x[i];
        if (planes[4]) b->py[i] += V[4 * (size_t)n + k] / energy / pow(rbeta, 2) / by * b->y[i];
    }
    free(cnt);
    free(idx);
}
