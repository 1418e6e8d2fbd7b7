// cts_growth.cpp -- coupled-bunch-mode growth rates from turn-by-turn bunch
// centroids (MPBeam::GetCBMGR, src/MPBeam.cpp:1496-1891: ideal modes, IQ
// decomposition, Hilbert analytic signal (MPBeam::GetAnalyticalWithFilter /
// GetHilbertAnalytical, src/MPBeam.cpp:1892-2011), with
// FittingGSL::FitALinear, src/FittingGSL.cpp:231-243).  Host code.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/cetasim_hip.h"

namespace {

// Weighted least-squares line y = c0 + c1 x (gsl_fit_wlinear's formulas:
// weighted means, then the weighted second moments about them).
void fit_wlinear(const double *x, const double *w, const double *y, int n, double &c0, double &c1)
{
    double W = 0, wm_x = 0, wm_y = 0;
    for (int i = 0; i < n; i++) {
        if (w[i] > 0) {
            W += w[i];
            wm_x += (x[i] - wm_x) * (w[i] / W);
            wm_y += (y[i] - wm_y) * (w[i] / W);
        }
    }
    W = 0;
    double wm_dx2 = 0, wm_dxdy = 0;
    for (int i = 0; i < n; i++) {
        if (w[i] > 0) {
            const double dx = x[i] - wm_x, dy = y[i] - wm_y;
            W += w[i];
            wm_dx2 += (dx * dx - wm_dx2) * (w[i] / W);
            wm_dxdy += (dx * dy - wm_dxdy) * (w[i] / W);
        }
    }
    c1 = wm_dxdy / wm_dx2;
    c0 = wm_y - wm_x * c1;
}

typedef std::complex<double> cplx;
constexpr double kPi = 3.141592653589793;

// Discrete Fourier transform of a[0..n) in place, sign -1 forward
// (e^{-2 pi i j k / n}), +1 backward without the 1/n: mixed radix
// (Cooley-Tukey on the smallest prime factor), direct sums for prime n.
void fft(cplx *a, int n, int sign, std::vector<cplx> &work)
{
    if (n <= 1) return;
    int p = 2;
    while (p * p <= n && n % p) p++;
    if (n % p) p = n;
    const int m = n / p;
    const size_t base = work.size();
    work.resize(base + (size_t)n);
    cplx *t = work.data() + base;
    if (p == n) {
        for (int k = 0; k < n; k++) {
            cplx acc(0.0, 0.0);
            for (int j = 0; j < n; j++) {
                const double ang = sign * 2.0 * kPi * (double)((int64_t)j * k % n) / n;
                acc += a[j] * cplx(std::cos(ang), std::sin(ang));
            }
            t[k] = acc;
        }
        for (int k = 0; k < n; k++) a[k] = t[k];
        work.resize(base);
        return;
    }
    // p interleaved subsequences of length m, each transformed
    for (int r = 0; r < p; r++)
        for (int q = 0; q < m; q++) t[(size_t)r * m + q] = a[(size_t)q * p + r];
    for (int r = 0; r < p; r++) {
        fft(work.data() + base + (size_t)r * m, m, sign, work);
        t = work.data() + base;  // the recursion may have reallocated
    }
    for (int k = 0; k < m; k++)
        for (int q = 0; q < p; q++) {
            const int kk = k + q * m;
            cplx acc(0.0, 0.0);
            for (int r = 0; r < p; r++) {
                const double ang = sign * 2.0 * kPi * (double)((int64_t)r * kk % n) / n;
                acc += t[(size_t)r * m + k] * cplx(std::cos(ang), std::sin(ang));
            }
            a[kk] = acc;
        }
    work.resize(base);
}

// Top-level transform: the scratch holds the recursion's copies (at most
// 2n values, reserved up front so pointers into it stay valid).
void fft_top(cplx *a, int n, int sign, std::vector<cplx> &work)
{
    work.clear();
    work.reserve(2 * (size_t)n + 16);
    fft(a, n, sign, work);
}

// GetCBMGR's normalized phasor of one bunch in plane pl (0 x, 1 y, 2 z),
// rotated to the observation point by -2 pi nu h_b / h.
cplx phasor(int pl, double q, double pq, const double *tunes, const double *twiss, int hb, int harmonics)
{
    const cplx li(0.0, 1.0);
    const double ph = -2.0 * kPi * tunes[pl] * hb / harmonics;
    if (pl < 2) {
        const double al = twiss[2 * pl], be = twiss[2 * pl + 1];
        return (q / std::sqrt(be) - li * (std::sqrt(be) * pq + al / std::sqrt(be) * q)) * std::exp(li * ph);
    }
    const double bz = twiss[4];  // alpha_z = 0 (src/MPBeam.cpp:1502)
    return (q / std::sqrt(bz) + li * (std::sqrt(bz) * pq + 0.0 / std::sqrt(bz) * q)) * std::exp(li * ph);
}

// slope of log(v) against n * interval over samples [i0, i0 + dim), / t0
double fit_rate(const std::vector<double> &v, int i0, int dim, int interval, double t0)
{
    std::vector<double> fx(dim), fw(dim, 1.0), fy(dim);
    for (int n = 0; n < dim; n++) {
        fx[n] = (double)n * interval;
        fy[n] = std::log(v[n + i0]);
    }
    double c0, c1;
    fit_wlinear(fx.data(), fw.data(), fy.data(), dim, c0, c1);
    return c1 / t0;
}

}  // namespace

extern "C" {

int cts_cbm_analysis(const cts_moments *m, int32_t nsamples, int32_t nb, const int32_t *harm_num, int32_t harmonics,
                     const double *tunes, const double *twiss, int32_t interval, int32_t fit_start, int32_t fit_end,
                     double t0, const cts_cbm_result *out)
{
    if (!m || nsamples < 1 || nb < 1 || !harm_num || harmonics < 1 || !tunes || !twiss || interval < 1 || !out ||
        !(t0 > 0))
        return CTS_ERR_ARG;
    const int i0 = fit_start / interval, i1 = fit_end / interval;
    const int dim = i1 - i0;
    const bool fit = out->rate_ideal || out->rate_iq || out->rate_hil || out->rate_bunch;
    if (fit && (i0 < 0 || dim < 2 || i1 > nsamples)) return CTS_ERR_ARG;
    const size_t NS = (size_t)nsamples, NB = (size_t)nb;
    auto at = [&](size_t n, int pl, size_t b) { return (n * 3 + (size_t)pl) * NB + b; };
    std::vector<double> ideal_amp(NS * 3 * NB), iq_amp(NS * 3 * NB), hil_amp(NS * 3 * NB);
    std::vector<cplx> ana(NS * 3 * NB);
    std::vector<cplx> buf(std::max(NS, NB));
    std::vector<cplx> work;
    // (1) ideal modes: phasors of (centroid, momentum), DFT over the bunches
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++) {
            for (size_t b = 0; b < NB; b++) {
                const cts_moments &q = m[n * NB + b];
                buf[b] = phasor(pl, q.mean[2 * pl], q.mean[2 * pl + 1], tunes, twiss, harm_num[b], harmonics);
            }
            fft_top(buf.data(), nb, -1, work);
            for (size_t k = 0; k < NB; k++) {
                const double re = buf[k].real(), im = buf[k].imag();
                ideal_amp[at(n, pl, k)] = std::sqrt(std::pow(re, 2) + std::pow(im, 2));
                if (out->ideal_amp) out->ideal_amp[at(n, pl, k)] = ideal_amp[at(n, pl, k)];
                if (out->ideal_arg) out->ideal_arg[at(n, pl, k)] = std::atan2(re, im);  // (real, imag) as printed
            }
        }
    // (2) IQ decomposition of the centroids at f_j = (j + nu) f0, bunch i at
    // time h_i t0 / h
    const double f0 = 1 / t0, trf = t0 / harmonics;
    const cplx li(0.0, 1.0);
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++)
            for (size_t j = 0; j < NB; j++) {
                const double fj = ((double)j + tunes[pl]) * f0;
                cplx acc(0.0, 0.0);
                for (size_t i = 0; i < NB; i++) {
                    const double time = harm_num[i] * trf;
                    acc += m[n * NB + i].mean[2 * pl] * std::exp(-li * 2.0 * kPi * fj * time);
                }
                acc /= (double)nb;
                iq_amp[at(n, pl, j)] = std::abs(acc);
                if (out->iq_amp) out->iq_amp[at(n, pl, j)] = iq_amp[at(n, pl, j)];
                if (out->iq_arg) out->iq_arg[at(n, pl, j)] = std::arg(acc);
            }
    // (3) analytic signal of every bunch's centroid history (one-sided
    // spectrum, DC kept, inverse with 1/N)
    std::vector<cplx> sig(NS);
    for (int pl = 0; pl < 3; pl++)
        for (size_t b = 0; b < NB; b++) {
            for (size_t n = 0; n < NS; n++) sig[n] = cplx(m[n * NB + b].mean[2 * pl], 0.0);
            fft_top(sig.data(), nsamples, -1, work);
            for (size_t i = 1; i <= NS / 2; i++) sig[i] *= 2.0;
            if (NS % 2 == 0) sig[NS / 2] /= 2.0;
            for (size_t i = NS / 2 + 1; i < NS; i++) sig[i] = cplx(0.0, 0.0);
            fft_top(sig.data(), nsamples, +1, work);
            for (size_t n = 0; n < NS; n++) {
                ana[at(n, pl, b)] = sig[n] / (double)nsamples;
                if (out->ana_re) out->ana_re[at(n, pl, b)] = ana[at(n, pl, b)].real();
                if (out->ana_im) out->ana_im[at(n, pl, b)] = ana[at(n, pl, b)].imag();
            }
        }
    // (4) modes of the analytic signal: x = Re, px = -Im
    for (size_t n = 0; n < NS; n++)
        for (int pl = 0; pl < 3; pl++) {
            for (size_t b = 0; b < NB; b++) {
                const cplx a = ana[at(n, pl, b)];
                buf[b] = phasor(pl, a.real(), -a.imag(), tunes, twiss, harm_num[b], harmonics);
            }
            fft_top(buf.data(), nb, -1, work);
            for (size_t k = 0; k < NB; k++) {
                const double re = buf[k].real(), im = buf[k].imag();
                hil_amp[at(n, pl, k)] = std::sqrt(std::pow(re, 2) + std::pow(im, 2));
                if (out->hil_amp) out->hil_amp[at(n, pl, k)] = hil_amp[at(n, pl, k)];
                if (out->hil_arg) out->hil_arg[at(n, pl, k)] = std::atan2(re, im);
            }
        }
    // (5) growth rates: slope of log amplitude against turn
    if (fit) {
        std::vector<double> v(NS);
        for (int pl = 0; pl < 3; pl++)
            for (size_t k = 0; k < NB; k++) {
                if (out->rate_ideal) {
                    for (size_t n = 0; n < NS; n++) v[n] = ideal_amp[at(n, pl, k)];
                    out->rate_ideal[(size_t)pl * NB + k] = fit_rate(v, i0, dim, interval, t0);
                }
                if (out->rate_iq) {
                    for (size_t n = 0; n < NS; n++) v[n] = iq_amp[at(n, pl, k)];
                    out->rate_iq[(size_t)pl * NB + k] = fit_rate(v, i0, dim, interval, t0);
                }
                if (out->rate_hil) {
                    for (size_t n = 0; n < NS; n++) v[n] = hil_amp[at(n, pl, k)];
                    out->rate_hil[(size_t)pl * NB + k] = fit_rate(v, i0, dim, interval, t0);
                }
                if (out->rate_bunch) {
                    for (size_t n = 0; n < NS; n++) v[n] = std::abs(ana[at(n, pl, k)]);
                    out->rate_bunch[(size_t)pl * NB + k] = fit_rate(v, i0, dim, interval, t0);
                }
            }
    }
    return CTS_OK;
}


int cts_fit_wlinear(const double *x, const double *w, const double *y, int32_t n, double *c0, double *c1)
{
    if (!x || !w || !y || n < 2 || !c0 || !c1) return CTS_ERR_ARG;
    fit_wlinear(x, w, y, n, *c0, *c1);
    return CTS_OK;
}

int cts_cbm_growth_from_moments(const cts_moments *m, int32_t nsamples, int32_t nb, const int32_t *harm_num,
                                int32_t harmonics, const double *tunes, const double *twiss, int32_t interval,
                                int32_t fit_start, int32_t fit_end, double t0, double *rates, double *amps)
{
    if (!rates || nsamples < 2) return CTS_ERR_ARG;
    cts_cbm_result r;
    std::memset(&r, 0, sizeof r);
    r.ideal_amp = amps;
    r.rate_ideal = rates;
    return cts_cbm_analysis(m, nsamples, nb, harm_num, harmonics, tunes, twiss, interval, fit_start, fit_end, t0, &r);
}

}  // extern "C"
