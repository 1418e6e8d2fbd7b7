// cts_growth.cpp -- coupled-bunch-mode growth rates from turn-by-turn bunch
// centroids (MPBeam::GetCBMGR, src/MPBeam.cpp:1496-1891, "ideal" modes,
// with FittingGSL::FitALinear, src/FittingGSL.cpp:231-243).  Host code.
#include <cmath>
#include <complex>
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

}  // namespace

extern "C" {

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
    if (!m || nsamples < 2 || nb < 1 || !harm_num || harmonics < 1 || !tunes || !twiss || interval < 1 ||
        !rates || !(t0 > 0))
        return CTS_ERR_ARG;
    const int i0 = fit_start / interval, i1 = fit_end / interval;
    const int dim = i1 - i0;
    if (i0 < 0 || dim < 2 || i1 > nsamples) return CTS_ERR_ARG;
    const double kPi = 3.141592653589793;
    const double ax = twiss[0], bx = twiss[1], ay = twiss[2], by = twiss[3], bz = twiss[4];
    const std::complex<double> li(0.0, 1.0);
    // |mode amplitude| [sample][plane][mode]
    std::vector<double> amp((size_t)nsamples * 3 * nb);
    std::vector<std::complex<double>> zb((size_t)3 * nb);
    for (int n = 0; n < nsamples; n++) {
        for (int b = 0; b < nb; b++) {
            const cts_moments &q = m[(size_t)n * nb + b];
            const double x = q.mean[0], px = q.mean[1], y = q.mean[2], py = q.mean[3], z = q.mean[4], pz = q.mean[5];
            // rotation from bunch b to the observation point (src/MPBeam.cpp:1532-1538)
            const double phx = -2.0 * kPi * tunes[0] * harm_num[b] / harmonics;
            const double phy = -2.0 * kPi * tunes[1] * harm_num[b] / harmonics;
            const double phz = -2.0 * kPi * tunes[2] * harm_num[b] / harmonics;
            zb[b] = (x / std::sqrt(bx) - li * (std::sqrt(bx) * px + ax / std::sqrt(bx) * x)) * std::exp(li * phx);
            zb[nb + b] = (y / std::sqrt(by) - li * (std::sqrt(by) * py + ay / std::sqrt(by) * y)) * std::exp(li * phy);
            zb[2 * nb + b] = (z / std::sqrt(bz) + li * (std::sqrt(bz) * pz)) * std::exp(li * phz);
        }
        // forward DFT over the bunches, e^{-2 pi i j k / nb}
        for (int pl = 0; pl < 3; pl++)
            for (int k = 0; k < nb; k++) {
                std::complex<double> s(0.0, 0.0);
                for (int j = 0; j < nb; j++) {
                    const double a = -2.0 * kPi * (double)((int64_t)j * k % nb) / nb;
                    s += zb[(size_t)pl * nb + j] * std::complex<double>(std::cos(a), std::sin(a));
                }
                amp[((size_t)n * 3 + pl) * nb + k] = std::abs(s);
            }
    }
    if (amps) for (size_t i = 0; i < amp.size(); i++) amps[i] = amp[i];
    // slope of log|amplitude| against turn, unit weights, per mode
    std::vector<double> fx(dim), fw(dim, 1.0), fy(dim);
    for (int n = 0; n < dim; n++) fx[n] = (double)n * interval;
    for (int pl = 0; pl < 3; pl++)
        for (int k = 0; k < nb; k++) {
            for (int n = 0; n < dim; n++) fy[n] = std::log(amp[((size_t)(n + i0) * 3 + pl) * nb + k]);
            double c0, c1;
            fit_wlinear(fx.data(), fw.data(), fy.data(), dim, c0, c1);
            rates[(size_t)pl * nb + k] = c1 / t0;
        }
    return CTS_OK;
}

}  // extern "C"
