/*
 * fit.c -- weighted least-squares line, as FittingGSL::FitALinear calls it
 * (src/FittingGSL.cpp:231-243 -> gsl_fit_wlinear).  Test infrastructure
 * only (see cts_oracle.h).
 *
 * gsl_fit_wlinear (GSL 2.8 fit/linear.c) forms running weighted means of
 * x and y, then running weighted means of dx^2 and dx*dy about them;
 * slope = <dx dy>/<dx^2>, intercept = <y> - <x>*slope.  Points with
 * non-positive weight are skipped.
 */
#include "cts_oracle.h"

void orc_fit_wlinear(const double *x, const double *w, const double *y, int n,
                     double *c0, double *c1)
{
    double W = 0, wm_x = 0, wm_y = 0, wm_dx2 = 0, wm_dxdy = 0;
    for (int i = 0; i < n; i++) {
        const double wi = w[i];
        if (wi > 0) {
            W += wi;
            wm_x += (x[i] - wm_x) * (wi / W);
            wm_y += (y[i] - wm_y) * (wi / W);
        }
    }
    W = 0;
    for (int i = 0; i < n; i++) {
        const double wi = w[i];
        if (wi > 0) {
            const double dx = x[i] - wm_x;
            const double dy = y[i] - wm_y;
            W += wi;
            wm_dx2 += (dx * dx - wm_dx2) * (wi / W);
            wm_dxdy += (dx * dy - wm_dxdy) * (wi / W);
        }
    }
    const double b = wm_dxdy / wm_dx2;
    *c1 = b;
    *c0 = wm_y - wm_x * b;
}
