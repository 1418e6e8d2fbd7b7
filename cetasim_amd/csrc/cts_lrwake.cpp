// cts_lrwake.cpp -- host set-up of the long-range (coupled-bunch) wake:
// the wake table of WakeFunction::InitialLRWake (src/WakeFunction.cpp:49-205)
// from the resistive wall of the pipe sectors (GetRWLRWakeFun,
// src/WakeFunction.cpp:499-555, with the Yokoya factors of elliptic
// sectors, src/WakeFunction.cpp:292-350 and 723-880) and from broadband
// resonators (GetBBRWakeFun, src/WakeFunction.cpp:655-720).
//
// The Yokoya factors sum the reference's 50 x 50 double series; its GSL
// special functions are evaluated here as Gamma(m + 1/2) / Gamma(1/2) =
// prod_{k<m} (k + 1/2), the power series of 2F1(1/2, b; c; x) for the
// arguments x = exp(-4 mu) < 1 the series takes, and the complete elliptic
// integral K by the arithmetic-geometric mean.
#include <cmath>
#include <string>
#include <vector>

#include "cts_internal.hpp"

namespace cts {
namespace {

constexpr double kPiL = 3.14159265358979323846;
// include/Global.h:37-39: Z0 = 120 pi, FactorGaussSI = Z0 c / (4 pi)
constexpr double kGaussSI = 120.0 * kPiL * 299792458.0 / (4.0 * kPiL);

double gamma_half_ratio(int m)
{
    double r = 1.0;
    for (int k = 0; k < m; k++) r *= k + 0.5;
    return r;
}

double fact(int m)
{
    double r = 1.0;
    for (int k = 2; k <= m; k++) r *= (double)k;
    return r;
}

// 2F1(1/2, b; c; x), 0 <= x < 1
double f21_half(double b, double c, double x)
{
    double sum = 1.0, term = 1.0;
    for (int k = 0; k < 2000000; k++) {
        term *= (k + 0.5) * (b + k) / ((c + k) * (k + 1.0)) * x;
        sum += term;
        if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
    }
    return sum;
}

double ellk(double k)
{
    double a = 1.0, g = std::sqrt(1.0 - k * k);
    for (int i = 0; i < 64 && std::fabs(a - g) > 1e-16 * a; i++) {
        const double m = 0.5 * (a + g);
        g = std::sqrt(a * g);
        a = m;
    }
    return kPiL / (2.0 * a);
}

// sqrt(2) pi e^{-(2 m + 1 + extra) mu} Gamma(m + 1/2 + e)/Gamma(1/2)/(m + e)! 2F1(...)
// with e = 0 (the L functions' terms) or e = 1 (Ldfunction2)
double lterm(int m, int e, double mu)
{
    return std::sqrt(2.) * kPiL * std::exp(-(2 * m + 1 + 2 * e) * mu) * gamma_half_ratio(m + e) / fact(m + e) *
           f21_half(m + e + 0.5, m + e + 1, std::exp(-4 * mu));
}

// GetyokoyaFactor(a, b), a > b: longitudinal, dipole x, dipole y, quadrupole x, y
void yokoya(double a, double b, double out[5])
{
    const double f = std::sqrt(std::fabs(a * a - b * b));
    const double mu = std::acosh(a / f);
    const double c0 = 2 * std::sqrt(2) / kPiL * b / f, c1 = std::sqrt(2) / kPiL * std::pow(b / f, 3);
    const double coef[5] = {c0, c1, c1, -c1, c1};
    double s[5] = {0, 0, 0, 0, 0};
    for (int L = 0; L < 50; L++)
        for (int P = 0; P < 50; P++) {
            const int d = std::abs(P - L);
            const double l1 = P != L ? lterm(d, 0, mu) : 2 * std::sqrt(2) * std::exp(-mu) * ellk(std::sqrt(std::exp(-4 * mu)));
            const double lf = l1 + lterm(P + L, 0, mu);
            const double ld1 = lterm(d, 0, mu), ld2 = lterm(P + L, 1, mu);
            const double sg = ((P + L) & 1) ? -1.0 : 1.0;
            const double kp = P == 0 ? 2 : 1, kl = L == 0 ? 2 : 1;
            const double ch0 = std::cosh(2 * P * mu) * std::cosh(2 * L * mu);
            const double q = (2 * P + 1) * (2 * L + 1);
            s[0] += lf * sg * 1. / ch0 / kp / kl;
            s[1] += (ld1 + ld2) * sg * 1. / (std::cosh((2 * P + 1) * mu) * std::cosh((2 * L + 1) * mu)) * q;
            s[2] += (ld1 - ld2) * sg * 1. / (std::sinh((2 * P + 1) * mu) * std::sinh((2 * L + 1) * mu)) * q;
            s[3] += lf * sg * 1. / ch0 * std::pow(2 * P, 2) / kl;
            s[4] += lf * sg * 1. / ch0 * std::pow(2 * P, 2) / kl;
        }
    for (int i = 0; i < 5; i++) out[i] = coef[i] * s[i];
}

}  // namespace

void lrwake_yokoya(double rx, double ry, double out[5])
{
    for (int i = 0; i < 5; i++) out[i] = 1.0;
    if (rx > ry) {
        yokoya(rx, ry, out);
    } else if (rx < ry) {
        // the factors of the rotated sector, planes exchanged
        yokoya(ry, rx, out);
        std::swap(out[1], out[2]);
        std::swap(out[3], out[4]);
    }
}

int lrwake_table(const cts_lrwake_config &c, double t0, int harmonic, LrwTable &tab, std::string &err)
{
    const double trf = t0 / harmonic;
    const double dt = trf / 2;
    const int np = (int)(c.nturns * harmonic * trf / dt) + 2;
    if (np < 2 || np > (1 << 26)) {
        err = "cts_set_lrwake: wake table of " + std::to_string(np) + " points";
        return CTS_ERR_ARG;
    }
    // per sector: the tau-independent factors of GetRWLRWakeFun
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = 0; i < c.n_sectors; i++) {
        const double *s = c.sectors + 7 * i;
        double yk[5];
        lrwake_yokoya(s[0], s[1], yk);
        const double radius = s[0] >= s[1] ? s[1] : s[0];
        const double rs = std::sqrt(1.E0 / (s[6] * kGaussSI));
        ax += 1. / std::pow(radius, 3) * rs * s[2] * s[5] * yk[1] * s[3];
        ay += 1. / std::pow(radius, 3) * rs * s[2] * s[5] * yk[2] * s[4];
        az += 1. / radius * rs * s[2] * s[5] * yk[0];
    }
    tab.np = np;
    tab.dt = dt;
    tab.trf = trf;
    tab.time.assign(np, 0.0);
    tab.w.assign(3 * (size_t)np, 0.0);
    tab.rw.assign(3 * (size_t)np, 0.0);
    tab.bbr.assign(3 * (size_t)np, 0.0);
    for (int i = 0; i < np; i++) {
        const double tau = -i * dt;
        double rw[3] = {0, 0, 0}, bb[3] = {0, 0, 0};
        if (c.n_sectors > 0 && tau != 0) {
            rw[0] = ax * (-2.) / kPiL / std::pow(std::fabs(tau), 0.5) * kGaussSI / c.beta_x;
            rw[1] = ay * (-2.) / kPiL / std::pow(std::fabs(tau), 0.5) * kGaussSI / c.beta_y;
            rw[2] = az / (2.) / kPiL / kClight / std::pow(std::fabs(tau), 1.5) * kGaussSI;
        }
        for (int r = 0; r < c.n_res; r++) {
            const double *q = c.res + 9 * r;
            for (int pl = 0; pl < 2; pl++) {
                const double Rs = q[3 + 3 * pl], Q = q[4 + 3 * pl], om = q[5 + 3 * pl] * 2 * kPiL;
                const double alpha = om / 2.0 / Q;
                const double omb = std::sqrt(om * om - alpha * alpha);
                bb[pl] += kClight * Rs * om / Q / omb * std::exp(alpha * tau) * std::sin(omb * tau) /
                          (pl == 0 ? c.beta_x : c.beta_y);
            }
        }
        for (int r = 0; r < c.n_res; r++) {
            const double *q = c.res + 9 * r;
            const double om = q[2] * 2 * kPiL;
            const double alpha = om / 2.0 / q[1];
            const double omb = std::sqrt(om * om - alpha * alpha);
            const double e = 2 * alpha * q[0] * std::exp(alpha * tau);
            bb[2] += tau == 0 ? e / 2 : e * (std::cos(omb * tau) + std::sin(omb * tau) * alpha / omb);
        }
        tab.time[i] = -tau;
        for (int k = 0; k < 3; k++) {
            tab.rw[(size_t)k * np + i] = rw[k];
            tab.bbr[(size_t)k * np + i] = bb[k];
            tab.w[(size_t)k * np + i] = bb[k] + rw[k];
        }
    }
    return CTS_OK;
}

}  // namespace cts
