/*
 * rng.c -- oracle replica of the project's counter-based noise stream.
 * Test infrastructure only (see cts_oracle.h).
 *
 * The reference draws quantum-excitation noise and new ion positions from
 * std::mt19937 seeded by std::random_device (src/Bunch.cpp:989-991,
 * src/LatticeInterActionPoint.cpp:990-995), a stream no run can reproduce.
 * The project replaces it with Philox4x32-10 (Salmon et al., "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11) and float Box-Muller
 * transforms whose log, sin and cos are written with +,-,*,/, fmaf and
 * sqrt only, so the HIP path and this replica produce identical bits:
 * a series form for the new ions (orc_box_muller) and, for the
 * quantum-excitation normals, a form that reads ln and sincos from two
 * 256-entry tables (orc_qe_normals).
 */
#include "cts_oracle.h"

#include <math.h>
#include <string.h>

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float bits_to_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t float_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ln(u) for u in (0, 1]: u = m * 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^9. */
static float cts_logf(float u)
{
    uint32_t b = float_to_bits(u);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = bits_to_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float p = 1.0f + s2 * (0.333333333f + s2 * (0.2f + s2 * (0.142857143f + s2 * 0.111111111f)));
    return (float)e * 0.693147181f + 2.0f * s * p;
}

/* sin and cos of 2*pi*u2 for u2 = k*2^-24 in [0, 1): quadrant and octant
 * reduction are exact, Taylor series on [0, pi/4]. */
static void cts_sincos2pif(float u2, float *s_out, float *c_out)
{
    const float t = u2 * 4.0f;
    const int q = (int)t;
    float f = t - (float)q;
    int swap = 0;
    if (f > 0.5f) { f = 1.0f - f; swap = 1; }
    const float th = f * 1.57079633f;
    const float t2 = th * th;
    const float s = th * (1.0f - t2 * (0.166666667f - t2 * (0.00833333333f - t2 * 0.000198412698f)));
    const float c = 1.0f - t2 * (0.5f - t2 * (0.0416666667f - t2 * (0.00138888889f - t2 * 2.48015873e-5f)));
    float sn = swap ? c : s, cs = swap ? s : c;
    float rs, rc;
    switch (q & 3) {
    case 0: rs = sn; rc = cs; break;
    case 1: rs = cs; rc = -sn; break;
    case 2: rs = -sn; rc = -cs; break;
    default: rs = -cs; rc = sn; break;
    }
    *s_out = rs;
    *c_out = rc;
}

/* Box-Muller pair from two 24-bit integers: u1 = (a+1)*2^-24 in (0, 1],
 * u2 = b*2^-24 in [0, 1). */
static void box_muller24(uint32_t a, uint32_t b, float *n0, float *n1)
{
    const float u1 = (float)(a + 1u) * 0x1p-24f;
    const float u2 = (float)b * 0x1p-24f;
    const float r = sqrtf(-2.0f * cts_logf(u1));
    float s, c;
    cts_sincos2pif(u2, &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

void orc_box_muller(uint32_t a, uint32_t b, float *n0, float *n1)
{
    box_muller24(a >> 8, b >> 8, n0, n1);
}

/* Tables of the quantum-excitation transform, as the HIP library builds
 * them: sc[j] = {sin, cos}(2 pi j/256), lg[j] = {ln(1 + j/256),
 * 256/(256 + j)}, float roundings of the double values. */
static float g_sc[256][2], g_lg[256][2];
static int g_tab_ready;

static void normal_tables(void)
{
    if (g_tab_ready) return;
    for (int j = 0; j < 256; j++) {
        const double a = 6.283185307179586 * (double)j / 256.0;
        g_sc[j][0] = (float)sin(a);
        g_sc[j][1] = (float)cos(a);
        g_lg[j][0] = (float)log1p((double)j / 256.0);
        g_lg[j][1] = (float)(256.0 / (256.0 + (double)j));
    }
    g_tab_ready = 1;
}

void orc_normal_tables(float *out)
{
    normal_tables();
    for (int j = 0; j < 256; j++) {
        out[2 * j] = g_sc[j][0];
        out[2 * j + 1] = g_sc[j][1];
        out[512 + 2 * j] = g_lg[j][0];
        out[512 + 2 * j + 1] = g_lg[j][1];
    }
}

/* -2 ln(u1), u1 = (a + 1) 2^-21: k = a + 1 = 2^e m, m in [1, 2); ln m =
 * ln m_j + ln(1 + t), m_j = 1 + j/256 (top 8 mantissa bits), t = (m -
 * m_j)/m_j to third order; for u1 > 1 - 2^-8, -ln(1 - y) with y = 1 - u1
 * to fourth order. */
static float neg2ln_u21(uint32_t a)
{
    const uint32_t k = a + 1u;
    const uint32_t bits = float_to_bits((float)k);
    const int e = (int)(bits >> 23) - 127;
    const float *L = g_lg[(bits >> 15) & 0xffu];
    const float t = (float)(bits & 0x7fffu) * 0x1p-23f * L[1];
    const float q = fmaf(fmaf(t, 0x1.555556p-2f, -0.5f), t, 1.0f);
    float lnu = fmaf((float)(e - 21), 0x1.62e43p-1f, L[0] + t * q);
    if (k > 0x1fe000u) {
        const float y = (float)(0x200000u - k) * 0x1p-21f;
        const float r = fmaf(fmaf(fmaf(y, 0.25f, 0x1.555556p-2f), y, 0.5f), y, 1.0f);
        lnu = -(y * r);
    }
    return -2.0f * lnu;
}

/* sin and cos of 2 pi b 2^-21: table angle 2 pi (b >> 13)/256 plus
 * delta < 2 pi/256, sin to delta^3 and cos to delta^4. */
static void sincos2pi_u21(uint32_t b, float *s, float *c)
{
    const float d = (float)(b & 0x1fffu) * 0x1.921fb6p-19f;
    const float d2 = d * d;
    const float sd = d * fmaf(d2, -0x1.555556p-3f, 1.0f);
    const float cd = fmaf(d2, fmaf(d2, 0x1.555556p-5f, -0.5f), 1.0f);
    const float *T = g_sc[b >> 13];
    *s = fmaf(T[0], cd, T[1] * sd);
    *c = fmaf(T[1], cd, -(T[0] * sd));
}

static void box_muller_u21(uint32_t a, uint32_t b, float *n0, float *n1)
{
    const float r = sqrtf(neg2ln_u21(a));
    float s, c;
    sincos2pi_u21(b, &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

void orc_box_muller_u21(uint32_t a, uint32_t b, float *n0, float *n1)
{
    normal_tables();
    box_muller_u21(a, b, n0, n1);
}

/* One Philox block {index, turn, 0} cut into six 21-bit uniforms (bits
 * 0-20, 21-41, 42-62 of each 64-bit half), three Box-Muller pairs; the
 * sixth normal is unused. */
void orc_qe_normals(uint64_t seed, uint32_t turn, uint64_t index, double out[5])
{
    normal_tables();
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), turn, 0u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint64_t lo = ((uint64_t)o[1] << 32) | o[0], hi = ((uint64_t)o[3] << 32) | o[2];
    const uint32_t m21 = 0x1fffffu;
    float n0, n1, n2, n3, n4, n5;
    box_muller_u21((uint32_t)lo & m21, (uint32_t)(lo >> 21) & m21, &n0, &n1);
    box_muller_u21((uint32_t)(lo >> 42) & m21, (uint32_t)hi & m21, &n2, &n3);
    box_muller_u21((uint32_t)(hi >> 21) & m21, (uint32_t)(hi >> 42) & m21, &n4, &n5);
    out[0] = n0; out[1] = n1; out[2] = n2; out[3] = n3; out[4] = n4;
}

void orc_qe_fill(uint64_t seed, uint32_t turn, uint64_t base, int64_t n, double *noise)
{
    double r[5];
    for (int64_t i = 0; i < n; i++) {
        orc_qe_normals(seed, turn, base + (uint64_t)i, r);
        for (int j = 0; j < 5; j++) noise[j * n + i] = r[j];
    }
}
