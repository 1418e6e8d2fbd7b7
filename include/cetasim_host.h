/*
 * cetasim_host.h -- host-side set-up helpers of the C ABI (same library as
 * cetasim_hip.h).  They compute, from plain numbers, the structures the
 * tracking entry points take, as the reference's initialisation does.
 */
#ifndef CETASIM_HOST_H
#define CETASIM_HOST_H

#include <stdint.h>

#include "cetasim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Twiss values at a section start (twiss file columns,
 * src/LatticeInterActionPoint.cpp:225-243); beta_z = natural bunch length /
 * natural energy spread (243). */
typedef struct {
    double beta_x, alpha_x, eta_x, etap_x;
    double beta_y, alpha_y, eta_y, etap_y;
    double beta_z;
} cts_twiss;

/* Replaces LatticeInterActionPoint::InitialLatticeSympMat for one section
 * (src/LatticeInterActionPoint.cpp:326-431): B1H1, InvH1InvB1, InvH2InvB2
 * from the twiss at this section (t1) and the next (t2), plus the phase
 * advance of the section (264-280). */
int cts_build_section(const cts_twiss *t1, const cts_twiss *t2, double phase_adv_x,
                      double phase_adv_y, cts_section *out);

/* input.dat quantities the turn depends on (src/ReadInputSettings.cpp
 * keys ringcircring, ringelectronbeamenergy, ringu0, ringsdelta0,
 * ringalphac, ringworkq, ringchrom, ringadtx/y, ringskewquadk,
 * ringsynchraddamptime, ringnaturalemit, ringcoupling, rfringharm,
 * rfresvol, ringpipeaperaturex/y). */
typedef struct {
    double circ, energy, u0, sdelta0;
    double alphac[3];
    double nux, nuy;
    double chrom[2];
    double adtx[5], adty[5];
    double skew_k;
    double synrad_tau[3];
    double natural_emit, coupling;
    int32_t harmonic;
    double rf_volt;            /* single ideal cavity, rfMode 0 */
    double ap_x, ap_y;
} cts_ring_input;

typedef struct {
    double eta, workqz, natural_bunch_length, beta_z;
} cts_ring_derived;

/* Replaces the derived-quantity block of ReadInputSettings::ParamRead
 * (src/ReadInputSettings.cpp:986-1077) and, for one cavity,
 * CavityResonator::Initial's phasor (src/CavityResonator.cpp:121-132). */
int cts_derive_ring(const cts_ring_input *in, cts_ring *ring, cts_ring_derived *derived);

#ifdef __cplusplus
}
#endif
#endif
