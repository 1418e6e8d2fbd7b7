// C entry points onto the reference's Faddeeva::w / erfcx / w_im
// (include/Faddeeva.h:43-60), so the oracle tests can call the reference
// build in oracle/_ref through ctypes.  Test infrastructure only.
#include <complex>
#include "Faddeeva.h"

extern "C" void ref_faddeeva_w(double re, double im, double *wr, double *wi)
{
    std::complex<double> w = Faddeeva::w(std::complex<double>(re, im), 0.);
    *wr = w.real();
    *wi = w.imag();
}

extern "C" double ref_erfcx(double x) { return Faddeeva::erfcx(x); }
extern "C" double ref_w_im(double x) { return Faddeeva::w_im(x); }
