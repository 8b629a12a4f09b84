// sfb_ylm.hpp -- Legendre x azimuth rows for one (point, l), FP64, O(1) per m.
//
// The reference evaluates gsl_sf_legendre_sphPlm(l, m, cos t) for every m
// (sfb_raw.cpp:72-74), an O(l-m) recurrence each, i.e. O(l^2) per point and l.
// Here the rows are generated in m from m = l down to 0 with the three-term
// recurrence of the orthonormal functions at fixed l,
//   sqrt((l-m)(l+m+1)) P^{m+1} + 2m cot(t) P^m + sqrt((l+m)(l-m+1)) P^{m-1} = 0,
// started from the sectoral value P_l^l = c_l sin^l t and P_l^{l+1} = 0.
// Downward in m the wanted solution dominates in the evanescent region
// (m > l sin t) and is oscillatory below it, so the recurrence is stable.
// sin^l t underflows FP64 near the poles, so the state carries a binary
// exponent E (value = v * 2^E) and is renormalised when |v| exceeds 2^600
// (2^300 in the branch-free form that runs every eight steps).
// sin t is clamped to >= 2^-80: the rows of such points differ from the
// pole values by O(l^2 2^-80).
#pragma once

#include "sfb_common.hpp"

namespace sfb {

struct YlmDown {
  double v, vp;  // P^m, P^{m+1} without the 2^E factor
  int E;
  double cm, sm;  // cos(m phi), sin(m phi)
  double cp, sp;  // cos((m+1) phi), sin((m+1) phi)
  double cot, cph, sph;

  __device__ __forceinline__ void init(int l, double x, double s, double phi, double cph_, double sph_,
                                       double2 sect) {
    init(l, l, x, s, phi, cph_, sph_, sect);
  }

  // The Legendre pair at m = l and the azimuth at m = m0 <= l: for a caller
  // that first takes the Legendre pair alone down to m0 (step_p).
  __device__ __forceinline__ void init(int l, int m0, double x, double s, double phi, double cph_, double sph_,
                                       double2 sect) {
    const double se = fmax(s, 0x1p-80);
    cot = x / se;
    cph = cph_;
    sph = sph_;
    // se^l by binary powering, kept normalised with frexp
    int e = 0, be, t;
    double m = 1.0;
    double b = frexp(se, &be);
    for (int k = l; k; k >>= 1) {
      if (k & 1) {
        m *= b;
        e += be;
        m = frexp(m, &t);
        e += t;
      }
      b *= b;
      be *= 2;
      b = frexp(b, &t);
      be += t;
    }
    v = m * sect.x;
    E = e + (int)sect.y;
    vp = 0.0;
    // e^{i m0 phi}; m0*phi is split exactly into hi + lo
    const double hi = (double)m0 * phi;
    const double lo = fma((double)m0, phi, -hi);
    double S, C;
    sincos(hi, &S, &C);
    cm = C - lo * S;
    sm = S + lo * C;
    cp = fma(cm, cph, -sm * sph);
    sp = fma(sm, cph, cm * sph);
  }

  // Y*_lm = P_l^m (cos m phi - i sin m phi)
  __device__ __forceinline__ void value(double &re, double &im) const {
    const double P = ldexp(v, E);
    re = P * cm;
    im = -P * sm;
  }

  // m -> m-1 with ab = (a_m, b_m); the azimuth by the Chebyshev recurrence
  // cos((m-1) phi) = 2 cos(phi) cos(m phi) - cos((m+1) phi) (and sin alike)
  __device__ __forceinline__ void step(double2 ab) {
    const double vn = -fma(ab.x * cot, v, ab.y * vp);
    vp = v;
    v = vn;
    const double c2 = cph + cph;
    const double cn = fma(c2, cm, -cp);
    const double sn = fma(c2, sm, -sp);
    cp = cm;
    sp = sm;
    cm = cn;
    sm = sn;
  }

  // m -> m-1 for the Legendre pair alone (the azimuth stays where it is)
  __device__ __forceinline__ void step_p(double2 ab) {
    const double vn = -fma(ab.x * cot, v, ab.y * vp);
    vp = v;
    v = vn;
  }

  // Branch-free renormalisation for callers that renormalise every eight
  // steps: one step grows max(|v|, |vp|) by at most |a_m cot| + |b_m| < 2^85
  // (|cot| <= 2^80, a_m <= sqrt(2m) <= 2^5, b_m <= 1), so from |v| <= 2^300
  // eight steps stay below 2^981.
  __device__ __forceinline__ void rescale_nb() {
    const bool big = fabs(v) > 0x1p+300;
    const double f = big ? 0x1p-300 : 1.0;
    v *= f;
    vp *= f;
    E += big ? 300 : 0;
  }

  __device__ __forceinline__ void rescale() {
    if (fabs(v) > 0x1p+600) {
      v = ldexp(v, -600);
      vp = ldexp(vp, -600);
      E += 600;
    }
  }
};

// Pixel-mode generator: Y* at a HEALPix pixel as the reference's ring FFT path
// sees it (sfb_reverse_fft.cpp:192-207).  For m with (m mod np) > np/2 the
// reference reads the folded bin X[np - s] without conjugating it, which equals
// the pixel sum with azimuth phi' = 2 phi0 - phi instead of phi; with
// alias_compat that azimuth is used for those m, otherwise the pixel centre.
struct YlmDownPix {
  YlmDown y;
  double cm2, sm2, cph2, sph2;  // cos/sin(m phi') and the step rotation
  int np, smod;                 // ring length, m mod np
  bool quirk;

  __device__ __forceinline__ void init(int l, double x, double s, double phi, double cph, double sph,
                                       double phi2, double cph2_, double sph2_, int np_, bool alias,
                                       double2 sect) {
    y.init(l, x, s, phi, cph, sph, sect);
    const double hi = (double)l * phi2;
    const double lo = fma((double)l, phi2, -hi);
    double S, C;
    sincos(hi, &S, &C);
    cm2 = C - lo * S;
    sm2 = S + lo * C;
    cph2 = cph2_;
    sph2 = sph2_;
    np = np_;
    smod = l % np_;
    quirk = alias;
  }
  __device__ __forceinline__ void value(int m, double &re, double &im) const {
    const double P = ldexp(y.v, y.E);
    const bool flip = quirk && m > 0 && 2 * smod > np;
    const double c = flip ? cm2 : y.cm, s = flip ? sm2 : y.sm;
    re = P * c;
    im = -P * s;
  }
  __device__ __forceinline__ void step(double2 ab) {
    y.step(ab);
    const double cn = fma(cm2, cph2, sm2 * sph2);
    const double sn = fma(sm2, cph2, -cm2 * sph2);
    cm2 = cn;
    sm2 = sn;
    smod = smod == 0 ? np - 1 : smod - 1;
  }
  __device__ __forceinline__ void rescale() { y.rescale(); }
};

}  // namespace sfb
