/*
 * sfb_oracle.c -- CPU restatement of CatAna's SFB decomposition hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU timing baseline).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function names the reference file:line it restates.  GSL, FFTW3 and
 * HEALPix C++ are not available, so their arithmetic is re-derived here from
 * the mathematics:
 *   - spherical Bessel j_l: Taylor series / upward recurrence / continued
 *     fraction + downward (Miller) recurrence, the same regions GSL's
 *     gsl_sf_bessel_jl uses for l < 400;
 *   - orthonormal associated Legendre with Condon-Shortley phase: sectoral
 *     start + three-term recurrence in l (GSL gsl_sf_legendre_sphPlm);
 *   - HEALPix RING geometry: Gorski et al. 2005, ApJ 622, 759;
 *   - the ring FFT of the HEALPix path: a direct DFT evaluated at the folded
 *     frequency the reference reads out of FFTW's half spectrum.
 */
#include "sfb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* spherical Bessel functions j_l(x), x >= 0                                */
/* ------------------------------------------------------------------------ */

static double sj0(double x) {
  if (fabs(x) < 1e-3) {
    double x2 = x * x;
    return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0));
  }
  return sin(x) / x;
}

static double sj1(double x) {
  if (fabs(x) < 0.25) {
    double x2 = x * x;
    return x / 3.0 * (1.0 - x2 / 10.0 * (1.0 - x2 / 28.0 * (1.0 - x2 / 54.0 * (1.0 - x2 / 88.0))));
  }
  return (sin(x) / x - cos(x)) / x;
}

/* j_l(x) = x^l/(2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1)) */
static double sjl_series(int l, double x) {
  double pre = 1.0;
  for (int k = 1; k <= l; ++k) pre *= x / (2.0 * k + 1.0);
  if (pre == 0.0) return 0.0;
  double h = -0.5 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 1000; ++k) {
    term *= h / (k * (2.0 * l + 2.0 * k + 1.0));
    sum += term;
    if (fabs(term) <= 1e-17 * fabs(sum)) break;
  }
  return pre * sum;
}

/* x < l: R_l = j_l/j_{l-1} from the continued fraction
 *   R_l = 1/((2l+1)/x - 1/((2l+3)/x - ...))    (modified Lentz),
 * then u_{k-1} = (2k+1)/x u_k - u_{k+1} down to k = 0 and normalisation by
 * j_0 or j_1, whichever recursion value is larger. */
static double sjl_downward(int l, double x) {
  const double tiny = 1e-300;
  double f = (2.0 * l + 1.0) / x, C = f, D = 0.0;
  for (int k = 1; k < 200000; ++k) {
    double b = (2.0 * (l + k) + 1.0) / x;
    D = b - D;
    if (D == 0.0) D = tiny;
    D = 1.0 / D;
    C = b - 1.0 / C;
    if (C == 0.0) C = tiny;
    double delta = C * D;
    f *= delta;
    if (fabs(delta - 1.0) < 1e-16) break;
  }
  double R = 1.0 / f;
  double up = R, u = 1.0; /* u_l (scaled) and u_{l-1} */
  int scale = 0;
  for (int k = l - 1; k >= 1; --k) {
    double um = (2.0 * k + 1.0) / x * u - up;
    up = u;
    u = um;
    if (fabs(u) > 0x1p+600) {
      u = ldexp(u, -600);
      up = ldexp(up, -600);
      scale += 600;
    }
  }
  double norm = (fabs(u) >= fabs(up)) ? sj0(x) / u : sj1(x) / up;
  return ldexp(R * norm, -scale);
}

/* gsl_sf_bessel_jl semantics as used at sfb_raw.cpp:51 (error handler off:
 * underflow returns 0). */
double orc_sbessel_jl(int l, double x) {
  if (l < 0 || x < 0.0) return NAN;
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  if (l == 0) return sj0(x);
  if (x * x < 10.0 * (l + 0.5) / 2.718281828459045235) return sjl_series(l, x);
  if (x >= (double)l) {
    double jm = sj0(x), j = sj1(x);
    for (int k = 1; k < l; ++k) {
      double jp = (2.0 * k + 1.0) / x * j - jm;
      jm = j;
      j = jp;
    }
    return j;
  }
  return sjl_downward(l, x);
}

/* ------------------------------------------------------------------------ */
/* orthonormal associated Legendre functions (gsl_sf_legendre_sphPlm)        */
/*   sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m(x), P_l^m with (-1)^m phase      */
/* ------------------------------------------------------------------------ */

double orc_sph_plm(int l, int m, double x) {
  if (m < 0 || m > l || !(x >= -1.0 && x <= 1.0)) return NAN;
  double omx2 = (1.0 - x) * (1.0 + x);
  double s = sqrt(omx2);
  int e = 0;
  double pmm = sqrt((2.0 * m + 1.0) / (4.0 * ORC_PI));
  for (int k = 1; k <= m; ++k) {
    pmm *= sqrt((2.0 * k - 1.0) / (2.0 * k)) * s;
    if (pmm == 0.0) return 0.0;
    if (fabs(pmm) < 0x1p-500) {
      pmm = ldexp(pmm, 500);
      e -= 500;
    }
  }
  if (m & 1) pmm = -pmm;
  if (l == m) return ldexp(pmm, e);
  double p0 = pmm, p1 = x * sqrt(2.0 * m + 3.0) * pmm;
  for (int ll = m + 2; ll <= l; ++ll) {
    double dl = ll, dm = m;
    double a = sqrt((4.0 * dl * dl - 1.0) / (dl * dl - dm * dm));
    double c = sqrt(((dl - 1.0) * (dl - 1.0) - dm * dm) / (4.0 * (dl - 1.0) * (dl - 1.0) - 1.0));
    double p2 = a * (x * p1 - c * p0);
    p0 = p1;
    p1 = p2;
    if (fabs(p1) > 0x1p+500) {
      p1 = ldexp(p1, -500);
      p0 = ldexp(p0, -500);
      e += 500;
    }
  }
  return ldexp(p1, e);
}

/* ------------------------------------------------------------------------ */
/* zeros of j_l                                                             */
/* ------------------------------------------------------------------------ */

/* Newton iteration with the reference's stopping rule
 * (SphericalBesselZeros.cpp:49-61): the loop continues while delta/z > 1e-15,
 * a SIGNED test, so it ends on the first step that moves z upward.  In
 * converged mode the test is on |delta|/z. */
static double newton(int l, double z, int mode) {
  for (int it = 0; it < 200; ++it) {
    double jl = orc_sbessel_jl(l, z);
    double d = jl / ((double)l / z * jl - orc_sbessel_jl(l + 1, z));
    z -= d;
    double t = (mode == ORC_ZEROS_REFERENCE) ? d / z : fabs(d) / z;
    if (!(t > 1e-15)) break;
  }
  return z;
}

/* asymptotic first/second zero guesses, SphericalBesselZeros.cpp:17-46 */
static double zero_guess(int l, int which) {
  if (l == 0) return which == 0 ? 3.141592653589793238463 : 6.283185307179586476925;
  double a = l + 0.5;
  if (which == 0)
    return a + 1.8557571 * pow(a, 1. / 3.) + 1.033150 * pow(a, -1. / 3.) - 0.00397 * pow(a, -1.) -
           0.0908 * pow(a, -5. / 3.) + 0.043 * pow(a, -7. / 3.);
  return a + 3.2446076 * pow(a, 1. / 3.) + 3.158244 * pow(a, -1. / 3.) - 0.08331 * pow(a, -1.) -
         0.8437 * pow(a, -5. / 3.) + 0.864 * pow(a, -7. / 3.);
}

/* SphericalBesselZeros(l)[0..n-1]: two asymptotic zeros refined by Newton,
 * then z_{k+1} = Newton(2 z_k - z_{k-1})   (SphericalBesselZeros.cpp:11-15, 63-68, 95-100) */
int orc_bessel_zeros(int l, int n, int mode, double *z) {
  if (l < 0 || n < 0 || !z) return ORC_EINVAL;
  double z0 = newton(l, zero_guess(l, 0), mode);
  double z1 = newton(l, zero_guess(l, 1), mode);
  if (n > 0) z[0] = z0;
  if (n > 1) z[1] = z1;
  for (int k = 2; k < n; ++k) z[k] = newton(l, 2.0 * z[k - 1] - z[k - 2], mode);
  return ORC_OK;
}

/* compute_kln, return_types.cpp:10-19 */
int orc_kln(int lmax, int nmax, double rmax, int mode, double *k_ln) {
  if (lmax < 0 || nmax < 0 || !(rmax > 0.0) || !k_ln) return ORC_EINVAL;
  for (int l = 0; l < lmax; ++l) {
    int rc = orc_bessel_zeros(l, nmax, mode, k_ln + (size_t)l * nmax);
    if (rc) return rc;
    for (int n = 0; n < nmax; ++n) k_ln[(size_t)l * nmax + n] /= rmax;
  }
  return ORC_OK;
}

/* n_l = #{n < nmax : z_ln < k_max r_max}; k_max <= 0 means no cut */
int orc_count_nl(int lmax, int nmax, double rmax, double k_max, int mode, int *nl) {
  if (lmax < 0 || nmax < 0 || !nl) return ORC_EINVAL;
  double *z = (double *)malloc(sizeof(double) * (nmax > 0 ? nmax : 1));
  if (!z) return ORC_ENOMEM;
  for (int l = 0; l < lmax; ++l) {
    orc_bessel_zeros(l, nmax, mode, z);
    int c = 0;
    if (k_max > 0.0) {
      while (c < nmax && z[c] < k_max * rmax) ++c;
    } else {
      c = nmax;
    }
    nl[l] = c;
  }
  free(z);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* reference linear interpolation table (SBesselInterpolator_uptozero)      */
/* ------------------------------------------------------------------------ */

int orc_interp_table(int l, int nmax, int mode, double *y, int *npts, double *zmax, double *dx_inv) {
  if (nmax < 1) return ORC_EINVAL;
  double *z = (double *)malloc(sizeof(double) * (nmax + 1));
  if (!z) return ORC_ENOMEM;
  orc_bessel_zeros(l, nmax + 1, mode, z);
  int P = 1000 * nmax;
  double zm = z[nmax];
  free(z);
  double dx = zm / (P - 1);
  if (npts) *npts = P;
  if (zmax) *zmax = zm;
  if (dx_inv) *dx_inv = 1.0 / dx;
  if (y) {
    /* FunctionInterpolator.cpp:15 fills the table with an OpenMP loop */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) y[i] = orc_sbessel_jl(l, 0.0 + i * dx);
  }
  return ORC_OK;
}

typedef struct {
  double *y;
  int P;
  double zmax, dx_inv;
} interp_t;

static int interp_eval(const interp_t *t, double z, double *out) {
  if (z >= t->zmax || z < 0.0) return ORC_EDOMAIN; /* SBesselInterpolator.cpp:20-22 */
  double ia = z * t->dx_inv;
  int i = (int)ia;
  if (i > t->P - 2) i = t->P - 2;
  *out = t->y[i] + (t->y[i + 1] - t->y[i]) * (ia - i);
  return ORC_OK;
}

static int interp_build(interp_t *t, int l, int nmax, int mode) {
  t->y = (double *)malloc(sizeof(double) * 1000 * (size_t)nmax);
  if (!t->y) return ORC_ENOMEM;
  return orc_interp_table(l, nmax, mode, t->y, &t->P, &t->zmax, &t->dx_inv);
}

/* ------------------------------------------------------------------------ */
/* direct point-mode SFB, sfb_raw.cpp:13-108                                 */
/* ------------------------------------------------------------------------ */

static int set_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

static double norm_factor(double rmax, double volume, double npts) {
  if (volume > 0.0) return sqrt(2. / ORC_PI) * volume / npts;
  double nf = sqrt(2. / ORC_PI) * (4. / 3.) * ORC_PI; /* sfb_raw.cpp:29-30 */
  nf *= rmax * rmax * rmax / npts;
  return nf;
}

int orc_sfb_raw(const double *pts, long long npts, int lmax, int nmax, double rmax, const int *nl,
                double volume, int interpolated, int zeros_mode, int nthreads, double *k_ln,
                double *c_ln, double *f_lmn) {
  if (!pts || npts <= 0 || lmax < 0 || nmax < 1 || !(rmax > 0.0) || !k_ln || !c_ln) return ORC_EINVAL;
  int rc = orc_kln(lmax, nmax, rmax, zeros_mode, k_ln);
  if (rc) return rc;
  memset(c_ln, 0, sizeof(double) * (size_t)lmax * nmax);
  const double norm = norm_factor(rmax, volume, (double)npts);
  const int nt = set_threads(nthreads);
  size_t foff = 0;
  int status = ORC_OK;
  for (int l = 0; l < lmax; ++l) {
    const int n_l = nl ? (nl[l] < nmax ? nl[l] : nmax) : nmax;
    const size_t blk = (size_t)(l + 1) * nmax; /* complex entries of this l */
    const double *kl = k_ln + (size_t)l * nmax;
    interp_t tab = {0};
    if (interpolated) {
      rc = interp_build(&tab, l, nmax, zeros_mode);
      if (rc) { free(tab.y); return rc; }
    }
    double *priv = (double *)calloc((size_t)nt * blk * 2, sizeof(double));
    if (!priv) { free(tab.y); return ORC_ENOMEM; }
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      double *f = priv + (size_t)tid * blk * 2;
      double *jv = (double *)malloc(sizeof(double) * (n_l > 0 ? n_l : 1));
#pragma omp for schedule(static)
      for (long long i = 0; i < npts; ++i) {
        const double r = pts[3 * i], th = pts[3 * i + 1], ph = pts[3 * i + 2];
        for (int n = 0; n < n_l; ++n) {
          if (interpolated) {
            if (interp_eval(&tab, kl[n] * r, &jv[n])) {
#pragma omp atomic write
              err = 1;
              jv[n] = 0.0;
            }
          } else {
            jv[n] = orc_sbessel_jl(l, kl[n] * r);
          }
        }
        for (int m = 0; m <= l; ++m) {
          double p = orc_sph_plm(l, m, cos(th));
          double yr = p * cos(-m * ph), yi = p * sin(-m * ph);
          double *fm = f + (size_t)m * nmax * 2;
          for (int n = 0; n < n_l; ++n) {
            double kj = kl[n] * jv[n];
            fm[2 * n] += kj * yr;
            fm[2 * n + 1] += kj * yi;
          }
        }
      }
      free(jv);
    }
    free(tab.y);
    if (err) { free(priv); status = ORC_EDOMAIN; return status; }
    /* deterministic merge in thread order (the reference merges in omp critical) */
    double *fl = priv;
    for (int t = 1; t < nt; ++t) {
      const double *ft = priv + (size_t)t * blk * 2;
      for (size_t q = 0; q < blk * 2; ++q) fl[q] += ft[q];
    }
    for (size_t q = 0; q < blk * 2; ++q) fl[q] *= norm;
    for (int n = 0; n < n_l; ++n) {
      double s = 0.0;
      for (int m = -l; m <= l; ++m) {
        const double *v = fl + ((size_t)abs(m) * nmax + n) * 2;
        s += v[0] * v[0] + v[1] * v[1];
      }
      c_ln[(size_t)l * nmax + n] = s / (2 * l + 1);
    }
    if (f_lmn) memcpy(f_lmn + foff * 2, fl, sizeof(double) * blk * 2);
    foff += blk;
    free(priv);
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* HEALPix RING scheme                                                       */
/* ------------------------------------------------------------------------ */

static long long isqrt_ll(long long v) {
  long long r = (long long)sqrt((double)v + 0.5);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

long long orc_ang2pix_ring(int nside, double theta, double phi) {
  const long long ns = nside, npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  const double z = cos(theta), za = fabs(z);
  double tt = fmod(phi * (2.0 / ORC_PI), 4.0);
  if (tt < 0) tt += 4.0;
  if (tt >= 4.0) tt = 0.0;
  if (za <= 2.0 / 3.0) {
    const long long nl4 = 4 * ns;
    double t1 = ns * (0.5 + tt), t2 = ns * z * 0.75;
    long long jp = (long long)(t1 - t2), jm = (long long)(t1 + t2);
    long long ir = ns + 1 + jp - jm;
    long long kshift = 1 - (ir & 1);
    long long ip = (jp + jm - ns + kshift + 1 + nl4 + nl4) / 2;
    ip %= nl4;
    return ncap + (ir - 1) * nl4 + ip;
  }
  /* polar caps; near the poles use sin(theta) for the edge-line distance */
  double tp = tt - (long long)tt;
  double tmp;
  if (theta < 0.01 || theta > 3.14159 - 0.01)
    tmp = ns * sin(theta) / sqrt((1.0 + za) / 3.0);
  else
    tmp = ns * sqrt(3.0 * (1.0 - za));
  long long jp = (long long)(tp * tmp), jm = (long long)((1.0 - tp) * tmp);
  long long ir = jp + jm + 1;
  long long ip = (long long)(tt * ir);
  if (ip >= 4 * ir) ip -= 4 * ir;
  return (z > 0) ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

void orc_ring_info(int nside, int ring, long long *startpix, int *ringpix, int *shifted, double *z,
                   double *phi0) {
  const long long ns = nside, npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  long long sp;
  int rp, sh;
  double zz;
  if (ring < nside) {
    sh = 1;
    rp = 4 * ring;
    sp = 2LL * ring * (ring - 1);
    zz = 1.0 - (double)ring * ring / (3.0 * ns * ns);
  } else if (ring < 3 * nside) {
    sh = ((ring - nside) & 1) == 0;
    rp = 4 * nside;
    sp = ncap + (long long)(ring - nside) * rp;
    zz = (2.0 * ns - ring) * 2.0 / (3.0 * ns);
  } else {
    long long nr = 4 * ns - ring;
    sh = 1;
    rp = (int)(4 * nr);
    sp = npix - 2 * nr * (nr + 1);
    zz = -1.0 + (double)nr * nr / (3.0 * ns * ns);
  }
  if (startpix) *startpix = sp;
  if (ringpix) *ringpix = rp;
  if (shifted) *shifted = sh;
  if (z) *z = zz;
  if (phi0) *phi0 = (sh ? 0.5 : 0.0) * (2.0 * ORC_PI / rp);
}

void orc_pix2ang_ring(int nside, long long pix, double *theta, double *phi) {
  const long long ns = nside, npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  double z, ph;
  if (pix < ncap) {
    long long iring = (1 + isqrt_ll(1 + 2 * pix)) >> 1;
    long long iphi = (pix + 1) - 2 * iring * (iring - 1);
    z = 1.0 - (double)iring * iring / (3.0 * ns * ns);
    ph = (iphi - 0.5) * (ORC_PI / 2.0) / iring;
  } else if (pix < npix - ncap) {
    long long ip = pix - ncap, nl4 = 4 * ns;
    long long tmp = ip / nl4;
    long long iring = tmp + ns, iphi = ip - nl4 * tmp + 1;
    double fodd = ((iring + ns) & 1) ? 1.0 : 0.5;
    z = (2.0 * ns - iring) * 2.0 / (3.0 * ns);
    ph = (iphi - fodd) * ORC_PI / (2.0 * ns);
  } else {
    long long ip = npix - pix;
    long long iring = (1 + isqrt_ll(2 * ip - 1)) >> 1;
    long long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    z = -1.0 + (double)iring * iring / (3.0 * ns * ns);
    ph = (iphi - 0.5) * (ORC_PI / 2.0) / iring;
  }
  if (theta) *theta = acos(z);
  if (phi) *phi = ph;
}

/* PixelizedPointContainer(nside, pc): RING ang2pix per point, radii appended
 * in input order (PixelizedPointContainer.cpp:22-30) */
int orc_pixelize(const double *pts, long long npts, int nside, long long *offsets, double *radii) {
  if (nside < 1 || (nside & (nside - 1)) || !offsets || (npts > 0 && (!pts || !radii))) return ORC_EINVAL;
  const long long npix = 12LL * nside * nside;
  long long *pix = (long long *)malloc(sizeof(long long) * (npts > 0 ? npts : 1));
  if (!pix) return ORC_ENOMEM;
  memset(offsets, 0, sizeof(long long) * (npix + 1));
  for (long long i = 0; i < npts; ++i) {
    pix[i] = orc_ang2pix_ring(nside, pts[3 * i + 1], pts[3 * i + 2]);
    offsets[pix[i] + 1]++;
  }
  for (long long p = 0; p < npix; ++p) offsets[p + 1] += offsets[p];
  long long *fill = (long long *)malloc(sizeof(long long) * npix);
  if (!fill) { free(pix); return ORC_ENOMEM; }
  memcpy(fill, offsets, sizeof(long long) * npix);
  for (long long i = 0; i < npts; ++i) radii[fill[pix[i]]++] = pts[3 * i];
  free(fill);
  free(pix);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* pixelised SFB                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  long long start;
  int npix;
  double z, phi0;
} ring_t;

static double bessel_or_table(int l, double x, int interpolated, const interp_t *t, int *err) {
  if (!interpolated) return orc_sbessel_jl(l, x);
  double v;
  if (interp_eval(t, x, &v)) {
    *err = 1;
    return 0.0;
  }
  return v;
}

/* _sfb_reverse_fft, sfb_reverse_fft.cpp:43-209 */
int orc_sfb_reverse_fft(const long long *offsets, const double *radii, int nside, int lmax, int nmax,
                        double rmax, const int *nl, double volume, int interpolated, int zeros_mode,
                        int alias_compat, int nthreads, double *k_ln, double *c_ln, double *f_lmn) {
  if (!offsets || nside < 1 || (nside & (nside - 1)) || lmax < 0 || nmax < 1 || !(rmax > 0.0) || !k_ln || !c_ln)
    return ORC_EINVAL;
  const long long npixtot = 12LL * nside * nside;
  const long long npts = offsets[npixtot];
  if (npts <= 0) return ORC_EINVAL;
  int rc = orc_kln(lmax, nmax, rmax, zeros_mode, k_ln);
  if (rc) return rc;
  memset(c_ln, 0, sizeof(double) * (size_t)lmax * nmax);
  const double norm = norm_factor(rmax, volume, (double)npts);
  const int nrings = 4 * nside - 1;
  ring_t *rings = (ring_t *)malloc(sizeof(ring_t) * nrings);
  for (int i = 0; i < nrings; ++i) {
    int rp;
    orc_ring_info(nside, i + 1, &rings[i].start, &rp, NULL, &rings[i].z, &rings[i].phi0);
    rings[i].npix = rp;
  }
  const int nt = set_threads(nthreads);
  size_t foff = 0;
  for (int l = 0; l < lmax; ++l) {
    const int n_l = nl ? (nl[l] < nmax ? nl[l] : nmax) : nmax;
    interp_t tab = {0};
    if (interpolated && (rc = interp_build(&tab, l, nmax, zeros_mode))) { free(tab.y); free(rings); return rc; }
    /* ring-wise Y*_lm at the first pixel of each ring (:115-125) */
    double *ylm = (double *)malloc(sizeof(double) * 2 * (size_t)nrings * (l + 1));
    for (int i = 0; i < nrings; ++i)
      for (int m = 0; m <= l; ++m) {
        double p = orc_sph_plm(l, m, rings[i].z);
        ylm[2 * ((size_t)i * (l + 1) + m)] = p * cos(-m * rings[i].phi0);
        ylm[2 * ((size_t)i * (l + 1) + m) + 1] = p * sin(-m * rings[i].phi0);
      }
    double *fl = f_lmn ? f_lmn + foff * 2 : NULL;
    if (fl) memset(fl, 0, sizeof(double) * 2 * (size_t)(l + 1) * nmax);
    int err = 0;
    double *priv = (double *)malloc(sizeof(double) * 2 * (size_t)nt * (l + 1));
    for (int n = 0; n < n_l; ++n) {
      const double k = k_ln[(size_t)l * nmax + n];
      memset(priv, 0, sizeof(double) * 2 * (size_t)nt * (l + 1));
#pragma omp parallel num_threads(nt)
      {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *f = priv + 2 * (size_t)tid * (l + 1);
        double *map = (double *)malloc(sizeof(double) * 4 * nside);
#pragma omp for schedule(static)
        for (int i = 0; i < nrings; ++i) {
          const ring_t *rg = &rings[i];
          const int np = rg->npix;
          for (int j = 0; j < np; ++j) { /* ring map (:180-186) */
            double s = 0.0;
            for (long long q = offsets[rg->start + j]; q < offsets[rg->start + j + 1]; ++q)
              s += bessel_or_table(l, k * radii[q], interpolated, &tab, &err);
            map[j] = s;
          }
          for (int m = 0; m <= l; ++m) {
            /* frequency read out of the r2c half spectrum (:202-206) */
            int s = m % np;
            int mf = np / 2 - abs(np / 2 - s);
            double xr = 0.0, xi = 0.0;
            for (int j = 0; j < np; ++j) {
              long long a = ((long long)j * mf) % np;
              double ang = -2.0 * ORC_PI * (double)a / np;
              xr += map[j] * cos(ang);
              xi += map[j] * sin(ang);
            }
            if (m == 0) xi = 0.0;
            else if (!alias_compat && s > np / 2) xi = -xi;
            const double *y = ylm + 2 * ((size_t)i * (l + 1) + m);
            f[2 * m] += xr * y[0] - xi * y[1];
            f[2 * m + 1] += xr * y[1] + xi * y[0];
          }
        }
        free(map);
      }
      double *fm = priv;
      for (int t = 1; t < nt; ++t)
        for (int q = 0; q < 2 * (l + 1); ++q) fm[q] += priv[2 * (size_t)t * (l + 1) + q];
      double sq = 0.0;
      for (int m = 0; m <= l; ++m) sq += fm[2 * m] * fm[2 * m] + fm[2 * m + 1] * fm[2 * m + 1];
      double c = 2.0 * sq - (fm[0] * fm[0] + fm[1] * fm[1]); /* :147 */
      c *= (norm * k) * (norm * k) / (2 * l + 1);             /* :150 */
      c_ln[(size_t)l * nmax + n] = c;
      if (fl)
        for (int m = 0; m <= l; ++m) {
          fl[2 * ((size_t)m * nmax + n)] = fm[2 * m] * k * norm;
          fl[2 * ((size_t)m * nmax + n) + 1] = fm[2 * m + 1] * k * norm;
        }
    }
    free(priv);
    free(ylm);
    free(tab.y);
    if (err) { free(rings); return ORC_EDOMAIN; }
    foff += (size_t)(l + 1) * nmax;
  }
  free(rings);
  return ORC_OK;
}

/* _sfb_reverse, sfb_reverse.cpp:13-113: f = Y (l+1 x npix) * F (npix x nmax) */
int orc_sfb_reverse(const long long *offsets, const double *radii, int nside, int lmax, int nmax, double rmax,
                    const int *nl, double volume, int interpolated, int zeros_mode, int nthreads,
                    double *k_ln, double *c_ln, double *f_lmn) {
  if (!offsets || nside < 1 || (nside & (nside - 1)) || lmax < 0 || nmax < 1 || !(rmax > 0.0) || !k_ln || !c_ln)
    return ORC_EINVAL;
  const long long npix = 12LL * nside * nside;
  const long long npts = offsets[npix];
  if (npts <= 0) return ORC_EINVAL;
  int rc = orc_kln(lmax, nmax, rmax, zeros_mode, k_ln);
  if (rc) return rc;
  memset(c_ln, 0, sizeof(double) * (size_t)lmax * nmax);
  const double norm = norm_factor(rmax, volume, (double)npts);
  const int nt = set_threads(nthreads);
  double *ct = (double *)malloc(sizeof(double) * npix), *ph = (double *)malloc(sizeof(double) * npix);
  for (long long p = 0; p < npix; ++p) {
    double th;
    orc_pix2ang_ring(nside, p, &th, &ph[p]);
    ct[p] = cos(th);
  }
  size_t foff = 0;
  for (int l = 0; l < lmax; ++l) {
    const int n_l = nl ? (nl[l] < nmax ? nl[l] : nmax) : nmax;
    interp_t tab = {0};
    if (interpolated && (rc = interp_build(&tab, l, nmax, zeros_mode))) { free(tab.y); free(ct); free(ph); return rc; }
    const size_t blk = (size_t)(l + 1) * nmax;
    double *priv = (double *)calloc((size_t)nt * blk * 2, sizeof(double));
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      double *f = priv + (size_t)tid * blk * 2;
      double *F = (double *)malloc(sizeof(double) * (n_l > 0 ? n_l : 1));
#pragma omp for schedule(static)
      for (long long p = 0; p < npix; ++p) {
        if (offsets[p] == offsets[p + 1]) continue;
        for (int n = 0; n < n_l; ++n) {
          const double k = k_ln[(size_t)l * nmax + n];
          double s = 0.0;
          for (long long q = offsets[p]; q < offsets[p + 1]; ++q)
            s += bessel_or_table(l, k * radii[q], interpolated, &tab, &err);
          F[n] = s;
        }
        for (int m = 0; m <= l; ++m) {
          double pl = orc_sph_plm(l, m, ct[p]);
          double yr = pl * cos(-m * ph[p]), yi = pl * sin(-m * ph[p]);
          double *fm = f + (size_t)m * nmax * 2;
          for (int n = 0; n < n_l; ++n) {
            fm[2 * n] += yr * F[n];
            fm[2 * n + 1] += yi * F[n];
          }
        }
      }
      free(F);
    }
    free(tab.y);
    if (err) { free(priv); free(ct); free(ph); return ORC_EDOMAIN; }
    double *fl = priv;
    for (int t = 1; t < nt; ++t)
      for (size_t q = 0; q < blk * 2; ++q) fl[q] += priv[(size_t)t * blk * 2 + q];
    for (int n = 0; n < n_l; ++n) {
      const double k = k_ln[(size_t)l * nmax + n];
      double s = 0.0;
      for (int m = -l; m <= l; ++m) {
        const double *v = fl + ((size_t)abs(m) * nmax + n) * 2;
        s += v[0] * v[0] + v[1] * v[1];
      }
      c_ln[(size_t)l * nmax + n] = s * (norm * k) * (norm * k) / (2 * l + 1); /* :95 */
    }
    if (f_lmn) {
      for (int m = 0; m <= l; ++m)
        for (int n = 0; n < nmax; ++n) {
          const double kn = (n < n_l) ? k_ln[(size_t)l * nmax + n] * norm : 0.0;
          f_lmn[(foff + (size_t)m * nmax + n) * 2] = fl[((size_t)m * nmax + n) * 2] * kn;
          f_lmn[(foff + (size_t)m * nmax + n) * 2 + 1] = fl[((size_t)m * nmax + n) * 2 + 1] * kn;
        }
    }
    foff += blk;
    free(priv);
  }
  free(ct);
  free(ph);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* KClkk::savetxt format, return_types.cpp:25-37                            */
/* ------------------------------------------------------------------------ */
/* double spherical-Bessel integral (sbessel_integrator.cpp:45-138)          */
/* ------------------------------------------------------------------------ */

typedef struct {
  int l, kind;
  double k1, k2, a, b, xmin, xmax, dxinv;
  const double *tab;
  int ntab, err;
} orc_dsb_t;

static double dsb_weight(orc_dsb_t *c, double r) {
  if (c->kind == ORC_W_TABLE) {
    /* FunctionInterpolator::operator() (FunctionInterpolator.cpp:21-30) */
    if (r < c->xmin || r >= c->xmax) {
      c->err = ORC_EDOMAIN;
      return 0.0;
    }
    double u = (r - c->xmin) * c->dxinv;
    int i = (int)u;
    if (i > c->ntab - 2) i = c->ntab - 2;
    return c->tab[i] + (c->tab[i + 1] - c->tab[i]) * (u - i);
  }
  double w = c->a == 0.0 ? 1.0 : pow(r, c->a);
  if (c->b > 0.0) w *= exp(-(r / c->b) * (r / c->b));
  return w;
}

static double dsb_f(orc_dsb_t *c, double r) {
  return dsb_weight(c, r) * orc_sbessel_jl(c->l, c->k1 * r) * orc_sbessel_jl(c->l, c->k2 * r);
}

/* Gauss-Legendre nodes/weights on [-1, 1] by Newton on P_n (textbook) */
#define DSB_GL 10
static double dsb_x[DSB_GL], dsb_w[DSB_GL];
static int dsb_init_done = 0;
static void dsb_init(void) {
  if (dsb_init_done) return;
  const int n = DSB_GL;
  for (int i = 0; i < n; ++i) {
    double x = cos(ORC_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    dsb_x[i] = x;
    dsb_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
  dsb_init_done = 1;
}

/* Gauss-Legendre on [a, b]; *abs_out gets the same rule applied to |f| */
static double dsb_gl(orc_dsb_t *c, double a, double b, double *abs_out) {
  double h = 0.5 * (b - a), m = 0.5 * (a + b), s = 0.0, sa = 0.0;
  for (int i = 0; i < DSB_GL; ++i) {
    double f = dsb_f(c, m + h * dsb_x[i]);
    s += dsb_w[i] * f;
    sa += dsb_w[i] * fabs(f);
  }
  if (abs_out) *abs_out = sa * fabs(h);
  return s * h;
}

/* bisection until the two halves agree with the whole to tol (halved with
 * each bisection): 1e-13 of the piece's integral of |f|, plus 1e-14 of the
 * width times the envelope |w| min(1, 1/(k1 r)) min(1, 1/(k2 r)) -- near the
 * zeros of j_l the integrand's own rounding sits ~1e-16 of that envelope, not
 * of the (small) local |f| */
static double dsb_adapt(orc_dsb_t *c, double a, double b, double whole, double tol, int depth) {
  double m = 0.5 * (a + b), L = dsb_gl(c, a, m, NULL), R = dsb_gl(c, m, b, NULL);
  if (depth >= 18 || fabs(L + R - whole) <= tol) return L + R;
  return dsb_adapt(c, a, m, L, 0.5 * tol, depth + 1) + dsb_adapt(c, m, b, R, 0.5 * tol, depth + 1);
}

static double dsb_piece(orc_dsb_t *c, double a, double b) {
  double ga, g = dsb_gl(c, a, b, &ga);
  const double m = 0.5 * (a + b), x1 = c->k1 * m, x2 = c->k2 * m;
  const double env = fabs(dsb_weight(c, m)) * (x1 > 1.0 ? 1.0 / x1 : 1.0) * (x2 > 1.0 ? 1.0 / x2 : 1.0);
  return dsb_adapt(c, a, b, g, 1e-13 * ga + 1e-14 * (b - a) * env + 1e-300, 0);
}

/* one range; a table weight is split at its sample points */
static double dsb_range(orc_dsb_t *c, double a, double b) {
  if (!(b > a)) return 0.0;
  if (c->kind == ORC_W_TABLE) {
    double s = 0.0, lo = a;
    int i = (int)floor((a - c->xmin) * c->dxinv) + 1;
    for (;; ++i) {
      double x = c->xmin + i / c->dxinv, hi = x < b ? x : b;
      if (hi > lo) s += dsb_piece(c, lo, hi);
      lo = hi;
      if (!(x < b)) break;
    }
    return s;
  }
  return dsb_piece(c, a, b);
}

/* zeros of j_l as a growing array (SphericalBesselZeros: operator[], compute_up_to) */
typedef struct {
  int l, n, cap;
  double *z;
} dsb_zeros_t;

static int dsb_zeros_push(dsb_zeros_t *zs) {
  if (zs->n == zs->cap) {
    int cap = zs->cap ? 2 * zs->cap : 64;
    double *nz = (double *)realloc(zs->z, sizeof(double) * cap);
    if (!nz) return ORC_ENOMEM;
    zs->z = nz;
    zs->cap = cap;
  }
  if (zs->n < 2)
    zs->z[zs->n] = newton(zs->l, zero_guess(zs->l, zs->n), ORC_ZEROS_REFERENCE);
  else
    zs->z[zs->n] = newton(zs->l, 2.0 * zs->z[zs->n - 1] - zs->z[zs->n - 2], ORC_ZEROS_REFERENCE);
  zs->n++;
  return ORC_OK;
}

/* zeros up to index idx (operator[], SphericalBesselZeros.cpp:95-100) */
static int dsb_zeros_index(dsb_zeros_t *zs, int idx) {
  while (zs->n <= idx) {
    int rc = dsb_zeros_push(zs);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* compute_up_to(zmax): until the last zero is >= zmax (:70-74) */
static int dsb_zeros_reach(dsb_zeros_t *zs, double zmax) {
  while (zs->n < 2 || zs->z[zs->n - 1] < zmax) {
    int rc = dsb_zeros_push(zs);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* SBesselIntegrationRangeGenerator (sbessel_integrator.cpp:113-138) */
typedef struct {
  dsb_zeros_t *zs;
  double k, rmax, low;
  int hi;
} dsb_gen_t;

static int dsb_gen_set(dsb_gen_t *g, double rmin, double rmax) {
  g->rmax = rmax;
  g->low = rmin;
  /* next_zero(rmin k): index of the first zero above rmin k (:76-80) */
  int rc = dsb_zeros_reach(g->zs, rmin * g->k);
  if (rc) return rc;
  int ix = 0;
  while (ix < g->zs->n && !(g->zs->z[ix] > rmin * g->k)) ++ix;
  rc = dsb_zeros_index(g->zs, ix);
  if (rc) return rc;
  g->hi = (g->zs->z[ix] / g->k > rmin) ? ix : ix + 1;
  return ORC_OK;
}

static int dsb_gen_next(dsb_gen_t *g, double *lo, double *hi) {
  int rc = dsb_zeros_index(g->zs, g->hi);
  if (rc) return rc;
  double z = g->zs->z[g->hi] / g->k;
  *lo = g->low;
  *hi = z < g->rmax ? z : g->rmax;
  g->low = *hi;
  g->hi++;
  return ORC_OK;
}

int orc_sbessel_integral(int l, double rmax, double k1, double k2, int wkind, double a, double b,
                         const double *tab, int ntab, double xmin, double xmax, double *out) {
  if (l < 0 || !(rmax > 0.0) || !(k1 >= 0.0) || !(k2 >= 0.0) || !out) return ORC_EINVAL;
  if (wkind == ORC_W_TABLE && (!tab || ntab < 2 || !(xmax > xmin))) return ORC_EINVAL;
  dsb_init();
  orc_dsb_t c = {l, wkind, k1, k2, a, b, xmin, xmax, (ntab - 1) / (xmax - xmin), tab, ntab, 0};
  if (k1 > k2) {
    double t = k1;
    k1 = k2;
    k2 = t;
  }
  dsb_zeros_t zs = {l, 0, 0, NULL};
  dsb_gen_t gl = {&zs, k1, 0, 0, 0}, gs = {&zs, k2, 0, 0, 0};
  double result = 0.0, lo, hi, slo, shi;
  int rc = dsb_gen_set(&gl, 0.0, rmax);
  if (!rc) rc = dsb_gen_next(&gl, &lo, &hi);
  while (!rc && lo != hi) {
    double inter = 0.0;
    rc = dsb_gen_set(&gs, lo, hi);
    if (!rc) rc = dsb_gen_next(&gs, &slo, &shi);
    while (!rc && slo != shi) {
      inter += dsb_range(&c, slo, shi);
      rc = dsb_gen_next(&gs, &slo, &shi);
    }
    if (rc) break;
    double old = result;
    result += inter;
    if (fabs(result - old) < result * 1e-10) break;
    rc = dsb_gen_next(&gl, &lo, &hi);
  }
  free(zs.z);
  if (rc) return rc;
  if (c.err) return c.err;
  *out = result;
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */

int orc_savetxt(const char *path, const double *a, int rows, int cols) {
  FILE *f = fopen(path, "w");
  if (!f) return ORC_EINVAL;
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) {
      fprintf(f, "%.18e", a[(size_t)i * cols + j]);
      if (j != cols - 1) fputc(' ', f);
    }
    fputc('\n', f);
  }
  fclose(f);
  return ORC_OK;
}
