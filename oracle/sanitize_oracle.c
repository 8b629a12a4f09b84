/*
 * sanitize_oracle.c -- drives every function of the CPU restatement
 * (sfb_oracle.c) on small seeded inputs, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile, target sanitize).  Test
 * infrastructure only.  Exits non-zero when a consistency check fails; the
 * sanitizers abort on the first memory or undefined-behaviour error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "sfb_oracle.h"

static int failures = 0;

static void check(int ok, const char *what) {
  if (!ok) {
    fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

/* uniform points in a ball of radius R, AoS (r, theta, phi), from a 64-bit LCG */
static void uniform_ball(double *pts, int n, double R, unsigned long long seed) {
  unsigned long long s = seed;
  for (int i = 0; i < n; ++i) {
    double u[3];
    for (int k = 0; k < 3; ++k) {
      s = s * 6364136223846793005ULL + 1442695040888963407ULL;
      u[k] = (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
    pts[3 * i] = R * cbrt(u[0]);
    pts[3 * i + 1] = acos(1.0 - 2.0 * u[1]);
    pts[3 * i + 2] = 2.0 * M_PI * u[2];
  }
}

int main(void) {
  const double pi = M_PI;
  /* special functions */
  for (double x = 0.25; x < 200.0; x *= 1.7) {
    check(fabs(orc_sbessel_jl(0, x) - sin(x) / x) <= 1e-14 * (1.0 + fabs(sin(x) / x)) + 1e-16, "j_0 closed form");
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    check(fabs(orc_sbessel_jl(1, x) - j1) <= 1e-12 * fabs(sin(x) / x) + 1e-15, "j_1 closed form");
  }
  for (int l = 0; l < 40; l += 3) {
    double s = 0.0;
    for (int m = 0; m <= l; ++m) {
      const double p = orc_sph_plm(l, m, 0.3);
      s += (m ? 2.0 : 1.0) * p * p;
    }
    check(fabs(s - (2.0 * l + 1.0) / (4.0 * pi)) < 1e-12, "addition theorem");
  }
  double z[8];
  check(orc_bessel_zeros(3, 8, ORC_ZEROS_CONVERGED, z) == ORC_OK, "zeros");
  for (int n = 0; n < 8; ++n) check(fabs(orc_sbessel_jl(3, z[n])) < 1e-12, "zero of j_3");

  /* direct SFB, exact and table, with f_lmn; k_ln and n_l */
  enum { N = 1500, L = 10, NM = 8 };
  const double R = 300.0;
  double *pts = malloc(sizeof(double) * 3 * N);
  uniform_ball(pts, N, R, 7);
  double kln[L * NM], c1[L * NM], c2[L * NM];
  int nl[L];
  check(orc_count_nl(L, NM, R, 0.08, ORC_ZEROS_REFERENCE, nl) == ORC_OK, "count_nl");
  double *f1 = calloc(2 * NM * (L * (L + 1) / 2), sizeof(double));
  double *f2 = calloc(2 * NM * (L * (L + 1) / 2), sizeof(double));
  check(orc_sfb_raw(pts, N, L, NM, R, nl, 0.0, 0, ORC_ZEROS_REFERENCE, 2, kln, c1, f1) == ORC_OK, "sfb_raw exact");
  check(orc_sfb_raw(pts, N, L, NM, R, NULL, 0.0, 1, ORC_ZEROS_REFERENCE, 2, kln, c2, f2) == ORC_OK, "sfb_raw table");
  for (int l = 0; l < L; ++l)
    for (int n = 0; n < nl[l]; ++n)
      check(fabs(c2[l * NM + n] / c1[l * NM + n] - 1.0) < 1e-3, "table vs exact c_ln");
  double *y = malloc(sizeof(double) * 1000 * NM + 8);
  int np;
  double zmax, dxi;
  check(orc_interp_table(4, NM, ORC_ZEROS_REFERENCE, y, &np, &zmax, &dxi) == ORC_OK, "interp table");
  check(np == 1000 * NM, "table size");
  /* a radius past the table is a domain error, as in SBesselInterpolator */
  double far[3] = {2.5 * R, 1.0, 1.0};
  check(orc_sfb_raw(far, 1, 3, 4, R, NULL, 0.0, 1, ORC_ZEROS_REFERENCE, 1, kln, c2, NULL) == ORC_EDOMAIN,
        "domain error");

  /* HEALPix geometry, binning, pixelised transforms */
  const int nside = 8;
  const long long npix = 12LL * nside * nside;
  for (long long p = 0; p < npix; ++p) {
    double th, ph;
    orc_pix2ang_ring(nside, p, &th, &ph);
    check(orc_ang2pix_ring(nside, th, ph) == p, "pix2ang/ang2pix round trip");
  }
  long long *off = malloc(sizeof(long long) * (npix + 1));
  double *rad = malloc(sizeof(double) * N);
  check(orc_pixelize(pts, N, nside, off, rad) == ORC_OK, "pixelize");
  check(off[0] == 0 && off[npix] == N, "CSR ends");
  double c3[L * NM], c4[L * NM], c5[L * NM];
  double *f3 = calloc(2 * NM * (L * (L + 1) / 2), sizeof(double));
  check(orc_sfb_reverse_fft(off, rad, nside, L, NM, R, nl, 0.0, 0, ORC_ZEROS_REFERENCE, 1, 2, kln, c3, f3) == ORC_OK,
        "reverse_fft alias");
  check(orc_sfb_reverse_fft(off, rad, nside, L, NM, R, nl, 0.0, 0, ORC_ZEROS_REFERENCE, 0, 2, kln, c4, NULL) ==
            ORC_OK,
        "reverse_fft conjugated");
  check(orc_sfb_reverse(off, rad, nside, L, NM, R, nl, 0.0, 0, ORC_ZEROS_REFERENCE, 2, kln, c5, NULL) == ORC_OK,
        "reverse");
  for (int i = 0; i < L * NM; ++i)
    if (c5[i] != 0.0) check(fabs(c4[i] / c5[i] - 1.0) < 1e-10, "conjugated FFT path equals the exact transform");

  /* double spherical-Bessel integrals: power weight, and a table weight */
  double I1, I2;
  check(orc_sbessel_integral(2, R, 0.05, 0.05, ORC_W_POWER_GAUSS, 2.0, 0.0, NULL, 0, 0.0, 0.0, &I1) == ORC_OK,
        "integral power");
  {
    const double x = 0.05 * R, j = orc_sbessel_jl(2, x), jm = orc_sbessel_jl(1, x), jp = orc_sbessel_jl(3, x);
    check(fabs(I1 / (R * R * R / 2.0 * (j * j - jm * jp)) - 1.0) < 1e-9, "integral closed form");
  }
  double tab[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
  check(orc_sbessel_integral(1, R, 0.03, 0.07, ORC_W_TABLE, 0.0, 0.0, tab, 5, 0.0, R * 1.01, &I2) == ORC_OK,
        "integral table");

  /* savetxt */
  check(orc_savetxt("/tmp/sanitize_oracle_savetxt.k_ln", kln, L, NM) == ORC_OK, "savetxt");
  remove("/tmp/sanitize_oracle_savetxt.k_ln");

  free(pts);
  free(f1);
  free(f2);
  free(f3);
  free(y);
  free(off);
  free(rad);
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("sanitize_oracle: all checks passed\n");
  return 0;
}
