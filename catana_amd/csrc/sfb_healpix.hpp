// sfb_healpix.hpp -- HEALPix RING geometry and binning on the device.
//
// Formulas from Gorski et al. 2005 (ApJ 622, 759), the RING scheme the
// reference uses through Healpix_Base (PixelizedPointContainer.cpp:11-30):
//   rings i = 1 .. 4 nside - 1; north cap i < nside has 4i pixels starting at
//   2i(i-1), z = 1 - i^2/(3 nside^2), phi_j = (j + 1/2) pi/(2i); the equatorial
//   belt has 4 nside pixels per ring, z = (2 nside - i) 2/(3 nside), shifted by
//   half a pixel when i - nside is even; the south cap mirrors the north.
// ang2pix keeps the reference library's branch structure, including the use of
// sin(theta) near the poles (theta < 0.01 or > pi - 0.01).
#pragma once

#include "sfb_common.hpp"

namespace sfb {

constexpr double HP_PI = 3.14159265358979323846;

__host__ __device__ inline int64_t hp_isqrt(int64_t v) {
  int64_t r = (int64_t)sqrt((double)v + 0.5);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

// ring index (1-based), first pixel, ring length, shifted flag and 1 - z (exact
// in the caps) of pixel p.
__host__ __device__ inline void hp_ring_of(int64_t ns, int64_t p, int64_t &ring, int64_t &start, int64_t &np,
                                           bool &shifted) {
  const int64_t npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  if (p < ncap) {
    ring = (1 + hp_isqrt(1 + 2 * p)) >> 1;
    np = 4 * ring;
    start = 2 * ring * (ring - 1);
    shifted = true;
  } else if (p < npix - ncap) {
    ring = (p - ncap) / (4 * ns) + ns;
    np = 4 * ns;
    start = ncap + (ring - ns) * 4 * ns;
    shifted = ((ring - ns) & 1) == 0;
  } else {
    const int64_t nr = (1 + hp_isqrt(2 * (npix - p) - 1)) >> 1;
    ring = 4 * ns - nr;
    np = 4 * nr;
    start = npix - 2 * nr * (nr + 1);
    shifted = true;
  }
}

// Pixel data for the pixelised SFB (see PixelsSoA).
__global__ void pixel_setup_kernel(int nside, double *__restrict__ x, double *__restrict__ s,
                                   double *__restrict__ phi, double *__restrict__ cphi, double *__restrict__ sphi,
                                   double *__restrict__ phi2, double *__restrict__ cphi2, double *__restrict__ sphi2,
                                   int32_t *__restrict__ npr) {
  const int64_t ns = nside, npix = 12 * ns * ns;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  int64_t ring, start, np;
  bool shifted;
  hp_ring_of(ns, p, ring, start, np, shifted);
  double z, omz, opz;  // z, 1 - z, 1 + z
  if (ring < ns) {
    omz = (double)(ring * ring) / (3.0 * ns * ns);
    z = 1.0 - omz;
    opz = 2.0 - omz;
  } else if (ring < 3 * ns) {
    z = (2.0 * ns - ring) * 2.0 / (3.0 * ns);
    omz = 1.0 - z;
    opz = 1.0 + z;
  } else {
    const int64_t nr = 4 * ns - ring;
    opz = (double)(nr * nr) / (3.0 * ns * ns);
    z = opz - 1.0;
    omz = 2.0 - opz;
  }
  const int64_t j = p - start;
  const double dphi = 2.0 * HP_PI / np;
  const double phi0 = shifted ? 0.5 * dphi : 0.0;
  const double ph = phi0 + j * dphi, ph2 = phi0 - j * dphi;
  x[p] = z;
  s[p] = sqrt(omz * opz);
  phi[p] = ph;
  phi2[p] = ph2;
  double sn, cs;
  sincos(ph, &sn, &cs);
  cphi[p] = cs;
  sphi[p] = sn;
  sincos(ph2, &sn, &cs);
  cphi2[p] = cs;
  sphi2[p] = sn;
  npr[p] = (int32_t)np;
}

__host__ __device__ inline int64_t hp_ang2pix(int64_t ns, double theta, double phi) {
  const int64_t npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  const double z = cos(theta), za = fabs(z);
  double tt = fmod(phi * (2.0 / HP_PI), 4.0);
  if (tt < 0) tt += 4.0;
  if (tt >= 4.0) tt = 0.0;
  if (za <= 2.0 / 3.0) {
    const int64_t nl4 = 4 * ns;
    const double t1 = ns * (0.5 + tt), t2 = ns * z * 0.75;
    const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
    const int64_t ir = ns + 1 + jp - jm;
    const int64_t kshift = 1 - (ir & 1);
    int64_t ip = (jp + jm - ns + kshift + 1 + nl4 + nl4) / 2;
    ip %= nl4;
    return ncap + (ir - 1) * nl4 + ip;
  }
  const double tp = tt - (double)(int64_t)tt;
  const double tmp = (theta < 0.01 || theta > 3.14159 - 0.01) ? ns * sin(theta) / sqrt((1.0 + za) / 3.0)
                                                              : ns * sqrt(3.0 * (1.0 - za));
  const int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
  const int64_t ir = jp + jm + 1;
  int64_t ip = (int64_t)(tt * ir);
  if (ip >= 4 * ir) ip -= 4 * ir;
  return (z > 0) ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

// Point -> (pixel key, radius) for the stable sort that builds the CSR.
__global__ void ang2pix_kernel(const double *__restrict__ aos, int64_t n, int nside, uint32_t *__restrict__ keys,
                               double *__restrict__ radii, int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = aos[3 * i], th = aos[3 * i + 1], ph = aos[3 * i + 2];
  if (!(r >= 0.0) || !isfinite(r) || !isfinite(th) || !isfinite(ph)) {
    atomicOr(err, ERR_BADPOINT);
    keys[i] = 0;
    radii[i] = 0.0;
    return;
  }
  keys[i] = (uint32_t)hp_ang2pix(nside, th, ph);
  radii[i] = r;
}

// off[p] = first position of key >= p in the sorted keys (p = 0 .. npix).
__global__ void csr_offsets_kernel(const uint32_t *__restrict__ keys, int64_t n, int64_t npix,
                                   int64_t *__restrict__ off) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > npix) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)keys[mid] < p) lo = mid + 1;
    else hi = mid;
  }
  off[p] = lo;
}

// Checks a caller's CSR before the pixelised contraction reads through it:
// off[0] = 0, off non-decreasing up to off[npix] = npts (read by the host),
// every radius finite and >= 0.  Grid-stride over max(npix, npts).  Also
// err[1] = the largest pixel count and err[2..3] = the bits of the largest
// radius (of a non-negative double, ordered like the value).
__global__ __launch_bounds__(256) void csr_check_kernel(const int64_t *__restrict__ off, int64_t npix,
                                                        const double *__restrict__ radii,
                                 int64_t npts, int *__restrict__ err) {
  const int64_t n = npix > npts ? npix : npts;
  int bad = 0, cmax = 0;
  double rmx = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < npix) {
      const int64_t c = off[i + 1] - off[i];
      if (c < 0 || (i == 0 && off[0] != 0)) bad |= ERR_BADCSR;
      cmax = max(cmax, (int)min<int64_t>(c, 0x7fffffff));
    }
    if (i < npts) {
      const double r = radii[i];
      if (!(r >= 0.0) || !isfinite(r)) bad |= ERR_BADPOINT;
      else rmx = fmax(rmx, r);
    }
  }
  // block reduction, then one atomic of each kind per block
  __shared__ int sb[256], sc[256];
  __shared__ double sr[256];
  sb[threadIdx.x] = bad;
  sc[threadIdx.x] = cmax;
  sr[threadIdx.x] = rmx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sb[threadIdx.x] |= sb[threadIdx.x + w];
      sc[threadIdx.x] = max(sc[threadIdx.x], sc[threadIdx.x + w]);
      sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (sb[0]) atomicOr(err, sb[0]);
    atomicMax(err + 1, sc[0]);
    atomicMax(reinterpret_cast<unsigned long long *>(err + 2), (unsigned long long)__double_as_longlong(sr[0]));
  }
}

// Sort keys of the pixelised contraction's pixel order: the point count of
// pixel p (a stable sort keeps ties in pixel order).  Each lane of the
// contraction loops over its pixel's points, so a wave of pixels with about
// equal counts does not wait on its fullest pixel.
__global__ void pixel_count_keys_kernel(const int64_t *__restrict__ off, int64_t npix, uint32_t *__restrict__ keys,
                                        int32_t *__restrict__ idx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const int64_t c = off[p + 1] - off[p];
  keys[p] = (uint32_t)(c < 0xffffffffLL ? c : 0xffffffffLL);
  idx[p] = (int32_t)p;
}

}  // namespace sfb
