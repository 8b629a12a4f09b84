// sfb_support.hpp -- the small kernels around the contraction: point
// preparation, Bessel zeros, the reference interpolation table, the
// deterministic reduction of partial blocks and the C_l assembly.
#pragma once

#include "sfb_common.hpp"
#include "sfb_math.hpp"

namespace sfb {

// AoS (r, theta, phi) -> SoA (r, cos t, sin t, phi, cos phi, sin phi); flags
// r < 0 and non-finite input; *rmax_bits receives the largest r (the bit
// pattern of a non-negative double orders like the value).
__global__ void prep_points_kernel(const double *__restrict__ aos, int64_t n, double *__restrict__ r,
                                   double *__restrict__ x, double *__restrict__ s, double *__restrict__ phi,
                                   double *__restrict__ cphi, double *__restrict__ sphi, int *__restrict__ err,
                                   unsigned long long *__restrict__ rmax_bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (i < n) rr = aos[3 * i];
  // block maximum, then one atomic per block
  __shared__ double red[256];
  red[threadIdx.x] = (rr >= 0.0 && rr <= 1.7e308) ? rr : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(rmax_bits, (unsigned long long)__double_as_longlong(red[0]));
  if (i >= n) return;
  const double th = aos[3 * i + 1], ph = aos[3 * i + 2];
  double st, ctt, sp, cp;
  sincos(th, &st, &ctt);
  sincos(ph, &sp, &cp);
  r[i] = rr;
  x[i] = ctt;
  s[i] = fabs(st);
  phi[i] = ph;
  cphi[i] = cp;
  sphi[i] = sp;
  if (!(rr >= 0.0) || !isfinite(rr) || !isfinite(th) || !isfinite(ph)) atomicOr(err, ERR_BADPOINT);
}

// 32-bit radius keys of a single call: the FP32 pattern of r (monotone in r >= 0;
// radii equal to FP32 rounding keep their input order in the stable sort).
__global__ void radius_keys32_kernel(const double *__restrict__ r, int64_t n, unsigned *__restrict__ keys,
                                     int32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = __float_as_uint((float)(r[i] > 0.0 ? r[i] : 0.0));  // -0.0 (and NaN) key as +0.0
  idx[i] = (int32_t)i;
}

// Keys of a batched call: catalogue index in the top 7 bits, then the top 57
// bits of r's pattern (order within a catalogue only needs to be approximate).
__global__ void radius_keys_cat_kernel(const double *__restrict__ r, int64_t n, int64_t base, int cat,
                                       unsigned long long *__restrict__ keys, int32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long b = (unsigned long long)__double_as_longlong(r[base + i] > 0.0 ? r[base + i] : 0.0);
  keys[base + i] = ((unsigned long long)cat << 57) | (b >> 7);
  idx[base + i] = (int32_t)(base + i);
}

// out[i] = in[idx[i]] for the six arrays of the point table.
__global__ void gather_points_kernel(PointsSoA in, const int32_t *__restrict__ idx, int64_t n,
                                     double *__restrict__ r, double *__restrict__ x, double *__restrict__ s,
                                     double *__restrict__ phi, double *__restrict__ cphi,
                                     double *__restrict__ sphi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = idx[i];
  r[i] = in.r[j];
  x[i] = in.x[j];
  s[i] = in.s[j];
  phi[i] = in.phi[j];
  cphi[i] = in.cphi[j];
  sphi[i] = in.sphi[j];
}

// One thread per l: zeros z_{l,0..nz-1} of j_l, row-major [l][nz].
__global__ void bessel_zeros_kernel(int lmax, int nz, int mode, double *__restrict__ z) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lmax) return;
  bessel_zeros(l, nz, mode, z + (size_t)l * nz);
}

// Reference linear table (SBesselInterpolator.cpp:10-17, FunctionInterpolator.cpp:10-17):
// P = 1000*nmax nodes x_i = i*dx on [0, z_{l,nmax}], dx = z_{l,nmax}/(P-1).
__global__ void bessel_table_kernel(int lmax, int P, const double *__restrict__ zmax, double *__restrict__ y) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)lmax * P) return;
  const int l = (int)(idx / P);
  const int i = (int)(idx - (int64_t)l * P);
  const double dx = zmax[l] / (P - 1);
  y[idx] = sbessel_jl(l, 0.0 + i * dx);
}

// Sum the partial blocks of a group and scale by norm * k_ln: f[l][m][n]
// (complex, m >= 0, n < nmax) of the group's catalogue.  Block (g, c) sums
// elements c, c + gridDim.y*256, ... of group g.  The splits are summed in
// eight interleaved chains (split s into chain s mod 8, in split order),
// added pairwise in a fixed order: the result depends only on the plan, and
// eight loads are in flight per element (a group of few elements and many
// splits is otherwise one long dependent chain per thread).
__global__ void reduce_partials_kernel(const ReduceGroup *__restrict__ groups, const double *__restrict__ partial,
                                       const double *__restrict__ kln, double *__restrict__ flmn) {
  const ReduceGroup g = groups[blockIdx.x];
  const int ncol = 2 * (g.l + 1);
  const int total = g.nrows * ncol;
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < total; e += gridDim.y * blockDim.x) {
    const int row = e / ncol, col = e - row * ncol;
    const double *src = partial + g.base + (size_t)row * ncol + col;
    double c[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int s = 0;
    for (; s + 8 <= g.nsplit; s += 8) {
#pragma unroll
      for (int k = 0; k < 8; ++k) c[k] += src[(size_t)(s + k) * g.stride];
    }
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (s + k < g.nsplit) c[k] += src[(size_t)(s + k) * g.stride];
    const double acc = ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
    const int n = g.n0 + row;
    const int m = col >> 1;
    flmn[(g.fofs + (size_t)m * g.nmax + n) * 2 + (col & 1)] = acc * g.norm * kln[g.kofs + row];
  }
}

// c_ln = (|f_l0n|^2 + 2 sum_{m>=1} |f_lmn|^2) / (2l+1)   (sfb_raw.cpp:91-96);
// columns n >= n_l are zero.
__global__ void cln_kernel(int lmax, int nmax, const int32_t *__restrict__ nl, const int64_t *__restrict__ flmn_off,
                           const double *__restrict__ flmn, double *__restrict__ cln) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= lmax * nmax) return;
  const int l = idx / nmax, n = idx - l * nmax;
  double c = 0.0;
  if (n < nl[l]) {
    const double *f = flmn + flmn_off[l] * 2;
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      const double re = f[((size_t)am * nmax + n) * 2], im = f[((size_t)am * nmax + n) * 2 + 1];
      c += re * re + im * im;
    }
    c /= (2 * l + 1);
  }
  cln[idx] = c;
}

// One catalogue of a batch plan: its AoS points, their count, where its rows
// start in the concatenated point table, and the radius its Bessel tables cover.
struct CatPrep {
  const double *aos;
  int64_t n, base;
  double R;
};

// Point table of a batch plan (blockIdx.y = catalogue), in input order, and
// each point's radius bucket floor(r 2^16 / R) (keys).  No point leaves the
// tables the plan sized: r > R is clamped to R and flagged ERR_DOMAIN; r < 0
// or a non-finite coordinate becomes (0, 0, 0), flagged ERR_BADPOINT.  The
// caller reads the flags after the launch.
__global__ void prep_points_batch_kernel(const CatPrep *__restrict__ cats, double *__restrict__ soa, int64_t ntot,
                                         uint16_t *__restrict__ keys, int *__restrict__ err) {
  const CatPrep c = cats[blockIdx.y];
  int bits = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (int64_t)gridDim.x * blockDim.x) {
    double rr = c.aos[3 * i], th = c.aos[3 * i + 1], ph = c.aos[3 * i + 2];
    if (!(rr >= 0.0) || !isfinite(rr) || !isfinite(th) || !isfinite(ph)) {
      bits |= ERR_BADPOINT;
      rr = th = ph = 0.0;
    } else if (rr > c.R) {
      bits |= ERR_DOMAIN;
      rr = c.R;
    }
    double st, ctt, sp, cp;
    sincos(th, &st, &ctt);
    sincos(ph, &sp, &cp);
    const int64_t j = c.base + i;
    keys[j] = (uint16_t)min((int)(rr * (65536.0 / c.R)), 65535);
    soa[j] = rr;
    soa[ntot + j] = ctt;
    soa[2 * ntot + j] = fabs(st);
    soa[3 * ntot + j] = ph;
    soa[4 * ntot + j] = cp;
    soa[5 * ntot + j] = sp;
  }
  if (bits) atomicOr(err, bits);
}

// Outputs of one catalogue of a batch plan: its k_ln, c_ln and (optional)
// f_lmn buffers, its shape and where its rows sit in the plan's arrays.
struct CatOut {
  double *k_ln, *c_ln, *f_lmn;
  int32_t lmax, nmax;
  int64_t kbase, nlbase, foffbase, fbase;
};

// blockIdx.y = catalogue: c_ln as cln_kernel computes it, k_ln, and a copy of
// f_lmn out of the plan's f workspace.
__global__ void batch_outputs_kernel(const CatOut *__restrict__ outs, const double *__restrict__ f,
                                     const double *__restrict__ kln, const int32_t *__restrict__ nl,
                                     const int64_t *__restrict__ foff) {
  const CatOut o = outs[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double *fc = f + 2 * o.fbase;
  for (int64_t idx = t0; idx < (int64_t)o.lmax * o.nmax; idx += stride) {
    const int l = (int)(idx / o.nmax), n = (int)(idx - (int64_t)l * o.nmax);
    double c = 0.0;
    if (n < nl[o.nlbase + l]) {
      const double *fl = fc + foff[o.foffbase + l] * 2;
      for (int m = -l; m <= l; ++m) {
        const int am = m < 0 ? -m : m;
        const double re = fl[((size_t)am * o.nmax + n) * 2], im = fl[((size_t)am * o.nmax + n) * 2 + 1];
        c += re * re + im * im;
      }
      c /= (2 * l + 1);
    }
    o.c_ln[idx] = c;
    o.k_ln[idx] = kln[o.kbase + idx];
  }
  if (o.f_lmn) {
    const int64_t nf = (int64_t)o.lmax * (o.lmax + 1) / 2 * o.nmax * 2;
    for (int64_t idx = t0; idx < nf; idx += stride) o.f_lmn[idx] = fc[idx];
  }
}

__global__ void fill_kernel(double *__restrict__ p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_int_kernel(int *__restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace sfb
