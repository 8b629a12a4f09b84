// sfb_direct_f64.hpp -- FP64 SFB kernel: the drop-in path and the GPU oracle,
// for points (direct sum) and for HEALPix pixels (pixelised path).
//
// Per l the reference computes f_lmn = sum_i k_ln j_l(k_ln r_i) Y*_lm(t_i, p_i)
// (sfb_raw.cpp:61-80), i.e. a product of the (n x points) Bessel matrix with
// the (points x 2(l+1)) matrix of real/imaginary Y* columns.  The pixelised
// path (sfb_reverse_fft.cpp:174-209, sfb_reverse.cpp:63-88) is the same
// product over pixels with F_l(p, n) = sum_{r in p} j_l(k_ln r) in place of the
// Bessel value and the pixel's Y* (YlmDownPix) in place of the point's.
//
// Pixel slots are taken in the order pix.perm (pixels sorted by point count):
// a lane loops over its pixel's points, so the 64 pixels of a wave should
// hold about equally many.
// A work item is (l, 32 rows of n, a range of points or pixels).  Per 128-wide
// tile the workgroup
//   1. evaluates the 32 x 128 Bessel (or pixel-sum) tile into LDS (all 256 threads),
//   2. walks m from l down to 0 in chunks of 16 (32 columns): threads 0..127
//      advance their point's Legendre/azimuth recurrence and write the chunk
//      into LDS, then the four waves contract it on the FP64 matrix cores
//      (v_mfma_f64_16x16x4f64), each wave owning one 16 x 16 output tile of
//      the 32 x 32 chunk block.
// Accumulators stay in registers for the whole range (QMAX chunks of double4
// per lane), and the item's 32 x 2(l+1) block is written once.  An item covers
// m chunks [qlo, qhi) counted from m = 0; for l >= 16*QMAX the columns are
// split over several items, each of which first advances the recurrence from
// m = l down to its own top chunk.
#pragma once

#include "sfb_common.hpp"
#include "sfb_math.hpp"
#include "sfb_ylm.hpp"

namespace sfb {

constexpr int F64_NT = 32;           // rows (n) per item
constexpr int F64_TP = 128;          // points per tile
constexpr int F64_MC = 16;           // m values per chunk (32 real columns)
constexpr int F64_LDP = F64_TP + 1;  // LDS row pitch (doubles)

enum { JL_EXACT = 0, JL_REFERENCE_TABLE = 1, JL_QUINTIC = 2 };

// Quintic Hermite spacing of the FP64 j_l table of the mixed path's low
// multipoles: interpolation error <= h^6/46080 max|j_l^(6)| ~ 6e-11 of the
// amplitude at h = 1/8.
constexpr double Q5_H = 0.125;
constexpr double Q5_HINV = 8.0;

struct BesselTableRef {
  const double *y;       // [lmax][P] values of j_l at i*dx
  const double *zmax;    // [lmax]
  const double *dx_inv;  // [lmax]
  int32_t P;             // nodes per l
  // quintic table: per l, nint[l] intervals [i h, (i+1) h] of 6 coefficients in t = x/h - i, from off[l]
  const double *q5;
  const int32_t *q5_off;
  const int32_t *q5_nint;
};

// j_l, j_l' and j_l'' at x (FP64).  j_l' = (l/x) j_l - j_{l+1}; j_l'' from
// the differential equation x^2 j'' + 2x j' + (x^2 - l(l+1)) j = 0.
__device__ inline void jl_d012(int l, double x, double &f, double &d, double &s) {
  if (x == 0.0) {
    f = l == 0 ? 1.0 : 0.0;
    d = l == 1 ? 1.0 / 3.0 : 0.0;
    s = l == 0 ? -1.0 / 3.0 : (l == 2 ? 2.0 / 15.0 : 0.0);
    return;
  }
  f = sbessel_jl(l, x);
  const double ix = 1.0 / x;
  d = l * ix * f - sbessel_jl(l + 1, x);
  s = -2.0 * ix * d - (1.0 - l * (l + 1.0) * ix * ix) * f;
}

// One thread per interval: the quintic Hermite polynomial matching j, h j',
// h^2 j'' at both ends, as coefficients of t^0 .. t^5.
__global__ void quintic_table_kernel(int lmax, const int32_t *__restrict__ off, const int32_t *__restrict__ nint,
                                     double *__restrict__ c) {
  const int l = blockIdx.y;
  if (l >= lmax) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nint[l]; i += gridDim.x * blockDim.x) {
    double f0, d0, s0, f1, d1, s1;
    jl_d012(l, i * Q5_H, f0, d0, s0);
    jl_d012(l, (i + 1) * Q5_H, f1, d1, s1);
    d0 *= Q5_H;
    d1 *= Q5_H;
    s0 *= Q5_H * Q5_H;
    s1 *= Q5_H * Q5_H;
    const double df = f1 - f0;
    double *o = c + 6 * ((size_t)off[l] + i);
    o[0] = f0;
    o[1] = d0;
    o[2] = 0.5 * s0;
    o[3] = 10.0 * df - 6.0 * d0 - 4.0 * d1 - 1.5 * s0 + 0.5 * s1;
    o[4] = -15.0 * df + 8.0 * d0 + 7.0 * d1 + 1.5 * s0 - s1;
    o[5] = 6.0 * df - 3.0 * d0 - 3.0 * d1 - 0.5 * s0 + 0.5 * s1;
  }
}

// j_l(z): exact (gsl_sf_bessel_jl semantics), the reference's linear table
// (SBesselInterpolator.cpp:20-25, domain error past z_max), or the quintic table.
__device__ __forceinline__ double bessel_f64(int l, double z, int use_table, const BesselTableRef &tab, int &err) {
  if (use_table == JL_EXACT) return sbessel_jl(l, z);
  if (use_table == JL_QUINTIC) {
    // the index is clamped on both sides, so no argument reads outside the table
    // (arguments are checked before a launch; this guards the reads regardless)
    const double u = z * Q5_HINV;
    int i = (int)u;
    const int last = tab.q5_nint[l] - 1;
    i = i < last ? i : last;
    i = i > 0 ? i : 0;
    const double t = u - (double)i;
    const double2 *cc = reinterpret_cast<const double2 *>(tab.q5 + 6 * ((size_t)tab.q5_off[l] + i));
    const double2 a = cc[0], b = cc[1], e = cc[2];
    return fma(t, fma(t, fma(t, fma(t, fma(t, e.y, e.x), b.y), b.x), a.y), a.x);
  }
  // the table covers [0, z_max): anything else (including NaN) is a domain error
  if (!(z >= 0.0 && z < tab.zmax[l])) {
    err |= ERR_DOMAIN;
    return 0.0;
  }
  const double *y = tab.y + (size_t)l * tab.P;
  const double ia = z * tab.dx_inv[l];
  int i = (int)ia;
  i = i > tab.P - 2 ? tab.P - 2 : i;
  i = i > 0 ? i : 0;
  return y[i] + (y[i + 1] - y[i]) * (ia - i);
}

template <int QMAX, bool PIX>
__global__ __launch_bounds__(256) void sfb_direct_f64_kernel(
    PointsSoA pts, PixelsSoA pix, const DirectItem *__restrict__ items, const double *__restrict__ kln, int nmax,
    LegendreTables leg, int use_table, BesselTableRef tab, double *__restrict__ partial, int *__restrict__ err) {
  __shared__ double Js[F64_NT * F64_LDP];
  __shared__ double Ys[2 * F64_MC * F64_LDP];

  const DirectItem it = items[blockIdx.x];
  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int rt = wave >> 1, ct = wave & 1;

  const double *kl = kln + it.kofs;
  const double2 sect = leg.sect[l];
  const double2 *coef = leg.coef + leg.coef_off[l];

  v4d acc[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};

  int myerr = 0;
  for (int64_t base = it.p0; base < it.p1; base += F64_TP) {
    // ---- 1. Bessel (or pixel-sum) tile: thread -> point tid & 127, rows (tid >> 7) + 2j
    {
      const int p = tid & (F64_TP - 1);
      const int64_t gp = base + p;
      double accj[F64_NT / 2];
#pragma unroll
      for (int j = 0; j < F64_NT / 2; ++j) accj[j] = 0.0;
      if (gp < it.p1) {
        int64_t q0 = gp, q1 = gp + 1;
        if constexpr (PIX) {
          const int64_t pp = pix.perm[gp];
          q0 = pix.off[pp];
          q1 = pix.off[pp + 1];
        }
        for (int64_t q = q0; q < q1; ++q) {
          const double r = PIX ? pix.radii[q] : pts.r[q];
#pragma unroll
          for (int j = 0; j < F64_NT / 2; ++j) {
            const int row = (tid >> 7) + 2 * j;
            if (row < it.nrows) accj[j] += bessel_f64(l, kl[row] * r, use_table, tab, myerr);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < F64_NT / 2; ++j) Js[((tid >> 7) + 2 * j) * F64_LDP + p] = accj[j];
    }
    // ---- 2. Legendre/azimuth rows, chunk by chunk from high m to low m
    YlmDown g;
    YlmDownPix gx;
    const bool ygen = tid < F64_TP;
    const int64_t gp = base + tid;
    if (ygen) {
      if (gp < it.p1) {
        if constexpr (PIX) {
          const int64_t pp = pix.perm[gp];
          gx.init(l, pix.x[pp], pix.s[pp], pix.phi[pp], pix.cphi[pp], pix.sphi[pp], pix.phi2[pp], pix.cphi2[pp],
                  pix.sphi2[pp], pix.np[pp], pix.alias_compat != 0, sect);
          for (int m = l; m >= F64_MC * it.qhi; --m) {
            gx.step(coef[m]);
            gx.rescale();
          }
        } else {
          // the azimuth starts at the top of the highest chunk; the Legendre
          // pair alone is taken there from m = l
          const int mstart = min(l, F64_MC * it.qhi - 1);
          g.init(l, mstart, pts.x[gp], pts.s[gp], pts.phi[gp], pts.cphi[gp], pts.sphi[gp], sect);
          for (int m = l; m > mstart; --m) {
            g.step_p(coef[m]);
            g.rescale();
          }
        }
      } else {
        g.v = g.vp = 0.0;
        g.E = 0;
        g.cm = g.sm = g.cp = g.sp = g.cot = g.cph = g.sph = 0.0;
        gx.y = g;
        gx.cm2 = gx.sm2 = gx.cph2 = gx.sph2 = 0.0;
        gx.np = 1;
        gx.smod = 0;
        gx.quirk = false;
      }
    }
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qhi - 1 - qq;
      if (q >= it.qlo) {
        __syncthreads();  // previous readers of Ys (and Js writers) are done
        if (ygen) {
          const int mlo = q * F64_MC;
          for (int mm = F64_MC - 1; mm >= 0; --mm) {
            const int m = mlo + mm;
            double re = 0.0, im = 0.0;
            if (m <= l) {
              if constexpr (PIX) {
                gx.value(m, re, im);
                if (m > 0) {
                  gx.step(coef[m]);
                  gx.rescale();
                }
              } else {
                g.value(re, im);
                if (m > 0) {
                  g.step(coef[m]);
                  g.rescale();
                }
              }
            }
            Ys[(2 * mm) * F64_LDP + tid] = re;
            Ys[(2 * mm + 1) * F64_LDP + tid] = im;
          }
        }
        __syncthreads();
        // contraction of the 32 x 32 chunk block: wave (rt, ct) owns a 16 x 16 tile
        const double *Arow = Js + (16 * rt + (lane & 15)) * F64_LDP + (lane >> 4);
        const double *Brow = Ys + (16 * ct + (lane & 15)) * F64_LDP + (lane >> 4);
        v4d c = acc[qq];
#pragma unroll 8
        for (int s = 0; s < F64_TP / 4; ++s) {
          c = __builtin_amdgcn_mfma_f64_16x16x4f64(Arow[4 * s], Brow[4 * s], c, 0, 0, 0);
        }
        acc[qq] = c;
      }
    }
    __syncthreads();  // Js is rewritten by the next tile
  }
  if (myerr) atomicOr(err, myerr);

  // ---- write the item's partial block (rows x 2(l+1), row-major)
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qhi - 1 - qq;
    if (q >= it.qlo) {
      const int col = q * 2 * F64_MC + 16 * ct + (lane & 15);
      // v_mfma_f64_16x16x4f64 result: element j of lane i is row i/16 + 4j, column i%16
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 16 * rt + (lane >> 4) + 4 * j;
        if (col < ncol) out[(size_t)row * ncol + col] = acc[qq][j];
      }
    }
  }
}

}  // namespace sfb
