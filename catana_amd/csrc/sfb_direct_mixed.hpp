// sfb_direct_mixed.hpp -- mixed-precision fused direct SFB kernel.
//
// Same contraction as the FP64 kernel (per l: Bessel rows x Y* columns summed
// over points, sfb_raw.cpp:61-80), with the operands generated cheaply:
//   * j_l(k_ln r) from a per-l table of (j_l(x_i), h j_l'(x_i)) on the grid
//     x_i = i h (h = 1/16) held in LDS as FP32 and evaluated by cubic Hermite
//     interpolation (error <= h^4/384 max|j_l''''| ~ 4e-8 of the local
//     amplitude).  The argument t = r k_ln / h and its split into node index
//     and fraction are formed in FP64, so the phase error does not grow with
//     k r.  Arguments past the table (r > rmax) fall back to FP64 j_l.
//   * Y*_lm from the same FP64 downward-m recurrence as the FP64 kernel,
//     rounded to FP32 when written to LDS.
//   * the contraction runs on v_mfma_f32_16x16x4f32; each 128-point tile is
//     accumulated in FP32 and then added into FP64 accumulators, so FP32
//     rounding never sees more than 128 terms.
// Work item: (l, 64 rows of n, m chunks [qlo, qhi), a point range).  Wave w
// owns rows 16w..16w+15 and both 16-column halves of every 32-column chunk.
#pragma once

#include "sfb_common.hpp"
#include "sfb_math.hpp"
#include "sfb_ylm.hpp"

namespace sfb {

constexpr int MX_NT = 64;            // rows (n) per item
constexpr int MX_TP = 128;           // points per tile
constexpr int MX_MC = 16;            // m values per chunk (32 real columns)
constexpr int MX_LDP = MX_TP + 4;    // LDS row pitch (floats): rows land 4 banks apart
constexpr double MX_H = 1.0 / 16.0;  // Bessel table spacing
constexpr int MX_TAB_MAX = 13568;    // table nodes that fit in LDS next to the tiles (x up to ~848)

// Per-l Hermite tables: node i of l at tab[off[l] + i] = (j_l(i h), h j_l'(i h)).
struct HermiteTables {
  const float2 *tab;
  const int32_t *off;
  const int32_t *nnodes;
};

// j_l and h*j_l' at x = i*h, FP64, rounded to FP32.
__global__ void hermite_table_kernel(int lmax, const int32_t *__restrict__ off, const int32_t *__restrict__ nn,
                                     float2 *__restrict__ tab) {
  const int l = blockIdx.y;
  if (l >= lmax) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn[l]; i += gridDim.x * blockDim.x) {
    const double x = i * MX_H;
    double j, dj;
    if (x == 0.0) {
      j = (l == 0) ? 1.0 : 0.0;
      dj = (l == 1) ? 1.0 / 3.0 : 0.0;
    } else {
      j = sbessel_jl(l, x);
      dj = (double)l / x * j - sbessel_jl(l + 1, x);
    }
    tab[off[l] + i] = make_float2((float)j, (float)(MX_H * dj));
  }
}

template <int QMAX>
__global__ __launch_bounds__(256) void sfb_direct_mixed_kernel(
    PointsSoA pts, const DirectItem *__restrict__ items, const double *__restrict__ kln, int nmax,
    LegendreTables leg, HermiteTables ht, double *__restrict__ partial) {
  extern __shared__ float2 T[];  // Bessel table of this item's l (dynamic: largest table of the launch)
  __shared__ float Js[MX_NT * MX_LDP];
  __shared__ float Ys[2 * MX_MC * MX_LDP];
  __shared__ double kh[MX_NT];

  const DirectItem it = items[blockIdx.x];
  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;

  // ---- item setup: Bessel table of this l and k_ln / h for the rows
  const int nn = ht.nnodes[l];
  const float2 *gt = ht.tab + ht.off[l];
  for (int i = tid; i < nn; i += 256) T[i] = gt[i];
  if (tid < MX_NT) kh[tid] = (tid < it.nrows) ? kln[(size_t)l * nmax + it.n0 + tid] * (1.0 / MX_H) : 0.0;
  const double2 sect = leg.sect[l];
  const double2 *coef = leg.coef + leg.coef_off[l];

  v4f acc[QMAX][2];
  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[q][c] = v4f{0.f, 0.f, 0.f, 0.f};
      sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  __syncthreads();

  for (int64_t base = it.p0; base < it.p1; base += MX_TP) {
    // ---- 1. Bessel tile (64 rows x 128 points): thread -> point tid & 127, rows (tid >> 7) + 2j
    {
      const int p = tid & (MX_TP - 1);
      const int64_t gp = base + p;
      const bool ok = gp < it.p1;
      const double r = ok ? pts.r[gp] : 0.0;
#pragma unroll 4
      for (int j = 0; j < MX_NT / 2; ++j) {
        const int row = (tid >> 7) + 2 * j;
        float val = 0.f;
        if (ok && row < it.nrows) {
          const double t = r * kh[row];
          const int i = (int)t;
          if (i < nn - 1) {
            const float f = (float)(t - (double)i);
            const float2 a = T[i], b = T[i + 1];
            const float d = b.x - a.x;
            const float c2 = 3.f * d - 2.f * a.y - b.y;
            const float c3 = -2.f * d + a.y + b.y;
            val = fmaf(f, fmaf(f, fmaf(f, c3, c2), a.y), a.x);
          } else {
            val = (float)sbessel_jl(l, t * MX_H);
          }
        }
        Js[row * MX_LDP + p] = val;
      }
    }
    // ---- 2. Legendre/azimuth rows (threads 0..127), chunks from high m to low m
    YlmDown g;
    const bool ygen = tid < MX_TP;
    const int64_t gp = base + tid;
    if (ygen) {
      if (gp < it.p1) {
        g.init(l, pts.x[gp], pts.s[gp], pts.phi[gp], pts.cphi[gp], pts.sphi[gp], sect);
        for (int m = l; m >= MX_MC * it.qhi; --m) {
          g.step(coef[m]);
          g.rescale();
        }
      } else {
        g.v = g.vp = 0.0;
        g.E = 0;
        g.cm = g.sm = g.cot = g.cph = g.sph = 0.0;
      }
    }
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qhi - 1 - qq;
      if (q >= it.qlo) {
        __syncthreads();
        if (ygen) {
          const int mlo = q * MX_MC;
          for (int mm = MX_MC - 1; mm >= 0; --mm) {
            const int m = mlo + mm;
            double re = 0.0, im = 0.0;
            if (m <= l) {
              g.value(re, im);
              if (m > 0) {
                g.step(coef[m]);
                g.rescale();
              }
            }
            Ys[(2 * mm) * MX_LDP + tid] = (float)re;
            Ys[(2 * mm + 1) * MX_LDP + tid] = (float)im;
          }
        }
        __syncthreads();
        const float *Ap = Js + (16 * wave + (lane & 15)) * MX_LDP + (lane >> 4);
        const float *B0 = Ys + (lane & 15) * MX_LDP + (lane >> 4);
        const float *B1 = B0 + 16 * MX_LDP;
        v4f c0 = acc[qq][0], c1 = acc[qq][1];
#pragma unroll 8
        for (int s = 0; s < MX_TP / 4; ++s) {
          const float a = Ap[4 * s];
          c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, B0[4 * s], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, B1[4 * s], c1, 0, 0, 0);
        }
        acc[qq][0] = c0;
        acc[qq][1] = c1;
      }
    }
    // ---- 3. fold the tile's FP32 sums into the FP64 accumulators
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int j = 0; j < 4; ++j) sum[qq][c][j] += (double)acc[qq][c][j];
        acc[qq][c] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    __syncthreads();
  }

  // ---- write the item's partial block (64 rows x its columns), FP64
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qhi - 1 - qq;
    if (q >= it.qlo) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int col = q * 2 * MX_MC + 16 * c + (lane & 15);
        // v_mfma_f32_16x16x4f32 result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = 16 * wave + 4 * (lane >> 4) + j;
          if (col < ncol) out[(size_t)row * ncol + col] = sum[qq][c][j];
        }
      }
    }
  }
}

}  // namespace sfb
