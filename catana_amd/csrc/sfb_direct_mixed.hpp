// sfb_direct_mixed.hpp -- mixed-precision fused direct SFB kernel.
//
// Same contraction as the FP64 kernel (per l: Bessel rows x Y* columns summed
// over points, sfb_raw.cpp:61-80), with the operands generated cheaply:
//   * j_l(k_ln r) from a per-l table of (j_l(x_i), h j_l'(x_i)) on the grid
//     x_i = i h (h = 1/16) held in LDS as FP32 and evaluated by cubic Hermite
//     interpolation (error <= h^4/384 max|j_l''''| ~ 4e-8 of the local
//     amplitude).  The argument t = r k_ln / h and its split into node index
//     and fraction are formed in FP64, so the phase error does not grow with
//     k r.  The table of l reaches k_{l,n_l-1} max(rmax, max_i r_i), so every
//     argument of the catalogue lies inside it.
//   * Y*_lm from the same FP64 downward-m recurrence as the FP64 kernel.
//   * each wave evaluates its Bessel values straight into MFMA A-operand
//     registers (it owns 16 rows of n, so every value is computed once and
//     never staged); the Y* chunk is staged in LDS as the B operand.
//   * contraction on the matrix cores, in one of two operand formats:
//       F16X3 = false: FP32 operands, v_mfma_f32_16x16x4f32;
//       F16X3 = true : each operand split as v*2^12 = hi + lo (two FP16
//                      values, relative residual 2^-22) and
//                      hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16.
//     Each 256-point tile is summed in FP32 and then added into FP64
//     accumulators, so FP32 rounding never sees more than 256 terms.
// Work item: (l, 64 rows of n, m chunks [qlo, qhi), a point range).  Chunks
// are counted from the top: chunk c holds m = l - 16c - mm, mm = 0..15, so
// only the lowest chunk is partial.  All 256 threads generate the Legendre rows
// (one point each); wave w owns rows 16w..16w+15 and both 16-column halves of
// every 32-column chunk.  Waves and column halves that only cover padding skip
// their MFMAs.
#pragma once

#include "sfb_common.hpp"
#include "sfb_math.hpp"
#include "sfb_ylm.hpp"

namespace sfb {

constexpr int MX_NT = 64;            // rows (n) per item
constexpr int MX_TP = 256;           // points per tile
constexpr int MX_MC = 16;            // m values per chunk (32 real columns)
constexpr double MX_H = 1.0 / 16.0;  // Bessel table spacing
constexpr float MX_SCALE = 4096.f;   // FP16 operands carry v * 2^12
constexpr int MX_LMAX = 512;         // recurrence coefficients of l < MX_LMAX fit in LDS
// LDS: Y* chunk (<= 2 x 32 x 288 halves = 36 KB) + radii (2 KB) + coefficients (8 KB) + table
constexpr int MX_TAB_MAX = 14336;    // table nodes (x up to 896)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

// (a, b) -> FP16 pairs (hi, lo) with a ~= hi.x + lo.x, b ~= hi.y + lo.y: two
// v_cvt_pk_f16_f32 (round to nearest) and the FP32 remainders.
__device__ __forceinline__ void split_pair(float a, float b, h2 &hi, h2 &lo) {
  hi = __builtin_convertvector((f2){a, b}, h2);
  const f2 back = __builtin_convertvector(hi, f2);
  lo = __builtin_convertvector((f2){a - back.x, b - back.y}, h2);
}

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

// j_l(t h) from the LDS table (0 <= t < nn - 1 in units of h; the index is
// clamped so that a rounding at the last node stays inside the table).
__device__ __forceinline__ float jl_hermite(const float2 *T, int nn, double t) {
  int i = (int)t;
  i = i < nn - 2 ? i : nn - 2;
  const float f = (float)(t - (double)i);
  const float2 a = T[i], b = T[i + 1];
  const float d = b.x - a.x;
  const float c2 = 3.f * d - 2.f * a.y - b.y;
  const float c3 = -2.f * d + a.y + b.y;
  return fmaf(f, fmaf(f, fmaf(f, c3, c2), a.y), a.x);
}

// Write Y*_lm of the generator's current m into local columns (2mm, 2mm+1).
template <bool F16X3, int LDP>
__device__ __forceinline__ void emit_y(const YlmDown &y, float *Yf, _Float16 *Yhi, _Float16 *Ylo, int mm, int tid) {
  if constexpr (F16X3) {
    // Y* * 2^12 formed in FP64 and rounded once to FP32, then split into two FP16 halves
    const double P = ldexp(y.v, y.E + 12);
    h2 hi, lo;
    split_pair((float)(P * y.cm), (float)(-P * y.sm), hi, lo);
    Yhi[(2 * mm) * LDP + tid] = hi.x;
    Yhi[(2 * mm + 1) * LDP + tid] = hi.y;
    Ylo[(2 * mm) * LDP + tid] = lo.x;
    Ylo[(2 * mm + 1) * LDP + tid] = lo.y;
  } else {
    double re, im;
    y.value(re, im);
    Yf[(2 * mm) * LDP + tid] = (float)re;
    Yf[(2 * mm + 1) * LDP + tid] = (float)im;
  }
}

template <bool F16X3, int LDP>
__device__ __forceinline__ void emit_zero(float *Yf, _Float16 *Yhi, _Float16 *Ylo, int mm, int tid) {
  if constexpr (F16X3) {
    Yhi[(2 * mm) * LDP + tid] = (_Float16)0.f;
    Yhi[(2 * mm + 1) * LDP + tid] = (_Float16)0.f;
    Ylo[(2 * mm) * LDP + tid] = (_Float16)0.f;
    Ylo[(2 * mm + 1) * LDP + tid] = (_Float16)0.f;
  } else {
    Yf[(2 * mm) * LDP + tid] = 0.f;
    Yf[(2 * mm + 1) * LDP + tid] = 0.f;
  }
}

template <int QMAX, bool F16X3>
__global__ __launch_bounds__(256) void sfb_direct_mixed_kernel(
    PointsSoA pts, const DirectItem *__restrict__ items, const double *__restrict__ kln, int nmax,
    LegendreTables leg, HermiteTables ht, double *__restrict__ partial) {
  constexpr int LDP = F16X3 ? MX_TP + 32 : MX_TP + 4;  // row pitch of the Y* chunk (elements)
  constexpr int YBYTES = F16X3 ? 2 * 2 * MX_MC * LDP * 2 : 2 * MX_MC * LDP * 4;
  extern __shared__ float2 T[];  // Bessel table of this item's l (dynamic: largest of the launch)
  __shared__ __attribute__((aligned(16))) char ybuf[YBYTES];
  __shared__ double rs[MX_TP];
  __shared__ double2 cf[MX_LMAX];  // recurrence coefficients (a_m, b_m) of this l
  float *Yf = reinterpret_cast<float *>(ybuf);
  _Float16 *Yhi = reinterpret_cast<_Float16 *>(ybuf);
  _Float16 *Ylo = Yhi + 2 * MX_MC * LDP;

  const DirectItem it = items[blockIdx.x];
  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4;
  const int row = 16 * wave + (lane & 15);
  const bool wave_rows = 16 * wave < it.nrows;  // wave-uniform
  const bool rowok = row < it.nrows;

  const int nn = ht.nnodes[l];
  const float2 *gt = ht.tab + ht.off[l];
  for (int i = tid; i < nn; i += 256) T[i] = gt[i];
  const double khr = rowok ? kln[(size_t)l * nmax + it.n0 + row] * (1.0 / MX_H) : 0.0;
  const double2 sect = leg.sect[l];
  {
    const double2 *gc = leg.coef + leg.coef_off[l];
    for (int m = tid; m <= l; m += 256) cf[m] = gc[m];
  }

  v4f acc[QMAX][2];
  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[q][c] = v4f{0.f, 0.f, 0.f, 0.f};
      sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  for (int64_t base = it.p0; base < it.p1; base += MX_TP) {
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    rs[tid] = pok ? pts.r[gp] : -1.0;
    __syncthreads();  // rs and (first tile) T ready; previous tile's readers done

    // ---- 1. Bessel values into A-operand registers
    float af[F16X3 ? 1 : MX_TP / 4];
    h8 ahi[F16X3 ? MX_TP / 32 : 1], alo[F16X3 ? MX_TP / 32 : 1];
    if (wave_rows) {
      if constexpr (F16X3) {
#pragma unroll
        for (int s = 0; s < MX_TP / 32; ++s) {
          h2 hp[4], lp[4];
#pragma unroll
          for (int e = 0; e < 8; e += 2) {
            const double r0 = rs[32 * s + 8 * g + e], r1 = rs[32 * s + 8 * g + e + 1];
            const float v0 = (rowok && r0 >= 0.0) ? jl_hermite(T, nn, r0 * khr) * MX_SCALE : 0.f;
            const float v1 = (rowok && r1 >= 0.0) ? jl_hermite(T, nn, r1 * khr) * MX_SCALE : 0.f;
            split_pair(v0, v1, hp[e / 2], lp[e / 2]);
          }
          ahi[s] = __builtin_shufflevector(__builtin_shufflevector(hp[0], hp[1], 0, 1, 2, 3),
                                           __builtin_shufflevector(hp[2], hp[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5,
                                           6, 7);
          alo[s] = __builtin_shufflevector(__builtin_shufflevector(lp[0], lp[1], 0, 1, 2, 3),
                                           __builtin_shufflevector(lp[2], lp[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5,
                                           6, 7);
        }
      } else {
#pragma unroll
        for (int s = 0; s < MX_TP / 4; ++s) {
          const double r = rs[4 * s + g];
          af[s] = (rowok && r >= 0.0) ? jl_hermite(T, nn, r * khr) : 0.f;
        }
      }
    }

    // ---- 2. Legendre/azimuth rows of this thread's point, chunks from high m to low m
    YlmDown y;
    if (pok) {
      y.init(l, pts.x[gp], pts.s[gp], pts.phi[gp], pts.cphi[gp], pts.sphi[gp], sect);
      for (int m = l; m > l - MX_MC * it.qlo; --m) {  // down to the top of the first chunk
        y.step(cf[m]);
        y.rescale_nb();
      }
    } else {
      y.v = y.vp = 0.0;
      y.E = 0;
      y.cm = y.sm = y.cot = y.cph = y.sph = 0.0;
    }
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qlo + qq;  // chunk index from the top
      if (q < it.qhi) {
        __syncthreads();  // previous chunk's readers are done
        const int mtop = l - MX_MC * q;
        if (mtop >= MX_MC) {
          // full chunk: every m of it is >= 1, no per-step guards
#pragma unroll
          for (int mm = 0; mm < MX_MC; ++mm) {
            emit_y<F16X3, LDP>(y, Yf, Yhi, Ylo, mm, tid);
            y.step(cf[mtop - mm]);
            if ((mm & 3) == 3) y.rescale_nb();
          }
        } else {
          // lowest chunk: m = mtop .. 0 are generated, the rest are zero columns
          for (int mm = 0; mm < MX_MC; ++mm) {
            const int m = mtop - mm;
            if (m >= 0) {
              emit_y<F16X3, LDP>(y, Yf, Yhi, Ylo, mm, tid);
              if (m > 0) {
                y.step(cf[m]);
                y.rescale_nb();
              }
            } else {
              emit_zero<F16X3, LDP>(Yf, Yhi, Ylo, mm, tid);
            }
          }
        }
        __syncthreads();
        if (wave_rows) {
          // two independent accumulator chains (column halves); the second half
          // is skipped when it only covers padding (wave-uniform)
          const bool c1 = mtop >= 8;  // local columns 16..31 hold m = mtop-8 .. mtop-15
          const int col = lane & 15;
          v4f a0 = acc[qq][0], a1 = acc[qq][1];
          if constexpr (F16X3) {
#pragma unroll
            for (int s = 0; s < MX_TP / 32; ++s) {
              const int o = col * LDP + 32 * s + 8 * g;
              const h8 bh0 = *reinterpret_cast<const h8 *>(Yhi + o);
              const h8 bl0 = *reinterpret_cast<const h8 *>(Ylo + o);
              a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[s], bh0, a0, 0, 0, 0);
              a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[s], bl0, a0, 0, 0, 0);
              a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[s], bh0, a0, 0, 0, 0);
              if (c1) {
                const h8 bh1 = *reinterpret_cast<const h8 *>(Yhi + o + 16 * LDP);
                const h8 bl1 = *reinterpret_cast<const h8 *>(Ylo + o + 16 * LDP);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[s], bh1, a1, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[s], bl1, a1, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[s], bh1, a1, 0, 0, 0);
              }
            }
          } else {
            const float *bp = Yf + col * LDP + g;
#pragma unroll
            for (int s = 0; s < MX_TP / 4; ++s) {
              a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(af[s], bp[4 * s], a0, 0, 0, 0);
              if (c1) a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(af[s], bp[4 * s + 16 * LDP], a1, 0, 0, 0);
            }
          }
          acc[qq][0] = a0;
          acc[qq][1] = a1;
        }
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
  }

  // ---- write the item's partial block (64 rows x its columns), FP64
  const double scale = F16X3 ? 1.0 / ((double)MX_SCALE * (double)MX_SCALE) : 1.0;
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);       // local column: (mm, re/im)
        const int m = l - MX_MC * q - (lc >> 1);
        const int col = m >= 0 ? 2 * m + (lc & 1) : ncol;
        // v_mfma_f32_16x16x{4f32,32_f16} result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int orow = 16 * wave + 4 * g + j;
          if (col < ncol) out[(size_t)orow * ncol + col] = sum[qq][c][j] * scale;
        }
      }
    }
  }
}

}  // namespace sfb
