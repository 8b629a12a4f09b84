// sfb_direct_split16.hpp -- split-FP16 direct SFB kernel, 256 threads, Bessel
// values from one cubic per (row, 8-point group).
//
// The contraction is that of the FP64 kernel (per l: Bessel rows x Y* columns
// summed over points, sfb_raw.cpp:61-80) on v_mfma_f32_16x16x32_f16 with each
// operand v 2^12 split into two FP16 pieces (p0 q0 + p0 q1 + p1 q0, FP32
// partial sums folded into FP64 every 64 points), as in sfb_direct_mixed.hpp.
// What differs is how the operands are produced:
//   * Bessel values.  A lane of the A fragment holds one row n and, per 32-point
//     step, eight consecutive points of the radius-sorted tile.  Their
//     arguments t = k_n r / h (h = 1/16) lie close together: the lane forms t
//     in FP64 once, at the group's middle point, takes that interval's cubic
//     coefficients (c0..c3 of j_l on [i h, (i+1) h], one 16-byte load), and
//     evaluates the eight values at f = f_mid + (r_e - r_mid) k_n / h with the
//     offsets r_e - r_mid rounded once to FP32 (staged per tile in LDS): one
//     FMA for f and three for the cubic per value, instead of the FP64
//     argument split and an LDS gather per value.  The cubic Hermite error
//     f^2 (f-1)^2 h^4 |j''''| / 24 stays within its in-interval maximum for
//     f in [-0.2, 1.2], so the group is evaluated this way when its offsets
//     keep f there; otherwise (sparse radii) each value takes the per-point
//     FP64 path from the same table.  The choice is per wave and step.
//   * Y* rows.  Legendre functions by the two-operation scaled recurrence
//     Q^{m-1} = g_m cot(t) Q^m - Q^{m+1} (P^m = d_m Q^m, d_m applied to the
//     FP64 sums at the end), carried binary exponent, FP64; azimuth by the
//     Chebyshev recurrence in FP64; Y* 2^12 rounded once to FP32 and split,
//     each FP16 remainder formed and rounded in one v_fma_mix{lo,hi}_f16.
// Work item: (l, 64 rows of n, m chunks [qlo, qhi), point range), as for the
// 256-thread kernel of sfb_direct_mixed.hpp, including the point split KS for
// items of at most 32 (16) rows.
#pragma once

#include "sfb_common.hpp"
#include "sfb_direct_mixed.hpp"
#include "sfb_direct_mixed2.hpp"

namespace sfb {

constexpr int S16_GPAD = 16;   // g_m table padding below m = 0
constexpr int S16_LDP = MX_TP + 8;             // row pitch of a Y* column (FP16 elements)
constexpr int S16_PLANE = 2 * MX_MC * S16_LDP;  // one FP16 piece of a chunk
constexpr float S16_FLO = -0.2f, S16_FHI = 1.2f;  // range of f served by one interval's cubic

// The block's LDS areas.
struct S16Shared {
  unsigned short *y;  // Y* chunk: planes 0 (p0) and 1 (p1) of 32 columns x S16_LDP points
  double *rs;         // tile radii
  float *dg;          // r - r_mid of the point's 8-point group, FP32
  float *gs;          // per 8-point group: max |r - r_mid|
  double *cg;         // g_m of this l at cg[m] (cg[m] = 0 for m <= 0)
};

constexpr int S16_YBUF = 2 * S16_PLANE;  // one Y* chunk buffer (both pieces), FP16 elements

// One phase of a tile.  MF: contract the chunk in Yr (this wave's 16 rows x 32
// columns over its SPW 32-point steps: leading products p0 q0 in (a0, a1),
// folded into the FP64 sums every 64 points, correction products in (c0, c1),
// folded once).  GEN: generate a chunk (16 m from mtop down) of this thread's
// point into Yw, MX_MC / SPW m values per step.  All MFMAs and the reads of
// their results form one basic block.
template <bool MF, bool GEN, int SPW>
__device__ __forceinline__ void s16_phase(const h8 (&ah)[2][SPW], const unsigned short *Yr, v4d (&sum)[2],
                                          YlmScaled &y, const double *cg, int mtop, unsigned short *Yw, int tid,
                                          int lane, int kp, int g) {
  constexpr int LDP = S16_LDP;
  constexpr int PLANE = S16_PLANE;
  constexpr int MPS = MX_MC / SPW;  // m values generated per contraction step
  constexpr int FOLD = 64;          // points per FP32 leading-product partial sum
  const int col = lane & 15;
  v4f a0 = v4f{0.f, 0.f, 0.f, 0.f}, a1 = v4f{0.f, 0.f, 0.f, 0.f};
  v4f c0 = v4f{0.f, 0.f, 0.f, 0.f}, c1 = v4f{0.f, 0.f, 0.f, 0.f};
  const _Float16 *Yh = reinterpret_cast<const _Float16 *>(Yr);
#pragma unroll
  for (int st = 0; st < SPW; ++st) {
    if constexpr (MF) {
      const int o = col * LDP + 32 * (kp * SPW + st) + 8 * g;
      const h8 b00 = *reinterpret_cast<const h8 *>(Yh + o);
      const h8 b01 = *reinterpret_cast<const h8 *>(Yh + PLANE + o);
      const h8 b10 = *reinterpret_cast<const h8 *>(Yh + o + 16 * LDP);
      const h8 b11 = *reinterpret_cast<const h8 *>(Yh + PLANE + o + 16 * LDP);
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][st], b01, c0, 0, 0, 0);
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][st], b00, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][st], b11, c1, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][st], b10, c1, 0, 0, 0);
      a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][st], b00, a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][st], b10, a1, 0, 0, 0);
      if ((32 * (st + 1)) % FOLD == 0 && st + 1 < SPW) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          sum[0][j] += (double)a0[j];
          sum[1][j] += (double)a1[j];
        }
        a0 = v4f{0.f, 0.f, 0.f, 0.f};
        a1 = v4f{0.f, 0.f, 0.f, 0.f};
      }
    }
    if constexpr (GEN) {
      // m = mtop - mm; in the lowest chunk the steps past m = 0 use g = 0 and
      // the columns of m < 0 are computed and discarded
#pragma unroll
      for (int k = 0; k < MPS; ++k) {
        const int mm = st * MPS + k;
        const double P = ldexp(y.v, y.E + 12);
        const float re = (float)(P * y.cm), im = (float)(-P * y.sm);
        unsigned hi, lo;
        split2(re, im, hi, lo);
        Yw[(2 * mm) * LDP + tid] = (unsigned short)hi;
        Yw[(2 * mm + 1) * LDP + tid] = (unsigned short)(hi >> 16);
        Yw[PLANE + (2 * mm) * LDP + tid] = (unsigned short)lo;
        Yw[PLANE + (2 * mm + 1) * LDP + tid] = (unsigned short)(lo >> 16);
        y.step(cg[mtop - mm]);
        if ((mm & 7) == 7) y.rescale_nb();
      }
    }
  }
  if constexpr (MF) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sum[0][j] += (double)a0[j] + (double)c0[j];
      sum[1][j] += (double)a1[j] + (double)c1[j];
    }
  }
}

template <int QMAX, int KS>
__device__ __forceinline__ void split16_item(const DirectItem &it, const PointsSoA &pts,
                                             const double *__restrict__ kln, const LegendreScaled &leg,
                                             const CubicTables &ct, double *__restrict__ partial,
                                             const S16Shared &sh) {
  static_assert(KS == 1 || KS == 2 || KS == 4, "point split");
  constexpr int RBN = 4 / KS;           // row blocks of 16 per item
  constexpr int SPW = MX_TP / 32 / KS;  // 32-point steps of a tile per wave

  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4;
  const int rb = wave % RBN, kp = wave / RBN;
  const int row = 16 * rb + (lane & 15);
  const bool wave_rows = 16 * rb < it.nrows;  // wave-uniform
  const bool rowok = row < it.nrows;

  const int nint = ct.nint[l];
  const float4 *__restrict__ T = ct.tab + ct.off[l];
  // rows past nrows take the first row's k (their results are not read)
  const double khr = kln[it.kofs + (rowok ? row : 0)] * (1.0 / MX_H);
  const float khr32 = (float)khr;
  const double rlo = pts.r[it.p0];
  const double2 sect = leg.sect[l];
  double *cg = sh.cg + S16_GPAD;
  {
    const double *gg = leg.g + leg.off[l];
    for (int m = tid - S16_GPAD; m <= l; m += 256) cg[m] = m > 0 ? gg[m] : 0.0;
  }
  unsigned short *Y = sh.y;

  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int64_t base = it.p0; base < it.p1; base += MX_TP) {
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    // padding points take the item's first radius (finite Bessel values; their
    // Y* columns are zero, so they add nothing)
    const double r = pok ? pts.r[gp] : rlo;
    // offset from the middle point (lane 4) of this thread's 8-point group and
    // the group's largest |offset|; the group is 8 consecutive lanes of one wave
    const double rmid = __shfl(r, (lane & ~7) | 4);
    const float d = (float)(r - rmid);
    float s = fabsf(d);
    s = fmaxf(s, __shfl_xor(s, 1));
    s = fmaxf(s, __shfl_xor(s, 2));
    s = fmaxf(s, __shfl_xor(s, 4));
    sh.rs[tid] = r;
    sh.dg[tid] = d;
    if ((lane & 7) == 0) sh.gs[tid >> 3] = s;
    __syncthreads();  // tile data ready; previous tile's readers done

    // ---- 1. Bessel values of this lane's row into A fragments (pieces 0, 1)
    h8 ah[2][SPW];
    if (wave_rows) {
#pragma unroll
      for (int st = 0; st < SPW; ++st) {
        const int j = 4 * (kp * SPW + st) + g;  // 8-point group: tile points 8j .. 8j+7
        const double t = sh.rs[8 * j + 4] * khr;
        int i = (int)t;
        i = i < nint - 1 ? i : nint - 1;
        const float fm = (float)(t - (double)i);
        const float sp = sh.gs[j] * khr32;
        const float4 cf = T[i];
        unsigned hw[4], lw[4];
        if (__all((fm - sp >= S16_FLO) && (fm + sp <= S16_FHI))) {
          const float c0 = cf.x * MX_SCALE, c1 = cf.y * MX_SCALE, c2 = cf.z * MX_SCALE, c3 = cf.w * MX_SCALE;
          const float4 da = *reinterpret_cast<const float4 *>(sh.dg + 8 * j);
          const float4 db = *reinterpret_cast<const float4 *>(sh.dg + 8 * j + 4);
          const float dd[8] = {da.x, da.y, da.z, da.w, db.x, db.y, db.z, db.w};
          float v[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float f = fmaf(dd[e], khr32, fm);
            v[e] = fmaf(fmaf(fmaf(c3, f, c2), f, c1), f, c0);
          }
#pragma unroll
          for (int e = 0; e < 8; e += 2) split2(v[e], v[e + 1], hw[e / 2], lw[e / 2]);
        } else {
          float v[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = jl_cubic_global(T, nint, sh.rs[8 * j + e] * khr);
#pragma unroll
          for (int e = 0; e < 8; e += 2) split2(v[e], v[e + 1], hw[e / 2], lw[e / 2]);
        }
        ah[0][st] = __builtin_bit_cast(h8, (u4v){hw[0], hw[1], hw[2], hw[3]});
        ah[1][st] = __builtin_bit_cast(h8, (u4v){lw[0], lw[1], lw[2], lw[3]});
      }
    }

    // ---- 2. Y* generator of this thread's point, positioned at the item's top chunk
    YlmScaled y;
    if (pok) {
      const int mstart = l - MX_MC * it.qlo;
      y.init(l, mstart, pts.x[gp], pts.s[gp], pts.phi[gp], pts.cphi[gp], pts.sphi[gp], sect);
      for (int m = l; m > mstart; m -= 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) y.step_p(cg[m - k]);
        y.rescale_nb();
      }
    } else {
      y.zero();
    }
    // per chunk: all threads generate it into LDS, then the waves with rows
    // contract it (generating the next chunk while contracting this one, from
    // a second buffer, measured 11 % slower: the two blocks per CU already
    // overlap one block's generation with the other's contraction)
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qlo + qq;  // chunk index from the top
      if (q < it.qhi) {           // block-uniform
        __syncthreads();          // previous chunk's readers are done
        s16_phase<false, true, SPW>(ah, Y, sum[qq], y, cg, l - MX_MC * q, Y, tid, lane, kp, g);
        __syncthreads();
        if (wave_rows) {
          // the contracting wave takes issue priority over the other block's
          // generating wave on its SIMD, so the matrix core is fed first
          __builtin_amdgcn_s_setprio(1);
          s16_phase<true, false, SPW>(ah, Y, sum[qq], y, cg, 0, Y, tid, lane, kp, g);
          __builtin_amdgcn_s_setprio(0);
        }
      }
    }
  }

  if constexpr (KS > 1) {
    // parts 1 .. KS-1 of each row block hand their sums to part 0 through LDS
    // (the Y* area), one chunk at a time; part 0 adds them in part order
    double *red = reinterpret_cast<double *>(Y);
    static_assert((KS - 1) * RBN * 64 * 8 * (int)sizeof(double) <= S16_YBUF * 2, "reduction area");
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      if (it.qlo + qq < it.qhi) {  // block-uniform
        __syncthreads();
        if (kp > 0 && wave_rows) {
          double *dd = red + (size_t)(((kp - 1) * RBN + rb) * 64 + lane) * 8;
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) dd[4 * c + j] = sum[qq][c][j];
        }
        __syncthreads();
        if (kp == 0 && wave_rows) {
#pragma unroll
          for (int k = 1; k < KS; ++k) {
            const double *dd = red + (size_t)(((k - 1) * RBN + rb) * 64 + lane) * 8;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int j = 0; j < 4; ++j) sum[qq][c][j] += dd[4 * c + j];
          }
        }
      }
    }
  }

  // ---- the item's partial block (64 rows x its columns), FP64, with d_m applied
  const double *dm = leg.d + leg.off[l];
  constexpr double unscale = 1.0 / ((double)MX_SCALE * (double)MX_SCALE);
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi && kp == 0) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);  // local column: (mm, re/im)
        const int m = l - MX_MC * q - (lc >> 1);
        if (m >= 0) {
          const int cc = 2 * m + (lc & 1);
          const double sc = dm[m] * unscale;
          // v_mfma_f32_16x16x32_f16 result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int orow = 16 * rb + 4 * g + j;
            out[(size_t)orow * ncol + cc] = sum[qq][c][j] * sc;
          }
        }
      }
    }
  }
}

// One item per block (blockIdx.x), all of one class (QMAX, KS).
template <int QMAX, int KS>
__global__ __launch_bounds__(256, 2) void sfb_direct_split16_kernel(PointsSoA pts, const DirectItem *__restrict__ items,
                                                                 const double *__restrict__ kln, LegendreScaled leg,
                                                                 CubicTables ct, double *__restrict__ partial) {
  __shared__ __attribute__((aligned(16))) unsigned short ybuf[S16_YBUF];
  __shared__ double rs[MX_TP];
  __shared__ __attribute__((aligned(16))) float dg[MX_TP];
  __shared__ float gs[MX_TP / 8];
  __shared__ double cgs[S16_GPAD + MX_LMAX];
  const DirectItem it = items[blockIdx.x];
  split16_item<QMAX, KS>(it, pts, kln, leg, ct, partial, S16Shared{ybuf, rs, dg, gs, cgs});
}

}  // namespace sfb
