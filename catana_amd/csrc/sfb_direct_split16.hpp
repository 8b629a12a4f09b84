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
//     each FP16 remainder formed and rounded in one v_fma_mix{lo,hi}_f16; the
//     chunk is staged point-major (64-bit stores) and read by transposed LDS
//     loads.
// Measured and not kept: generating chunk q+1 into a second buffer while
// contracting chunk q in the same wave (11 % slower), a 512-thread form in
// which four waves generate Y* while four contract (1.5x slower), and
// 512-thread blocks for items of 65-128 rows, one Y* generation shared by all
// eight waves (2^25 points: 3.52 against 3.24 s): the two 256-thread blocks
// per CU already overlap one block's vector work with the other's
// matrix-core work, where a single block's waves meet at every barrier.
// Work item: (l, 64 rows of n, m chunks [qlo, qhi), point range), as for the
// 256-thread kernel of sfb_direct_mixed.hpp, including the point split KS for
// items of at most 32 (16) rows.
#pragma once

#include "sfb_common.hpp"
#include "sfb_direct_mixed.hpp"
#include "sfb_direct_mixed2.hpp"

namespace sfb {

constexpr int S16_GPAD = 16;  // g_m table padding below m = 0
// range of f served by one interval's cubic: [-0.2, 1.2], i.e. |f - 0.5| <= 0.7
constexpr float S16_FMID = 0.5f, S16_FHALF = 0.7f;

// Y* chunk in LDS, point-major: plane k (FP16 piece k) holds one row of 32
// columns (2 mm: real, 2 mm + 1: imaginary part of m = mtop - mm) per tile
// point, padded to S16_ROW elements (a 72-byte pitch keeps the row writes and
// the transposed fragment reads spread over the banks).  A generating thread
// writes its point's (re, im) pair of two m as one 64-bit store per plane
// (16-bit stores measured at about 4.5 cycles per wave instruction); the
// contraction reads each B fragment with two ds_read_b64_tr_b16, which hand
// lane 4a + c element c of the 8-byte rows addressed by lanes a, a+4, a+8,
// a+12 of its 16-lane group.
constexpr int S16_ROW = 36;                // FP16 elements per point row of a plane
constexpr int S16_PLANE = MX_TP * S16_ROW;  // one FP16 piece of a chunk
constexpr int S16_YBUF = 2 * S16_PLANE;     // one Y* chunk buffer (both pieces), FP16 elements

// The block's LDS areas.
struct S16Shared {
  unsigned short *y;  // Y* chunk: planes 0 (p0) and 1 (p1), point-major
  double *rs;         // tile radii
  float *dg;          // r - r_mid of the point's 8-point group, FP32
  float *gs;          // per 8-point group: max |r - r_mid|
  double *cg;         // g_m of this l at cg[m] (cg[m] = 0 for m <= 0)
};

typedef short s4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s4v lds_s4v;

// B fragment of v_mfma_f32_16x16x32_f16 from a point-major plane: lane 16 g + n
// receives column col0 + n of points p0 + 8 g + e, e = 0..7 (the A fragment's
// point order).  Lane 16 g + s supplies the row of point p0 + 8 g + 4 i + s / 4,
// columns col0 + 4 (s % 4) .. +3, for the two reads i = 0, 1.
__device__ __forceinline__ h8 s16_bfrag(const unsigned short *plane, int p0, int col0, int lane) {
  const int g = lane >> 4, s = lane & 15;
  const unsigned short *a = plane + (p0 + 8 * g + (s >> 2)) * S16_ROW + col0 + 4 * (s & 3);
  const s4v x0 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4v *)(a));
  const s4v x1 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4v *)(a + 4 * S16_ROW));
  return __builtin_bit_cast(h8, __builtin_shufflevector(x0, x1, 0, 1, 2, 3, 4, 5, 6, 7));
}

// Y* 2^12 of this thread's point for m = mtop - mm, mm = m0 .. m0 + 2 K - 1,
// into its rows of the two planes (64-bit stores of two m), stepping the
// generator.  In the lowest chunk the steps past m = 0 use g = 0 and the
// columns of m < 0 are computed and discarded.
template <int K>
__device__ __forceinline__ void s16_emit(YlmScaled &y, const double *cg, int mtop, int m0, unsigned short *Yw,
                                         int tid) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    unsigned hi[2], lo[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int mm = m0 + 2 * k + u;
      const double P = ldexp(y.v, y.E + 12);
      const float re = (float)(P * y.cm), im = (float)(-P * y.sm);
      split2(re, im, hi[u], lo[u]);
      y.step(cg[mtop - mm]);
      if ((mm & 7) == 7) y.rescale_nb();
    }
    const int e = tid * S16_ROW + 2 * (m0 + 2 * k);
    *reinterpret_cast<uint2 *>(Yw + e) = make_uint2(hi[0], hi[1]);
    *reinterpret_cast<uint2 *>(Yw + S16_PLANE + e) = make_uint2(lo[0], lo[1]);
  }
}

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
  constexpr int MPS = MX_MC / SPW;  // m values generated per contraction step
  static_assert(MPS % 2 == 0, "m values are stored in pairs");
  constexpr int FOLD = 64;          // points per FP32 leading-product partial sum
  v4f a0 = v4f{0.f, 0.f, 0.f, 0.f}, a1 = v4f{0.f, 0.f, 0.f, 0.f};
  v4f c0 = v4f{0.f, 0.f, 0.f, 0.f}, c1 = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < SPW; ++st) {
    if constexpr (MF) {
      const int p0 = 32 * (kp * SPW + st);
      const h8 b00 = s16_bfrag(Yr, p0, 0, lane);
      const h8 b01 = s16_bfrag(Yr + S16_PLANE, p0, 0, lane);
      const h8 b10 = s16_bfrag(Yr, p0, 16, lane);
      const h8 b11 = s16_bfrag(Yr + S16_PLANE, p0, 16, lane);
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
    if constexpr (GEN) s16_emit<MPS / 2>(y, cg, mtop, st * MPS, Yw, tid);
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
  const float4 *__restrict__ T = ct.tab16 + ct.off[l];  // coefficients of j_l 2^12
  // rows past nrows take the first row's k (their results are not read)
  const double khr = kln[it.kofs + (rowok ? row : 0)] * (1.0 / MX_H);
  const float khr32 = (float)khr;
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

  // padding points read the item's first point (finite Bessel values; their
  // Y* columns are zero, so they add nothing)
  double rn = pts.r[it.p0 + tid < it.p1 ? it.p0 + tid : it.p0];
  for (int64_t base = it.p0; base < it.p1; base += MX_TP) {
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    const int64_t gq = pok ? gp : it.p0;
    const double r = rn;
    {
      // the next tile's radius, loaded now so that it arrives during this tile
      const int64_t gn = gp + MX_TP;
      rn = pts.r[gn < it.p1 ? gn : it.p0];
    }
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
      // every step's interval and coefficients first, so that the SPW table
      // loads are in flight together (loaded inside each step's branch, each
      // load was waited for in turn)
      float4 cfs[SPW];
      float fms[SPW];
#pragma unroll
      for (int st = 0; st < SPW; ++st) {
        const int j = 4 * (kp * SPW + st) + g;  // 8-point group: tile points 8j .. 8j+7
        const double t = sh.rs[8 * j + 4] * khr;
        int i = (int)t;
        i = i < nint - 1 ? i : nint - 1;
        fms[st] = (float)(t - (double)i);
        cfs[st] = T[i];
      }
#pragma unroll
      for (int st = 0; st < SPW; ++st) {
        const int j = 4 * (kp * SPW + st) + g;
        const float fm = fms[st];
        const float sp = sh.gs[j] * khr32;
        const float4 cf = cfs[st];
        unsigned hw[4], lw[4];
        if (__all(fabsf(fm - S16_FMID) + sp <= S16_FHALF)) {
          const float c0 = cf.x, c1 = cf.y, c2 = cf.z, c3 = cf.w;
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
          for (int e = 0; e < 8; ++e) v[e] = jl_cubic_scaled(T, nint, sh.rs[8 * j + e] * khr);
#pragma unroll
          for (int e = 0; e < 8; e += 2) split2(v[e], v[e + 1], hw[e / 2], lw[e / 2]);
        }
        ah[0][st] = __builtin_bit_cast(h8, (u4v){hw[0], hw[1], hw[2], hw[3]});
        ah[1][st] = __builtin_bit_cast(h8, (u4v){lw[0], lw[1], lw[2], lw[3]});
      }
    }

    // ---- 2. Y* generator of this thread's point, positioned at the item's top chunk
    YlmScaled y;
    {
      const int mstart = l - MX_MC * it.qlo;
      y.init(l, mstart, pts.x[gq], pts.s[gq], pts.phi[gq], pts.cphi[gq], pts.sphi[gq], sect);
      for (int m = l; m > mstart; m -= 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) y.step_p(cg[m - k]);
        y.rescale_nb();
      }
      if (!pok) y.zero();
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

// DirectItem::flags of the grouped kernel's items: the point split KS
enum : int32_t { ITEM_SPLIT_KS2 = 2, ITEM_SPLIT_KS4 = 4 };

// Work lanes of a persistent grouped launch: lane x holds items
// [off[x], off[x + 1]) of the launch's list and counter x hands them out.
constexpr int GROUPED_LANES = 8;
struct GroupedLanes {
  int32_t off[GROUPED_LANES + 1];
};

// Persistent grouped launch of split16_item for the items of one point split
// KS.  The items come in GROUPED_LANES lanes, one per XCD (the dispatcher
// places workgroup b on XCD b mod 8, and lane x holds the items of every
// eighth point range, so the blocks of one XCD share their points in its L2):
// a block pulls items of its own lane from that lane's counter (zeroed before
// the launch), then helps the other lanes in turn, and leaves once it has
// found all of them exhausted.  Each item runs with the class its m-chunk
// count (1, 2 or up to 4) selects, so items of every chunk count and
// catalogue share one launch with no tail per class.  One kernel per KS: a
// single kernel over all nine (chunks, KS) classes needs more than the 256
// registers of two blocks per CU and spills (196 bytes per lane); one KS with
// its three chunk counts spills at most 12 bytes per lane (KS = 1).
template <int KS>
__global__ __launch_bounds__(256, 2) void sfb_direct_split16_grouped_kernel(
    PointsSoA pts, const DirectItem *__restrict__ items, GroupedLanes lanes, int *__restrict__ counters,
    const double *__restrict__ kln, LegendreScaled leg, CubicTables ct, double *__restrict__ partial) {
  __shared__ __attribute__((aligned(16))) unsigned short ybuf[S16_YBUF];
  __shared__ double rs[MX_TP];
  __shared__ __attribute__((aligned(16))) float dg[MX_TP];
  __shared__ float gs[MX_TP / 8];
  __shared__ double cgs[S16_GPAD + MX_LMAX];
  __shared__ int next;
  const S16Shared sh{ybuf, rs, dg, gs, cgs};
  int x = (int)(blockIdx.x % GROUPED_LANES);
  for (int tried = 0; tried < GROUPED_LANES;) {
    if (threadIdx.x == 0) next = atomicAdd(counters + x, 1);
    __syncthreads();  // next written; the previous item's LDS readers are done
    const int i = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // every thread has read next before thread 0 overwrites it
    if (i >= lanes.off[x + 1] - lanes.off[x]) {  // lane exhausted: the next one
      x = x + 1 == GROUPED_LANES ? 0 : x + 1;
      ++tried;
      continue;
    }
    const DirectItem it = items[lanes.off[x] + i];
    const int nq = it.qhi - it.qlo;
    if (nq == 1) split16_item<1, KS>(it, pts, kln, leg, ct, partial, sh);
    else if (nq == 2) split16_item<2, KS>(it, pts, kln, leg, ct, partial, sh);
    else split16_item<4, KS>(it, pts, kln, leg, ct, partial, sh);
  }
}

}  // namespace sfb
