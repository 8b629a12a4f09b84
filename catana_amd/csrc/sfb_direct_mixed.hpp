// sfb_direct_mixed.hpp -- mixed-precision fused direct SFB kernel.
//
// Same contraction as the FP64 kernel (per l: Bessel rows x Y* columns summed
// over points, sfb_raw.cpp:61-80), with the operands generated cheaply:
//   * j_l(k_ln r) by cubic Hermite interpolation of (j_l(x_i), h j_l'(x_i)) on
//     the grid x_i = i h (h = 1/16), FP32 (error <= h^4/384 max|j_l''''| ~ 4e-8
//     of the local amplitude).  The nodes an item's arguments reach are copied
//     into LDS (pre-scaled by the FP16 operand scale) when they fit two blocks
//     per CU; otherwise the item reads the global table, one float4 (j_i, h j_i',
//     j_{i+1}, h j_{i+1}') per interval.  The points are sorted by radius
//     beforehand, so the lanes of a row touch a few neighbouring intervals.  The
//     argument t = r k_ln / h and its split into interval index and fraction
//     are formed in FP64, so the phase error does not grow with k r.  The table
//     of l reaches k_{l,n_l-1} max(rmax, max_i r_i) plus two nodes, padding
//     points have r = 0 and padding rows k = 0, so a lookup needs no guard.
//   * Y*_lm from the same FP64 downward-m recurrence as the FP64 kernel (the
//     azimuth by the Chebyshev recurrence); the recurrence coefficients of l
//     are staged in LDS.
//   * each wave evaluates its Bessel values straight into MFMA A-operand
//     registers (it owns 16 rows of n, so every value is computed once and
//     never staged); the Y* chunk is staged in LDS as the B operand.
//   * contraction on the matrix cores, in one of three operand formats
//     (NP = number of FP16 pieces per operand):
//       NP = 0: FP32 operands, v_mfma_f32_16x16x4f32;
//       NP = 2: each operand split as v*2^12 = p0 + p1 (two FP16 values,
//               22 bits) and p0*p0 + p0*p1 + p1*p0 on v_mfma_f32_16x16x32_f16
//               (relative error per product ~2^-22);
//       NP = 3: three FP16 pieces (33 bits, exact for the FP32 values) and the
//               six products of order <= 2^-22 -- p0p0, p0p1, p1p0, p0p2, p1p1,
//               p2p0 -- so that a product is as accurate as with FP32 operands.
//     FP32 partial sums are added into FP64 accumulators every 64 points
//     (FP32 rounding in the matrix-core accumulation is the largest error
//     term otherwise); with FP16 pieces the leading products and the
//     correction products accumulate separately, so only the leading chain
//     needs the frequent fold.
// Work item: (l, 64 rows of n, m chunks [qlo, qhi), a point range).  Chunks
// are counted from the top: chunk c holds m = l - 16c - mm, mm = 0..15, so
// only the lowest chunk is partial.  All 256 threads generate the Legendre rows
// (one point each); wave w owns rows 16w..16w+15 and both 16-column halves of
// every 32-column chunk.  Waves whose rows are all padding skip their MFMAs.
// An item of at most 16 (32) rows runs with KS = 4 (2): the waves then split
// each tile's points instead of the rows -- wave w owns row block w % (4/KS)
// and the points of part w / (4/KS) -- so that every wave evaluates Bessel
// values and runs MFMAs; the parts' FP64 sums are added in LDS at the end, in
// a fixed order.
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
constexpr int MX_TAB_MAX = 1 << 24;  // table intervals per l (x up to 2^20)
constexpr int MX_LMAX = 512;         // recurrence coefficients of l < MX_LMAX fit in LDS (8 KB)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

// a - (float)h.x and b - (float)h.y, each in one v_fma_mix_f32 (the FP16
// operand read in place; the remainder of a rounding is exact in FP32)
__device__ __forceinline__ f2 rem_f16(float a, float b, h2 h) {
  const unsigned hb = __builtin_bit_cast(unsigned, h);
  float ra, rb;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(ra) : "v"(a), "v"(hb));
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(rb) : "v"(b), "v"(hb));
  return (f2){ra, rb};
}

// (a, b) -> FP16 pairs (hi, lo) with a ~= hi.x + lo.x, b ~= hi.y + lo.y: two
// v_cvt_pk_f16_f32 (round to nearest) and the FP32 remainders.
__device__ __forceinline__ void split_pair(float a, float b, h2 &hi, h2 &lo) {
  hi = __builtin_convertvector((f2){a, b}, h2);
  lo = __builtin_convertvector(rem_f16(a, b, hi), h2);
}

// (a, b) -> three FP16 pairs with a = p0.x + p1.x + p2.x (and b likewise) to
// 33 bits, i.e. exactly for FP32 inputs away from the FP16 range limits; the
// remainders after each rounding are exact in FP32.
__device__ __forceinline__ void split3_pair(float a, float b, h2 &p0, h2 &p1, h2 &p2) {
  p0 = __builtin_convertvector((f2){a, b}, h2);
  f2 back = __builtin_convertvector(p0, f2);
  const f2 r1 = (f2){a - back.x, b - back.y};
  p1 = __builtin_convertvector(r1, h2);
  back = __builtin_convertvector(p1, f2);
  p2 = __builtin_convertvector((f2){r1.x - back.x, r1.y - back.y}, h2);
}

// Per-l Hermite tables: interval i of l at tab[off[l] + i] =
// (j_l(i h), h j_l'(i h), j_l((i+1) h), h j_l'((i+1) h)), i < nint[l].
struct HermiteTables {
  const float4 *tab;
  const int32_t *off;
  const int32_t *nint;
};

__device__ inline float2 hermite_node(int l, int i) {
  const double x = i * MX_H;
  double j, dj;
  if (x == 0.0) {
    j = (l == 0) ? 1.0 : 0.0;
    dj = (l == 1) ? 1.0 / 3.0 : 0.0;
  } else {
    j = sbessel_jl(l, x);
    dj = (double)l / x * j - sbessel_jl(l + 1, x);
  }
  return make_float2((float)j, (float)(MX_H * dj));
}

// j_l and h*j_l' at both ends of each interval, FP64 rounded to FP32.
__global__ void hermite_table_kernel(int lmax, const int32_t *__restrict__ off, const int32_t *__restrict__ nint,
                                     float4 *__restrict__ tab) {
  const int l = blockIdx.y;
  if (l >= lmax) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nint[l]; i += gridDim.x * blockDim.x) {
    const float2 a = hermite_node(l, i), b = hermite_node(l, i + 1);
    tab[off[l] + i] = make_float4(a.x, a.y, b.x, b.y);
  }
}

// j_l(t h) from the table of l (0 <= t <= nint in units of h; the index is
// clamped so that a rounding at the last node stays inside the table).
__device__ __forceinline__ float jl_hermite(const float4 *__restrict__ T, int nint, double t) {
  int i = (int)t;
  i = i < nint - 1 ? i : nint - 1;
  i = i > 0 ? i : 0;
  const float f = (float)(t - (double)i);
  const float4 q = T[i];
  // p(f) = j0 + f (s0 + f (c2 + f c3)) with d = j1 - j0, c3 = s0 + s1 - 2 d,
  // c2 = 3 d - 2 s0 - s1, and c2 + f c3 = (d - s0) + (f - 1) c3
  const float d = q.z - q.x;
  const float c3 = fmaf(-2.f, d, q.y + q.w);
  return fmaf(f, fmaf(f, fmaf(f - 1.f, c3, d - q.y), q.y), q.x);
}

// The same from a window of nodes copied into LDS: T[j] = node lo + j, node
// i = (j_l(i h), h j_l'(i h)), at t = r khr - lo (one rounding).  No clamping:
// the caller guarantees 0 <= t and that nodes floor(t) and floor(t) + 1 are staged.
__device__ __forceinline__ float jl_hermite_lds(const float2 *T, double r, double khr, double lo) {
  const double t = fma(r, khr, -lo);
  const int i = (int)t;
  const float f = (float)__builtin_amdgcn_fract(t);  // t - floor(t), exact
  const float2 a = T[i], b = T[i + 1];
  const float d = b.x - a.x;
  const float c3 = fmaf(-2.f, d, a.y + b.y);
  return fmaf(f, fmaf(f, fmaf(f - 1.f, c3, d - a.y), a.y), a.x);
}

// Write Y*_lm of the generator's current m into local columns (2mm, 2mm+1)
// of each operand piece (Yf: FP32; Yh: NP FP16 planes of 2*MX_MC*LDP).
template <int NP, int LDP>
__device__ __forceinline__ void emit_y(const YlmDown &y, float *Yf, _Float16 *Yh, int mm, int tid) {
  if constexpr (NP == 0) {
    double re, im;
    y.value(re, im);
    Yf[(2 * mm) * LDP + tid] = (float)re;
    Yf[(2 * mm + 1) * LDP + tid] = (float)im;
  } else {
    // Y* * 2^12 formed in FP64 and rounded once to FP32, then split into FP16 pieces
    const double P = ldexp(y.v, y.E + 12);
    const float re = (float)(P * y.cm), im = (float)(-P * y.sm);
    h2 pc[3];
    if constexpr (NP == 2)
      split_pair(re, im, pc[0], pc[1]);
    else
      split3_pair(re, im, pc[0], pc[1], pc[2]);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      Yh[k * 2 * MX_MC * LDP + (2 * mm) * LDP + tid] = pc[k].x;
      Yh[k * 2 * MX_MC * LDP + (2 * mm + 1) * LDP + tid] = pc[k].y;
    }
  }
}

// Row pitch of the Y* chunk (elements): a column starts 4 banks after the
// previous one, so the 16-byte B-fragment reads of 8 consecutive columns
// (FP16) and the 4-byte reads of 16 columns x 4 points (FP32) hit distinct banks.
template <int NP>
struct MixedLds {
  static constexpr int LDP = NP ? MX_TP + 8 : MX_TP + 4;
  static constexpr int PLANE = 2 * MX_MC * LDP;  // elements of one piece of the chunk
  static constexpr int YBYTES = NP ? NP * PLANE * 2 : PLANE * 4;
};

// ItemShared: the block's LDS areas -- the Y* chunk (ybuf, YBYTES), the tile's
// radii (rs, MX_TP), the recurrence coefficients (cfs, MX_MC + MX_LMAX) and
// the Bessel node window (TN, lds_nodes, dynamic).
struct ItemShared {
  char *ybuf;
  double *rs;
  double2 *cfs;
  float2 *TN;
};

// One work item on a 256-thread block.  The Bessel intervals the item's
// arguments reach (it.tab) are copied into LDS (the dynamic window of
// lds_nodes nodes, sized for two blocks per CU) when they fit; otherwise the
// item reads the table from global memory (L1/L2).  The choice is per item,
// so one launch serves both kinds of items.  The caller separates two items
// run by one block with a barrier.
template <int QMAX, int NP, int KS>
__device__ __forceinline__ void mixed_item(const DirectItem &it, const PointsSoA &pts, const double *__restrict__ kln,
                                           const LegendreTables &leg, const HermiteTables &ht, int lds_nodes,
                                           double *__restrict__ partial, const ItemShared &sh) {
  static_assert(NP == 0 || NP == 2 || NP == 3, "operand pieces");
  static_assert(KS == 1 || ((KS == 2 || KS == 4) && NP != 0), "point split: FP16 operands only");
  constexpr int RBN = 4 / KS;                // row blocks of 16 per item
  constexpr int SPW = MX_TP / 32 / KS;       // 32-point blocks of a tile per wave (FP16 operands)
  constexpr int LDP = MixedLds<NP>::LDP;
  constexpr int PLANE = MixedLds<NP>::PLANE;
  constexpr int YBYTES = MixedLds<NP>::YBYTES;
  constexpr int FOLD = 64;                          // points per FP32 partial sum
  constexpr int NA = NP ? NP : 1;
  char *ybuf = sh.ybuf;
  double *rs = sh.rs;
  // recurrence coefficients (a_m, b_m) of this l at cf[MX_MC + m]; zero for m <= 0, so that
  // stepping past m = 0 leaves P = 0 and the lowest chunk needs no per-step guards
  double2 *cf = sh.cfs + MX_MC;
  float2 *TN = sh.TN;  // window of the nodes this item's arguments reach
  float *Yf = reinterpret_cast<float *>(ybuf);
  _Float16 *Yh = reinterpret_cast<_Float16 *>(ybuf);

  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4;
  const int rb = wave % RBN, kp = wave / RBN;  // row block, part of the tile's points
  const int row = 16 * rb + (lane & 15);
  const bool wave_rows = 16 * rb < it.nrows;  // wave-uniform
  const bool rowok = row < it.nrows;

  const int nint = ht.nint[l];
  const float4 *__restrict__ T = ht.tab + ht.off[l];
  // The item's points are sorted by radius and its rows by k, so every
  // argument k r / h lies in [k_first r_first, k_last r_last] / h: nodes lo ..
  // hi (two past the largest argument's interval) are staged when they fit.
  const double khr0 = kln[it.kofs] * (1.0 / MX_H);
  const double khr1 = kln[it.kofs + it.nrows - 1] * (1.0 / MX_H);
  const double rlo = pts.r[it.p0];
  // one node of margin below: a batched catalogue is sorted by a key that drops
  // the lowest 7 bits of r, so the first radius may exceed the smallest by an ulp.
  // Points unsorted (ITEM_FULL_WINDOW): every radius lies in [0, R] and the
  // window covers nodes 0 .. it.tab + 1 (it.tab = floor(k_last R / h) + 2)
  const bool full = (it.flags & ITEM_FULL_WINDOW) != 0;
  const int lo = full ? 0 : max((int)(rlo * khr0) - 1, 0);
  const double lod = (double)lo;
  const int nwin = full ? it.tab + 2 : (int)(pts.r[it.p1 - 1] * khr1) + 3 - lo;
  const bool tl = nwin <= lds_nodes;  // block-uniform
  if (tl) {
    // FP16 operands carry v * 2^12: the staged nodes carry the (exact) scale;
    // nodes past the table (never read) repeat its last node
    const float sc = NP ? MX_SCALE : 1.f;
    for (int j = tid; j < nwin; j += 256) {
      const int i = lo + j;
      const float4 q = T[i < nint ? i : nint - 1];
      TN[j] = i < nint ? make_float2(q.x * sc, q.y * sc) : make_float2(q.z * sc, q.w * sc);
    }
  }
  // rows past nrows take the first row's k and padding points the first
  // radius, so their arguments stay inside the window (their results are not read)
  const double khr = rowok ? kln[it.kofs + row] * (1.0 / MX_H) : khr0;
  const double2 sect = leg.sect[l];
  {
    const double2 *gc = leg.coef + leg.coef_off[l];
    for (int m = tid - MX_MC; m <= l; m += 256) cf[m] = m > 0 ? gc[m] : make_double2(0.0, 0.0);
  }

  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int64_t base = it.p0; base < it.p1; base += MX_TP) {
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    // a padding point gets the item's first radius: its Bessel arguments stay
    // inside the window, and its Y* columns are zero, so it adds nothing
    rs[tid] = pok ? pts.r[gp] : rlo;
    __syncthreads();  // rs ready; previous tile's readers done

    // ---- 1. Bessel values into A-operand registers
    float af[NP ? 1 : MX_TP / 4];
    h8 ah[NA][NP ? SPW : 1];
    if (wave_rows) {
      if (tl) {
        if constexpr (NP) {
#pragma unroll
          for (int s = 0; s < SPW; ++s) {
            const int sp = kp * SPW + s;
            h2 pc[3][4];
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
              const double r0 = rs[32 * sp + 8 * g + e], r1 = rs[32 * sp + 8 * g + e + 1];
              const float v0 = jl_hermite_lds(TN, r0, khr, lod);
              const float v1 = jl_hermite_lds(TN, r1, khr, lod);
              if constexpr (NP == 2)
                split_pair(v0, v1, pc[0][e / 2], pc[1][e / 2]);
              else
                split3_pair(v0, v1, pc[0][e / 2], pc[1][e / 2], pc[2][e / 2]);
            }
#pragma unroll
            for (int k = 0; k < NA; ++k)
              ah[k][s] = __builtin_shufflevector(__builtin_shufflevector(pc[k][0], pc[k][1], 0, 1, 2, 3),
                                                 __builtin_shufflevector(pc[k][2], pc[k][3], 0, 1, 2, 3), 0, 1, 2,
                                                 3, 4, 5, 6, 7);
          }
        } else {
#pragma unroll
          for (int s = 0; s < MX_TP / 4; ++s) {
            const double r = rs[4 * s + g];
            af[s] = jl_hermite_lds(TN, r, khr, lod);
          }
        }
      } else {
        if constexpr (NP) {
#pragma unroll
          for (int s = 0; s < SPW; ++s) {
            const int sp = kp * SPW + s;
            h2 pc[3][4];
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
              const double r0 = rs[32 * sp + 8 * g + e], r1 = rs[32 * sp + 8 * g + e + 1];
              const float v0 = jl_hermite(T, nint, r0 * khr) * MX_SCALE;
              const float v1 = jl_hermite(T, nint, r1 * khr) * MX_SCALE;
              if constexpr (NP == 2)
                split_pair(v0, v1, pc[0][e / 2], pc[1][e / 2]);
              else
                split3_pair(v0, v1, pc[0][e / 2], pc[1][e / 2], pc[2][e / 2]);
            }
#pragma unroll
            for (int k = 0; k < NA; ++k)
              ah[k][s] = __builtin_shufflevector(__builtin_shufflevector(pc[k][0], pc[k][1], 0, 1, 2, 3),
                                                 __builtin_shufflevector(pc[k][2], pc[k][3], 0, 1, 2, 3), 0, 1, 2,
                                                 3, 4, 5, 6, 7);
          }
        } else {
#pragma unroll
          for (int s = 0; s < MX_TP / 4; ++s) {
            const double r = rs[4 * s + g];
            af[s] = jl_hermite(T, nint, r * khr);
          }
        }
      }
    }

    // ---- 2. Legendre/azimuth rows of this thread's point, chunks from high m to low m
    YlmDown y;
    if (pok) {
      // the azimuth starts at the top of the first chunk; the Legendre pair
      // alone is taken there from m = l (16 qlo steps, renormalised every eight
      // as in the emitting loop below)
      const int mstart = l - MX_MC * it.qlo;
      y.init(l, mstart, pts.x[gp], pts.s[gp], pts.phi[gp], pts.cphi[gp], pts.sphi[gp], sect);
      for (int m = l; m > mstart; m -= 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) y.step_p(cf[m - k]);
        y.rescale_nb();
      }
    } else {
      y.v = y.vp = 0.0;
      y.E = 0;
      y.cm = y.sm = y.cp = y.sp = y.cot = y.cph = y.sph = 0.0;
    }
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qlo + qq;  // chunk index from the top
      if (q < it.qhi) {
        __syncthreads();  // previous chunk's readers are done
        const int mtop = l - MX_MC * q;
        // m = mtop - mm; in the lowest chunk the steps at m <= 0 use zero
        // coefficients, so the columns of m < 0 come out as zeros
#pragma unroll
        for (int mm = 0; mm < MX_MC; ++mm) {
          emit_y<NP, LDP>(y, Yf, Yh, mm, tid);
          y.step(cf[mtop - mm]);
          if ((mm & 7) == 7) y.rescale_nb();
        }
        __syncthreads();
        if (wave_rows) {
          // two independent accumulator chains (column halves).  Both always
          // run (padding columns hold zeros), so that the MFMA chain and the
          // VALU reads of its results below form one basic block, where the
          // compiler accounts for the MFMA result latency.
          const int col = lane & 15;
          v4f a0 = v4f{0.f, 0.f, 0.f, 0.f}, a1 = v4f{0.f, 0.f, 0.f, 0.f};
          if constexpr (NP) {
            // the leading products p0*q0 go to (a0, a1), folded every FOLD
            // points; the correction products, 2^-11 smaller, to (c0, c1),
            // whose FP32 rounding is correspondingly smaller
            v4f c0 = v4f{0.f, 0.f, 0.f, 0.f}, c1 = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < SPW; ++s) {
              const int o = col * LDP + 32 * (kp * SPW + s) + 8 * g;
              h8 b0[NA], b1[NA];
#pragma unroll
              for (int k = 0; k < NP; ++k) {
                b0[k] = *reinterpret_cast<const h8 *>(Yh + k * PLANE + o);
                b1[k] = *reinterpret_cast<const h8 *>(Yh + k * PLANE + o + 16 * LDP);
              }
              // correction products p_i q_j, 1 <= i + j <= NP - 1, smallest first
              if constexpr (NP == 3) {
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b0[2], c0, 0, 0, 0);
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b0[1], c0, 0, 0, 0);
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[2][s], b0[0], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b1[2], c1, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b1[1], c1, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[2][s], b1[0], c1, 0, 0, 0);
              }
              c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b0[1], c0, 0, 0, 0);
              c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b0[0], c0, 0, 0, 0);
              c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b1[1], c1, 0, 0, 0);
              c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b1[0], c1, 0, 0, 0);
              a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b0[0], a0, 0, 0, 0);
              a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b1[0], a1, 0, 0, 0);
              if ((32 * (s + 1)) % FOLD == 0 && s + 1 < SPW) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  sum[qq][0][j] += (double)a0[j];
                  sum[qq][1][j] += (double)a1[j];
                }
                a0 = v4f{0.f, 0.f, 0.f, 0.f};
                a1 = v4f{0.f, 0.f, 0.f, 0.f};
              }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              sum[qq][0][j] += (double)c0[j];
              sum[qq][1][j] += (double)c1[j];
            }
          } else {
            const float *bp = Yf + col * LDP + g;
#pragma unroll
            for (int s = 0; s < MX_TP / 4; ++s) {
              a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(af[s], bp[4 * s], a0, 0, 0, 0);
              a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(af[s], bp[4 * s + 16 * LDP], a1, 0, 0, 0);
              if ((4 * (s + 1)) % FOLD == 0 && 4 * (s + 1) < MX_TP) {
                // fold this FOLD-point partial sum into the FP64 accumulators
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  sum[qq][0][j] += (double)a0[j];
                  sum[qq][1][j] += (double)a1[j];
                }
                a0 = v4f{0.f, 0.f, 0.f, 0.f};
                a1 = v4f{0.f, 0.f, 0.f, 0.f};
              }
            }
          }
          // fold the chunk's FP32 partial sums into the FP64 accumulators
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            sum[qq][0][j] += (double)a0[j];
            sum[qq][1][j] += (double)a1[j];
          }
        }
      }
    }
  }

  if constexpr (KS > 1) {
    // parts 1 .. KS-1 of each row block hand their sums to part 0 through LDS
    // (the Y* area), one chunk at a time; part 0 adds them in part order
    double *red = reinterpret_cast<double *>(ybuf);
    static_assert((KS - 1) * RBN * 64 * 8 * (int)sizeof(double) <= YBYTES, "reduction area");
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      if (it.qlo + qq < it.qhi) {  // block-uniform
        __syncthreads();  // the Y* area is free (last chunk's readers, previous chunk's adders)
        if (kp > 0 && wave_rows) {
          double *d = red + (size_t)(((kp - 1) * RBN + rb) * 64 + lane) * 8;
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) d[4 * c + j] = sum[qq][c][j];
        }
        __syncthreads();
        if (kp == 0 && wave_rows) {
#pragma unroll
          for (int k = 1; k < KS; ++k) {
            const double *d = red + (size_t)(((k - 1) * RBN + rb) * 64 + lane) * 8;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int j = 0; j < 4; ++j) sum[qq][c][j] += d[4 * c + j];
          }
        }
      }
    }
  }

  // ---- write the item's partial block (64 rows x its columns), FP64
  const double scale = NP ? 1.0 / ((double)MX_SCALE * (double)MX_SCALE) : 1.0;
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi && kp == 0) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);       // local column: (mm, re/im)
        const int m = l - MX_MC * q - (lc >> 1);
        const int col = m >= 0 ? 2 * m + (lc & 1) : ncol;
        // v_mfma_f32_16x16x{4f32,32_f16} result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int orow = 16 * rb + 4 * g + j;
          if (col < ncol) out[(size_t)orow * ncol + col] = sum[qq][c][j] * scale;
        }
      }
    }
  }
}

// One item per block (blockIdx.x), all of one class (QMAX, NP, KS).
template <int QMAX, int NP, int KS = 1>
__global__ __launch_bounds__(256) void sfb_direct_mixed_kernel(
    PointsSoA pts, const DirectItem *__restrict__ items, const double *__restrict__ kln, int nmax,
    LegendreTables leg, HermiteTables ht, int lds_nodes, double *__restrict__ partial) {
  __shared__ __attribute__((aligned(16))) char ybuf[MixedLds<NP>::YBYTES];
  __shared__ double rs[MX_TP];
  __shared__ double2 cfs[MX_MC + MX_LMAX];
  extern __shared__ float2 TN[];
  const DirectItem it = items[blockIdx.x];
  mixed_item<QMAX, NP, KS>(it, pts, kln, leg, ht, lds_nodes, partial, ItemShared{ybuf, rs, cfs, TN});
}

}  // namespace sfb
