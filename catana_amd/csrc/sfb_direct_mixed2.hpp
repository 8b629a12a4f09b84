// sfb_direct_mixed2.hpp -- mixed-precision fused direct SFB kernel, 512-thread form.
//
// The contraction is that of the FP64 kernel (per l: Bessel rows x Y* columns
// summed over points, sfb_raw.cpp:61-80) on the matrix cores with split-FP16
// operands (each operand v 2^12 = p0 + p1, products p0q0 + p0q1 + p1q0 on
// v_mfma_f32_16x16x32_f16, FP32 partial sums folded into FP64 every 64 points).
// What differs from the 256-thread kernel (sfb_direct_mixed.hpp) is how the
// operands are produced:
//   * a block is 8 waves, one block per CU (two waves per SIMD), so it can
//     hold in LDS the cubic coefficients (c0..c3 of j_l(x) on [i h, (i+1) h],
//     h = 1/16, FP32 from FP64 values and slopes) of every interval an item's
//     arguments reach: a Bessel value costs the argument split (FP64), one
//     16-byte LDS read and three FMAs;
//   * an item covers up to 128 rows of n (RB = 8 row blocks of 16, one per
//     wave): the Y* chunk that waves 0-3 generate is then contracted by eight
//     waves, so for l with n_l > 64 each Y* value is generated once instead of
//     once per 64 rows.  Items of at most 64 (32, 16) rows split each tile's
//     points over KP = 8 / RB parts instead, every thread generating the Y*
//     of its own point;
//   * the Y* chunk (16 m) is double-buffered by column halves (8 m): in each
//     phase a wave contracts one half while the generating waves write the
//     next, so the matrix-core and vector work of a phase are independent;
//   * the Legendre recurrence runs on Q^m = P^m / d_m with d_m chosen so that
//     Q^{m-1} = g_m cot(t) Q^m - Q^{m+1} (two FP64 operations a step); the
//     per-column factor d_m is applied to the FP64 sums when they are written.
//     The azimuth is the Chebyshev recurrence in FP64 as before; the value
//     Q 2^{E+12} (cos m phi, -sin m phi) is rounded once to FP32 and split;
//   * the FP16 remainders of a split are formed and rounded in one
//     v_fma_mixlo_f16 / v_fma_mixhi_f16 each.
// Work item: (l, up to 16 RB rows of n, m chunks [qlo, qhi), point range);
// chunks are counted from the top (chunk q holds m = l - 16q - mm), so only the
// lowest is partial; its columns of m < 0 are computed and discarded.
#pragma once

#include "sfb_common.hpp"
#include "sfb_direct_mixed.hpp"
#include "sfb_math.hpp"

namespace sfb {

constexpr int M2_NT = 512;    // threads per block
constexpr int M2_HM = 8;      // m values per column half (16 real columns)
constexpr int M2_GPAD = 16;   // g_m table padding below m = 0

// Cubic coefficients per interval: tab[off[l] + i] = (c0, c1, c2, c3) with
// j_l((i + f) h) ~ c0 + f (c1 + f (c2 + f c3)), 0 <= f < 1 (cubic Hermite on
// the values and slopes at both ends), i < nint[l]; tab16 holds the same
// coefficients times 2^12 (MX_SCALE), the scale of the split-FP16 operands.
struct CubicTables {
  const float4 *tab;
  const int32_t *off;
  const int32_t *nint;
  const float4 *tab16;
};

// Legendre constants of the scaled recurrence, per l at off[l] = l (l+1) / 2:
//   P_l^m = d_m Q^m,  Q^l = sect 2^E sin^l t,  Q^{m-1} = g_m cot(t) Q^m - Q^{m+1}.
// g[off[l]] (m = 0) is 0; the steps below m = 0 only produce discarded columns.
struct LegendreScaled {
  const double2 *sect;  // (mantissa, exponent) of P_l^l(cos t) / (d_l sin^l t)
  const double *g;
  const double *d;
  const int32_t *off;
};

__device__ inline void jl_value_slope(int l, double x, double &j, double &dj) {
  if (x == 0.0) {
    j = (l == 0) ? 1.0 : 0.0;
    dj = (l == 1) ? 1.0 / 3.0 : 0.0;
  } else {
    j = sbessel_jl(l, x);
    dj = (double)l / x * j - sbessel_jl(l + 1, x);
  }
}

__global__ void cubic_table_kernel(int lmax, const int32_t *__restrict__ off, const int32_t *__restrict__ nint,
                                   float4 *__restrict__ tab, float4 *__restrict__ tab16) {
  const int l = blockIdx.y;
  if (l >= lmax) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nint[l]; i += gridDim.x * blockDim.x) {
    double j0, d0, j1, d1;
    jl_value_slope(l, i * MX_H, j0, d0);
    jl_value_slope(l, (i + 1) * MX_H, j1, d1);
    const double s0 = MX_H * d0, s1 = MX_H * d1, dj = j1 - j0;
    const float4 c = make_float4((float)j0, (float)s0, (float)(3.0 * dj - 2.0 * s0 - s1), (float)(s0 + s1 - 2.0 * dj));
    tab[off[l] + i] = c;
    tab16[off[l] + i] = make_float4(c.x * MX_SCALE, c.y * MX_SCALE, c.z * MX_SCALE, c.w * MX_SCALE);
  }
}

// (a, b) -> FP16 pairs hi, lo (as 32-bit words, a in the low half) with
// a ~= hi.x + lo.x, b ~= hi.y + lo.y: hi by v_cvt_pk_f16_f32, each remainder
// (exact in FP32) rounded to FP16 by one v_fma_mix{lo,hi}_f16.
__device__ __forceinline__ void split2(float a, float b, unsigned &hi, unsigned &lo) {
  hi = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){a, b}, h2));
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(lo) : "v"(a), "v"(hi));
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(lo) : "v"(b), "v"(hi));
}

// j_l(k r) 2^12 from the LDS window: C[j] = coefficients of interval lo + j
// (pre-scaled), t = r (k/h) - lo.  The caller guarantees 0 <= t and that
// interval floor(t) is staged.
__device__ __forceinline__ float jl_cubic_lds(const float4 *C, double r, double khr, double lo) {
  const double t = fma(r, khr, -lo);
  const int i = (int)t;
  const float f = (float)__builtin_amdgcn_fract(t);
  const float4 c = C[i];
  return fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
}

// j_l 2^12 at t = r k / h from the scaled table (tab16): the cubic of interval
// floor(t), clamped to the table.
__device__ __forceinline__ float jl_cubic_scaled(const float4 *__restrict__ T16, int nint, double t) {
  int i = (int)t;
  i = i < nint - 1 ? i : nint - 1;
  i = i > 0 ? i : 0;
  const float f = (float)(t - (double)i);
  const float4 c = T16[i];
  return fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
}

// The same from the global table (unscaled; index clamped to the table).
__device__ __forceinline__ float jl_cubic_global(const float4 *__restrict__ T, int nint, double t) {
  int i = (int)t;
  i = i < nint - 1 ? i : nint - 1;
  i = i > 0 ? i : 0;
  const float f = (float)(t - (double)i);
  const float4 c = T[i];
  return fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x) * MX_SCALE;
}

typedef unsigned u4v __attribute__((ext_vector_type(4)));

// A fragments (pieces 0 and 1) of SPW 32-point steps: lane value e of step s
// is j(r[32 s + e]) (the caller offsets r by the lane's 8-point group).
template <int SPW, class F>
__device__ __forceinline__ void bessel_frags(h8 (&ah)[2][SPW], const double *r, F jl) {
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    unsigned hw[4], lw[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) split2(jl(r[32 * s + e]), jl(r[32 * s + e + 1]), hw[e / 2], lw[e / 2]);
    ah[0][s] = __builtin_bit_cast(h8, (u4v){hw[0], hw[1], hw[2], hw[3]});
    ah[1][s] = __builtin_bit_cast(h8, (u4v){lw[0], lw[1], lw[2], lw[3]});
  }
}

// Legendre (scaled) and azimuth rows of one point, downward in m.
struct YlmScaled {
  double v, vp;  // Q^m, Q^{m+1} without the 2^E factor
  int E;
  double cm, sm, cp, sp;  // cos/sin(m phi), cos/sin((m+1) phi)
  double cot, c2;         // cot(t), 2 cos(phi)

  // Q at m = l, the azimuth at m0 <= l (the caller steps Q alone down to m0).
  __device__ __forceinline__ void init(int l, int m0, double x, double s, double phi, double cph, double sph,
                                       double2 sect) {
    // |cot| <= 2^70 bounds the growth of one step by 2^78, so eight steps
    // from |v| <= 2^300 stay below 2^924
    const double se = fmax(s, 0x1p-70);
    cot = x / se;
    c2 = cph + cph;
    // se^l = b^l 2^(l be) with b = frexp(se) in [0.5, 1): b^l >= 2^-l is a
    // normal FP64 number for l < MX_LMAX, so plain square-and-multiply needs
    // no renormalisation, and one frexp puts the result in [0.5, 1)
    static_assert(MX_LMAX < 1000, "b^l must stay a normal FP64 number");
    int be, t;
    double b = frexp(se, &be), m = 1.0;
    for (int k = l; k; k >>= 1) {
      if (k & 1) m *= b;
      b *= b;
    }
    m = frexp(m, &t);
    v = m * sect.x;
    E = l * be + t + (int)sect.y;
    vp = 0.0;
    const double hi = (double)m0 * phi;
    const double lo = fma((double)m0, phi, -hi);
    double S, C;
    sincos(hi, &S, &C);
    cm = C - lo * S;
    sm = S + lo * C;
    cp = fma(cm, cph, -sm * sph);
    sp = fma(sm, cph, cm * sph);
  }
  __device__ __forceinline__ void zero() {
    v = vp = 0.0;
    E = 0;
    cm = sm = cp = sp = cot = c2 = 0.0;
  }
  __device__ __forceinline__ void step_p(double g) {
    const double vn = fma(g * cot, v, -vp);
    vp = v;
    v = vn;
  }
  __device__ __forceinline__ void step(double g) {
    step_p(g);
    const double cn = fma(c2, cm, -cp);
    const double sn = fma(c2, sm, -sp);
    cp = cm;
    sp = sm;
    cm = cn;
    sm = sn;
  }
  __device__ __forceinline__ void rescale_nb() {
    const bool big = fabs(v) > 0x1p+300;
    const double f = big ? 0x1p-300 : 1.0;
    v *= f;
    vp *= f;
    E += big ? 300 : 0;
  }
};

// Eight m values of Y* 2^12 (from the generator's current m down) into one
// column-half buffer: planes k = 0, 1 of 16 columns (2mm: real, 2mm+1:
// imaginary) x LDP points, FP16.  Leaves the generator 8 steps lower.
template <int LDP>
__device__ __forceinline__ void gen_half(YlmScaled &y, const double *cg, int mtop, unsigned short *Yb, int t) {
#pragma unroll
  for (int mm = 0; mm < M2_HM; ++mm) {
    const double P = ldexp(y.v, y.E + 12);
    const float re = (float)(P * y.cm), im = (float)(-P * y.sm);
    unsigned hi, lo;
    split2(re, im, hi, lo);
    Yb[(2 * mm) * LDP + t] = (unsigned short)hi;
    Yb[(2 * mm + 1) * LDP + t] = (unsigned short)(hi >> 16);
    Yb[(16 + 2 * mm) * LDP + t] = (unsigned short)lo;
    Yb[(16 + 2 * mm + 1) * LDP + t] = (unsigned short)(lo >> 16);
    y.step(cg[mtop - mm]);
  }
  y.rescale_nb();
}

// One column half on the matrix cores: this wave's 16 rows x 16 columns over
// its SPW 32-point steps of the tile.  The leading products go to `a`, folded
// into the FP64 sums every 64 points; the correction products (2^-11 smaller)
// to `c`, folded once.
template <int LDP, int SPW>
__device__ __forceinline__ void mfma_half(const h8 (&ah)[2][SPW], const unsigned short *Yb, int col, int pbase, int g,
                                          v4d &sum) {
  v4f a = v4f{0.f, 0.f, 0.f, 0.f}, c = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    const int o = col * LDP + pbase + 32 * s + 8 * g;
    const h8 b0 = *reinterpret_cast<const h8 *>(Yb + o);
    const h8 b1 = *reinterpret_cast<const h8 *>(Yb + 16 * LDP + o);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b1, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b0, c, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b0, a, 0, 0, 0);
    if ((s & 1) && s + 1 < SPW) {
#pragma unroll
      for (int j = 0; j < 4; ++j) sum[j] += (double)a[j];
      a = v4f{0.f, 0.f, 0.f, 0.f};
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) sum[j] += (double)a[j] + (double)c[j];
}

// RB row blocks of 16 (one per wave when RB = 8), KP = 8 / RB point parts.
template <int QMAX, int RB>
__global__ __launch_bounds__(M2_NT) void sfb_direct_mixed2_kernel(
    PointsSoA pts, const DirectItem *__restrict__ items, const double *__restrict__ kln, LegendreScaled leg,
    CubicTables ct, int lds_nodes, double *__restrict__ partial) {
  static_assert(RB == 8 || RB == 4 || RB == 2 || RB == 1, "row blocks");
  constexpr int KP = 8 / RB;
  constexpr int TP = RB == 8 ? 256 : 512;  // points per tile (one generating thread each)
  constexpr int PW = TP / KP;              // points of a wave per tile
  constexpr int SPW = PW / 32;             // 32-point MFMA steps of a wave per tile
  constexpr int LDP = TP + 8;              // row pitch of a column in the Y* buffers (FP16)
  constexpr int HBUF = 2 * 16 * LDP;       // one column half: two planes of 16 columns
  __shared__ __attribute__((aligned(16))) unsigned short ybuf[2 * HBUF];
  __shared__ double rs[TP];
  __shared__ double cgs[M2_GPAD + MX_LMAX];
  extern __shared__ float4 TC[];  // cubic coefficients of the intervals this item's arguments reach
  double *cg = cgs + M2_GPAD;

  const DirectItem it = items[blockIdx.x];
  const int l = it.l;
  const int ncol = 2 * (l + 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4;
  const int rb = wave % RB, kp = wave / RB;
  const int row = 16 * rb + (lane & 15);
  const bool wave_rows = 16 * rb < it.nrows;  // wave-uniform
  const bool rowok = row < it.nrows;
  const bool gen_wave = wave < TP / 64;       // wave-uniform: this wave generates Y*

  const int nint = ct.nint[l];
  const float4 *__restrict__ T = ct.tab + ct.off[l];
  // the item's points are sorted by radius and its rows by k: every argument
  // k r / h lies in [k_first r_first, k_last r_last] / h
  const double khr0 = kln[it.kofs] * (1.0 / MX_H);
  const double khr1 = kln[it.kofs + it.nrows - 1] * (1.0 / MX_H);
  const double rlo = pts.r[it.p0];
  // points unsorted (ITEM_FULL_WINDOW): every radius lies in [0, R] and the
  // window covers nodes 0 .. it.tab + 1 (it.tab = floor(k_last R / h) + 2)
  const bool full = (it.flags & ITEM_FULL_WINDOW) != 0;
  const int lo = full ? 0 : max((int)(rlo * khr0) - 1, 0);
  const double lod = (double)lo;
  const int nwin = full ? it.tab + 2 : (int)(pts.r[it.p1 - 1] * khr1) + 3 - lo;
  const bool tl = nwin <= lds_nodes;  // block-uniform
  if (tl) {
    for (int j = tid; j < nwin; j += M2_NT) {
      const int i = lo + j;
      const float4 q = T[i < nint ? i : nint - 1];
      TC[j] = make_float4(q.x * MX_SCALE, q.y * MX_SCALE, q.z * MX_SCALE, q.w * MX_SCALE);
    }
  }
  const double khr = rowok ? kln[it.kofs + row] * (1.0 / MX_H) : khr0;
  const double2 sect = leg.sect[l];
  {
    const double *gg = leg.g + leg.off[l];
    for (int m = tid - M2_GPAD; m <= l; m += M2_NT) cg[m] = m > 0 ? gg[m] : 0.0;
  }

  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};

  const int nh = 2 * (it.qhi - it.qlo);  // column halves of this item
  const int col = lane & 15;
  const int pbase = PW * kp;
  for (int64_t base = it.p0; base < it.p1; base += TP) {
    const int64_t gp = base + tid;
    const bool pok = tid < TP && gp < it.p1;
    if (tid < TP) rs[tid] = pok ? pts.r[gp] : rlo;
    __syncthreads();  // rs (and on the first tile the window and g_m) ready; previous tile done

    // ---- Bessel values of this wave's rows and points into A fragments
    h8 ah[2][SPW];
    if (wave_rows) {
      if (tl)
        bessel_frags<SPW>(ah, rs + pbase + 8 * g, [&](double r) { return jl_cubic_lds(TC, r, khr, lod); });
      else
        bessel_frags<SPW>(ah, rs + pbase + 8 * g, [&](double r) { return jl_cubic_global(T, nint, r * khr); });
    }

    // ---- Y* generator of this thread's point, positioned at the item's top chunk
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
    const int mtop0 = l - MX_MC * it.qlo;
    if (gen_wave) gen_half<LDP>(y, cg, mtop0, ybuf, tid);
    __syncthreads();
    // phase h contracts column half h (buffer h & 1) and generates half h + 1
#pragma unroll
    for (int h = 0; h < 2 * QMAX; ++h) {
      if (h < nh) {  // block-uniform
        const unsigned short *Yr = ybuf + (h & 1) * HBUF;
        unsigned short *Yw = ybuf + ((h + 1) & 1) * HBUF;
        const int mtop = mtop0 - M2_HM * (h + 1);
        const bool more = h + 1 < nh;  // block-uniform
        if (wave_rows && gen_wave && more) {
          mfma_half<LDP, SPW>(ah, Yr, col, pbase, g, sum[h >> 1][h & 1]);
          gen_half<LDP>(y, cg, mtop, Yw, tid);
        } else if (wave_rows) {
          mfma_half<LDP, SPW>(ah, Yr, col, pbase, g, sum[h >> 1][h & 1]);
        } else if (gen_wave && more) {
          gen_half<LDP>(y, cg, mtop, Yw, tid);
        }
        __syncthreads();
      }
    }
  }

  if constexpr (KP > 1) {
    // parts 1 .. KP-1 of each row block hand their sums to part 0 through LDS
    // (the Y* buffers), one chunk at a time; part 0 adds them in part order
    double *red = reinterpret_cast<double *>(ybuf);
    static_assert((KP - 1) * RB * 64 * 8 * (int)sizeof(double) <= 2 * HBUF * 2, "reduction area");
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      if (it.qlo + qq < it.qhi) {  // block-uniform
        __syncthreads();
        if (kp > 0 && wave_rows) {
          double *dd = red + (size_t)(((kp - 1) * RB + rb) * 64 + lane) * 8;
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) dd[4 * c + j] = sum[qq][c][j];
        }
        __syncthreads();
        if (kp == 0 && wave_rows) {
#pragma unroll
          for (int k = 1; k < KP; ++k) {
            const double *dd = red + (size_t)(((k - 1) * RB + rb) * 64 + lane) * 8;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
              for (int j = 0; j < 4; ++j) sum[qq][c][j] += dd[4 * c + j];
          }
        }
      }
    }
  }

  // ---- the item's partial block (16 RB rows x its columns), FP64, with d_m applied
  const double *dm = leg.d + leg.off[l];
  constexpr double unscale = 1.0 / ((double)MX_SCALE * (double)MX_SCALE);
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi && kp == 0 && wave_rows) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);
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

}  // namespace sfb
