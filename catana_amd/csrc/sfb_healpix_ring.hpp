// sfb_healpix_ring.hpp -- the HEALPix SFB path on the ring structure of the map.
//
// The reference (sfb_reverse_fft.cpp:115-153, :174-209) evaluates Y_lm once per
// ring, and for every (l, n) builds the map F(p) = sum_{r in p} j_l(k_ln r),
// Fourier-transforms each ring and sums ring by ring.  Here the same steps are
// these kernels:
//   1. ring_legendre_kernel: P_l^m(cos theta) once per (l, ring) -- 4 nside - 1
//      FP64 recurrences per l instead of one per pixel;
//   2. ring_maps_kernel: the FP32 maps F_l(n, p), streamed from the CSR radii,
//      j_l from the cubic-coefficient table of l staged in LDS (FP64 argument
//      split, three FP32 FMAs per value), summed per pixel in FP64, written
//      once as FP32, coalesced by pixel;
//   3. healpix_ring_contract_kernel: per l the ring DFT and the ring-wise
//      Legendre sum as one matrix-core product over the pixels,
//        f_lmn = sum_rings P_l^m(ring) sum_{j in ring} F(n, j) e^{-i m phi_j},
//      with the maps (A operand, read from HBM) and Y* = P_ring e^{-i m phi_j}
//      (B operand, the ring table times the pixel azimuth, FP64 Chebyshev
//      recurrence) split into FP16 pieces: v_mfma_f32_16x16x32_f16, three
//      products, FP32 partial sums folded into FP64 every 64 pixels;
//   4. healpix_ring_eq_kernel: the same contraction for the equatorial tiles
//      (nside a multiple of 128: every 256-pixel tile there lies in one ring of
//      4 nside pixels at phi_j = (2 j + s) pi / (4 nside)): Y* is e^{-i m phi}
//      alone, read from an LDS table of the 8 nside factors e^{-2 pi i k /
//      (8 nside)}, and the ring's Legendre values scale the tile's FP64 sums
//      per column -- no per-pixel recurrence.
// With alias_compat the m whose folded FFT bin the reference reads without
// conjugation ((m mod n_ring) > n_ring / 2) use the mirrored azimuth
// phi' = 2 phi_0 - phi, as the FP64 pixel kernel does (YlmDownPix).
#pragma once

#include "sfb_common.hpp"
#include "sfb_direct_mixed.hpp"
#include "sfb_direct_mixed2.hpp"
#include "sfb_direct_split16.hpp"
#include "sfb_healpix.hpp"
#include "sfb_ylm.hpp"

namespace sfb {

constexpr int HR_NB = 16;   // map rows (n) per map item
constexpr int HR_TP = 256;  // pixels per contraction tile
constexpr int HR_NT = 64;   // rows (n) per contraction item
constexpr int HR_RINGS = 16;  // rings of one contraction tile whose Legendre values are staged in LDS

// z = cos(theta), 1 - z and 1 + z of HEALPix ring i (1-based), exact in the caps.
__host__ __device__ inline void hp_ring_z(int64_t ns, int64_t ring, double &z, double &omz, double &opz) {
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
}

// Ring cos/sin theta (ring rho = i - 1) and the ring index of every pixel.
__global__ void ring_setup_kernel(int nside, double *__restrict__ rx, double *__restrict__ rs,
                                  int32_t *__restrict__ ring_of) {
  const int64_t ns = nside, npix = 12 * ns * ns, nrings = 4 * ns - 1;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nrings) {
    double z, omz, opz;
    hp_ring_z(ns, p + 1, z, omz, opz);
    rx[p] = z;
    rs[p] = sqrt(omz * opz);
  }
  if (p < npix) {
    int64_t ring, start, np;
    bool shifted;
    hp_ring_of(ns, p, ring, start, np, shifted);
    ring_of[p] = (int32_t)(ring - 1);
  }
}

// Slot order of the maps and of the contraction: within each tile of HR_TP
// consecutive pixels (the tiles of the map and contraction items), slot
// HR_TP b + i holds the pixel of rank i by decreasing point count (ties in
// pixel order).  A wave's 64 lanes then loop over about equally many radii in
// the map kernel, and a tile still spans the same few rings.
__global__ __launch_bounds__(256) void ring_slot_order_kernel(const int64_t *__restrict__ off, int64_t npix,
                                                            int32_t *__restrict__ perm) {
  __shared__ int64_t cnt[HR_TP];
  const int64_t b0 = (int64_t)blockIdx.x * HR_TP;
  const int n = (int)min<int64_t>(HR_TP, npix - b0);
  const int t = threadIdx.x;
  if (t < n) cnt[t] = off[b0 + t + 1] - off[b0 + t];
  __syncthreads();
  if (t < n) {
    const int64_t c = cnt[t];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (cnt[k] > c) || (cnt[k] == c && k < t);
    perm[b0 + rank] = (int32_t)(b0 + t);
  }
}

// ptab[poff[l] + rho (l + 1) + m] = P_l^m(cos theta_rho) 2^12 for m = 0..l (the
// orthonormal functions with the Condon-Shortley phase, gsl_sf_legendre_sphPlm
// as in sfb_reverse_fft.cpp:119), by the downward recurrence of YlmDown;
// values below the FP64 range come out as 0.
__global__ void ring_legendre_kernel(int lmax, int nrings, const double *__restrict__ rx,
                                     const double *__restrict__ rs, LegendreTables leg,
                                     const int64_t *__restrict__ poff, double *__restrict__ ptab) {
  const int l = blockIdx.y;
  const int rho = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lmax || rho >= nrings) return;
  YlmDown y;
  y.init(l, rx[rho], rs[rho], 0.0, 1.0, 0.0, leg.sect[l]);
  const double2 *coef = leg.coef + leg.coef_off[l];
  double *o = ptab + poff[l] + (int64_t)rho * (l + 1);
  for (int m = l; m >= 0; --m) {
    o[m] = ldexp(y.v, y.E + 12);
    if (m > 0) {
      y.step_p(coef[m]);
      y.rescale();
    }
  }
}

// One map item: rows n0 .. n0 + nrows - 1 (<= HR_NB) of multipole l over
// pixels [p0, p1); map row n of l starts at maps + moff + n npix.
struct MapItem {
  int32_t l, n0, nrows, pad;
  int64_t p0, p1;
  int64_t moff;
  int64_t kofs;  // k_{l, n0} in the k_ln array
};

// j_l(t h) from a cubic table (t = k r / h >= 0 and floor(t) inside the table).
__device__ __forceinline__ float cubic_at(const float4 *C, double t) {
  const int i = (int)t;
  const float f = (float)__builtin_amdgcn_fract(t);
  const float4 c = C[i];
  return fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
}

// The map values of one item: one slot (pixel) per lane, each lane looping
// over its pixel's radii and evaluating the item's rows for each.
template <class F>
__device__ __forceinline__ void map_pixels(const MapItem &it, const int64_t *__restrict__ off,
                                           const int32_t *__restrict__ perm, const double *__restrict__ radii,
                                           int64_t npix,
                                           const double (&khr)[HR_NB], double scale, float *__restrict__ maps,
                                           F jl) {
  // one slot per lane: the slots of a 256-pixel tile hold its pixels by
  // decreasing point count, so a wave's lanes loop over about equally many
  // radii (12.0 -> 11.0 ms of maps at nside 256 against pixel order), and the
  // map writes stay contiguous.  Measured slower: sixteen lanes per pixel, one
  // row each, with the rows staged in LDS for the writes (12.1 -> 23 ms, one
  // evaluation per radius load instead of 16); skipping the rows past the
  // item's by a wave-uniform branch per row instead of evaluating them at
  // k = 0 (maps 9.9 -> 10.8 ms).
  for (int64_t p = it.p0 + threadIdx.x; p < it.p1; p += blockDim.x) {
    // FP64 pixel sums (a pixel may hold hundreds of radii whose j_l values
    // cancel), rounded to FP32 once; slot p holds pixel perm[p]
    double acc[HR_NB];
#pragma unroll
    for (int j = 0; j < HR_NB; ++j) acc[j] = 0.0;
    const int64_t px = perm[p];
    const int64_t q1 = off[px + 1];
    for (int64_t q = off[px]; q < q1; ++q) {
      const double r = radii[q];
#pragma unroll
      for (int j = 0; j < HR_NB; ++j) acc[j] += (double)jl(r * khr[j]);
    }
#pragma unroll
    for (int j = 0; j < HR_NB; ++j)
      if (j < it.nrows) maps[it.moff + (int64_t)(it.n0 + j) * npix + p] = (float)(acc[j] * scale);
  }
}

// maps[moff + n npix + p] = scale sum_{r in pixel p} j_l(k_ln r).  The table
// of l covers k_{l, n_l - 1} max(rmax, max r) (so no index is clamped); it is
// staged in LDS when it has at most lds_nodes intervals.  Rows past nrows use
// k = 0 (t = 0) and are not written.
__global__ __launch_bounds__(256) void ring_maps_kernel(const MapItem *__restrict__ items,
                                                        const int64_t *__restrict__ off,
                                                        const int32_t *__restrict__ perm,
                                                        const double *__restrict__ radii, int64_t npix,
                                                        const double *__restrict__ kln, CubicTables ct,
                                                        int lds_nodes, double scale, float *__restrict__ maps) {
  extern __shared__ float4 TM[];
  const MapItem it = items[blockIdx.x];
  const int l = it.l;
  const int nint = ct.nint[l];
  const float4 *__restrict__ T = ct.tab + ct.off[l];
  const bool tl = nint <= lds_nodes;  // block-uniform
  if (tl)
    for (int j = threadIdx.x; j < nint; j += blockDim.x) TM[j] = T[j];
  __syncthreads();
  double khr[HR_NB];
#pragma unroll
  for (int j = 0; j < HR_NB; ++j) khr[j] = j < it.nrows ? kln[it.kofs + j] * (1.0 / MX_H) : 0.0;
  if (tl)
    map_pixels(it, off, perm, radii, npix, khr, scale, maps, [&](double t) { return cubic_at(TM, t); });
  else
    map_pixels(it, off, perm, radii, npix, khr, scale, maps, [&](double t) { return cubic_at(T, t); });
}

// Per-pixel data of the ring contraction.
struct RingPixels {
  const double *phi, *cphi, *sphi;     // pixel azimuth phi_j and its cos/sin
  const double *phi2, *cphi2, *sphi2;  // mirrored azimuth 2 phi_0 - phi_j
  const int32_t *np;                   // ring length of the pixel
  const int32_t *ring;                 // ring index (0-based) of the pixel
  const double *ptab;                  // ring Legendre table (ring_legendre_kernel)
  const int64_t *poff;
  const int32_t *perm;                 // slot order (ring_slot_order_kernel): slot i holds pixel perm[i]
  int32_t alias_compat;
  // equatorial tiles (ring_eq_table_kernel): eq_ns = nside when the path is on
  // (0 otherwise); tiles of [eq_lo, eq_hi) take their azimuth factors from
  // eqtab (8 nside entries, staged in the kernel's dynamic LDS)
  const uint2 *eqtab;
  int32_t eq_ns;
  int64_t eq_lo, eq_hi;
};

// Azimuth factors of the equatorial rings: E[k] = e^{-i 2 pi k / (8 nside)} 2^12
// as its two FP16 pieces, .x = (re, im) of piece 0, .y = (re, im) of piece 1.
// Every equatorial pixel's azimuth is phi_j = (2 j + s) pi / (4 nside) (s = 1 on
// shifted rings), so e^{-i m phi_j} = E[(m (2 j + s)) mod 8 nside] exactly.
__global__ void ring_eq_table_kernel(int nside, uint2 *__restrict__ E) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 8 * nside) return;
  double sn, cs;
  sincospi((double)k / (4.0 * nside), &sn, &cs);
  unsigned hi, lo;
  split2((float)(cs * 4096.0), (float)(-sn * 4096.0), hi, lo);
  E[k] = make_uint2(hi, lo);
}

// Y* 2^12 of one pixel for the 16 m of a chunk (from mtop down) into the
// chunk buffer: two point-major FP16 planes (row tid: 32 columns, 2 mm: real,
// 2 mm + 1: imaginary; the layout of sfb_direct_split16.hpp), stored as one
// 64-bit word per plane and two m.  Leaves the azimuth 16 steps lower.
// Columns of m < 0 and padding pixels are zero.
// STAGED: Pq[mm] is P_l^m at m = mtop - mm from the LDS copy (already zero for
// m < 0); otherwise Pq[-mm] from the global table, read at max(m, 0).  The two
// are separate instantiations so that each read has one address space: a
// pointer that may be either becomes a flat load, and waiting for a flat load
// waits for every global load in flight (here the tile's map values).
template <bool QUIRK, bool STAGED>
__device__ __forceinline__ void ring_emit(unsigned short *Yw, int tid, int mtop, bool pok, const double *Pq,
                                          int nr, int &smod, bool quirk, double &cm, double &sm, double &cp,
                                          double &sp, double c2, double &cm2, double &sm2, double cph2, double sph2) {
#pragma unroll
  for (int mp = 0; mp < MX_MC; mp += 2) {
    unsigned hi[2], lo[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int mm = mp + u;
      const int m = mtop - mm;
      // staged: index mm (the LDS copy holds zeros for m < 0, all inside it);
      // global: index -(mm) = P at m, or P at m = 0 for m < 0 (masked below)
      const double pr = STAGED ? Pq[mm] : Pq[-(m >= 0 ? mm : mtop)];
      const double pv = (pok && m >= 0) ? pr : 0.0;
      double c = cm, sn = sm;
      if constexpr (QUIRK) {
        const bool flip = quirk && m > 0 && 2 * smod > nr;
        c = flip ? cm2 : cm;
        sn = flip ? sm2 : sm;
      }
      split2((float)(pv * c), (float)(-pv * sn), hi[u], lo[u]);
      // m -> m - 1
      const double cn = fma(c2, cm, -cp), snn = fma(c2, sm, -sp);
      cp = cm;
      sp = sm;
      cm = cn;
      sm = snn;
      if constexpr (QUIRK) {
        const double cn2 = fma(cm2, cph2, sm2 * sph2), sn2 = fma(sm2, cph2, -cm2 * sph2);
        cm2 = cn2;
        sm2 = sn2;
        smod = smod == 0 ? nr - 1 : smod - 1;
      }
    }
    const int e = tid * S16_ROW + 2 * mp;
    *reinterpret_cast<uint2 *>(Yw + e) = make_uint2(hi[0], hi[1]);
    *reinterpret_cast<uint2 *>(Yw + S16_PLANE + e) = make_uint2(lo[0], lo[1]);
  }
}

// Equatorial tile: the chunk's Y* columns are e^{-i m phi} 2^12 alone (the
// tile's pixels share one ring; its Legendre values scale the FP64 sums per
// column afterwards), read from the LDS table E at index (m a) mod 8 nside,
// a = 2 j + s.  Columns of m < 0 and padding pixels are zero.
__device__ __forceinline__ void ring_emit_eq(unsigned short *Yw, int tid, int mtop, bool pok, int a, int mask,
                                             const uint2 *E) {
  int idx = (mtop * a) & mask;
#pragma unroll
  for (int mp = 0; mp < MX_MC; mp += 2) {
    unsigned hi[2], lo[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int m = mtop - mp - u;
      const uint2 e = E[idx];
      const bool ok = pok && m >= 0;
      hi[u] = ok ? e.x : 0u;
      lo[u] = ok ? e.y : 0u;
      idx = (idx - a) & mask;
    }
    const int e = tid * S16_ROW + 2 * mp;
    *reinterpret_cast<uint2 *>(Yw + e) = make_uint2(hi[0], hi[1]);
    *reinterpret_cast<uint2 *>(Yw + S16_PLANE + e) = make_uint2(lo[0], lo[1]);
  }
}

// A fragments of a tile from its map values (this lane: row, and points
// base + 32 s + 8 g + 4 hq .. + 3), zero past the item's pixels or rows.
template <int SPW>
__device__ __forceinline__ void ring_split_a(const float4 (&araw)[SPW][2], h8 (&ah)[2][SPW], int64_t base, int g,
                                             bool rowok, int64_t p1) {
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    unsigned hw[4], lw[4];
#pragma unroll
    for (int hq = 0; hq < 2; ++hq) {
      const int64_t p = base + 32 * s + 8 * g + 4 * hq;
      const bool ok = rowok && p < p1;
      const float4 v = ok ? araw[s][hq] : make_float4(0.f, 0.f, 0.f, 0.f);
      split2(v.x, v.y, hw[2 * hq], lw[2 * hq]);
      split2(v.z, v.w, hw[2 * hq + 1], lw[2 * hq + 1]);
    }
    ah[0][s] = __builtin_bit_cast(h8, (u4v){hw[0], hw[1], hw[2], hw[3]});
    ah[1][s] = __builtin_bit_cast(h8, (u4v){lw[0], lw[1], lw[2], lw[3]});
  }
}

// The contraction of one chunk: this wave's 16 rows x 32 columns over the
// tile's SPW 32-point steps, three products per step (leading p0 q0 in (a0,
// a1), folded every 64 points; corrections in (c0, c1), folded once).  EQ: the
// folds go to FP64 tile sums t, added to sum scaled by the column's Legendre
// value pc (one per column half); otherwise straight into sum.  All MFMAs and
// the reads of their results form one basic block.
template <int SPW, bool EQ>
__device__ __forceinline__ void ring_mfma_chunk(const h8 (&ah)[2][SPW], const unsigned short *Yh, int lane,
                                                v4d (&sum)[2], double pc0, double pc1) {
  constexpr int FOLD = 64;
  v4f a0 = v4f{0.f, 0.f, 0.f, 0.f}, a1 = a0, c0 = a0, c1 = a0;
  v4d t[2] = {v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}};
  v4d(&acc)[2] = EQ ? t : sum;
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    const h8 b00 = s16_bfrag(Yh, 32 * s, 0, lane);
    const h8 b01 = s16_bfrag(Yh, 32 * s, 16, lane);
    const h8 b10 = s16_bfrag(Yh + S16_PLANE, 32 * s, 0, lane);
    const h8 b11 = s16_bfrag(Yh + S16_PLANE, 32 * s, 16, lane);
    c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b10, c0, 0, 0, 0);
    c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b00, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b11, c1, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b01, c1, 0, 0, 0);
    a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b00, a0, 0, 0, 0);
    a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b01, a1, 0, 0, 0);
    if ((32 * (s + 1)) % FOLD == 0 && s + 1 < SPW) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[0][j] += (double)a0[j];
        acc[1][j] += (double)a1[j];
      }
      a0 = v4f{0.f, 0.f, 0.f, 0.f};
      a1 = a0;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    acc[0][j] += (double)a0[j] + (double)c0[j];
    acc[1][j] += (double)a1[j] + (double)c1[j];
  }
  if constexpr (EQ) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sum[0][j] = fma(pc0, t[0][j], sum[0][j]);
      sum[1][j] = fma(pc1, t[1][j], sum[1][j]);
    }
  }
}

// Work item: DirectItem (l, rows n0 .. n0 + nrows - 1 <= 64, m chunks [qlo,
// qhi) counted from the top, pixels [p0, p1), partial block at out).  Wave w
// owns rows 16w .. 16w+15; map rows of l start at maps + moff[l].
template <int QMAX>
__global__ __launch_bounds__(256, 2) void healpix_ring_contract_kernel(RingPixels rp,
                                                                    const DirectItem *__restrict__ items,
                                                                    const float *__restrict__ maps, int64_t npix,
                                                                    const int64_t *__restrict__ moff, double unscale,
                                                                    double *__restrict__ partial) {
  constexpr int FOLD = 64;
  constexpr int SPW = HR_TP / 32;
  static_assert(HR_TP == MX_TP, "the point-major Y* planes hold MX_TP rows");
  __shared__ __attribute__((aligned(16))) unsigned short Yh[S16_YBUF];
  // ring Legendre values of the tile's rings for the item's m range:
  // Ps[(rho - rho_lo) 16 QMAX + (mtop0 - m)]
  __shared__ double Ps[HR_RINGS * MX_MC * QMAX];

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
  const float *__restrict__ M = maps + moff[l] + (int64_t)(it.n0 + (rowok ? row : 0)) * npix;
  const double *__restrict__ P = rp.ptab + rp.poff[l];
  const bool alias = rp.alias_compat != 0;

  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};

  const int mtop0 = l - MX_MC * it.qlo;
  for (int64_t base = it.p0; base < it.p1; base += HR_TP) {
    // ---- 1. map values of this wave's rows into A fragments (zero past the
    // item's pixels or rows; item ranges and npix are multiples of 4)
    h8 ah[2][SPW];
    if (wave_rows) {
      // loads stay inside the item (clamped) and are zeroed past it, without branches
      const int64_t last = it.p1 - 4;
#pragma unroll
      for (int s = 0; s < SPW; ++s) {
        unsigned hw[4], lw[4];
#pragma unroll
        for (int hq = 0; hq < 2; ++hq) {
          const int64_t p = base + 32 * s + 8 * g + 4 * hq;
          float4 v = *reinterpret_cast<const float4 *>(M + (p < last ? p : last));
          const bool ok = rowok && p < it.p1;
          v = ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
          split2(v.x, v.y, hw[2 * hq], lw[2 * hq]);
          split2(v.z, v.w, hw[2 * hq + 1], lw[2 * hq + 1]);
        }
        ah[0][s] = __builtin_bit_cast(h8, (u4v){hw[0], hw[1], hw[2], hw[3]});
        ah[1][s] = __builtin_bit_cast(h8, (u4v){lw[0], lw[1], lw[2], lw[3]});
      }
    }

    // ---- 2. the tile's ring Legendre values into LDS (visible after the
    // first chunk's barrier; the previous tile's readers are past their last
    // chunk barrier), unless the tile spans more than HR_RINGS rings
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    const int64_t plast = min(base + HR_TP, it.p1) - 1;
    const int rho_lo = rp.ring[base], nrt = rp.ring[plast] - rho_lo + 1;
    const bool staged = nrt <= HR_RINGS;  // block-uniform
    const int mspan = MX_MC * (it.qhi - it.qlo);
    if (staged) {
      for (int e = tid; e < nrt * MX_MC * QMAX; e += 256) {
        const int rr = e / (MX_MC * QMAX), k = e - rr * (MX_MC * QMAX);
        const int m = mtop0 - k;
        Ps[e] = (k < mspan && m >= 0) ? P[(int64_t)(rho_lo + rr) * (l + 1) + m] : 0.0;
      }
    }
    // ---- 3. azimuth of this thread's pixel at the item's top m
    const double *Pr = P;
    const double *Pl = Ps;
    int nr = 1, smod = 0;
    double cm = 0.0, sm = 0.0, cp = 0.0, sp = 0.0, c2 = 0.0;
    double cm2 = 0.0, sm2 = 0.0, cph2 = 0.0, sph2 = 0.0;
    bool quirk = false;
    if (pok) {
      // slot gp holds pixel px (the maps are stored in slot order); the slots of
      // a tile hold the tile's pixels, so its rings are those of [base, plast]
      const int64_t px = rp.perm[gp];
      const int rho = rp.ring[px];
      Pr = P + (int64_t)rho * (l + 1);
      if (staged) Pl = Ps + (rho - rho_lo) * MX_MC * QMAX;
      nr = rp.np[px];
      const double phi = rp.phi[px], cph = rp.cphi[px], sph = rp.sphi[px];
      c2 = cph + cph;
      const double hi = (double)mtop0 * phi, lo = fma((double)mtop0, phi, -hi);
      double S, C;
      sincos(hi, &S, &C);
      cm = C - lo * S;
      sm = S + lo * C;
      cp = fma(cm, cph, -sm * sph);
      sp = fma(sm, cph, cm * sph);
      // the mirrored azimuth only matters where some m <= l folds past n_ring / 2
      quirk = alias && 2 * l > nr;
      if (quirk) {
        const double phi2 = rp.phi2[px];
        const double h2 = (double)mtop0 * phi2, l2 = fma((double)mtop0, phi2, -h2);
        sincos(h2, &S, &C);
        cm2 = C - l2 * S;
        sm2 = S + l2 * C;
        cph2 = rp.cphi2[px];
        sph2 = rp.sphi2[px];
      }
      smod = mtop0 % nr;
    }

    // waves without a polar pixel whose m fold skip the mirrored azimuth
    const bool wave_quirk = __builtin_amdgcn_ballot_w64(quirk) != 0;
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qlo + qq;
      if (q < it.qhi) {  // block-uniform
        __syncthreads();  // previous chunk's readers are done
        const int mtop = l - MX_MC * q;
        // Legendre values from LDS (staged) at index (mtop0 - m), else from the global table at m
        if (staged) {
          if (wave_quirk)
            ring_emit<true, true>(Yh, tid, mtop, pok, Pl + MX_MC * qq, nr, smod, quirk, cm, sm, cp, sp, c2, cm2, sm2,
                                  cph2, sph2);
          else
            ring_emit<false, true>(Yh, tid, mtop, pok, Pl + MX_MC * qq, nr, smod, quirk, cm, sm, cp, sp, c2, cm2, sm2,
                                   cph2, sph2);
        } else {
          if (wave_quirk)
            ring_emit<true, false>(Yh, tid, mtop, pok, Pr + mtop, nr, smod, quirk, cm, sm, cp, sp, c2, cm2, sm2, cph2,
                                   sph2);
          else
            ring_emit<false, false>(Yh, tid, mtop, pok, Pr + mtop, nr, smod, quirk, cm, sm, cp, sp, c2, cm2, sm2, cph2,
                                    sph2);
        }
        __syncthreads();
        if (wave_rows) {
          v4f a0 = v4f{0.f, 0.f, 0.f, 0.f}, a1 = a0, c0 = a0, c1 = a0;
#pragma unroll
          for (int s = 0; s < SPW; ++s) {
            const h8 b00 = s16_bfrag(Yh, 32 * s, 0, lane);
            const h8 b01 = s16_bfrag(Yh, 32 * s, 16, lane);
            const h8 b10 = s16_bfrag(Yh + S16_PLANE, 32 * s, 0, lane);
            const h8 b11 = s16_bfrag(Yh + S16_PLANE, 32 * s, 16, lane);
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b10, c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b00, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b11, c1, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1][s], b01, c1, 0, 0, 0);
            a0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b00, a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0][s], b01, a1, 0, 0, 0);
            if ((32 * (s + 1)) % FOLD == 0 && s + 1 < SPW) {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                sum[qq][0][j] += (double)a0[j];
                sum[qq][1][j] += (double)a1[j];
              }
              a0 = v4f{0.f, 0.f, 0.f, 0.f};
              a1 = a0;
            }
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            sum[qq][0][j] += (double)a0[j] + (double)c0[j];
            sum[qq][1][j] += (double)a1[j] + (double)c1[j];
          }
        }
      }
    }
  }

  // ---- the item's partial block (64 rows x its columns), FP64
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi && wave_rows) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);
        const int m = l - MX_MC * q - (lc >> 1);
        if (m >= 0) {
          const int cc = 2 * m + (lc & 1);
          // v_mfma_f32_16x16x32_f16 result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
          for (int j = 0; j < 4; ++j) out[(size_t)(16 * wave + 4 * g + j) * ncol + cc] = sum[qq][c][j] * unscale;
        }
      }
    }
  }
}

// The ring contraction of equatorial items (every tile in [eq_lo, eq_hi), so
// each 256-pixel tile lies in one ring of 4 nside pixels; rp.eq_ns != 0).  Per
// tile the B operand is e^{-i m phi} 2^12 alone, from the LDS table of 8 nside
// azimuth factors: no per-pixel recurrence or set-up.  The tile's FP64 sums are
// scaled per column by the ring's Legendre value and added to the item's
// sums.  Same items, layout and output as healpix_ring_contract_kernel.
// Dynamic LDS: the azimuth table (8 eq_ns uint2).
template <int QMAX>
__global__ __launch_bounds__(256, 2) void healpix_ring_eq_kernel(RingPixels rp, const DirectItem *__restrict__ items,
                                                              const float *__restrict__ maps, int64_t npix,
                                                              const int64_t *__restrict__ moff, double unscale,
                                                              double *__restrict__ partial) {
  constexpr int SPW = HR_TP / 32;
  static_assert(HR_TP == MX_TP, "the point-major Y* planes hold MX_TP rows");
  __shared__ __attribute__((aligned(16))) unsigned short Yh[S16_YBUF];
  __shared__ double Pe[MX_MC * QMAX];  // the tile ring's P_l^m 2^-12 at index mtop0 - m (0 for m < 0)
  extern __shared__ uint2 Eq[];

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
  const float *__restrict__ M = maps + moff[l] + (int64_t)(it.n0 + (rowok ? row : 0)) * npix;
  const double *__restrict__ P = rp.ptab + rp.poff[l];
  const int64_t ns = rp.eq_ns;
  const int mask = (int)(8 * ns - 1);
  for (int k = tid; k < 8 * ns; k += 256) Eq[k] = rp.eqtab[k];  // visible after the first chunk's barrier

  v4d sum[QMAX][2];
#pragma unroll
  for (int q = 0; q < QMAX; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) sum[q][c] = v4d{0.0, 0.0, 0.0, 0.0};

  const int mtop0 = l - MX_MC * it.qlo;
  const int mspan = MX_MC * (it.qhi - it.qlo);
  for (int64_t base = it.p0; base < it.p1; base += HR_TP) {
    // this thread's pixel (slot gp holds pixel perm[gp], inside the tile) and
    // the tile's ring, then its Legendre values, then the map values (issued
    // last: a wait on a global load also waits on those issued before it)
    const int64_t gp = base + tid;
    const bool pok = gp < it.p1;
    const int64_t px = rp.perm[pok ? gp : base];
    const int rho = rp.ring[base];
    const int mk = mtop0 - tid;
    const bool pk = tid < mspan && mk >= 0;
    const double pv = P[pk ? (int64_t)rho * (l + 1) + mk : 0];
    // the tile's A fragments (the wait for the map values is exposed once
    // per tile: the table generation below is too short to cover it)
    h8 ah[2][SPW];
    {
      float4 araw[SPW][2];
      const int64_t last = it.p1 - 4;
#pragma unroll
      for (int s = 0; s < SPW; ++s)
#pragma unroll
        for (int hq = 0; hq < 2; ++hq) {
          const int64_t p = base + 32 * s + 8 * g + 4 * hq;
          araw[s][hq] = *reinterpret_cast<const float4 *>(M + (p < last ? p : last));
        }
      if (wave_rows) ring_split_a(araw, ah, base, g, rowok, it.p1);
    }
    // (the previous tile's readers of Pe are past its last chunk barrier)
    if (tid < MX_MC * QMAX) Pe[tid] = pk ? pv * 0x1p-12 : 0.0;
    // azimuth phi_j = (2 j + s) pi / (4 nside) of the pixel: table step a = 2 j + s
    const int64_t r1 = (int64_t)rho + 1 - ns;  // ring number - nside
    const int64_t start = 2 * ns * (ns - 1) + r1 * 4 * ns;
    const int ea = (int)(2 * (px - start)) + ((r1 & 1) == 0 ? 1 : 0);
#pragma unroll
    for (int qq = 0; qq < QMAX; ++qq) {
      const int q = it.qlo + qq;
      if (q < it.qhi) {  // block-uniform
        __syncthreads();  // previous chunk's readers are done; Pe and Eq written
        const int mtop = l - MX_MC * q;
        ring_emit_eq(Yh, tid, mtop, pok, ea, mask, Eq);
        // the Legendre values of this lane's two result columns, read before
        // the barrier below (after which the next tile may rewrite Pe)
        const int m0 = mtop - ((lane & 15) >> 1), m1 = m0 - MX_MC / 2;
        const double pc0 = m0 >= 0 ? Pe[mtop0 - m0] : 0.0;
        const double pc1 = m1 >= 0 ? Pe[mtop0 - m1] : 0.0;
        __syncthreads();
        if (wave_rows) ring_mfma_chunk<SPW, true>(ah, Yh, lane, sum[qq], pc0, pc1);
      }
    }
  }

  // ---- the item's partial block (64 rows x its columns), FP64
  double *out = partial + it.out;
#pragma unroll
  for (int qq = 0; qq < QMAX; ++qq) {
    const int q = it.qlo + qq;
    if (q < it.qhi && wave_rows) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lc = 16 * c + (lane & 15);
        const int m = l - MX_MC * q - (lc >> 1);
        if (m >= 0) {
          const int cc = 2 * m + (lc & 1);
          // v_mfma_f32_16x16x32_f16 result: element j of lane i is row 4(i/16) + j, column i%16
#pragma unroll
          for (int j = 0; j < 4; ++j) out[(size_t)(16 * wave + 4 * g + j) * ncol + cc] = sum[qq][c][j] * unscale;
        }
      }
    }
  }
}

}  // namespace sfb
