// sfb_lowl.hpp -- FP64 kernel for the lowest multipoles of the mixed path
// (l < mixed_fp64_below, at most 8; by default l = 0 and 1).
//
// For l < 8 the 2(l+1) real columns of Y* fit one 16-wide MFMA column tile,
// so each wave runs independently: per step of 4 points, lane (k = lane/16,
// c = lane%16) evaluates Y*_l,c/2 of point k (column c, real or imaginary part)
// -- the B fragment of v_mfma_f64_16x16x4f64 -- and, for each 16-row tile of
// n, j_l(k_ln r) of row lane%16 and point k -- the A fragment -- from the
// quintic FP64 table (error ~1e-10 of the amplitude).  An item is one 16-row
// tile (many small items keep enough waves in flight for these two or so
// multipoles); the point data of each 256-point tile is staged in LDS, and the
// four waves' blocks are summed, in a fixed order, at the end of the item.  (The FP32 cubic Hermite table of
// the mixed kernel is not accurate enough here: l = 0 and 1 have C_l with one
// or three real degrees of freedom, so some c_ln are far below their noise
// level and their relative error is magnified.)
//
// Y*_lm = Pbar_l^m(cos t) (cos m phi - i sin m phi) with the orthonormal
// Pbar (Condon-Shortley phase, gsl_sf_legendre_sphPlm as at sfb_raw.cpp:73):
//   Pbar_m^m = (-1)^m sqrt((2m+1)/(4 pi)) sqrt((2m-1)!!/(2m)!!) sin^m t,
//   Pbar_L^m = a_L (x Pbar_{L-1}^m - Pbar_{L-2}^m / a_{L-1}),  a_L = sqrt((4L^2-1)/(L^2-m^2)),
// at most 7 steps here, so no exponent scaling is needed.
#pragma once

#include "sfb_common.hpp"
#include "sfb_direct_f64.hpp"

namespace sfb {

constexpr int LW_LMAX = 8;   // multipoles handled: l < LW_LMAX
constexpr int LW_ROWS = 16;  // rows of n per item (one 16-row MFMA tile: many small items fill the chip)
constexpr int LW_TILES = LW_ROWS / 16;
constexpr int LW_STEP = 16;  // points per block step (4 waves x 4 points)
constexpr int LW_TILE = 256; // points staged in LDS per block iteration

__global__ __launch_bounds__(256) void sfb_lowl_f64_kernel(PointsSoA pts, const DirectItem *__restrict__ items,
                                                           const double *__restrict__ kln, BesselTableRef tab,
                                                           double *__restrict__ partial) {
  __shared__ double2 acoef[LW_LMAX * LW_LMAX];  // [m][L]: (a_L, 1 / a_{L-1})
  __shared__ double sect[LW_LMAX];
  // one LDS area: the tile's point data during the sweep, the waves' blocks at the end
  constexpr int LW_LDS = 5 * LW_TILE > 3 * LW_ROWS * 16 ? 5 * LW_TILE : 3 * LW_ROWS * 16;
  __shared__ double lds[LW_LDS];
  double(*pt)[LW_TILE] = reinterpret_cast<double(*)[LW_TILE]>(lds);  // r (or -1), cos t, sin t, cos phi, sin phi
  double(*red)[LW_ROWS * 16] = reinterpret_cast<double(*)[LW_ROWS * 16]>(lds);

  const DirectItem it = items[blockIdx.x];
  const int l = it.l;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < LW_LMAX * LW_LMAX) {
    const int m = tid / LW_LMAX, L = tid % LW_LMAX;
    const double a = L > m ? sqrt((4.0 * L * L - 1.0) / ((double)L * L - (double)m * m)) : 0.0;
    const int K = L - 1;
    const double ia = K > m ? sqrt(((double)K * K - (double)m * m) / (4.0 * K * K - 1.0)) : 0.0;
    acoef[tid] = make_double2(a, ia);
  }
  if (tid < LW_LMAX) {
    const int m = tid;
    double q = 1.0;
    for (int k = 1; k <= m; ++k) q *= (2.0 * k - 1.0) / (2.0 * k);
    const double v = sqrt((2.0 * m + 1.0) / (4.0 * 3.14159265358979323846) * q);
    sect[m] = (m & 1) ? -v : v;
  }
  __syncthreads();

  const int col = lane & 15, m = col >> 1, kq = lane >> 4;
  const bool live_col = m <= l;
  const double *kl = kln + it.kofs;
  double kr[LW_TILES];
#pragma unroll
  for (int t = 0; t < LW_TILES; ++t) {
    const int row = 16 * t + (lane & 15);
    kr[t] = row < it.nrows ? kl[row] : 0.0;
  }
  const int ntile = (it.nrows + 15) >> 4;
  const double2 *tb = reinterpret_cast<const double2 *>(tab.q5 + 6 * (size_t)tab.q5_off[l]);
  const int last = tab.q5_nint[l] - 1;
  const double s0 = sect[m];

  v4d acc[LW_TILES];
#pragma unroll
  for (int t = 0; t < LW_TILES; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};

  // 256-point tiles: the block stages the tile's point data in LDS with one
  // coalesced load per thread, then wave w takes points 64w .. 64w+63, four per step
  for (int64_t tb0 = it.p0; tb0 < it.p1; tb0 += LW_TILE) {
    {
      const int64_t gp = tb0 + tid;
      const bool ok = gp < it.p1;
      pt[0][tid] = ok ? pts.r[gp] : -1.0;
      pt[1][tid] = ok ? pts.x[gp] : 0.0;
      pt[2][tid] = ok ? pts.s[gp] : 0.0;
      pt[3][tid] = ok ? pts.cphi[gp] : 1.0;
      pt[4][tid] = ok ? pts.sphi[gp] : 0.0;
    }
    __syncthreads();
#pragma unroll 2
    for (int st = 0; st < LW_TILE / 16; ++st) {
      const int pl = 64 * wave + 4 * st + kq;
      const double r0 = pt[0][pl];
      const bool ok = r0 >= 0.0;
      const double r = ok ? r0 : 0.0, x = pt[1][pl], s = pt[2][pl], cp = pt[3][pl], sp = pt[4][pl];
      // B fragment: Y* of point kq, column col
      double P = s0, c = 1.0, sn = 0.0;
      for (int k = 0; k < m; ++k) {
        P *= s;
        const double c2 = c * cp - sn * sp;
        sn = fma(sn, cp, c * sp);
        c = c2;
      }
      double prev = 0.0;
      for (int L = m + 1; L <= l; ++L) {
        const double2 a = acoef[m * LW_LMAX + L];
        const double nx = a.x * fma(x, P, -a.y * prev);
        prev = P;
        P = nx;
      }
      double y = (col & 1) ? -P * sn : P * c;
      y = (ok && live_col) ? y : 0.0;
      // A fragments: j_l(k r) of rows 16t + lane%16, point kq
#pragma unroll
      for (int t = 0; t < LW_TILES; ++t) {
        if (t < ntile) {
          const double u = kr[t] * r * Q5_HINV;
          int i = (int)u;
          i = i < last ? i : last;
          i = i > 0 ? i : 0;
          const double w = u - (double)i;
          const double2 a0 = tb[3 * i], a1 = tb[3 * i + 1], a2 = tb[3 * i + 2];
          const double jv = fma(w, fma(w, fma(w, fma(w, fma(w, a2.y, a2.x), a1.y), a1.x), a0.y), a0.x);
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(jv, y, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();  // the next tile overwrites pt
  }

  // fixed-order sum of the four waves' blocks; v_mfma_f64_16x16x4f64 result:
  // element j of lane i is row i/16 + 4j, column i%16
  if (wave > 0) {
#pragma unroll
    for (int t = 0; t < LW_TILES; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) red[wave - 1][(16 * t + (lane >> 4) + 4 * j) * 16 + col] = acc[t][j];
  }
  __syncthreads();
  if (wave == 0) {
    const int ncol = 2 * (l + 1);
    double *out = partial + it.out;
#pragma unroll
    for (int t = 0; t < LW_TILES; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 16 * t + (lane >> 4) + 4 * j;
        const int e = row * 16 + col;
        const double v = acc[t][j] + red[0][e] + red[1][e] + red[2][e];
        if (row < it.nrows && col < ncol) out[(size_t)row * ncol + col] = v;
      }
  }
}

}  // namespace sfb
