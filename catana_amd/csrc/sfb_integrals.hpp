// sfb_integrals.hpp -- batched double spherical-Bessel integrals
//
//   I_t = int_0^R w(r) j_l(k1 r) j_l(k2 r) dr      for many triples t = (l, k1, k2)
//
// (besseltools::double_sbessel_integrator, sbessel_integrator.cpp:45-110; the
// shot-noise term C_l^sn(k,k) = (V/N)(2/pi) k^2 int r^2 w(r) j_l(k r)^2 dr is
// one such integral per (l, n)).
//
// Ranges follow the reference's generator (SBesselIntegrationRangeGenerator,
// sbessel_integrator.cpp:113-138): with ka = min(k1,k2) and kb = max, outer
// ranges run between consecutive zeros of j_l(ka r) ("A" breakpoints
// z_i / ka), each split at the zeros of j_l(kb r) ("C" breakpoints z_j / kb);
// the last range ends at R.  The merged, sorted breakpoints E_0 < E_1 < ...
// are found per interval by a merge-path binary search over the two scaled
// copies of one zero row, so every interval is addressable without a
// sequential walk.  Because kb >= ka, C_0 <= A_0 and the first interval is
// [0, C_0]: it holds the whole rise of j_l(kb r) (x < l), so it is cut into
// S0 pieces of at most 8 units of kb r (and at most R/32).  Every later
// interval lies between two consecutive C zeros, at most the first gap
// z_1 - z_0 long; it is cut into P pieces of at most R/32.  Each piece is
// integrated by 16-point Gauss-Legendre.
//
// Stopping rule (sbessel_integrator.cpp:97-100): after each outer range,
// result += range sum; stop once |result - old| < result * 1e-10 (a signed
// test: a non-positive running sum never stops).  One wave per triple walks
// its pieces four at a time (16 lanes per piece), reduces each piece with a
// fixed butterfly, and applies the rule in wave-uniform code, so the early
// exit also skips the work past the stopping point.
#pragma once

#include "sfb_common.hpp"
#include "sfb_math.hpp"

namespace sfb {

constexpr int IG_NODES = 16;     // Gauss-Legendre nodes per piece
constexpr double IG_XSTEP = 8.0; // first interval: pieces of at most 8 units of kb r
constexpr int IG_RFRAC = 32;     // pieces of at most R / 32

enum : int { W_NODES = 0, W_TABLE = 1 };

// Per-triple plan, written by sbessel_plan_kernel.
struct IntegralPlan {
  int32_t l;
  int32_t c1, c2;  // A and C breakpoints below R
  int32_t s0, p;   // pieces of the first interval and of each later one
  int32_t pad;
  int64_t npieces;
  double ka, kb;
};

struct WeightRef {
  int32_t kind;
  int32_t ntab;
  int32_t r_power;  // w(r) r^r_power
  int32_t pad;
  const double *values;  // W_NODES: w at every node of the call; W_TABLE: samples
  const int64_t *node_off;  // W_NODES: first node of each triple
  double x_min, dx_inv, x_max;
};

// number of zeros z_0..z_{nz-1} of the row with z_i / k < R
__device__ __forceinline__ int ig_count_below(const double *__restrict__ z, int nz, double k, double R) {
  if (!(k > 0.0)) return 0;
  int lo = 0, hi = nz;  // first index with z_i / k >= R
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (z[mid] / k < R)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void sbessel_plan_kernel(int64_t n, const int32_t *__restrict__ ls, const double *__restrict__ k1s,
                                    const double *__restrict__ k2s, double R, const double *__restrict__ zeros,
                                    int nz, IntegralPlan *__restrict__ plan, int *__restrict__ err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  IntegralPlan P;
  P.l = ls[t];
  P.ka = fmin(k1s[t], k2s[t]);
  P.kb = fmax(k1s[t], k2s[t]);
  P.pad = 0;
  const double *z = zeros + (size_t)P.l * nz;
  P.c1 = ig_count_below(z, nz, P.ka, R);
  P.c2 = ig_count_below(z, nz, P.kb, R);
  // the zero row must reach past the last breakpoint (the host sizes it so)
  if (P.c2 >= nz - 1) atomicOr(err, 1);
  const double e0 = P.c2 > 0 ? z[0] / P.kb : R;
  const int s_x = P.kb > 0.0 ? (int)ceil(e0 * P.kb / IG_XSTEP) : 1;
  const int s_r = (int)ceil(IG_RFRAC * e0 / R);
  P.s0 = max(1, max(s_x, s_r));
  const int64_t M = (int64_t)P.c1 + P.c2 + 1;
  P.p = 1;
  if (M > 1) P.p = max(1, (int)ceil(IG_RFRAC * (z[1] - z[0]) / (P.kb * R)));
  P.npieces = P.s0 + (int64_t)P.p * (M - 1);
  plan[t] = P;
}

// E_g, the (g+1)-th smallest of A = z[0..c1)/ka and C = z[0..c2)/kb (A first
// on ties), and whether it is an A breakpoint (it then closes an outer range).
__device__ __forceinline__ double ig_breakpoint(const double *__restrict__ z, const IntegralPlan &P, int g,
                                                bool *from_a) {
  const int k = g + 1;
  int lo = max(0, k - P.c2), hi = min(k, P.c1);  // i = number of A in the first k
  // largest i in [lo, hi] with i == lo or (k - i == c2) or A[i-1] <= C[k-i]
  while (lo < hi) {
    const int i = (lo + hi + 1) >> 1;
    const bool ok = (k - i >= P.c2) || (z[i - 1] / P.ka <= z[k - i] / P.kb);
    if (ok)
      lo = i;
    else
      hi = i - 1;
  }
  const int i = lo, j = k - i;
  const double a = i > 0 ? z[i - 1] / P.ka : -1.0;
  const double c = j > 0 ? z[j - 1] / P.kb : -1.0;
  *from_a = i > 0 && (j == 0 || a > c);
  return fmax(a, c);
}

// Piece p of the triple: its range [lo, hi], the interval it belongs to, and
// whether it is the last piece of an interval that closes an outer range.
__device__ __forceinline__ void ig_piece(const double *__restrict__ z, const IntegralPlan &P, double R, int64_t p,
                                         double *lo, double *hi, bool *closes) {
  const int M = P.c1 + P.c2 + 1;
  int g, sub, ns;
  if (p < P.s0) {
    g = 0;
    sub = (int)p;
    ns = P.s0;
  } else {
    const int64_t q = p - P.s0;
    g = 1 + (int)(q / P.p);
    sub = (int)(q - (int64_t)(g - 1) * P.p);
    ns = P.p;
  }
  bool fa = false, dummy;
  const double b = g < M - 1 ? ig_breakpoint(z, P, g, &fa) : R;
  const double a = g > 0 ? ig_breakpoint(z, P, g - 1, &dummy) : 0.0;
  const double w = (b - a) / ns;
  *lo = a + w * sub;
  *hi = sub == ns - 1 ? b : a + w * (sub + 1);
  *closes = (sub == ns - 1) && (g == M - 1 || fa);
}

__device__ __forceinline__ double ig_weight(const WeightRef &W, int64_t node, double r, int *err) {
  double rp = 1.0;
  for (int i = 0; i < W.r_power; ++i) rp *= r;
  if (W.kind == W_NODES) return rp * W.values[node];
  // FunctionInterpolator::operator() (FunctionInterpolator.cpp:21-30)
  if (!(r >= W.x_min && r < W.x_max)) {
    atomicOr(err, ERR_DOMAIN);
    return 0.0;
  }
  const double u = (r - W.x_min) * W.dx_inv;
  int i = (int)u;
  i = min(i, W.ntab - 2);
  return rp * (W.values[i] + (W.values[i + 1] - W.values[i]) * (u - (double)i));
}

// Quadrature nodes of every piece, for a caller that evaluates w itself:
// node (p, q) of triple t at r_nodes[off[t] + 16 p + q].
__global__ __launch_bounds__(256) void sbessel_nodes_kernel(int64_t n, const IntegralPlan *__restrict__ plan,
                                                            const double *__restrict__ zeros, int nz, double R,
                                                            const double *__restrict__ gl,
                                                            const int64_t *__restrict__ off,
                                                            double *__restrict__ r_nodes) {
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n) return;
  const int lane = threadIdx.x & 63, q = lane & 15;
  const IntegralPlan P = plan[t];
  const double *z = zeros + (size_t)P.l * nz;
  const double x = gl[q];
  for (int64_t p = lane >> 4; p < P.npieces; p += 4) {
    double lo, hi;
    bool closes;
    ig_piece(z, P, R, p, &lo, &hi, &closes);
    r_nodes[off[t] + IG_NODES * p + q] = 0.5 * (lo + hi) + 0.5 * (hi - lo) * x;
  }
}

// Sort keys of the order in which the integral kernel takes the triples:
// most pieces first (ascending keys), so the long triples do not form the tail.
__global__ void sbessel_order_keys_kernel(int64_t n, const IntegralPlan *__restrict__ plan, uint32_t *__restrict__ keys,
                                          int32_t *__restrict__ idx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t np = plan[t].npieces;
  keys[t] = 0xffffffffu - (uint32_t)(np < 0xffffffffLL ? np : 0xffffffffLL);
  idx[t] = (int32_t)t;
}

// order (nullable): wave slot s integrates triple order[s]; out[t] is written
// for the triple t itself.
__global__ __launch_bounds__(256) void sbessel_integral_kernel(int64_t n, const IntegralPlan *__restrict__ plan,
                                                               const double *__restrict__ zeros, int nz, double R,
                                                               const double *__restrict__ gl, WeightRef W,
                                                               const int32_t *__restrict__ order,
                                                               double *__restrict__ out, int *__restrict__ err) {
  const int64_t slot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slot >= n) return;  // whole waves leave together
  const int64_t t = order ? (int64_t)order[slot] : slot;
  const int lane = threadIdx.x & 63, q = lane & 15, grp = lane >> 4;
  const IntegralPlan P = plan[t];
  const double *z = zeros + (size_t)P.l * nz;
  const double x = gl[q], wq = gl[IG_NODES + q];
  const int64_t nbase = W.kind == W_NODES ? W.node_off[t] : 0;
  const bool same = P.ka == P.kb;
  double result = 0.0, cur = 0.0;
  bool stop = false;
  for (int64_t base = 0; base < P.npieces && !stop; base += 4) {
    const int64_t p = base + grp;
    double v = 0.0;
    bool closes = false;
    if (p < P.npieces) {
      double lo, hi;
      ig_piece(z, P, R, p, &lo, &hi, &closes);
      const double h = 0.5 * (hi - lo), r = 0.5 * (lo + hi) + h * x;
      const double ja = sbessel_jl(P.l, P.ka * r);
      const double jb = same ? ja : sbessel_jl(P.l, P.kb * r);
      v = wq * h * ig_weight(W, nbase + IG_NODES * p + q, r, err) * ja * jb;
    }
    // fixed-order sum over the 16 lanes of the piece
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    // wave-uniform: the four pieces in order, with the stopping rule
    for (int j = 0; j < 4; ++j) {
      if (base + j >= P.npieces) break;
      const double vj = __shfl(v, 16 * j, 64);
      const bool cj = __shfl((int)closes, 16 * j, 64) != 0;
      cur += vj;
      if (cj) {
        const double old = result;
        result += cur;
        cur = 0.0;
        if (fabs(result - old) < result * 1e-10) {
          stop = true;
          break;
        }
      }
    }
  }
  if (lane == 0) out[t] = result;
}

}  // namespace sfb
