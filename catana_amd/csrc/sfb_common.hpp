// sfb_common.hpp -- types shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sfb {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Device-side error bits (OR-ed into one int by any thread that hits them).
enum : int {
  ERR_DOMAIN = 1,   // Bessel argument outside the reference table range
  ERR_BADPOINT = 2,  // r < 0 or a non-finite coordinate
  ERR_BADCSR = 4     // pixel offsets that are not a CSR of the radii
};

// Per-point data derived once from the AoS (r, theta, phi) input; SoA, FP64.
struct PointsSoA {
  const double *r, *x, *s, *phi, *cphi, *sphi;
  int64_t n;
};

// Per-pixel data of a HEALPix RING map (the "points" of the pixelised path):
// ring cos(theta) and sin(theta), pixel-centre azimuth phi = phi0 + 2 pi j/np
// and its cos/sin, the mirrored azimuth 2 phi0 - phi used where the reference
// reads an unconjugated folded FFT bin, the ring length np and the CSR of radii.
struct PixelsSoA {
  const double *x, *s, *phi, *cphi, *sphi, *phi2, *cphi2, *sphi2;
  const int32_t *np;
  const int64_t *off;
  const double *radii;
  const int32_t *perm;  // contraction order: slot i holds pixel perm[i]
  int64_t n;  // pixels
  int32_t alias_compat;
};

// One work item of the direct kernels: multipole l, rows [n0, n0+nrows) of
// the mode index n, points [p0, p1), and where its partial sums go.
struct DirectItem {
  int32_t l;
  int32_t n0;
  int32_t nrows;     // valid rows (<= tile height)
  int32_t qlo, qhi;  // m chunks [qlo, qhi) handled by this item
  int32_t cat;       // catalogue of a batched call (0 otherwise)
  int32_t tab;       // mixed path: Bessel-table intervals the item's arguments reach
  int32_t flags;     // mixed path: ITEM_FULL_WINDOW -- points not sorted by radius, window from x = 0
  int64_t p0, p1;    // point (or pixel) range
  int64_t out;       // offset (in doubles) of this item's partial block
  int64_t kofs;      // offset of k_{l, n0} in the k_ln array of the call
};

// DirectItem::flags
enum : int32_t {
  ITEM_FULL_WINDOW = 1  // the item's points are in any order (radii in [0, R]): its Bessel
                        // window covers nodes 0 .. tab + 1, not [k_first r_first, k_last r_last]
};

// One (catalogue, l, row-tile) group: its items are `nsplit` consecutive point
// ranges whose partial blocks sit `stride` doubles apart starting at `base`.
// The sum lands in f (complex, interleaved) at ((fofs + m nmax + n) * 2 + re/im),
// scaled by norm * k_ln with k_ln at kofs + row.
struct ReduceGroup {
  int32_t l;
  int32_t n0;
  int32_t nrows;
  int32_t nsplit;
  int64_t base;
  int64_t stride;
  int64_t fofs;
  int64_t kofs;
  double norm;
  int32_t nmax;
  int32_t pad;
};

// Per-l constants for the Legendre rows:
//   sect[l] = (mantissa, exponent) of (-1)^l sqrt((2l+1)/(4 pi) prod_{k<=l} (2k-1)/(2k)),
//   so that P_l^l(cos t) = sect * sin^l t;
//   coef[coef_off[l] + m] = (a_m, b_m) of the downward recurrence
//   P_l^{m-1} = -(a_m cot t P_l^m + b_m P_l^{m+1}).
struct LegendreTables {
  const double2 *sect;
  const double2 *coef;
  const int32_t *coef_off;
};

}  // namespace sfb
