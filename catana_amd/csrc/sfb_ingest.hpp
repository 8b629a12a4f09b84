// sfb_ingest.hpp -- the ingestion step before the path on the device: record
// conversion, the reference's point filters and an order-preserving
// compaction into the AoS (r, theta, phi) catalogue the SFB kernels read.
//
// Three passes: a flag per record (conversion as far as the filters need it),
// an exclusive scan of the flags (hipcub), and the conversion of the kept
// records straight into their place in the output.
// Records:
//   POINT          (r, theta, phi) f64 -- catana::Point as read by a Source
//   CARTESIAN_F32 / _F64   (x, y, z): point_from_box_position(x, y, z, shift)
//                  (record_types.hpp CartesianRecord::point, Point.cpp:5-21):
//                  r = |x|, theta = atan2(sqrt(x^2 + y^2), z), phi = atan2(y, x) in [0, 2 pi)
//   SPHERICAL_F32 / _F64   (r, theta, phi) (SphericalRecordBase::point)
// Filters, applied in order, a point is kept when it passes all
// (radial_filters.cpp:10-33, angular_filters.cpp:9-12):
//   TOPHAT         r < a
//   GAUSSIAN       w(r) > u with w = exp(-(r/a)^2)
//   WINDOW_TABLE   w(r) > u with w linearly interpolated from samples on
//                  [x_min, x_max] (FunctionInterpolator; outside: domain error)
//   WINDOW_VALUES  values[i] > u (w evaluated by the caller at record i)
//   ANGULAR_MASK   mask[ang2pix_ring(nside, theta, phi)] == 1.0f
// u is uniform in [0, 1): the reference draws it from a global std::mt19937
// in the order points reach the filter; here it is a counter-based hash of
// (seed, filter slot, record index) -- SplitMix64's finaliser (Steele, Lea,
// Flood 2014) -- so every record's draw is independent of the launch shape
// and of the other filters.
#pragma once

#include "sfb_common.hpp"
#include "sfb_healpix.hpp"

namespace sfb {

enum : int { REC_POINT = 0, REC_CART_F32 = 1, REC_CART_F64 = 2, REC_SPH_F32 = 3, REC_SPH_F64 = 4 };
enum : int { FLT_TOPHAT = 0, FLT_GAUSSIAN = 1, FLT_WINDOW_TABLE = 2, FLT_WINDOW_VALUES = 3, FLT_ANGULAR_MASK = 4 };
constexpr int IN_MAX_FILTERS = 8;

struct FilterDesc {
  int32_t kind, nside;
  double a, x_min, x_max, dx_inv;
  int64_t n;
  const void *values;
};

struct FilterSet {
  FilterDesc f[IN_MAX_FILTERS];
  int32_t count;
};

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// the uniform draw of filter slot `slot` for record `index`
__host__ __device__ inline double filter_uniform(uint64_t seed, int slot, uint64_t index) {
  const uint64_t s = splitmix64(seed ^ (0x9e3779b97f4a7c15ULL * (uint64_t)(slot + 1)));
  return (double)(splitmix64(s + index) >> 11) * 0x1.0p-53;
}

struct Rec3F {
  float v[3];
};
struct Rec3D {
  double v[3];
};

__device__ inline void load_record(const void *rec, int kind, int64_t i, double *a, double *b, double *c) {
  if (kind == REC_CART_F32 || kind == REC_SPH_F32) {
    const Rec3F q = static_cast<const Rec3F *>(rec)[i];
    *a = q.v[0];
    *b = q.v[1];
    *c = q.v[2];
  } else {
    const Rec3D q = static_cast<const Rec3D *>(rec)[i];
    *a = q.v[0];
    *b = q.v[1];
    *c = q.v[2];
  }
}

// the direction of a shifted cartesian record (Point.cpp:5-10)
__device__ inline void cart_angles(double x, double y, double z, double *th, double *ph) {
  *th = atan2(sqrt(x * x + y * y), z);
  double f = (x == 0.0 && y == 0.0) ? 0.0 : atan2(y, x);
  if (f < 0.0) f += 6.283185307179586476925;
  if (f >= 6.283185307179586476925) f -= 6.283185307179586476925;
  *ph = f;
}

// Pass 1: the keep flag of every record.  The radius comes first; the
// direction of a cartesian record is computed only when a mask needs it.
__global__ void ingest_flag_kernel(const void *__restrict__ rec, int64_t n, int kind, double shift, FilterSet fs,
                                   uint64_t seed, uint64_t index_base, int32_t *__restrict__ flag,
                                   int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a, b, c;
  load_record(rec, kind, i, &a, &b, &c);
  const bool cart = kind == REC_CART_F32 || kind == REC_CART_F64;
  double r, th = b, ph = c;
  if (cart) {
    a += shift;
    b += shift;
    c += shift;
    r = sqrt(a * a + b * b + c * c);
  } else {
    r = a;
  }
  bool have_dir = !cart, ok = true;
  for (int k = 0; k < fs.count && ok; ++k) {
    const FilterDesc &F = fs.f[k];
    if (F.kind == FLT_TOPHAT) {
      ok = r < F.a;
    } else if (F.kind == FLT_ANGULAR_MASK) {
      if (!have_dir) {
        cart_angles(a, b, c, &th, &ph);
        have_dir = true;
      }
      const int64_t pix = (th == th && ph == ph) ? hp_ang2pix(F.nside, th, ph) : -1;  // NaN: no pixel
      ok = pix >= 0 && pix < F.n && static_cast<const float *>(F.values)[pix] == 1.0f;
    } else {
      double w;
      if (F.kind == FLT_GAUSSIAN) {
        const double q = r / F.a;
        w = exp(-q * q);
      } else if (F.kind == FLT_WINDOW_VALUES) {
        w = static_cast<const double *>(F.values)[i];
      } else {
        if (!(r >= F.x_min && r < F.x_max)) {
          atomicOr(err, ERR_DOMAIN);
          w = 0.0;
        } else {
          const double *y = static_cast<const double *>(F.values);
          const double u = (r - F.x_min) * F.dx_inv;
          int j = (int)u;
          j = j < (int)F.n - 2 ? j : (int)F.n - 2;
          w = y[j] + (y[j + 1] - y[j]) * (u - (double)j);
        }
      }
      ok = w > filter_uniform(seed, k, index_base + (uint64_t)i);
    }
  }
  flag[i] = ok ? 1 : 0;
}

// Pass 3: kept record i goes, converted, to out[pos[i]] (pos: exclusive
// prefix sum of the flags); only kept records are read again.
__global__ void compact_convert_kernel(const void *__restrict__ rec, const int32_t *__restrict__ flag,
                                       const int64_t *__restrict__ pos, int64_t n, int kind, double shift,
                                       double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  double a, b, c;
  load_record(rec, kind, i, &a, &b, &c);
  double r, th, ph;
  if (kind == REC_CART_F32 || kind == REC_CART_F64) {
    a += shift;
    b += shift;
    c += shift;
    r = sqrt(a * a + b * b + c * c);
    cart_angles(a, b, c, &th, &ph);
  } else {
    r = a;
    th = b;
    ph = c;
  }
  const int64_t o = pos[i];
  out[3 * o] = r;
  out[3 * o + 1] = th;
  out[3 * o + 2] = ph;
}

// keys of a uniform random subset: the k smallest keys win
__global__ void subset_keys_kernel(int64_t n, uint64_t seed, uint64_t *__restrict__ keys,
                                   int64_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = splitmix64(splitmix64(seed ^ 0x5851f42d4c957f2dULL) + (uint64_t)i);
  idx[i] = i;
}

__global__ void gather_aos_kernel(const double *__restrict__ in, const int64_t *__restrict__ idx, int64_t k,
                                  double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int64_t j = idx[i];
  out[3 * i] = in[3 * j];
  out[3 * i + 1] = in[3 * j + 1];
  out[3 * i + 2] = in[3 * j + 2];
}

}  // namespace sfb
