// sfb_capi.hip -- the C-ABI of libcatana_hip.so (declared in include/catana_hip.h).
//
// Host side of the drop-in boundary: validates arguments the way the
// reference does (exceptions become status codes), keeps a per-device
// workspace and cached k_ln / Legendre / interpolation tables in the handle,
// plans the work items of the direct and HEALPix kernels and launches them on
// the caller's stream.
#include "catana_hip.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <memory>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sfb_common.hpp"
#include "sfb_direct_f64.hpp"
#include "sfb_direct_mixed.hpp"
#include "sfb_direct_mixed2.hpp"
#include "sfb_direct_split16.hpp"
#include "sfb_lowl.hpp"
#include "sfb_healpix.hpp"
#include "sfb_healpix_ring.hpp"
#include "sfb_integrals.hpp"
#include "sfb_ingest.hpp"
#include "sfb_math.hpp"
#include "sfb_sort.hpp"
#include "sfb_support.hpp"

using namespace sfb;

namespace {

thread_local std::string g_last_error;

sfbhip_status fail(sfbhip_status st, const std::string &msg) {
  g_last_error = msg;
  return st;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(SFBHIP_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

void default_tiles(sfbhip_tile_table *t) {
  for (int b = 0; b < SFBHIP_TILE_BANDS; ++b) {
    t->chunks[b] = 4;
    t->point_split[b] = 4;
  }
  t->items_per_cu = 64;
}

// band of the tile table: [0,16) [16,32) [32,64) [64,128) [128,256) [256,512)
int tile_band(int l) {
  int b = 0;
  while (b < SFBHIP_TILE_BANDS - 1 && l >= (16 << b)) ++b;
  return b;
}

}  // namespace

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct KlnCache {
  bool valid = false;
  int lmax = -1, nmax = -1, mode = -1;
  double rmax = 0, kmax = 0;
  std::vector<double> zeros;  // [lmax][nmax+1]
  std::vector<double> kln;    // [lmax][nmax]
  std::vector<int32_t> nl;
  bool table_valid = false;   // reference interpolation table on the device
  bool herm_valid = false;    // Hermite tables of the mixed path on the device
  std::vector<int32_t> herm_nint;  // their interval counts per l (host copy)
  bool q5_valid = false;      // quintic FP64 tables of the low multipoles on the device
  int q5_lmax = 0;
  double q5_R = 0.0;
  double herm_R = 0.0;        // radius the tables cover: k_{l,n_l-1} * herm_R is the last argument
  bool cub_valid = false;     // cubic-coefficient Bessel tables of the 512-thread mixed kernel
  double cub_R = 0.0;
  std::vector<int32_t> cub_nint;
};

namespace {
struct RingPlanCache;
struct PointPlanCache;
}

struct sfbhip_handle {
  int device = 0;
  std::map<std::string, DevBuf> bufs;
  int *h_err = nullptr;  // pinned: [0] error bits, [2..3] largest radius of the last call (double)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // phase boundaries of the last call
  // the classes of the mixed and FP64 kernels run on the call's stream and these two (default
  // priority), forked and joined by events, so each launch's tail overlaps
  // another class's work
  hipStream_t mx_side[2] = {nullptr, nullptr};
  hipEvent_t mx_fork = nullptr, mx_join[2] = {nullptr, nullptr};
  float phase_ms[3] = {0.f, 0.f, 0.f};                      // FP64 contraction, mixed contraction, finalize
  float last_ms = 0.f;
  KlnCache kc;
  int leg_lmax = -1;
  int legs_lmax = -1;  // scaled Legendre constants of the 512-thread mixed kernel on the device
  int lds_max = 65536;  // largest LDS allocation of one block (bytes)
  int pix_nside = -1;  // pixel tables on the device are for this nside
  int ring_nside = -1;  // ring tables (cos/sin theta per ring, ring of each pixel) are for this nside
  std::shared_ptr<RingPlanCache> ring_plan;  // the last single-batch ring plan and its shape
  std::shared_ptr<PointPlanCache> point_plan;  // the last plan of sfbhip_direct and its shape
  int rleg_nside = -1, rleg_lmax = -1;  // ring Legendre table is for this (nside, lmax)
  int ig_L = -1, ig_nz = -1;  // zero rows of the sbessel integrals on the device
  int ncu = 256;  // compute units of the device
  sfbhip_tile_table tiles;  // tile table of batch plans
};

namespace {

template <class T>
sfbhip_status get_buf(sfbhip_handle *h, const std::string &name, size_t count, T **out) {
  DevBuf &b = h->bufs[name];
  const size_t need = std::max<size_t>(count * sizeof(T), 256);
  if (b.bytes < need) {
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.p, need);
    if (e != hipSuccess) {
      b.p = nullptr;
      return fail(SFBHIP_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(need) + " bytes) for " + name + ": " +
                                            hipGetErrorString(e));
    }
    b.bytes = need;
  }
  *out = static_cast<T *>(b.p);
  return SFBHIP_OK;
}

#define CHK(expr)                           \
  do {                                      \
    sfbhip_status s_ = (expr);              \
    if (s_ != SFBHIP_OK) return s_;         \
  } while (0)

sfbhip_status validate(const sfbhip_params *p) {
  if (!p) return fail(SFBHIP_INVALID_ARGUMENT, "params is NULL");
  if (p->lmax < 0 || p->lmax > 4096) return fail(SFBHIP_INVALID_ARGUMENT, "lmax must be in [0, 4096]");
  if (p->nmax < 1 || p->nmax > 65535) return fail(SFBHIP_INVALID_ARGUMENT, "nmax must be in [1, 65535]");
  if (!(p->rmax > 0.0) || !std::isfinite(p->rmax)) return fail(SFBHIP_INVALID_ARGUMENT, "rmax must be positive");
  if (p->precision != SFBHIP_PREC_FP64 && p->precision != SFBHIP_PREC_MIXED)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown precision mode");
  if (p->bessel < 0 || p->bessel > 2) return fail(SFBHIP_INVALID_ARGUMENT, "unknown bessel mode");
  if (p->zeros != SFBHIP_ZEROS_REFERENCE && p->zeros != SFBHIP_ZEROS_CONVERGED)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown zeros mode");
  if (p->mixed_chunks != 0 && p->mixed_chunks != 1 && p->mixed_chunks != 2 && p->mixed_chunks != 4 &&
      p->mixed_chunks != 8)
    return fail(SFBHIP_INVALID_ARGUMENT, "mixed_chunks must be 0, 1, 2, 4 or 8");
  if (p->mixed_kernel < SFBHIP_MIXED_KERNEL_AUTO || p->mixed_kernel > SFBHIP_MIXED_KERNEL_GROUPED)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown mixed kernel");
  if (p->mixed_operands != SFBHIP_OPERANDS_FP32 && p->mixed_operands != SFBHIP_OPERANDS_SPLIT_FP16 &&
      p->mixed_operands != SFBHIP_OPERANDS_SPLIT3_FP16)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown mixed operand format");
  return SFBHIP_OK;
}

// Bessel zeros on the device (one thread per l), then k_ln and n_l on the host.
sfbhip_status ensure_kln(sfbhip_handle *h, const sfbhip_params *p, hipStream_t st) {
  KlnCache &c = h->kc;
  const double kmax = p->k_max > 0 ? p->k_max : 0.0;
  if (c.valid && c.lmax == p->lmax && c.nmax == p->nmax && c.mode == p->zeros && c.rmax == p->rmax &&
      c.kmax == kmax)
    return SFBHIP_OK;
  c.valid = false;
  c.table_valid = false;
  c.herm_valid = false;
  c.q5_valid = false;
  c.cub_valid = false;
  const int L = p->lmax, nz = p->nmax + 1;
  c.zeros.assign((size_t)std::max(L, 0) * nz, 0.0);
  c.kln.assign((size_t)std::max(L, 0) * p->nmax, 0.0);
  c.nl.assign(std::max(L, 0), 0);
  if (L > 0) {
    double *dz;
    CHK(get_buf(h, "zeros", (size_t)L * nz, &dz));
    bessel_zeros_kernel<<<cdiv(L, 64), 64, 0, st>>>(L, nz, p->zeros, dz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c.zeros.data(), dz, sizeof(double) * L * nz, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int l = 0; l < L; ++l) {
      int cnt = 0;
      for (int n = 0; n < p->nmax; ++n) {
        const double z = c.zeros[(size_t)l * nz + n];
        c.kln[(size_t)l * p->nmax + n] = z / p->rmax;
        if (kmax <= 0.0 || z < kmax * p->rmax) cnt = n + 1;
      }
      // zeros increase with n, so the k_max cut keeps a prefix
      c.nl[l] = cnt;
    }
    double *dk;
    int32_t *dnl;
    CHK(get_buf(h, "kln", (size_t)L * p->nmax, &dk));
    CHK(get_buf(h, "nl", (size_t)L, &dnl));
    HIPCHK(hipMemcpyAsync(dk, c.kln.data(), sizeof(double) * L * p->nmax, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dnl, c.nl.data(), sizeof(int32_t) * L, hipMemcpyHostToDevice, st));
  }
  c.lmax = L;
  c.nmax = p->nmax;
  c.mode = p->zeros;
  c.rmax = p->rmax;
  c.kmax = kmax;
  c.valid = true;
  return SFBHIP_OK;
}

// Reference interpolation table: P = 1000*nmax nodes per l on [0, z_{l,nmax}].
sfbhip_status ensure_table(sfbhip_handle *h, const sfbhip_params *p, hipStream_t st, BesselTableRef *t) {
  KlnCache &c = h->kc;
  const int L = p->lmax, P = 1000 * p->nmax, nz = p->nmax + 1;
  double *y, *zm, *dxi;
  CHK(get_buf(h, "tab_y", (size_t)std::max(L, 1) * P, &y));
  CHK(get_buf(h, "tab_zmax", (size_t)std::max(L, 1), &zm));
  CHK(get_buf(h, "tab_dxinv", (size_t)std::max(L, 1), &dxi));
  if (!c.table_valid && L > 0) {
    std::vector<double> hz(L), hd(L);
    for (int l = 0; l < L; ++l) {
      hz[l] = c.zeros[(size_t)l * nz + p->nmax];
      hd[l] = 1.0 / (hz[l] / (P - 1));
    }
    HIPCHK(hipMemcpyAsync(zm, hz.data(), sizeof(double) * L, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dxi, hd.data(), sizeof(double) * L, hipMemcpyHostToDevice, st));
    const int64_t tot = (int64_t)L * P;
    bessel_table_kernel<<<cdiv(tot, 256), 256, 0, st>>>(L, P, zm, y);
    HIPCHK(hipGetLastError());
    c.table_valid = true;
  }
  t->y = y;
  t->zmax = zm;
  t->dx_inv = dxi;
  t->P = P;
  return SFBHIP_OK;
}

// Host arrays of the Legendre constants for l < lmax (see LegendreTables):
// hs, ho of max(lmax, 1) entries, hc of lmax (lmax + 1) / 2 + 1.
void legendre_host(int lmax, std::vector<double2> &hs, std::vector<double2> &hc, std::vector<int32_t> &ho) {
  hs.assign(std::max(lmax, 1), make_double2(0.0, 0.0));
  hc.assign((size_t)lmax * (lmax + 1) / 2 + 1, make_double2(0.0, 0.0));
  ho.assign(std::max(lmax, 1), 0);
  double prod = 1.0;
  int32_t o = 0;
  for (int l = 0; l < lmax; ++l) {
    if (l > 0) prod *= (2.0 * l - 1.0) / (2.0 * l);
    double c = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi) * prod);
    if (l & 1) c = -c;
    int e;
    const double mnt = std::frexp(c, &e);
    hs[l] = make_double2(mnt, (double)e);
    ho[l] = o;
    hc[o] = make_double2(0.0, 0.0);
    for (int m = 1; m <= l; ++m) {
      const double den = std::sqrt((double)(l + m) * (double)(l - m + 1));
      hc[o + m] = make_double2(2.0 * m / den, std::sqrt((double)(l - m) * (double)(l + m + 1)) / den);
    }
    o += l + 1;
  }
}

// Legendre constants for l < lmax (see LegendreTables).
sfbhip_status ensure_legendre(sfbhip_handle *h, int lmax, hipStream_t st, LegendreTables *lt) {
  double2 *sect, *coef;
  int32_t *off;
  const size_t ncoef = (size_t)lmax * (lmax + 1) / 2 + 1;
  CHK(get_buf(h, "leg_sect", (size_t)std::max(lmax, 1), &sect));
  CHK(get_buf(h, "leg_coef", ncoef, &coef));
  CHK(get_buf(h, "leg_off", (size_t)std::max(lmax, 1), &off));
  if (h->leg_lmax < lmax) {
    std::vector<double2> hs, hc;
    std::vector<int32_t> ho;
    legendre_host(lmax, hs, hc, ho);
    HIPCHK(hipMemcpyAsync(sect, hs.data(), sizeof(double2) * hs.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(coef, hc.data(), sizeof(double2) * hc.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(off, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // host vectors go out of scope
    h->leg_lmax = lmax;
  }
  lt->sect = sect;
  lt->coef = coef;
  lt->coef_off = off;
  return SFBHIP_OK;
}

double norm_factor(const sfbhip_params *p, double npts) {
  if (p->window_volume > 0.0) return std::sqrt(2. / kPi) * p->window_volume / npts;
  double nf = std::sqrt(2. / kPi) * (4. / 3.) * kPi;  // sfb_raw.cpp:29-30
  nf *= p->rmax * p->rmax * p->rmax / npts;
  return nf;
}

std::vector<int64_t> flmn_offsets(int lmax, int nmax) {
  std::vector<int64_t> off(std::max(lmax, 1));
  int64_t o = 0;
  for (int l = 0; l < lmax; ++l) {
    off[l] = o;
    o += (int64_t)(l + 1) * nmax;
  }
  return off;
}

sfbhip_status zero_err(sfbhip_handle *h, hipStream_t st, int **derr) {
  CHK(get_buf(h, "err", 4, derr));
  fill_int_kernel<<<1, 64, 0, st>>>(*derr, 4, 0);
  HIPCHK(hipGetLastError());
  return SFBHIP_OK;
}

sfbhip_status read_err(sfbhip_handle *h, hipStream_t st, int *derr, int *out) {
  HIPCHK(hipMemcpyAsync(h->h_err, derr, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *out = *h->h_err;
  return SFBHIP_OK;
}

int qclass(int q) {
  int c = 1;
  while (c < q) c <<= 1;
  return c;
}

// The m-chunk ranges [qlo, qhi) of one l's items, at most cap of its Q chunks
// each.  The mixed-precision kernels number chunks from the top (chunk q holds
// m = l - 16 q down) and step each point's Legendre recurrence from m = l to
// the item's first chunk, 16 qlo steps per point of every tile: there the
// short remainder range goes first (qlo = 0, no steps) and the full ranges
// follow, which saves the remainder's length in steps for each of them.  The
// FP64 kernel numbers chunks from m = 0 up and steps from l down to 16 qhi, so
// there the remainder goes last (from_top false: the highest chunks, no steps).
std::vector<std::pair<int, int>> chunk_ranges(int Q, int cap, bool from_top) {
  std::vector<std::pair<int, int>> v;
  const int r = from_top ? Q % cap : 0;
  if (r) v.emplace_back(0, r);
  for (int qlo = r; qlo < Q; qlo += cap) v.emplace_back(qlo, std::min(Q, qlo + cap));
  return v;
}

// Items of a row-tiled direct kernel: groups (l, n-tile), split over point
// ranges, with m chunks split into passes of at most qmax_cap chunks.
struct DirectPlan {
  // by QMAX class 1,2,4,8,16,32 (index ci), and for the mixed kernel by point
  // split KS = 1, 2, 4 (index ci + 6 * ksi)
  std::vector<DirectItem> items[18];
  std::vector<ReduceGroup> groups;
  int64_t partial_size = 0;
  int64_t chunk = 0;  // length of the point ranges (uniform splits; the last may be shorter)
  std::vector<int64_t> starts;  // first point of each range, ascending
};

// The last single-batch plan of the ring path: its work items are on the device
// (hr_lowl_items, hr_items, hr_mitems, hr_moff), so a call of the same shape
// plans and uploads nothing.  The shape: nside, lmax, nmax, the FP64 split,
// the equatorial path and n_l.
struct RingPlanCache {
  int32_t nside = 0, L = 0, nmax = 0, lsplit = 0;
  bool eq = false;
  std::vector<int32_t> nl;
  DirectPlan pl64;                   // the FP64 low-l classes (items on the device)
  std::vector<ReduceGroup> groups;   // every group of the call (norm set per call)
  int64_t partial_size = 0;
  std::array<int, 6> counts{};       // ring item classes on the device
  int64_t nmap = 0;                  // map items on the device
};

// The last plan of sfbhip_direct: its work items are on the device
// (px_items), so a call of the same shape plans and uploads nothing.  The
// shape: the point count, the k_ln parameters (lmax, nmax, rmax, k_max, zeros),
// the precision, the FP64 split, the kernel and chunk options, and the radius
// the Bessel tables cover (each item's table reach depends on it).
struct PointPlanCache {
  int64_t n = 0;  // 0: no plan kept
  int lmax = 0, nmax = 0, zeros = 0, lsplit = 0, llow = 0, kernel = 0, chunks = 0, operands = 0;
  bool mixed = false;
  double rmax = 0.0, kmax = 0.0, R = 0.0;
  DirectPlan pllw, pl64, plmx;
  std::vector<ReduceGroup> groups;  // every group of the call (norm set per call)
};

// Where one catalogue's items and groups point: its k_ln rows (k_{l,n} at
// kbase + l nmax + n), its f_lmn block (fbase, in complex entries), its first
// point in the point table (pbase), its normalisation and index.
struct PlanCtx {
  int nmax = 1;
  double norm = 1.0;
  int64_t kbase = 0, fbase = 0, pbase = 0;
  int cat = 0;
  const double *kln = nullptr;  // this catalogue's k_ln rows (host), for the table reach
  double R = 0.0;               // max(rmax, largest radius)
};

// ks_cap > 1 (mixed kernel, 64-row items): an item of at most 16 rows goes to
// point split KS = min(4, ks_cap), of at most 32 rows to KS = min(2, ks_cap).
// cuts (optional): segment boundaries 0 = c_0 < c_1 < ... = npts, multiples of
// tp; no point range of an item then crosses one (each segment gets its share
// of the ranges).  from_top: the kernel numbers m chunks from the top, so the
// remainder range of chunks comes first (chunk_ranges).
void plan_direct(const std::vector<int32_t> &nl, int lmin, int lmax, int64_t npts, int nt, int tp, int mc,
                 int qmax_cap, int64_t target_items, const PlanCtx &cx, DirectPlan *pl, int ks_cap = 1,
                 const std::vector<int64_t> *cuts = nullptr, bool from_top = false) {
  int64_t ngroups = 0, passes_total = 0;
  for (int l = lmin; l < lmax; ++l) {
    const int64_t g = cdiv(nl[l], nt);
    ngroups += g;
    passes_total += g * cdiv(l / mc + 1, qmax_cap);
  }
  if (ngroups == 0) return;
  int64_t S0 = std::max<int64_t>(1, cdiv(target_items, passes_total));
  S0 = std::min<int64_t>(S0, cdiv(npts, tp));
  // point ranges [r0[s], r1[s]) of the splits
  std::vector<int64_t> r0, r1;
  const std::vector<int64_t> whole{0, npts};
  const std::vector<int64_t> &cv = cuts ? *cuts : whole;
  for (size_t k = 0; k + 1 < cv.size(); ++k) {
    const int64_t a = cv[k], len = cv[k + 1] - cv[k];
    if (len <= 0) continue;
    const int64_t Sk = std::max<int64_t>(1, (S0 * len + npts / 2) / std::max<int64_t>(npts, 1));
    const int64_t chunk = cdiv(cdiv(len, Sk), tp) * tp;
    for (int64_t q = 0; q < len; q += chunk) {
      r0.push_back(a + q);
      r1.push_back(a + std::min(len, q + chunk));
    }
  }
  const int64_t S = (int64_t)r0.size();
  pl->chunk = cuts ? 0 : (S > 0 ? r1[0] - r0[0] : 0);
  pl->starts.resize(S);
  for (int64_t k = 0; k < S; ++k) pl->starts[k] = cx.pbase + r0[k];
  for (int l = lmax - 1; l >= lmin; --l) {
    const int Q = l / mc + 1;
    const int ncol = 2 * (l + 1);
    for (int n0 = 0; n0 < nl[l]; n0 += nt) {
      ReduceGroup g;
      g.l = l;
      g.n0 = n0;
      g.nrows = std::min(nt, nl[l] - n0);
      g.nsplit = (int32_t)S;
      g.base = pl->partial_size;
      g.stride = (int64_t)nt * ncol;
      g.fofs = cx.fbase + (int64_t)l * (l + 1) / 2 * cx.nmax;
      g.kofs = cx.kbase + (int64_t)l * cx.nmax + n0;
      g.norm = cx.norm;
      g.nmax = cx.nmax;
      g.pad = 0;
      pl->partial_size += S * g.stride;
      pl->groups.push_back(g);
      for (int64_t s = 0; s < S; ++s) {
        for (const auto &qr : chunk_ranges(Q, qmax_cap, from_top)) {
          DirectItem it;
          it.l = l;
          it.n0 = n0;
          it.nrows = g.nrows;
          it.qlo = qr.first;
          it.qhi = qr.second;
          it.cat = cx.cat;
          it.tab = cx.kln ? (int32_t)std::min<double>(
                                std::floor(cx.kln[(size_t)l * cx.nmax + n0 + g.nrows - 1] * cx.R / MX_H) + 2, 1e9)
                          : 0;
          it.flags = 0;
          it.p0 = cx.pbase + r0[s];
          it.p1 = cx.pbase + r1[s];
          it.out = g.base + s * g.stride;
          it.kofs = g.kofs;
          const int qc = qclass(it.qhi - it.qlo);
          int ci = 0;
          while ((1 << ci) < qc) ++ci;
          int ksi = 0;
          if (ks_cap >= 4 && it.nrows <= 16)
            ksi = 2;
          else if (ks_cap >= 2 && it.nrows <= 32)
            ksi = 1;
          pl->items[ci + 6 * ksi].push_back(it);
        }
      }
    }
  }
}

template <int QMAX>
void launch_f64(int nitems, const DirectItem *items, hipStream_t st, const PointsSoA &pts, const PixelsSoA *pix,
                const double *kln, int nmax, const LegendreTables &lt, int use_table, const BesselTableRef &tab,
                double *partial, int *err) {
  if (pix)
    sfb_direct_f64_kernel<QMAX, true><<<nitems, 256, 0, st>>>(pts, *pix, items, kln, nmax, lt, use_table, tab,
                                                                partial, err);
  else
    sfb_direct_f64_kernel<QMAX, false><<<nitems, 256, 0, st>>>(pts, PixelsSoA{}, items, kln, nmax, lt, use_table,
                                                                 tab, partial, err);
}

// The point table; also reads back the error bits and the catalogue's largest
// radius into *rdata (one stream synchronisation).  A catalogue with a point
// at r < 0 or a non-finite coordinate is refused here, before any kernel
// evaluates a Bessel function of it (the reference's interpolator rejects
// arguments outside its table, FunctionInterpolator.cpp:22-24).
sfbhip_status prep_points(sfbhip_handle *h, const double *d_points, int64_t n, hipStream_t st, int *err,
                          PointsSoA *pts, double *rdata) {
  double *buf;
  CHK(get_buf(h, "soa", (size_t)6 * n, &buf));
  double *r = buf, *x = buf + n, *s = buf + 2 * n, *ph = buf + 3 * n, *cp = buf + 4 * n, *sp = buf + 5 * n;
  unsigned long long *rbits = reinterpret_cast<unsigned long long *>(err + 2);
  prep_points_kernel<<<cdiv(n, 256), 256, 0, st>>>(d_points, n, r, x, s, ph, cp, sp, err, rbits);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->h_err, err, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(rdata, h->h_err + 2, sizeof(double));
  HIPCHK(hipGetLastError());
  if (h->h_err[0] & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  pts->r = r;
  pts->x = x;
  pts->s = s;
  pts->phi = ph;
  pts->cphi = cp;
  pts->sphi = sp;
  pts->n = n;
  return SFBHIP_OK;
}

// The point table reordered by radius (stable), so that consecutive points
// have close Bessel arguments k r: the table loads of a tile then stay within
// a few cache lines.  The sum over points does not depend on the order beyond
// rounding.  The keys are r rounded to FP32 (four 8-bit radix passes instead of
// the eight of its FP64 pattern).
sfbhip_status sort_points_by_radius(sfbhip_handle *h, const PointsSoA &in, hipStream_t st, PointsSoA *out) {
  const int64_t n = in.n;
  if (n < 2 || n > INT32_MAX) {
    *out = in;
    return SFBHIP_OK;
  }
  unsigned *k0, *k1;
  int32_t *i0, *i1;
  double *buf;
  CHK(get_buf(h, "sort_k32a", (size_t)n, &k0));
  CHK(get_buf(h, "sort_k32b", (size_t)n, &k1));
  CHK(get_buf(h, "sort_i0", (size_t)n, &i0));
  CHK(get_buf(h, "sort_i1", (size_t)n, &i1));
  CHK(get_buf(h, "soa_sorted", (size_t)6 * n, &buf));
  radius_keys32_kernel<<<cdiv(n, 256), 256, 0, st>>>(in.r, n, k0, i0);
  HIPCHK(hipGetLastError());
  size_t tmp_bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, i0, i1, (int)n, 0, 32, st));
  char *tmp;
  CHK(get_buf(h, "sort_tmp", tmp_bytes, &tmp));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i0, i1, (int)n, 0, 32, st));
  double *r = buf, *x = buf + n, *s = buf + 2 * n, *ph = buf + 3 * n, *cp = buf + 4 * n, *sp = buf + 5 * n;
  gather_points_kernel<<<cdiv(n, 256), 256, 0, st>>>(in, i1, n, r, x, s, ph, cp, sp);
  HIPCHK(hipGetLastError());
  *out = PointsSoA{r, x, s, ph, cp, sp, n};
  return SFBHIP_OK;
}

// f_lmn buffer (the caller's or the workspace), zeroed; reduce; c_ln; k_ln copy.
sfbhip_status finalize(sfbhip_handle *h, const sfbhip_params *p, const std::vector<ReduceGroup> &groups,
                       const double *partial, double norm, double *d_k_ln, double *d_c_ln, double *d_f_lmn,
                       hipStream_t st) {
  const int L = p->lmax, nmax = p->nmax;
  const size_t nf = sfbhip_flmn_count(L, nmax);
  double *f = d_f_lmn;
  if (!f) CHK(get_buf(h, "flmn", 2 * nf, &f));
  if (nf) {
    fill_kernel<<<cdiv(2 * nf, 256), 256, 0, st>>>(f, (int64_t)(2 * nf), 0.0);
    HIPCHK(hipGetLastError());
  }
  std::vector<int64_t> hoff = flmn_offsets(L, nmax);
  int64_t *doff;
  ReduceGroup *dg;
  CHK(get_buf(h, "flmn_off", hoff.size(), &doff));
  CHK(get_buf(h, "groups", std::max<size_t>(groups.size(), 1), &dg));
  HIPCHK(hipMemcpyAsync(doff, hoff.data(), sizeof(int64_t) * hoff.size(), hipMemcpyHostToDevice, st));
  if (!groups.empty()) {
    HIPCHK(hipMemcpyAsync(dg, groups.data(), sizeof(ReduceGroup) * groups.size(), hipMemcpyHostToDevice, st));
    double *dk;
    CHK(get_buf(h, "kln", (size_t)L * nmax, &dk));
    int chunks = 1;
    for (const ReduceGroup &g : groups) chunks = std::max<int>(chunks, (int)cdiv((int64_t)g.nrows * 2 * (g.l + 1), 256));
    dim3 grid((unsigned)groups.size(), (unsigned)std::min(chunks, 64));
    reduce_partials_kernel<<<grid, 256, 0, st>>>(dg, partial, dk, f);
    HIPCHK(hipGetLastError());
  }
  if (L > 0) {
    int32_t *dnl;
    double *dk;
    CHK(get_buf(h, "nl", (size_t)L, &dnl));
    CHK(get_buf(h, "kln", (size_t)L * nmax, &dk));
    cln_kernel<<<cdiv((int64_t)L * nmax, 256), 256, 0, st>>>(L, nmax, dnl, doff, f, d_c_ln);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(d_k_ln, dk, sizeof(double) * L * nmax, hipMemcpyDeviceToDevice, st));
  }
  return SFBHIP_OK;
}

// Quintic FP64 j_l tables for l < L (the mixed path's FP64 multipoles):
// intervals of Q5_H covering [0, k_{l,n_l-1} R], R = max(rmax, largest radius).
sfbhip_status ensure_quintic(sfbhip_handle *h, const sfbhip_params *p, int L, double rdata, hipStream_t st,
                             BesselTableRef *t) {
  KlnCache &c = h->kc;
  const double R = std::max(p->rmax, rdata);
  int32_t *doff, *dni;
  double *tab;
  if (!c.q5_valid || c.q5_R < R || c.q5_lmax < L) {
    std::vector<int32_t> off(std::max(L, 1), 0), ni(std::max(L, 1), 1);
    int32_t o = 0, mx = 1;
    for (int l = 0; l < L; ++l) {
      int32_t n = 1;
      if (c.nl[l] > 0) n = (int32_t)std::floor(c.kln[(size_t)l * p->nmax + c.nl[l] - 1] * R * Q5_HINV) + 2;
      off[l] = o;
      ni[l] = n;
      o += n;
      mx = std::max(mx, n);
    }
    CHK(get_buf(h, "q5_tab", (size_t)6 * std::max(o, 1), &tab));
    CHK(get_buf(h, "q5_off", off.size(), &doff));
    CHK(get_buf(h, "q5_nint", ni.size(), &dni));
    HIPCHK(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dni, ni.data(), sizeof(int32_t) * ni.size(), hipMemcpyHostToDevice, st));
    if (L > 0) {
      dim3 grid((unsigned)cdiv(mx, 256), (unsigned)L);
      quintic_table_kernel<<<grid, 256, 0, st>>>(L, doff, dni, tab);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    c.q5_valid = true;
    c.q5_R = R;
    c.q5_lmax = L;
  }
  CHK(get_buf(h, "q5_tab", 1, &tab));
  CHK(get_buf(h, "q5_off", 1, &doff));
  CHK(get_buf(h, "q5_nint", 1, &dni));
  t->q5 = tab;
  t->q5_off = doff;
  t->q5_nint = dni;
  return SFBHIP_OK;
}

// Hermite tables of the mixed path: intervals [i h, (i+1) h], i < nint_l =
// floor(k_{l,n_l-1} R / h) + 2, R = max(rmax, largest radius of the catalogue),
// so that no argument leaves the table.
sfbhip_status ensure_hermite(sfbhip_handle *h, const sfbhip_params *p, double rdata, hipStream_t st,
                             HermiteTables *ht) {
  KlnCache &c = h->kc;
  const int L = p->lmax;
  const double R = std::max(p->rmax, rdata);
  float4 *tab;
  int32_t *doff, *dni;
  if (!c.herm_valid || c.herm_R < R) {
    std::vector<int32_t> off(std::max(L, 1), 0), ni(std::max(L, 1), 1);
    int64_t o = 0;
    int32_t mx = 1;
    for (int l = 0; l < L; ++l) {
      int64_t n = 1;
      if (c.nl[l] > 0) n = (int64_t)std::floor(c.kln[(size_t)l * p->nmax + c.nl[l] - 1] * R / MX_H) + 2;
      if (n > MX_TAB_MAX)
        return fail(SFBHIP_NOT_SUPPORTED, "mixed path: largest Bessel argument k r exceeds " +
                                              std::to_string(MX_TAB_MAX * MX_H) + "; use precision FP64");
      off[l] = (int32_t)o;
      ni[l] = (int32_t)n;
      o += n;
      mx = std::max<int32_t>(mx, (int32_t)n);
      if (o > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "mixed path: Bessel tables too large");
    }
    CHK(get_buf(h, "herm_tab", (size_t)std::max<int64_t>(o, 1), &tab));
    CHK(get_buf(h, "herm_off", off.size(), &doff));
    CHK(get_buf(h, "herm_nint", ni.size(), &dni));
    HIPCHK(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dni, ni.data(), sizeof(int32_t) * ni.size(), hipMemcpyHostToDevice, st));
    if (L > 0) {
      dim3 grid((unsigned)cdiv(mx, 256), (unsigned)L);
      hermite_table_kernel<<<grid, 256, 0, st>>>(L, doff, dni, tab);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    c.herm_R = R;
    c.herm_nint = ni;
    c.herm_valid = true;
  }
  CHK(get_buf(h, "herm_tab", 1, &tab));
  CHK(get_buf(h, "herm_off", 1, &doff));
  CHK(get_buf(h, "herm_nint", 1, &dni));
  ht->tab = tab;
  ht->off = doff;
  ht->nint = dni;
  return SFBHIP_OK;
}

// Bessel tables of up to MX_LDS_NODES nodes (36 KB) are staged in LDS: the
// block then still fits twice per CU.
constexpr int MX_LDS_NODES = 4608;

template <int QMAX, int NP, int KS = 1>
sfbhip_status launch_mixed(int nitems, const DirectItem *items, hipStream_t st, const PointsSoA &pts,
                           const double *kln, int nmax, const LegendreTables &lt, const HermiteTables &ht,
                           int lds_nodes, double *partial) {
  const size_t shm = sizeof(float2) * (size_t)lds_nodes;
  if (shm > 0)
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&sfb_direct_mixed_kernel<QMAX, NP, KS>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  sfb_direct_mixed_kernel<QMAX, NP, KS><<<nitems, 256, shm, st>>>(pts, items, kln, nmax, lt, ht, lds_nodes,
                                                                 partial);
  HIPCHK(hipGetLastError());
  return SFBHIP_OK;
}

// Work items of a mixed call: many, so that items of unequal cost (l, row
// groups, m passes) even out over the 512 resident blocks and each class
// launch ends with a short tail; one per 256 points for large catalogues.
// Measured optimum: 16384 at 1M points (4096: +0.7 ms of 7.7), 65536 or more
// at 2^25 points; the partial blocks and their reduction grow with the count.
int64_t mixed_target_items(int64_t n_points) {
  return std::min<int64_t>(131072, std::max<int64_t>(16384, n_points / 256));
}

// The same for the FP64 kernel: its items cost in proportion to l + 1 (all m
// chunks in one item), so many short items keep the l = lmax - 1 items from
// forming the tail.  Measured: 1M points 43.6 ms (4096 items) -> 37.3 ms
// (16384); HEALPix nside 256 (786,432 pixels) 196 -> 126 (16384) -> 107 ms
// (65536 and more).
int64_t f64_target_items(int64_t n_points, bool pixels) {
  return std::min<int64_t>(131072, std::max<int64_t>(16384, pixels ? n_points / 8 : n_points / 64));
}

// the point split applies to the split-FP16 operands only
int mixed_ks_cap(int ops) { return ops == SFBHIP_OPERANDS_SPLIT_FP16 ? 4 : 1; }

// operand format SFBHIP_OPERANDS_* -> number of FP16 pieces (0: FP32 operands);
// ksi: point split KS = 1 << ksi (split-FP16 only)
template <int QMAX>
sfbhip_status launch_mixed_ops(int ops, int ksi, int nitems, const DirectItem *items, hipStream_t st,
                               const PointsSoA &pts, const double *kln, int nmax, const LegendreTables &lt,
                               const HermiteTables &ht, int lds_nodes, double *partial) {
  if (ops == SFBHIP_OPERANDS_SPLIT_FP16) {
    if (ksi == 2) return launch_mixed<QMAX, 2, 4>(nitems, items, st, pts, kln, nmax, lt, ht, lds_nodes, partial);
    if (ksi == 1) return launch_mixed<QMAX, 2, 2>(nitems, items, st, pts, kln, nmax, lt, ht, lds_nodes, partial);
    return launch_mixed<QMAX, 2>(nitems, items, st, pts, kln, nmax, lt, ht, lds_nodes, partial);
  }
  if (ksi != 0) return fail(SFBHIP_NOT_SUPPORTED, "point split planned for an operand format without it");
  if (ops == SFBHIP_OPERANDS_SPLIT3_FP16)
    return launch_mixed<QMAX, 3>(nitems, items, st, pts, kln, nmax, lt, ht, lds_nodes, partial);
  return launch_mixed<QMAX, 0>(nitems, items, st, pts, kln, nmax, lt, ht, lds_nodes, partial);
}

// Launch every class of a plan; items[] on the device holds the classes back to back from `*off`.
// The FP64 classes dealt in turn to st and the handle's two side streams
// (forked and joined by launch_plan_f64), or to the side streams only.
sfbhip_status launch_f64_classes(sfbhip_handle *h, const DirectPlan &pl, DirectItem *ditems, size_t *off,
                                 hipStream_t st, const PointsSoA &pts, const PixelsSoA *pix, const double *dk,
                                 int nmax, const LegendreTables &lt, int use_table, const BesselTableRef &tab,
                                 double *partial, int *derr, bool *used, bool sides_only = false) {
  const hipStream_t streams[3] = {st, h->mx_side[0], h->mx_side[1]};
  int next = sides_only ? 1 : 0;
  for (int ci = 0; ci < 5; ++ci) {
    const int n = (int)pl.items[ci].size();
    if (!n) continue;
    const DirectItem *di = ditems + *off;
    const int si = next;
    next = sides_only ? 3 - next : (next + 1) % 3;
    const hipStream_t s = streams[si];
    if (si > 0) used[si - 1] = true;
    switch (ci) {
      case 0: launch_f64<1>(n, di, s, pts, pix, dk, nmax, lt, use_table, tab, partial, derr); break;
      case 1: launch_f64<2>(n, di, s, pts, pix, dk, nmax, lt, use_table, tab, partial, derr); break;
      case 2: launch_f64<4>(n, di, s, pts, pix, dk, nmax, lt, use_table, tab, partial, derr); break;
      case 3: launch_f64<8>(n, di, s, pts, pix, dk, nmax, lt, use_table, tab, partial, derr); break;
      default: launch_f64<16>(n, di, s, pts, pix, dk, nmax, lt, use_table, tab, partial, derr); break;
    }
    HIPCHK(hipGetLastError());
    *off += n;
  }
  return SFBHIP_OK;
}

sfbhip_status fork_sides(sfbhip_handle *h, hipStream_t st) {
  HIPCHK(hipEventRecord(h->mx_fork, st));
  for (int k = 0; k < 2; ++k) HIPCHK(hipStreamWaitEvent(h->mx_side[k], h->mx_fork, 0));
  return SFBHIP_OK;
}

// st waits for the side streams that ran work; after a failure they are drained
// instead, so no launched work outlives the call.
sfbhip_status join_sides(sfbhip_handle *h, hipStream_t st, sfbhip_status rc, const bool *used) {
  if (rc != SFBHIP_OK) {
    for (int k = 0; k < 2; ++k) (void)hipStreamSynchronize(h->mx_side[k]);
    return rc;
  }
  for (int k = 0; k < 2; ++k)
    if (used[k]) {
      HIPCHK(hipEventRecord(h->mx_join[k], h->mx_side[k]));
      HIPCHK(hipStreamWaitEvent(st, h->mx_join[k], 0));
    }
  return SFBHIP_OK;
}

// All classes of an FP64 plan, forked from st and joined back to it.
sfbhip_status launch_plan_f64(sfbhip_handle *h, const DirectPlan &pl, DirectItem *ditems, size_t *off,
                              hipStream_t st, const PointsSoA &pts, const PixelsSoA *pix, const double *dk, int nmax,
                              const LegendreTables &lt, int use_table, const BesselTableRef &tab, double *partial,
                              int *derr) {
  CHK(fork_sides(h, st));
  bool used[2] = {false, false};
  return join_sides(h, st,
                    launch_f64_classes(h, pl, ditems, off, st, pts, pix, dk, nmax, lt, use_table, tab, partial, derr,
                                       used),
                    used);
}

// Bessel-table nodes an item needs in LDS
int lds_reach(const DirectItem &it, const std::vector<int32_t> &nint) { return std::min(it.tab, nint[it.l]); }

// One launch per class; its dynamic LDS holds the largest node window among
// the class's items, at most MX_LDS_NODES (an item's window, from one node
// below its first argument to its last, is at most its reach from 0 plus four
// nodes); an item
// whose window does not fit reads global memory.
sfbhip_status launch_mixed_classes(sfbhip_handle *h, const DirectPlan &pl, const std::vector<int32_t> &nint,
                                   DirectItem *ditems, size_t *off, hipStream_t st, const PointsSoA &pts,
                                   const double *dk, int nmax, const LegendreTables &lt, const HermiteTables &ht,
                                   int ops, double *partial, bool *used) {
  // classes in the order the items were uploaded (index ci + 6 ksi ascending),
  // dealt in turn to the call's stream and the two side streams
  const hipStream_t streams[3] = {st, h->mx_side[0], h->mx_side[1]};
  int next = 0;
  for (int idx = 0; idx < 18; ++idx) {
    const int ci = idx % 6, ksi = idx / 6;
    const auto &v = pl.items[idx];
    const int n = (int)v.size();
    if (!n) continue;
    if (ci > 3) return fail(SFBHIP_NOT_SUPPORTED, "mixed plan with more than 8 chunks per item");
    int nodes = 0;
    for (const DirectItem &it : v)
      nodes = std::max(nodes, std::min(lds_reach(it, nint) + 4, MX_LDS_NODES));
    const DirectItem *di = ditems + *off;
    const int si = next;
    next = (next + 1) % 3;
    const hipStream_t s = streams[si];
    if (si > 0) used[si - 1] = true;
    switch (ci) {
      case 0: CHK(launch_mixed_ops<1>(ops, ksi, n, di, s, pts, dk, nmax, lt, ht, nodes, partial)); break;
      case 1: CHK(launch_mixed_ops<2>(ops, ksi, n, di, s, pts, dk, nmax, lt, ht, nodes, partial)); break;
      case 2: CHK(launch_mixed_ops<4>(ops, ksi, n, di, s, pts, dk, nmax, lt, ht, nodes, partial)); break;
      default: CHK(launch_mixed_ops<8>(ops, ksi, n, di, s, pts, dk, nmax, lt, ht, nodes, partial)); break;
    }
    *off += n;
  }
  return SFBHIP_OK;
}

// All classes of the mixed plan, forked from st and joined back to it: st's
// later work (and its phase events) follows every class.  On a failure the
// side streams are drained before returning, so no launched work outlives the call.
sfbhip_status launch_plan_mixed(sfbhip_handle *h, const DirectPlan &pl, const std::vector<int32_t> &nint,
                                DirectItem *ditems, size_t *off, hipStream_t st, const PointsSoA &pts,
                                const double *dk, int nmax, const LegendreTables &lt, const HermiteTables &ht, int ops,
                                double *partial) {
  CHK(fork_sides(h, st));
  bool used[2] = {false, false};
  return join_sides(h, st,
                    launch_mixed_classes(h, pl, nint, ditems, off, st, pts, dk, nmax, lt, ht, ops, partial, used),
                    used);
}

// ---- The 512-thread mixed kernel: its tables, plan and launches.

// Scaled Legendre constants for l < lmax (LegendreScaled): with a_m, b_m of the
// downward recurrence P^{m-1} = -(a_m cot P^m + b_m P^{m+1}), d_l = d_{l-1} = 1,
// d_{m-1} = b_m d_{m+1}, all d divided by their minimum (so |Q| <= |P|), and
// g_m = -a_m d_m / d_{m-1}.
// Host arrays of LegendreScaled for l < lmax: sect per l, g and d at off[l] + m.
void legendre_scaled_host(int lmax, std::vector<double2> &hs, std::vector<double> &hg, std::vector<double> &hd,
                          std::vector<int32_t> &ho) {
  const size_t ncoef = (size_t)lmax * (lmax + 1) / 2 + 1;
  hs.assign(std::max(lmax, 1), make_double2(0.0, 0.0));
  hg.assign(ncoef, 0.0);
  hd.assign(ncoef, 1.0);
  ho.assign(std::max(lmax, 1), 0);
  std::vector<double> a, b, dd;
  double prod = 1.0;
  int32_t o = 0;
  for (int l = 0; l < lmax; ++l) {
    if (l > 0) prod *= (2.0 * l - 1.0) / (2.0 * l);
    double c = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi) * prod);
    if (l & 1) c = -c;
    a.assign(l + 2, 0.0);
    b.assign(l + 2, 0.0);
    dd.assign(l + 2, 1.0);
    for (int m = 1; m <= l; ++m) {
      const double den = std::sqrt((double)(l + m) * (double)(l - m + 1));
      a[m] = 2.0 * m / den;
      b[m] = std::sqrt((double)(l - m) * (double)(l + m + 1)) / den;
    }
    for (int m = l - 1; m >= 1; --m) dd[m - 1] = b[m] * dd[m + 1];
    double dmin = 1.0;
    for (int m = 0; m <= l; ++m) dmin = std::min(dmin, dd[m]);
    for (int m = 0; m <= l; ++m) dd[m] /= dmin;
    ho[l] = o;
    hg[o] = 0.0;
    for (int m = 1; m <= l; ++m) hg[o + m] = -a[m] * dd[m] / dd[m - 1];
    for (int m = 0; m <= l; ++m) hd[o + m] = dd[m];
    int e;
    const double mnt = std::frexp(c / dd[l], &e);
    hs[l] = make_double2(mnt, (double)e);
    o += l + 1;
  }
}

sfbhip_status ensure_legendre_scaled(sfbhip_handle *h, int lmax, hipStream_t st, LegendreScaled *ls) {
  double2 *sect;
  double *g, *d;
  int32_t *off;
  const size_t ncoef = (size_t)lmax * (lmax + 1) / 2 + 1;
  CHK(get_buf(h, "legs_sect", (size_t)std::max(lmax, 1), &sect));
  CHK(get_buf(h, "legs_g", ncoef, &g));
  CHK(get_buf(h, "legs_d", ncoef, &d));
  CHK(get_buf(h, "legs_off", (size_t)std::max(lmax, 1), &off));
  if (h->legs_lmax < lmax) {
    std::vector<double2> hs;
    std::vector<double> hg, hd;
    std::vector<int32_t> ho;
    legendre_scaled_host(lmax, hs, hg, hd, ho);
    HIPCHK(hipMemcpyAsync(sect, hs.data(), sizeof(double2) * hs.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(g, hg.data(), sizeof(double) * hg.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(off, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // host vectors go out of scope
    h->legs_lmax = lmax;
  }
  *ls = LegendreScaled{sect, g, d, off};
  return SFBHIP_OK;
}

// Cubic-coefficient Bessel tables for l < L: nint[l] = floor(xmax[l] / h) + 2
// intervals, so that no argument k r <= xmax[l] leaves the table.  `tag`
// names the workspace buffers.
sfbhip_status build_cubic(sfbhip_handle *h, int L, const std::vector<double> &xmax, const std::string &tag,
                          hipStream_t st, CubicTables *ct, std::vector<int32_t> *nint_out) {
  std::vector<int32_t> off(std::max(L, 1), 0), ni(std::max(L, 1), 1);
  int64_t o = 0;
  int32_t mx = 1;
  for (int l = 0; l < L; ++l) {
    const int64_t n = (int64_t)std::floor(xmax[l] / MX_H) + 2;
    if (n > MX_TAB_MAX)
      return fail(SFBHIP_NOT_SUPPORTED, "mixed path: largest Bessel argument k r exceeds " +
                                            std::to_string(MX_TAB_MAX * MX_H) + "; use precision FP64");
    off[l] = (int32_t)o;
    ni[l] = (int32_t)n;
    o += n;
    mx = std::max<int32_t>(mx, (int32_t)n);
    if (o > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "mixed path: Bessel tables too large");
  }
  float4 *tab, *tab16;
  int32_t *doff, *dni;
  CHK(get_buf(h, tag + "_tab", (size_t)std::max<int64_t>(o, 1), &tab));
  CHK(get_buf(h, tag + "_tab16", (size_t)std::max<int64_t>(o, 1), &tab16));
  CHK(get_buf(h, tag + "_off", off.size(), &doff));
  CHK(get_buf(h, tag + "_nint", ni.size(), &dni));
  HIPCHK(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dni, ni.data(), sizeof(int32_t) * ni.size(), hipMemcpyHostToDevice, st));
  if (L > 0) {
    cubic_table_kernel<<<dim3((unsigned)cdiv(mx, 256), (unsigned)L), 256, 0, st>>>(L, doff, dni, tab, tab16);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));  // host vectors go out of scope
  *ct = CubicTables{tab, doff, dni, tab16};
  if (nint_out) *nint_out = ni;
  return SFBHIP_OK;
}

// The tables of a single call, kept while k_ln and the covered radius stay.
sfbhip_status ensure_cubic(sfbhip_handle *h, const sfbhip_params *p, double rdata, hipStream_t st, CubicTables *ct) {
  KlnCache &c = h->kc;
  const int L = p->lmax;
  const double R = std::max(p->rmax, rdata);
  if (!c.cub_valid || c.cub_R < R) {
    std::vector<double> xmax(std::max(L, 1), 0.0);
    for (int l = 0; l < L; ++l)
      if (c.nl[l] > 0) xmax[l] = c.kln[(size_t)l * p->nmax + c.nl[l] - 1] * R;
    c.cub_valid = false;
    CHK(build_cubic(h, L, xmax, "cub", st, ct, &c.cub_nint));
    c.cub_R = R;
    c.cub_valid = true;
    return SFBHIP_OK;
  }
  float4 *tab, *tab16;
  int32_t *doff, *dni;
  CHK(get_buf(h, "cub_tab", 1, &tab));
  CHK(get_buf(h, "cub_tab16", 1, &tab16));
  CHK(get_buf(h, "cub_off", 1, &doff));
  CHK(get_buf(h, "cub_nint", 1, &dni));
  *ct = CubicTables{tab, doff, dni, tab16};
  return SFBHIP_OK;
}

// Row blocks of a 512-thread item of `rows` rows (<= 128).
int m2_rb(int rows) { return rows > 64 ? 8 : rows > 32 ? 4 : rows > 16 ? 2 : 1; }
int m2_rbi(int rb) { return rb == 8 ? 0 : rb == 4 ? 1 : rb == 2 ? 2 : 3; }

// Items of the 512-thread kernel, by class ci + 3 rbi (QMAX 1, 2, 4; RB 8, 4,
// 2, 1).  Per l the rows are cut into a segment of up to 128 rows while more
// than 64 remain, then one of at most 64; each segment is split over point
// ranges (multiples of 512 points) and into passes of at most qmax_cap chunks.
void plan_mixed2(const std::vector<int32_t> &nl, int lmin, int lmax, int64_t npts, int qmax_cap,
                 int64_t target_items, const PlanCtx &cx, DirectPlan *pl) {
  struct Seg {
    int l, n0, rows;
  };
  std::vector<Seg> segs;
  int64_t passes_total = 0;
  for (int l = lmax - 1; l >= lmin; --l) {
    for (int n0 = 0; n0 < nl[l];) {
      const int rem = nl[l] - n0;
      const int rows = rem > 64 ? std::min(128, rem) : rem;
      segs.push_back(Seg{l, n0, rows});
      passes_total += cdiv(l / MX_MC + 1, qmax_cap);
      n0 += rows;
    }
  }
  if (segs.empty()) return;
  constexpr int64_t TPM = 512;
  int64_t S = std::max<int64_t>(1, cdiv(target_items, passes_total));
  S = std::min<int64_t>(S, cdiv(npts, TPM));
  const int64_t chunk = cdiv(cdiv(npts, S), TPM) * TPM;
  S = cdiv(npts, chunk);
  for (const Seg &sg : segs) {
    const int l = sg.l;
    const int Q = l / MX_MC + 1;
    const int ncol = 2 * (l + 1);
    const int rb = m2_rb(sg.rows);
    ReduceGroup g;
    g.l = l;
    g.n0 = sg.n0;
    g.nrows = sg.rows;
    g.nsplit = (int32_t)S;
    g.base = pl->partial_size;
    g.stride = (int64_t)16 * rb * ncol;
    g.fofs = cx.fbase + (int64_t)l * (l + 1) / 2 * cx.nmax;
    g.kofs = cx.kbase + (int64_t)l * cx.nmax + sg.n0;
    g.norm = cx.norm;
    g.nmax = cx.nmax;
    g.pad = 0;
    pl->partial_size += S * g.stride;
    pl->groups.push_back(g);
    for (int64_t s = 0; s < S; ++s) {
      for (const auto &qr : chunk_ranges(Q, qmax_cap, true)) {
        DirectItem it;
        it.l = l;
        it.n0 = sg.n0;
        it.nrows = sg.rows;
        it.qlo = qr.first;
        it.qhi = qr.second;
        it.cat = cx.cat;
        it.tab = (int32_t)std::min<double>(
            std::floor(cx.kln[(size_t)l * cx.nmax + sg.n0 + sg.rows - 1] * cx.R / MX_H) + 2, 1e9);
        it.flags = 0;
        it.p0 = cx.pbase + s * chunk;
        it.p1 = cx.pbase + std::min(npts, (s + 1) * chunk);
        it.out = g.base + s * g.stride;
        it.kofs = g.kofs;
        const int qc = qclass(it.qhi - it.qlo);
        const int ci = qc == 1 ? 0 : qc == 2 ? 1 : 2;
        pl->items[ci + 3 * m2_rbi(rb)].push_back(it);
      }
    }
  }
}

// Intervals of cubic coefficients a 512-thread block stages in LDS at most.
constexpr int M2_LDS_NODES = 4608;

template <int QMAX, int RB>
sfbhip_status launch_mixed2(sfbhip_handle *h, int nitems, const DirectItem *items, hipStream_t st,
                            const PointsSoA &pts, const double *kln, const LegendreScaled &ls, const CubicTables &ct,
                            int nodes, double *partial) {
  const void *fn = reinterpret_cast<const void *>(&sfb_direct_mixed2_kernel<QMAX, RB>);
  hipFuncAttributes fa;
  HIPCHK(hipFuncGetAttributes(&fa, fn));
  const int cap = std::max(0, std::min(M2_LDS_NODES, (h->lds_max - (int)fa.sharedSizeBytes) / (int)sizeof(float4)));
  nodes = std::min(nodes, cap);
  const size_t shm = sizeof(float4) * (size_t)nodes;
  if (shm > 0) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  sfb_direct_mixed2_kernel<QMAX, RB><<<nitems, M2_NT, shm, st>>>(pts, items, kln, ls, ct, nodes, partial);
  HIPCHK(hipGetLastError());
  return SFBHIP_OK;
}

template <int RB>
sfbhip_status launch_mixed2_q(sfbhip_handle *h, int ci, int nitems, const DirectItem *items, hipStream_t st,
                              const PointsSoA &pts, const double *kln, const LegendreScaled &ls,
                              const CubicTables &ct, int nodes, double *partial) {
  if (ci == 0) return launch_mixed2<1, RB>(h, nitems, items, st, pts, kln, ls, ct, nodes, partial);
  if (ci == 1) return launch_mixed2<2, RB>(h, nitems, items, st, pts, kln, ls, ct, nodes, partial);
  return launch_mixed2<4, RB>(h, nitems, items, st, pts, kln, ls, ct, nodes, partial);
}

// All classes of a 512-thread plan (uploaded in class order), dealt in turn to
// the call's stream and the two side streams, forked from and joined back to st.
sfbhip_status launch_mixed2_classes(sfbhip_handle *h, const DirectPlan &pl, const std::vector<int32_t> &nint,
                                    DirectItem *ditems, size_t *off, hipStream_t st, const PointsSoA &pts,
                                    const double *dk, const LegendreScaled &ls, const CubicTables &ct,
                                    double *partial, bool *used) {
  const hipStream_t streams[3] = {st, h->mx_side[0], h->mx_side[1]};
  int next = 0;
  for (int idx = 0; idx < 12; ++idx) {
    const auto &v = pl.items[idx];
    const int n = (int)v.size();
    if (!n) continue;
    const int ci = idx % 3, rbi = idx / 3;
    int nodes = 0;
    for (const DirectItem &it : v) nodes = std::max(nodes, std::min(lds_reach(it, nint) + 4, M2_LDS_NODES));
    const DirectItem *di = ditems + *off;
    const int si = next;
    next = (next + 1) % 3;
    const hipStream_t s = streams[si];
    if (si > 0) used[si - 1] = true;
    switch (rbi) {
      case 0: CHK(launch_mixed2_q<8>(h, ci, n, di, s, pts, dk, ls, ct, nodes, partial)); break;
      case 1: CHK(launch_mixed2_q<4>(h, ci, n, di, s, pts, dk, ls, ct, nodes, partial)); break;
      case 2: CHK(launch_mixed2_q<2>(h, ci, n, di, s, pts, dk, ls, ct, nodes, partial)); break;
      default: CHK(launch_mixed2_q<1>(h, ci, n, di, s, pts, dk, ls, ct, nodes, partial)); break;
    }
    *off += n;
  }
  return SFBHIP_OK;
}

sfbhip_status launch_plan_mixed2(sfbhip_handle *h, const DirectPlan &pl, const std::vector<int32_t> &nint,
                                 DirectItem *ditems, size_t *off, hipStream_t st, const PointsSoA &pts,
                                 const double *dk, const LegendreScaled &ls, const CubicTables &ct,
                                 double *partial) {
  CHK(fork_sides(h, st));
  bool used[2] = {false, false};
  return join_sides(h, st, launch_mixed2_classes(h, pl, nint, ditems, off, st, pts, dk, ls, ct, partial, used),
                    used);
}

// The 512-thread kernel runs the mixed path's split-FP16 contraction when it
// is asked for; it was measured slower than the 256-thread kernel on MI355X
// (2^25 points, lmax 256: 4.34 s against 4.76 s; 1M clustered points: 6.70
// against 6.96 ms): the 512-thread kernel issues 17 % fewer vector
// instructions, but its eight waves per CU share one block's barriers and
// wait 54 % of their cycles, where two independent 256-thread blocks per CU
// cover each other's stalls.  QMAX <= 4 (its FP64 sums stay in registers).
bool use_mixed2(const sfbhip_params *p) {
  return p->precision == SFBHIP_PREC_MIXED && p->mixed_operands == SFBHIP_OPERANDS_SPLIT_FP16 &&
         p->mixed_kernel == SFBHIP_MIXED_KERNEL_512;
}
int mixed2_qcap(const sfbhip_params *p) { return std::min(p->mixed_chunks ? p->mixed_chunks : 4, 4); }

// The split-FP16 kernel with grouped Bessel evaluation (sfb_direct_split16.hpp)
// runs the mixed path's contraction by default (AUTO): measured on MI355X
// against the 256-thread kernel of sfb_direct_mixed.hpp, 2^25 points, lmax 256:
// 3.20 s against 4.29 s for the mixed phase; 1M clustered points: 4.6 against 5.6 ms.
bool use_split16(const sfbhip_params *p) {
  return p->precision == SFBHIP_PREC_MIXED && p->mixed_operands == SFBHIP_OPERANDS_SPLIT_FP16 &&
         (p->mixed_kernel == SFBHIP_MIXED_KERNEL_GROUPED || p->mixed_kernel == SFBHIP_MIXED_KERNEL_AUTO);
}
// chunks per item of a 256-thread-class plan
int mixed_qcap(const sfbhip_params *p) { return p->mixed_chunks ? p->mixed_chunks : 4; }

template <int QMAX, int KS>
sfbhip_status launch_split16(int nitems, const DirectItem *items, hipStream_t st, const PointsSoA &pts,
                             const double *kln, const LegendreScaled &ls, const CubicTables &ct, double *partial) {
  sfb_direct_split16_kernel<QMAX, KS><<<nitems, 256, 0, st>>>(pts, items, kln, ls, ct, partial);
  HIPCHK(hipGetLastError());
  return SFBHIP_OK;
}

template <int QMAX>
sfbhip_status launch_split16_ks(int ksi, int nitems, const DirectItem *items, hipStream_t st, const PointsSoA &pts,
                                const double *kln, const LegendreScaled &ls, const CubicTables &ct,
                                double *partial) {
  if (ksi == 2) return launch_split16<QMAX, 4>(nitems, items, st, pts, kln, ls, ct, partial);
  if (ksi == 1) return launch_split16<QMAX, 2>(nitems, items, st, pts, kln, ls, ct, partial);
  return launch_split16<QMAX, 1>(nitems, items, st, pts, kln, ls, ct, partial);
}

// All classes of a 256-thread mixed plan (plan_direct, uploaded in class
// order) on the grouped-Bessel kernel, dealt in turn to the call's stream and
// the two side streams.
// Items of one class reordered for the chip's eight XCDs (each with its own
// L2): the dispatcher hands workgroup i to XCD i mod 8, so position 8 k + x
// takes the k-th item of the point ranges given to XCD x (range rank mod 8),
// range after range, each range's items in their planned order (l
// descending).  The blocks resident on one XCD then work on one point range
// across many l, so its records are read from HBM about once per XCD instead
// of once per item.
constexpr int GFX950_XCDS = 8;
// Items whose point (pixel) range starts at p0: with chunk > 0 the ranges are
// [r chunk, (r + 1) chunk) and r = p0 / chunk; otherwise r is the index of p0
// in starts (the plan's ascending range starts) or, without them, its rank
// among the items' distinct starts.  Appends v to dst in XCD order.
template <class Item>
void append_xcd(std::vector<Item> &dst, const std::vector<Item> &v, int64_t chunk,
                const std::vector<int64_t> *plan_starts = nullptr) {
  const size_t n = v.size();
  if (n < 2) {
    dst.insert(dst.end(), v.begin(), v.end());
    return;
  }
  std::vector<int64_t> r(n);
  if (chunk > 0) {
    // p0 is a multiple of chunk: the rounded quotient in FP64 is exact
    const double inv = 1.0 / (double)chunk;
    for (size_t i = 0; i < n; ++i) r[i] = (int64_t)((double)v[i].p0 * inv + 0.5);
  } else {
    std::vector<int64_t> own;
    if (!plan_starts) {
      own.resize(n);
      for (size_t i = 0; i < n; ++i) own[i] = v[i].p0;
      std::sort(own.begin(), own.end());
      own.erase(std::unique(own.begin(), own.end()), own.end());
    }
    const std::vector<int64_t> &st = plan_starts ? *plan_starts : own;
    for (size_t i = 0; i < n; ++i) r[i] = std::lower_bound(st.begin(), st.end(), v[i].p0) - st.begin();
  }
  int64_t rmax = 0;
  for (size_t i = 0; i < n; ++i) rmax = std::max(rmax, r[i]);
  // the items of each range in plan order (counting sort): range q holds
  // byr[beg[q] .. beg[q + 1])
  std::vector<int64_t> beg(rmax + 2, 0);
  for (size_t i = 0; i < n; ++i) ++beg[r[i] + 1];
  for (int64_t q = 0; q <= rmax; ++q) beg[q + 1] += beg[q];
  std::vector<int32_t> byr(n);
  {
    std::vector<int64_t> fill(beg.begin(), beg.end() - 1);
    for (size_t i = 0; i < n; ++i) byr[fill[r[i]]++] = (int32_t)i;
  }
  // lane x: ranges x, x + 8, ... in order; the output interleaves the lanes
  // round-robin, skipping exhausted ones, so lane x's j-th item lands after
  // sum_y min(len_y, j) items and the lanes y < x that still have a j-th
  int64_t len[GFX950_XCDS] = {0};
  for (int64_t q = 0; q <= rmax; ++q) len[q % GFX950_XCDS] += beg[q + 1] - beg[q];
  const int64_t lmin = *std::min_element(len, len + GFX950_XCDS);
  std::vector<int32_t> perm(n);  // output position -> item
  for (int x = 0; x < GFX950_XCDS; ++x) {
    int64_t j = 0;
    for (int64_t q = x; q <= rmax; q += GFX950_XCDS)
      for (int64_t k = beg[q]; k < beg[q + 1]; ++k, ++j) {
        int64_t pos = GFX950_XCDS * j + x;  // while every lane still has items
        if (j >= lmin) {
          pos = 0;
          for (int y = 0; y < GFX950_XCDS; ++y) pos += std::min(len[y], j) + (y < x && len[y] > j ? 1 : 0);
        }
        perm[pos] = byr[k];
      }
  }
  dst.reserve(dst.size() + n);
  for (size_t k = 0; k < n; ++k) dst.push_back(v[perm[k]]);
}

sfbhip_status launch_split16_classes(sfbhip_handle *h, const DirectPlan &pl, DirectItem *ditems, size_t *off,
                                     hipStream_t st, const PointsSoA &pts, const double *dk, const LegendreScaled &ls,
                                     const CubicTables &ct, double *partial, bool *used) {
  const hipStream_t streams[3] = {st, h->mx_side[0], h->mx_side[1]};
  int next = 0;
  for (int idx = 0; idx < 18; ++idx) {
    const int ci = idx % 6, ksi = idx / 6;
    const int n = (int)pl.items[idx].size();
    if (!n) continue;
    if (ci > 3) return fail(SFBHIP_NOT_SUPPORTED, "mixed plan with more than 8 chunks per item");
    const DirectItem *di = ditems + *off;
    const int si = next;
    next = (next + 1) % 3;
    const hipStream_t s = streams[si];
    if (si > 0) used[si - 1] = true;
    switch (ci) {
      case 0: CHK(launch_split16_ks<1>(ksi, n, di, s, pts, dk, ls, ct, partial)); break;
      case 1: CHK(launch_split16_ks<2>(ksi, n, di, s, pts, dk, ls, ct, partial)); break;
      case 2: CHK(launch_split16_ks<4>(ksi, n, di, s, pts, dk, ls, ct, partial)); break;
      default: CHK(launch_split16_ks<8>(ksi, n, di, s, pts, dk, ls, ct, partial)); break;
    }
    *off += n;
  }
  return SFBHIP_OK;
}

sfbhip_status launch_plan_split16(sfbhip_handle *h, const DirectPlan &pl, DirectItem *ditems, size_t *off,
                                  hipStream_t st, const PointsSoA &pts, const double *dk, const LegendreScaled &ls,
                                  const CubicTables &ct, double *partial) {
  CHK(fork_sides(h, st));
  bool used[2] = {false, false};
  return join_sides(h, st, launch_split16_classes(h, pl, ditems, off, st, pts, dk, ls, ct, partial, used), used);
}

// The low-l FP64 items (all QMAX class 1), j_l from the quintic table.
sfbhip_status launch_lowl(const DirectPlan &pl, DirectItem *ditems, size_t *off, hipStream_t st,
                          const PointsSoA &pts, const double *dk, const BesselTableRef &tab, double *partial) {
  const int n = (int)pl.items[0].size();
  if (!n) return SFBHIP_OK;
  sfb_lowl_f64_kernel<<<n, 256, 0, st>>>(pts, ditems + *off, dk, tab, partial);
  HIPCHK(hipGetLastError());
  *off += n;
  return SFBHIP_OK;
}

void clear_timings(sfbhip_handle *h) {
  h->last_ms = 0.f;
  for (float &t : h->phase_ms) t = 0.f;
}

// After the stream has been synchronised: phase durations from the four
// events; the dominant kernel is the contraction of phase `dominant`.
sfbhip_status read_timings(sfbhip_handle *h, int dominant) {
  for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&h->phase_ms[i], h->ev[i], h->ev[i + 1]));
  h->last_ms = h->phase_ms[dominant];
  return SFBHIP_OK;
}

// ---- The HEALPix ring path (precision MIXED): FP32 maps, ring Legendre
// table, split-FP16 ring contraction (sfb_healpix_ring.hpp).

// Map workspace per batch of multipoles (bytes); l ranges whose maps exceed it
// run in several batches.
constexpr int64_t HR_MAP_CAP_BYTES = 16LL << 30;
// Cubic-table intervals the map kernel stages in LDS at most (64 KB).
constexpr int HR_LDS_NODES = 4096;

// One class of ring items: EQ, the equatorial kernel with the azimuth table
// (shm bytes) in dynamic LDS; otherwise the general ring kernel.
template <int QMAX, bool EQ>
sfbhip_status launch_ring_contract(int n, const DirectItem *di, hipStream_t s, const RingPixels &rp, const float *maps,
                                   int64_t npix, const int64_t *dmoff, double unscale, double *partial, size_t shm) {
  if constexpr (EQ) {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&healpix_ring_eq_kernel<QMAX>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    healpix_ring_eq_kernel<QMAX><<<n, 256, shm, s>>>(rp, di, maps, npix, dmoff, unscale, partial);
  } else {
    healpix_ring_contract_kernel<QMAX><<<n, 256, 0, s>>>(rp, di, maps, npix, dmoff, unscale, partial);
  }
  HIPCHK(hipGetLastError());
  return SFBHIP_OK;
}

sfbhip_status healpix_ring(sfbhip_handle *h, const int64_t *d_offsets, const double *d_radii, int32_t nside,
                           int64_t npts, int maxcount, double rdata, const sfbhip_params *p, const PixelsSoA &pix,
                           const LegendreTables &lt, double norm, double *dk, double *d_k_ln, double *d_c_ln,
                           double *d_f_lmn, int *derr, hipStream_t st) {
  const int L = p->lmax;
  // multipoles l < lsplit run in the FP64 pixel kernel with exact j_l: their
  // C_l have one to three degrees of freedom and sit closest to their noise
  const int lsplit = std::min(L, std::max(0, p->mixed_fp64_below));
  const int64_t npix = 12LL * nside * nside, nrings = 4LL * nside - 1;
  const std::vector<int32_t> &nl = h->kc.nl;
  // ring tables
  double *rx, *rsn;
  int32_t *ring_of;
  CHK(get_buf(h, "hr_rx", (size_t)nrings, &rx));
  CHK(get_buf(h, "hr_rs", (size_t)nrings, &rsn));
  CHK(get_buf(h, "hr_ring_of", (size_t)npix, &ring_of));
  // equatorial tiles take their azimuth factors from a table of 8 nside
  // entries when every such tile lies in one ring (nside a multiple of 128, so
  // that the cap pixel count and the ring length are multiples of the 256-pixel
  // tile), the table fits LDS beside the kernel's own (nside <= 512), and no m
  // folds on the equatorial rings (2 (lmax - 1) <= 4 nside)
  const bool eq_path = nside % 128 == 0 && nside <= 512 && 2 * (L - 1) <= 4 * nside;
  uint2 *eqtab;
  CHK(get_buf(h, "hr_eqtab", (size_t)8 * nside, &eqtab));
  if (h->ring_nside != nside) {
    ring_setup_kernel<<<cdiv(npix, 256), 256, 0, st>>>(nside, rx, rsn, ring_of);
    HIPCHK(hipGetLastError());
    ring_eq_table_kernel<<<cdiv(8 * nside, 256), 256, 0, st>>>(nside, eqtab);
    HIPCHK(hipGetLastError());
    h->ring_nside = nside;
  }
  // ring Legendre table: nrings (l + 1) values per l
  std::vector<int64_t> poff(std::max(L, 1), 0);
  int64_t ptot = 0;
  for (int l = 0; l < L; ++l) {
    poff[l] = ptot;
    ptot += nrings * (l + 1);
  }
  double *ptab;
  int64_t *dpoff;
  CHK(get_buf(h, "hr_ptab", (size_t)std::max<int64_t>(ptot, 1), &ptab));
  CHK(get_buf(h, "hr_poff", poff.size(), &dpoff));
  if (h->rleg_nside != nside || h->rleg_lmax < L) {
    HIPCHK(hipMemcpyAsync(dpoff, poff.data(), sizeof(int64_t) * poff.size(), hipMemcpyHostToDevice, st));
    if (L > 0) {
      ring_legendre_kernel<<<dim3((unsigned)cdiv(nrings, 64), (unsigned)L), 64, 0, st>>>(L, (int)nrings, rx, rsn, lt,
                                                                                         dpoff, ptab);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));  // poff goes out of scope
    h->rleg_nside = nside;
    h->rleg_lmax = L;
  }
  // Bessel tables up to k_{l, n_l - 1} max(rmax, largest radius)
  CubicTables ct{nullptr, nullptr, nullptr};
  if (L > 0) CHK(ensure_cubic(h, p, rdata, st, &ct));
  // map scale: |F| <= (largest pixel count) max |j_l| <= maxcount, and F 2^s <= 2^15
  int sexp = 0;
  while (sexp < 60 && std::ldexp((double)std::max(maxcount, 1), sexp + 1) <= 32768.0) ++sexp;
  while (sexp > -60 && std::ldexp((double)std::max(maxcount, 1), sexp) > 32768.0) --sexp;
  const double scale = std::ldexp(1.0, sexp);
  const double unscale = std::ldexp(1.0, -(12 + sexp));

  RingPixels rp;
  rp.phi = pix.phi;
  rp.cphi = pix.cphi;
  rp.sphi = pix.sphi;
  rp.phi2 = pix.phi2;
  rp.cphi2 = pix.cphi2;
  rp.sphi2 = pix.sphi2;
  rp.np = pix.np;
  rp.ring = ring_of;
  rp.ptab = ptab;
  rp.poff = dpoff;
  rp.alias_compat = p->alias_compat;
  rp.eqtab = eqtab;
  rp.eq_ns = eq_path ? nside : 0;
  rp.eq_lo = 2LL * nside * (nside - 1);  // the equatorial pixels [eq_lo, eq_hi)
  rp.eq_hi = npix - rp.eq_lo;
  const size_t eq_shm = eq_path ? sizeof(uint2) * 8 * (size_t)nside : 0;
  // slot order of the maps and the contraction: pixels by point count within each tile
  int32_t *slot_perm;
  CHK(get_buf(h, "hr_slot_perm", (size_t)std::max<int64_t>(npix, 1), &slot_perm));
  ring_slot_order_kernel<<<(unsigned)cdiv(npix, HR_TP), HR_TP, 0, st>>>(d_offsets, npix, slot_perm);
  HIPCHK(hipGetLastError());
  rp.perm = slot_perm;

  // batches of multipoles whose maps fit the workspace
  std::vector<std::pair<int, int>> batches;
  {
    int l0 = lsplit;
    int64_t bytes = 0;
    for (int l = lsplit; l < L; ++l) {
      const int64_t b = (int64_t)nl[l] * npix * 4;
      if (l > l0 && bytes + b > HR_MAP_CAP_BYTES) {
        batches.push_back({l0, l});
        l0 = l;
        bytes = 0;
      }
      bytes += b;
    }
    if (L > lsplit) batches.push_back({l0, L});
  }
  int64_t maxmap = 1;
  for (auto &bt : batches) {
    int64_t b = 0;
    for (int l = bt.first; l < bt.second; ++l) b += (int64_t)nl[l] * npix;
    maxmap = std::max(maxmap, b);
  }
  float *maps;
  int64_t *dmoff;
  CHK(get_buf(h, "hr_maps", (size_t)maxmap, &maps));
  CHK(get_buf(h, "hr_moff", (size_t)std::max(L, 1), &dmoff));

  // Map items of batch [lb0, lb1): per l and 16 rows, pixel ranges of one
  // chunk, appended in XCD order; moff[l] = offset of l's maps in the workspace.
  auto map_items = [&](int lb0, int lb1, std::vector<int64_t> &moff, std::vector<MapItem> &mitems) {
    moff.assign(std::max(L, 1), 0);
    mitems.clear();
    int64_t o = 0, nmaprows = 0;
    for (int l = lb0; l < lb1; ++l) {
      moff[l] = o;
      o += (int64_t)nl[l] * npix;
      nmaprows += cdiv(nl[l], HR_NB);
    }
    const int64_t target = 8192;
    int64_t S = std::max<int64_t>(1, cdiv(target, std::max<int64_t>(nmaprows, 1)));
    S = std::min<int64_t>(S, cdiv(npix, 256));
    const int64_t chunk = cdiv(cdiv(npix, S), 256) * 256;
    std::vector<MapItem> plain;
    for (int l = lb0; l < lb1; ++l)
      for (int n0 = 0; n0 < nl[l]; n0 += HR_NB)
        for (int64_t p0 = 0; p0 < npix; p0 += chunk) {
          MapItem mi;
          mi.l = l;
          mi.n0 = n0;
          mi.nrows = std::min(HR_NB, nl[l] - n0);
          mi.pad = 0;
          mi.p0 = p0;
          mi.p1 = std::min(npix, p0 + chunk);
          mi.moff = moff[l];
          mi.kofs = (int64_t)l * p->nmax + n0;
          plain.push_back(mi);
        }
    append_xcd(mitems, plain, chunk);
  };
  // cubic-table intervals the map kernel stages in LDS for batch [lb0, lb1)
  auto map_nodes = [&](int lb0, int lb1) {
    int nodes = 1;
    for (int l = lb0; l < lb1; ++l)
      if (nl[l] > 0) nodes = std::max(nodes, h->kc.cub_nint[l]);
    return std::min(nodes, HR_LDS_NODES);
  };

  // One plan over all multipoles (the partial blocks of every batch), items
  // launched per batch.  A single-batch plan and its device copies are kept in
  // the handle: a call of the same shape plans and uploads nothing.
  PlanCtx cx;
  cx.nmax = p->nmax;
  cx.norm = norm;
  const bool single = batches.size() == 1;
  if (!h->ring_plan) h->ring_plan = std::make_shared<RingPlanCache>();
  RingPlanCache &rc_plan = *h->ring_plan;
  const bool hit = single && rc_plan.nside == nside && rc_plan.L == L && rc_plan.nmax == p->nmax &&
                   rc_plan.lsplit == lsplit && rc_plan.eq == eq_path && rc_plan.nl == nl;
  DirectPlan pl64local, pl;
  std::vector<std::vector<DirectItem>> bitems(batches.size());
  std::vector<std::array<int, 6>> bcounts(batches.size());
  if (hit) {
    pl.groups = rc_plan.groups;
    for (ReduceGroup &g : pl.groups) g.norm = norm;
    pl.partial_size = rc_plan.partial_size;
    bcounts[0] = rc_plan.counts;
  } else {
    rc_plan.nside = 0;  // invalid until the new plan is uploaded
    if (lsplit > 0) plan_direct(nl, 0, lsplit, npix, F64_NT, F64_TP, F64_MC, 16, 4096, cx, &pl64local);
    pl.partial_size = pl64local.partial_size;
    pl.groups = pl64local.groups;
    // items never straddle the equatorial boundaries, so that each is all
    // equatorial (healpix_ring_eq_kernel) or all other (healpix_ring_contract_kernel);
    // per batch the classes are uploaded as [other Q = 1, 2, 4][equatorial Q = 1, 2, 4]
    const int64_t eq_lo = 2LL * nside * (nside - 1), eq_hi = npix - eq_lo;
    const std::vector<int64_t> cuts =
        eq_path ? std::vector<int64_t>{0, eq_lo, eq_hi, npix} : std::vector<int64_t>{0, npix};
    for (size_t b = 0; b < batches.size(); ++b) {
      DirectPlan pb;
      pb.partial_size = pl.partial_size;
      plan_direct(nl, batches[b].first, batches[b].second, npix, HR_NT, HR_TP, MX_MC, 4,
                  f64_target_items(npix, true) / 2, cx, &pb, 1, &cuts);
      pl.partial_size = pb.partial_size;
      pl.groups.insert(pl.groups.end(), pb.groups.begin(), pb.groups.end());
      if (!pb.items[3].empty() || !pb.items[4].empty()) return fail(SFBHIP_NOT_SUPPORTED, "ring plan class");
      for (int e = 0; e < 2; ++e)
        for (int ci = 0; ci < 3; ++ci) {
          std::vector<DirectItem> cls;
          for (const DirectItem &it : pb.items[ci]) {
            const bool is_eq = eq_path && it.p0 >= eq_lo && it.p1 <= eq_hi;
            if (is_eq == (e == 1)) cls.push_back(it);
          }
          append_xcd(bitems[b], cls, 0, &pb.starts);
          bcounts[b][3 * e + ci] = (int)cls.size();
        }
    }
  }
  const DirectPlan &pl64 = hit ? rc_plan.pl64 : pl64local;
  double *partial;
  CHK(get_buf(h, "partial", (size_t)std::max<int64_t>(pl.partial_size, 1), &partial));

  // host arrays the asynchronous copies read stay alive until the stream is synchronised
  std::vector<DirectItem> lo;
  int64_t nlo = 0;
  for (auto &v : pl64.items) nlo += (int64_t)v.size();
  DirectItem *dlo = nullptr;
  if (nlo > 0) CHK(get_buf(h, "hr_lowl_items", (size_t)nlo, &dlo));
  std::vector<std::vector<int64_t>> bmoff(batches.size());
  std::vector<std::vector<MapItem>> bmitems(batches.size());
  MapItem *dmi = nullptr;
  DirectItem *ditems = nullptr;
  int64_t nmap0 = hit ? rc_plan.nmap : 0;
  if (!hit) {
    for (auto &v : pl64.items) lo.insert(lo.end(), v.begin(), v.end());
    if (nlo > 0) HIPCHK(hipMemcpyAsync(dlo, lo.data(), sizeof(DirectItem) * lo.size(), hipMemcpyHostToDevice, st));
    if (single) {
      // the whole single-batch plan on the device now, and kept
      map_items(batches[0].first, batches[0].second, bmoff[0], bmitems[0]);
      nmap0 = (int64_t)bmitems[0].size();
      CHK(get_buf(h, "hr_mitems", std::max<size_t>(bmitems[0].size(), 1), &dmi));
      CHK(get_buf(h, "hr_items", std::max<size_t>(bitems[0].size(), 1), &ditems));
      HIPCHK(hipMemcpyAsync(dmoff, bmoff[0].data(), sizeof(int64_t) * bmoff[0].size(), hipMemcpyHostToDevice, st));
      if (nmap0)
        HIPCHK(hipMemcpyAsync(dmi, bmitems[0].data(), sizeof(MapItem) * bmitems[0].size(), hipMemcpyHostToDevice, st));
      if (!bitems[0].empty())
        HIPCHK(hipMemcpyAsync(ditems, bitems[0].data(), sizeof(DirectItem) * bitems[0].size(), hipMemcpyHostToDevice,
                              st));
      rc_plan.L = L;
      rc_plan.nmax = p->nmax;
      rc_plan.lsplit = lsplit;
      rc_plan.eq = eq_path;
      rc_plan.nl = nl;
      rc_plan.pl64 = pl64local;
      rc_plan.groups = pl.groups;
      rc_plan.partial_size = pl.partial_size;
      rc_plan.counts = bcounts[0];
      rc_plan.nmap = nmap0;
      rc_plan.nside = nside;  // valid
    }
  } else {
    CHK(get_buf(h, "hr_mitems", (size_t)std::max<int64_t>(nmap0, 1), &dmi));
    size_t nit = 0;
    for (int k = 0; k < 6; ++k) nit += (size_t)bcounts[0][k];
    CHK(get_buf(h, "hr_items", std::max<size_t>(nit, 1), &ditems));
  }

  // Streams: the maps and the largest contraction class run on st; l < lsplit
  // (FP64, exact j_l: some of these C_l sit at cancellation level, where even
  // the quintic table's 1e-10 of the amplitude shows) runs on the handle's two
  // side streams beside the maps; the smaller contraction classes run on the
  // side streams after the maps.  The side streams are forked from st here and
  // joined back before the reduction.
  HIPCHK(hipEventRecord(h->ev[0], st));
  CHK(fork_sides(h, st));
  bool used[2] = {false, false};
  sfbhip_status rc = SFBHIP_OK;
  if (nlo > 0) {
    size_t off = 0;
    rc = launch_f64_classes(h, pl64, dlo, &off, st, PointsSoA{}, &pix, dk, p->nmax, lt, JL_EXACT, BesselTableRef{},
                            partial, derr, used, true);
  }
  for (size_t b = 0; b < batches.size() && rc == SFBHIP_OK; ++b) {
    rc = [&]() -> sfbhip_status {
      const int lb0 = batches[b].first, lb1 = batches[b].second;
      int64_t nmap = nmap0;
      if (!single) {
        // several batches: each plans and uploads its own map items and work items
        map_items(lb0, lb1, bmoff[b], bmitems[b]);
        nmap = (int64_t)bmitems[b].size();
        CHK(get_buf(h, "hr_mitems", std::max<size_t>(bmitems[b].size(), 1), &dmi));
        CHK(get_buf(h, "hr_items", std::max<size_t>(bitems[b].size(), 1), &ditems));
        HIPCHK(hipMemcpyAsync(dmoff, bmoff[b].data(), sizeof(int64_t) * bmoff[b].size(), hipMemcpyHostToDevice, st));
        if (nmap)
          HIPCHK(hipMemcpyAsync(dmi, bmitems[b].data(), sizeof(MapItem) * bmitems[b].size(), hipMemcpyHostToDevice,
                                st));
        if (!bitems[b].empty())
          HIPCHK(hipMemcpyAsync(ditems, bitems[b].data(), sizeof(DirectItem) * bitems[b].size(),
                                hipMemcpyHostToDevice, st));
      }
      const int nodes = map_nodes(lb0, lb1);
      const size_t shm = sizeof(float4) * (size_t)nodes;
      if (nmap) {
        HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(&ring_maps_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        ring_maps_kernel<<<(unsigned)nmap, 256, shm, st>>>(dmi, d_offsets, slot_perm, d_radii, npix, dk, ct, nodes,
                                                           scale, maps);
        HIPCHK(hipGetLastError());
      }
      if (single) HIPCHK(hipEventRecord(h->ev[1], st));
      // the contraction classes start once the maps are written: the classes
      // of 4 chunks (the largest) on st, of 2 on side stream 0, of 1 on side
      // stream 1; equatorial classes after the others of their stream
      HIPCHK(hipEventRecord(h->mx_fork, st));
      for (int k = 0; k < 2; ++k) HIPCHK(hipStreamWaitEvent(h->mx_side[k], h->mx_fork, 0));
      size_t off = 0;
      for (int k = 0; k < 6; ++k) {
        const int n = bcounts[b][k];
        if (!n) continue;
        const DirectItem *di = ditems + off;
        const int ci = k % 3;
        const hipStream_t s = ci == 2 ? st : h->mx_side[1 - ci];
        if (ci < 2) used[1 - ci] = true;
        switch (k) {
          case 0: CHK((launch_ring_contract<1, false>(n, di, s, rp, maps, npix, dmoff, unscale, partial, 0))); break;
          case 1: CHK((launch_ring_contract<2, false>(n, di, s, rp, maps, npix, dmoff, unscale, partial, 0))); break;
          case 2: CHK((launch_ring_contract<4, false>(n, di, s, rp, maps, npix, dmoff, unscale, partial, 0))); break;
          case 3: CHK((launch_ring_contract<1, true>(n, di, s, rp, maps, npix, dmoff, unscale, partial, eq_shm))); break;
          case 4: CHK((launch_ring_contract<2, true>(n, di, s, rp, maps, npix, dmoff, unscale, partial, eq_shm))); break;
          default: CHK((launch_ring_contract<4, true>(n, di, s, rp, maps, npix, dmoff, unscale, partial, eq_shm))); break;
        }
        off += n;
      }
      if (b + 1 < batches.size()) {
        // the next batch rewrites the maps and items: every stream drains first
        for (int k = 0; k < 2; ++k) HIPCHK(hipStreamSynchronize(h->mx_side[k]));
        HIPCHK(hipStreamSynchronize(st));
      }
      return SFBHIP_OK;
    }();
  }
  CHK(join_sides(h, st, rc, used));
  if (batches.size() != 1) HIPCHK(hipEventRecord(h->ev[1], st));
  HIPCHK(hipEventRecord(h->ev[2], st));
  CHK(finalize(h, p, pl.groups, partial, norm, d_k_ln, d_c_ln, d_f_lmn, st));
  HIPCHK(hipEventRecord(h->ev[3], st));
  HIPCHK(hipStreamSynchronize(st));
  // phases (events on st): [0] maps (l < lsplit runs beside them on the side
  // streams), [1] ring contraction up to the join of every stream, [2]
  // reduction and C_l; the dominant kernels are the map and contraction kernels
  for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&h->phase_ms[i], h->ev[i], h->ev[i + 1]));
  h->last_ms = h->phase_ms[0] + h->phase_ms[1];
  return SFBHIP_OK;
}

}  // namespace

extern "C" {

sfbhip_status sfbhip_default_params(sfbhip_params *p) {
  if (!p) return fail(SFBHIP_INVALID_ARGUMENT, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->lmax = 0;
  p->nmax = 1;
  p->rmax = 1.0;
  p->k_max = 0.0;
  p->window_volume = 0.0;
  p->precision = SFBHIP_PREC_FP64;
  p->bessel = SFBHIP_BESSEL_AUTO;
  p->zeros = SFBHIP_ZEROS_REFERENCE;
  p->alias_compat = 1;
  p->verbose = 0;
  p->mixed_fp64_below = 2;
  p->mixed_operands = SFBHIP_OPERANDS_SPLIT_FP16;
  p->mixed_chunks = 4;
  p->mixed_kernel = SFBHIP_MIXED_KERNEL_AUTO;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_create(sfbhip_handle **out, int device) {
  if (!out) return fail(SFBHIP_INVALID_ARGUMENT, "out is NULL");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(SFBHIP_INVALID_ARGUMENT, "no such device");
  HIPCHK(hipSetDevice(device));
  sfbhip_handle *h = new sfbhip_handle();
  h->device = device;
  hipError_t e = hipHostMalloc((void **)&h->h_err, sizeof(int) * 4, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete h;
    return fail(SFBHIP_HIP_ERROR, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  {
    int a = 0, b = 0;
    if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) a = 0;
    if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, device) != hipSuccess) b = 0;
    (void)hipGetLastError();
    h->lds_max = std::max(65536, std::max(a, b));
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cu = 0;
    (void)hipGetLastError();
    h->ncu = cu > 0 ? cu : 256;
  }
  default_tiles(&h->tiles);
  for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventCreateWithFlags(&h->mx_fork, hipEventDisableTiming));
  for (int k = 0; k < 2; ++k) {
    HIPCHK(hipEventCreateWithFlags(&h->mx_join[k], hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&h->mx_side[k], hipStreamNonBlocking));
  }
  *out = h;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_destroy(sfbhip_handle *h) {
  if (!h) return SFBHIP_OK;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : h->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  if (h->h_err) (void)hipHostFree(h->h_err);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->mx_fork) (void)hipEventDestroy(h->mx_fork);
  for (int k = 0; k < 2; ++k) {
    if (h->mx_join[k]) (void)hipEventDestroy(h->mx_join[k]);
    if (h->mx_side[k]) (void)hipStreamDestroy(h->mx_side[k]);
  }
  delete h;
  return SFBHIP_OK;
}

size_t sfbhip_flmn_count(int32_t lmax, int32_t nmax) {
  if (lmax <= 0 || nmax <= 0) return 0;
  return (size_t)lmax * (lmax + 1) / 2 * (size_t)nmax;
}

sfbhip_status sfbhip_kln(sfbhip_handle *h, const sfbhip_params *p, double *k_ln_host, int32_t *nl_host) {
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  CHK(validate(p));
  HIPCHK(hipSetDevice(h->device));
  CHK(ensure_kln(h, p, nullptr));
  HIPCHK(hipStreamSynchronize(nullptr));
  if (k_ln_host) std::memcpy(k_ln_host, h->kc.kln.data(), sizeof(double) * h->kc.kln.size());
  if (nl_host) std::memcpy(nl_host, h->kc.nl.data(), sizeof(int32_t) * h->kc.nl.size());
  return SFBHIP_OK;
}

sfbhip_status sfbhip_last_kernel_ms(sfbhip_handle *h, float *ms) {
  if (!h || !ms) return fail(SFBHIP_INVALID_ARGUMENT, "NULL argument");
  *ms = h->last_ms;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_last_phase_ms(sfbhip_handle *h, float *ms3) {
  if (!h || !ms3) return fail(SFBHIP_INVALID_ARGUMENT, "NULL argument");
  for (int i = 0; i < 3; ++i) ms3[i] = h->phase_ms[i];
  return SFBHIP_OK;
}

const char *sfbhip_last_error(void) { return g_last_error.c_str(); }

sfbhip_status sfbhip_direct(sfbhip_handle *h, const double *d_points, int64_t n_points, const sfbhip_params *p,
                            double *d_k_ln, double *d_c_ln, double *d_f_lmn, void *stream) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  clear_timings(h);  // a call refused before its contraction reports zero phase times
  CHK(validate(p));
  if (n_points <= 0 || !d_points)
    return fail(SFBHIP_INVALID_ARGUMENT, "empty catalogue: the normalisation divides by the number of points");
  if (!d_k_ln || !d_c_ln) return fail(SFBHIP_INVALID_ARGUMENT, "k_ln / c_ln output is NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  if (p->verbose) {
    std::printf("SFB decomposition (raw method): lmax=%d nmax=%d Rmax=%g\n", p->lmax, p->nmax, p->rmax);
    std::printf("Catalog consists of %lld points.\n", (long long)n_points);
  }
  int *derr;
  CHK(zero_err(h, st, &derr));
  CHK(ensure_kln(h, p, st));
  LegendreTables lt;
  CHK(ensure_legendre(h, std::max(p->lmax, 1), st, &lt));
  PointsSoA pts;
  double rdata = 0.0;
  CHK(prep_points(h, d_points, n_points, st, derr, &pts, &rdata));
  const double norm = norm_factor(p, (double)n_points);
  double *dk;
  CHK(get_buf(h, "kln", (size_t)std::max(p->lmax, 1) * p->nmax, &dk));

  // Multipoles l < lsplit run in the FP64 kernel; the rest in the mixed kernel
  // when precision is MIXED.  The mixed path reads the catalogue sorted by radius.
  const bool mixed = p->precision == SFBHIP_PREC_MIXED;
  if (mixed && p->lmax > MX_LMAX)
    return fail(SFBHIP_NOT_SUPPORTED, "mixed path supports lmax <= " + std::to_string(MX_LMAX) + "; use precision FP64");
  // the mixed kernels rely on the radius order (their Bessel windows), which the sort provides up to 2^31 - 1 points
  if (mixed && n_points > INT32_MAX)
    return fail(SFBHIP_NOT_SUPPORTED, "mixed path: more than 2^31 - 1 points; split the catalogue or use precision FP64");
  if (mixed) CHK(sort_points_by_radius(h, pts, st, &pts));
  const int lsplit = mixed ? std::min(p->lmax, std::max(0, p->mixed_fp64_below)) : p->lmax;
  // FP64 kernel's j_l: in FP64 precision the reference's choice (exact or its
  // linear table); in the mixed path the quintic table (error ~1e-10 of the
  // amplitude, far below the mixed kernels' own).
  int use_table = JL_EXACT;
  if (mixed)
    use_table = JL_QUINTIC;
  else if (p->bessel == SFBHIP_BESSEL_REFERENCE || (p->bessel == SFBHIP_BESSEL_AUTO && n_points > 1000))
    use_table = JL_REFERENCE_TABLE;
  BesselTableRef tab{};
  if (use_table == JL_REFERENCE_TABLE) CHK(ensure_table(h, p, st, &tab));
  if (use_table == JL_QUINTIC && lsplit > 0) CHK(ensure_quintic(h, p, lsplit, rdata, st, &tab));
  HermiteTables ht{nullptr, nullptr, nullptr};
  CubicTables ct{nullptr, nullptr, nullptr};
  LegendreScaled ls{};
  const bool m2 = use_mixed2(p), s16 = use_split16(p);
  if (mixed && lsplit < p->lmax) {
    if (m2 || s16) {
      CHK(ensure_cubic(h, p, rdata, st, &ct));
      CHK(ensure_legendre_scaled(h, std::max(p->lmax, 1), st, &ls));
    } else {
      CHK(ensure_hermite(h, p, rdata, st, &ht));
    }
  }
  // mixed path: l < min(lsplit, LW_LMAX) in the low-l FP64 kernel, the rest of
  // l < lsplit in the general FP64 kernel (both with the quintic table)
  const int llow = mixed ? std::min(lsplit, LW_LMAX) : 0;
  if (!h->point_plan) h->point_plan = std::make_shared<PointPlanCache>();
  PointPlanCache &pc = *h->point_plan;
  const double R = std::max(p->rmax, rdata);
  const bool hit = pc.n == n_points && pc.lmax == p->lmax && pc.nmax == p->nmax && pc.zeros == p->zeros &&
                   pc.rmax == p->rmax && pc.kmax == p->k_max && pc.mixed == mixed && pc.lsplit == lsplit &&
                   pc.llow == llow && pc.kernel == p->mixed_kernel && pc.chunks == p->mixed_chunks &&
                   pc.operands == p->mixed_operands && pc.R == R;
  DirectPlan &pllw = pc.pllw, &pl64 = pc.pl64, &plmx = pc.plmx;
  std::vector<ReduceGroup> &groups = pc.groups;
  std::vector<DirectItem> all;  // the items of a new plan; copied to the device before the call returns
  DirectItem *ditems;
  if (!hit) {
    pc.n = 0;  // no plan kept until this one is on the device
    pllw = DirectPlan{};
    pl64 = DirectPlan{};
    plmx = DirectPlan{};
    PlanCtx cx;
    cx.nmax = p->nmax;
    cx.norm = norm;
    cx.kln = h->kc.kln.data();
    cx.R = R;
    // the low-l items: 1024 point ranges per (l, 16-row group) (the partial
    // reduction sums them in eight chains; 256 ranges: 0.04 ms slower at 1M points)
    if (llow > 0)
      plan_direct(h->kc.nl, 0, llow, n_points, LW_ROWS, LW_TILE, F64_MC, 1, 1024 * (int64_t)llow * (64 / LW_ROWS),
                  cx, &pllw);
    pl64.partial_size = pllw.partial_size;
    plan_direct(h->kc.nl, llow, lsplit, n_points, F64_NT, F64_TP, F64_MC, 16, f64_target_items(n_points, false),
                cx, &pl64);
    plmx.partial_size = pl64.partial_size;
    if (mixed) {
      const int64_t target = mixed_target_items(n_points);
      if (m2)
        plan_mixed2(h->kc.nl, lsplit, p->lmax, n_points, mixed2_qcap(p), target, cx, &plmx);
      else
        plan_direct(h->kc.nl, lsplit, p->lmax, n_points, MX_NT, MX_TP, MX_MC, mixed_qcap(p), target, cx, &plmx,
                    mixed_ks_cap(p->mixed_operands), nullptr, true);
    }
    groups = pllw.groups;
    groups.insert(groups.end(), pl64.groups.begin(), pl64.groups.end());
    groups.insert(groups.end(), plmx.groups.begin(), plmx.groups.end());
    for (auto &v : pllw.items) all.insert(all.end(), v.begin(), v.end());
    for (auto &v : pl64.items) all.insert(all.end(), v.begin(), v.end());
    // split-FP16 classes in XCD order (each class keeps its place in the list)
    for (auto &v : plmx.items)
      if (mixed && s16)
        append_xcd(all, v, plmx.chunk);
      else
        all.insert(all.end(), v.begin(), v.end());
    CHK(get_buf(h, "px_items", std::max<size_t>(all.size(), 1), &ditems));
    if (!all.empty())
      HIPCHK(hipMemcpyAsync(ditems, all.data(), sizeof(DirectItem) * all.size(), hipMemcpyHostToDevice, st));
    pc.lmax = p->lmax;
    pc.nmax = p->nmax;
    pc.zeros = p->zeros;
    pc.rmax = p->rmax;
    pc.kmax = p->k_max;
    pc.mixed = mixed;
    pc.lsplit = lsplit;
    pc.llow = llow;
    pc.kernel = p->mixed_kernel;
    pc.chunks = p->mixed_chunks;
    pc.operands = p->mixed_operands;
    pc.R = R;
    pc.n = n_points;
  } else {
    CHK(get_buf(h, "px_items", 1, &ditems));
  }
  for (ReduceGroup &g : groups) g.norm = norm;
  double *partial = nullptr;
  CHK(get_buf(h, "partial", (size_t)std::max<int64_t>(plmx.partial_size, 1), &partial));
  clear_timings(h);
  HIPCHK(hipEventRecord(h->ev[0], st));
  size_t off = 0;
  CHK(launch_lowl(pllw, ditems, &off, st, pts, dk, tab, partial));
  CHK(launch_plan_f64(h, pl64, ditems, &off, st, pts, nullptr, dk, p->nmax, lt, use_table, tab, partial, derr));
  HIPCHK(hipEventRecord(h->ev[1], st));
  if (mixed && m2)
    CHK(launch_plan_mixed2(h, plmx, h->kc.cub_nint, ditems, &off, st, pts, dk, ls, ct, partial));
  else if (mixed && s16)
    CHK(launch_plan_split16(h, plmx, ditems, &off, st, pts, dk, ls, ct, partial));
  else if (mixed)
    CHK(launch_plan_mixed(h, plmx, h->kc.herm_nint, ditems, &off, st, pts, dk, p->nmax, lt, ht,
                          p->mixed_operands, partial));
  HIPCHK(hipEventRecord(h->ev[2], st));
  CHK(finalize(h, p, groups, partial, norm, d_k_ln, d_c_ln, d_f_lmn, st));
  HIPCHK(hipEventRecord(h->ev[3], st));
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  CHK(read_timings(h, mixed && lsplit < p->lmax ? 1 : 0));
  if (e & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  if (e & ERR_DOMAIN) return fail(SFBHIP_DOMAIN_ERROR, "argument (z) outside of defined range.");
  return SFBHIP_OK;
}

sfbhip_status sfbhip_direct_batched(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                    const sfbhip_params *p, void *stream) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  clear_timings(h);  // a call refused before its contraction reports zero phase times
  if (!cats || ncat < 1 || ncat > 128) return fail(SFBHIP_INVALID_ARGUMENT, "ncat must be in [1, 128]");
  if (!p) return fail(SFBHIP_INVALID_ARGUMENT, "params is NULL");
  // per-catalogue parameters, each validated as a single call would be
  std::vector<sfbhip_params> cp(ncat, *p);
  int64_t ntot = 0;
  int L = 0, NZ = 1;
  for (int c = 0; c < ncat; ++c) {
    cp[c].lmax = cats[c].lmax;
    cp[c].nmax = cats[c].nmax;
    cp[c].rmax = cats[c].rmax;
    cp[c].k_max = cats[c].k_max;
    cp[c].window_volume = cats[c].window_volume;
    CHK(validate(&cp[c]));
    if (cats[c].n_points <= 0 || !cats[c].d_points)
      return fail(SFBHIP_INVALID_ARGUMENT, "catalogue " + std::to_string(c) +
                                               ": empty catalogue: the normalisation divides by the number of points");
    if (!cats[c].d_k_ln || !cats[c].d_c_ln)
      return fail(SFBHIP_INVALID_ARGUMENT, "catalogue " + std::to_string(c) + ": k_ln / c_ln output is NULL");
    ntot += cats[c].n_points;
    L = std::max(L, cats[c].lmax);
    NZ = std::max(NZ, cats[c].nmax + 1);
  }
  if (ntot > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "more than 2^31 points in one batch");
  const bool mixed = p->precision == SFBHIP_PREC_MIXED;
  if (mixed && L > MX_LMAX)
    return fail(SFBHIP_NOT_SUPPORTED, "mixed path supports lmax <= " + std::to_string(MX_LMAX) + "; use precision FP64");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  if (p->verbose) std::printf("SFB decomposition (raw method), batch of %d catalogues, %lld points\n", ncat, (long long)ntot);

  int *derr;
  CHK(zero_err(h, st, &derr));
  // zeros of j_l, l < L, shared by all catalogues (the first nmax_c + 1 of them are
  // what a single call with nmax_c computes)
  std::vector<double> zeros((size_t)std::max(L, 1) * NZ, 0.0);
  if (L > 0) {
    double *dz;
    CHK(get_buf(h, "b_zeros", (size_t)L * NZ, &dz));
    bessel_zeros_kernel<<<cdiv(L, 64), 64, 0, st>>>(L, NZ, p->zeros, dz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(zeros.data(), dz, sizeof(double) * L * NZ, hipMemcpyDeviceToHost, st));
  }
  // point tables, concatenated; largest radius per catalogue
  double *soa;
  unsigned long long *rbits;
  CHK(get_buf(h, "b_soa", (size_t)6 * ntot, &soa));
  CHK(get_buf(h, "b_rbits", (size_t)ncat, &rbits));
  fill_kernel<<<1, 128, 0, st>>>(reinterpret_cast<double *>(rbits), ncat, 0.0);
  HIPCHK(hipGetLastError());
  std::vector<int64_t> pbase(ncat);
  {
    int64_t o = 0;
    for (int c = 0; c < ncat; ++c) {
      pbase[c] = o;
      const int64_t n = cats[c].n_points;
      prep_points_kernel<<<cdiv(n, 256), 256, 0, st>>>(cats[c].d_points, n, soa + o, soa + ntot + o,
                                                       soa + 2 * ntot + o, soa + 3 * ntot + o, soa + 4 * ntot + o,
                                                       soa + 5 * ntot + o, derr, rbits + c);
      HIPCHK(hipGetLastError());
      o += n;
    }
  }
  std::vector<double> rdata(ncat);
  HIPCHK(hipMemcpyAsync(rdata.data(), rbits, sizeof(double) * ncat, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h->h_err, derr, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // no kernel evaluates a Bessel function of a bad point: refused before the tables and contractions
  if (h->h_err[0] & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  PointsSoA pts{soa, soa + ntot, soa + 2 * ntot, soa + 3 * ntot, soa + 4 * ntot, soa + 5 * ntot, ntot};
  if (mixed && ntot > 1) {
    // sort by (catalogue, radius): each catalogue stays in its range, ordered by r
    unsigned long long *k0, *k1;
    int32_t *i0, *i1;
    double *sb;
    CHK(get_buf(h, "sort_k0", (size_t)ntot, &k0));
    CHK(get_buf(h, "sort_k1", (size_t)ntot, &k1));
    CHK(get_buf(h, "sort_i0", (size_t)ntot, &i0));
    CHK(get_buf(h, "sort_i1", (size_t)ntot, &i1));
    CHK(get_buf(h, "b_soa_sorted", (size_t)6 * ntot, &sb));
    for (int c = 0; c < ncat; ++c) {
      radius_keys_cat_kernel<<<cdiv(cats[c].n_points, 256), 256, 0, st>>>(pts.r, cats[c].n_points, pbase[c], c, k0,
                                                                         i0);
      HIPCHK(hipGetLastError());
    }
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, i0, i1, (int)ntot, 0, 64, st));
    char *tmp;
    CHK(get_buf(h, "sort_tmp", tmp_bytes, &tmp));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i0, i1, (int)ntot, 0, 64, st));
    gather_points_kernel<<<cdiv(ntot, 256), 256, 0, st>>>(pts, i1, ntot, sb, sb + ntot, sb + 2 * ntot, sb + 3 * ntot,
                                                         sb + 4 * ntot, sb + 5 * ntot);
    HIPCHK(hipGetLastError());
    pts = PointsSoA{sb, sb + ntot, sb + 2 * ntot, sb + 3 * ntot, sb + 4 * ntot, sb + 5 * ntot, ntot};
  }

  // k_ln, n_l, f_lmn offsets per catalogue
  std::vector<double> kcat;
  std::vector<int32_t> nlcat;
  std::vector<int64_t> foffcat;
  std::vector<int64_t> kbase(ncat), nlbase(ncat), fbase(ncat), foffbase(ncat);
  std::vector<std::vector<int32_t>> nl(ncat);
  int64_t ftot = 0;
  for (int c = 0; c < ncat; ++c) {
    const int Lc = cats[c].lmax, Nc = cats[c].nmax;
    const double kmax = cats[c].k_max > 0 ? cats[c].k_max : 0.0;
    kbase[c] = (int64_t)kcat.size();
    nlbase[c] = (int64_t)nlcat.size();
    foffbase[c] = (int64_t)foffcat.size();
    fbase[c] = ftot;
    nl[c].assign(std::max(Lc, 1), 0);
    for (int l = 0; l < Lc; ++l) {
      int cnt = 0;
      for (int n = 0; n < Nc; ++n) {
        const double z = zeros[(size_t)l * NZ + n];
        kcat.push_back(z / cats[c].rmax);
        if (kmax <= 0.0 || z < kmax * cats[c].rmax) cnt = n + 1;
      }
      nl[c][l] = cnt;
    }
    nlcat.insert(nlcat.end(), nl[c].begin(), nl[c].end());
    std::vector<int64_t> fo = flmn_offsets(Lc, Nc);
    foffcat.insert(foffcat.end(), fo.begin(), fo.end());
    ftot += (int64_t)sfbhip_flmn_count(Lc, Nc);
  }
  double *dk;
  int32_t *dnl;
  int64_t *dfoff;
  CHK(get_buf(h, "b_kln", std::max<size_t>(kcat.size(), 1), &dk));
  CHK(get_buf(h, "b_nl", std::max<size_t>(nlcat.size(), 1), &dnl));
  CHK(get_buf(h, "b_foff", std::max<size_t>(foffcat.size(), 1), &dfoff));
  if (!kcat.empty()) HIPCHK(hipMemcpyAsync(dk, kcat.data(), sizeof(double) * kcat.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dnl, nlcat.data(), sizeof(int32_t) * nlcat.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dfoff, foffcat.data(), sizeof(int64_t) * foffcat.size(), hipMemcpyHostToDevice, st));

  // tables shared by all catalogues: Legendre constants for l < L; Bessel
  // tables per l long enough for the largest k_{l, n_l - 1} max(rmax, max r)
  LegendreTables lt;
  CHK(ensure_legendre(h, std::max(L, 1), st, &lt));
  const int lsplit = mixed ? std::min(L, std::max(0, p->mixed_fp64_below)) : L;
  std::vector<double> xmax(std::max(L, 1), 0.0);
  for (int c = 0; c < ncat; ++c)
    for (int l = 0; l < cats[c].lmax; ++l)
      if (nl[c][l] > 0)
        xmax[l] = std::max(xmax[l], kcat[kbase[c] + (int64_t)l * cats[c].nmax + nl[c][l] - 1] *
                                        std::max(cats[c].rmax, rdata[c]));
  BesselTableRef tab{};
  HermiteTables ht{nullptr, nullptr, nullptr};
  std::vector<int32_t> hnint(std::max(L, 1), 1);
  if (mixed && lsplit > 0) {
    std::vector<int32_t> off(lsplit), ni(lsplit);
    int64_t o = 0;
    int32_t mx = 1;
    for (int l = 0; l < lsplit; ++l) {
      ni[l] = (int32_t)std::floor(xmax[l] * Q5_HINV) + 2;
      off[l] = (int32_t)o;
      o += ni[l];
      mx = std::max(mx, ni[l]);
    }
    double *q5;
    int32_t *doff, *dni;
    CHK(get_buf(h, "b_q5_tab", (size_t)6 * o, &q5));
    CHK(get_buf(h, "b_q5_off", off.size(), &doff));
    CHK(get_buf(h, "b_q5_nint", ni.size(), &dni));
    HIPCHK(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dni, ni.data(), sizeof(int32_t) * ni.size(), hipMemcpyHostToDevice, st));
    quintic_table_kernel<<<dim3((unsigned)cdiv(mx, 256), (unsigned)lsplit), 256, 0, st>>>(lsplit, doff, dni, q5);
    HIPCHK(hipGetLastError());
    tab.q5 = q5;
    tab.q5_off = doff;
    tab.q5_nint = dni;
    HIPCHK(hipStreamSynchronize(st));  // host vectors go out of scope
  }
  const bool m2 = use_mixed2(p), s16 = use_split16(p);
  CubicTables ct{nullptr, nullptr, nullptr};
  LegendreScaled ls{};
  std::vector<int32_t> cnint;
  if (mixed && lsplit < L && (m2 || s16)) {
    CHK(build_cubic(h, L, xmax, "b_cub", st, &ct, &cnint));
    CHK(ensure_legendre_scaled(h, std::max(L, 1), st, &ls));
  }
  if (mixed && lsplit < L && !m2 && !s16) {
    std::vector<int32_t> off(L, 0);
    int64_t o = 0;
    int32_t mx = 1;
    for (int l = 0; l < L; ++l) {
      const int64_t n = (int64_t)std::floor(xmax[l] / MX_H) + 2;
      if (n > MX_TAB_MAX)
        return fail(SFBHIP_NOT_SUPPORTED, "mixed path: largest Bessel argument k r exceeds " +
                                              std::to_string(MX_TAB_MAX * MX_H) + "; use precision FP64");
      hnint[l] = (int32_t)n;
      off[l] = (int32_t)o;
      o += n;
      mx = std::max<int32_t>(mx, (int32_t)n);
      if (o > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "mixed path: Bessel tables too large");
    }
    float4 *tb;
    int32_t *doff, *dni;
    CHK(get_buf(h, "b_herm_tab", (size_t)o, &tb));
    CHK(get_buf(h, "b_herm_off", off.size(), &doff));
    CHK(get_buf(h, "b_herm_nint", hnint.size(), &dni));
    HIPCHK(hipMemcpyAsync(doff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dni, hnint.data(), sizeof(int32_t) * hnint.size(), hipMemcpyHostToDevice, st));
    hermite_table_kernel<<<dim3((unsigned)cdiv(mx, 256), (unsigned)L), 256, 0, st>>>(L, doff, dni, tb);
    HIPCHK(hipGetLastError());
    ht = HermiteTables{tb, doff, dni};
    HIPCHK(hipStreamSynchronize(st));
  }

  // one plan for all catalogues: work items in proportion to each catalogue's work
  double wtot = 0.0;
  std::vector<double> work(ncat, 0.0);
  for (int c = 0; c < ncat; ++c) {
    for (int l = 0; l < cats[c].lmax; ++l) work[c] += (double)(l + 1) * nl[c][l];
    work[c] *= (double)cats[c].n_points;
    wtot += work[c];
  }
  const int llow = mixed ? std::min(lsplit, LW_LMAX) : 0;
  DirectPlan pllw, pl64, plmx;
  auto ctx = [&](int c) {
    PlanCtx cx;
    cx.nmax = cats[c].nmax;
    cx.norm = norm_factor(&cp[c], (double)cats[c].n_points);
    cx.kbase = kbase[c];
    cx.fbase = fbase[c];
    cx.pbase = pbase[c];
    cx.cat = c;
    cx.kln = kcat.data() + kbase[c];
    cx.R = std::max(cats[c].rmax, rdata[c]);
    return cx;
  };
  const double ttot = (double)mixed_target_items(ntot);
  auto target = [&](int c) { return std::max<int64_t>(16, (int64_t)(ttot * work[c] / std::max(wtot, 1.0))); };
  for (int c = 0; c < ncat; ++c)
    if (llow > 0)
      plan_direct(nl[c], 0, std::min(llow, cats[c].lmax), cats[c].n_points, LW_ROWS, LW_TILE, F64_MC, 1, target(c),
                  ctx(c), &pllw);
  pl64.partial_size = pllw.partial_size;
  for (int c = 0; c < ncat; ++c)
    plan_direct(nl[c], std::min(llow, cats[c].lmax), std::min(lsplit, cats[c].lmax), cats[c].n_points, F64_NT,
                F64_TP, F64_MC, 16, target(c), ctx(c), &pl64);
  plmx.partial_size = pl64.partial_size;
  if (mixed) {
    for (int c = 0; c < ncat; ++c) {
      if (m2)
        plan_mixed2(nl[c], std::min(lsplit, cats[c].lmax), cats[c].lmax, cats[c].n_points, mixed2_qcap(p), target(c),
                    ctx(c), &plmx);
      else
        plan_direct(nl[c], std::min(lsplit, cats[c].lmax), cats[c].lmax, cats[c].n_points, MX_NT, MX_TP, MX_MC,
                    mixed_qcap(p), target(c), ctx(c), &plmx, mixed_ks_cap(p->mixed_operands), nullptr, true);
    }
  }
  std::vector<ReduceGroup> groups = pllw.groups;
  groups.insert(groups.end(), pl64.groups.begin(), pl64.groups.end());
  groups.insert(groups.end(), plmx.groups.begin(), plmx.groups.end());
  double *partial = nullptr;
  CHK(get_buf(h, "partial", (size_t)std::max<int64_t>(plmx.partial_size, 1), &partial));
  std::vector<DirectItem> all;
  for (auto &v : pllw.items) all.insert(all.end(), v.begin(), v.end());
  for (auto &v : pl64.items) all.insert(all.end(), v.begin(), v.end());
  for (auto &v : plmx.items) all.insert(all.end(), v.begin(), v.end());
  DirectItem *ditems;
  CHK(get_buf(h, "items", std::max<size_t>(all.size(), 1), &ditems));
  if (!all.empty())
    HIPCHK(hipMemcpyAsync(ditems, all.data(), sizeof(DirectItem) * all.size(), hipMemcpyHostToDevice, st));

  clear_timings(h);
  HIPCHK(hipEventRecord(h->ev[0], st));
  size_t off = 0;
  CHK(launch_lowl(pllw, ditems, &off, st, pts, dk, tab, partial));
  CHK(launch_plan_f64(h, pl64, ditems, &off, st, pts, nullptr, dk, 0, lt, mixed ? JL_QUINTIC : JL_EXACT, tab, partial,
                      derr));
  HIPCHK(hipEventRecord(h->ev[1], st));
  if (mixed && m2)
    CHK(launch_plan_mixed2(h, plmx, cnint, ditems, &off, st, pts, dk, ls, ct, partial));
  else if (mixed && s16)
    CHK(launch_plan_split16(h, plmx, ditems, &off, st, pts, dk, ls, ct, partial));
  else if (mixed)
    CHK(launch_plan_mixed(h, plmx, hnint, ditems, &off, st, pts, dk, 0, lt, ht, p->mixed_operands, partial));
  HIPCHK(hipEventRecord(h->ev[2], st));

  // f_lmn of all catalogues in one workspace: zero, reduce, then C_l, k_ln and f_lmn per catalogue
  double *f;
  CHK(get_buf(h, "b_flmn", (size_t)std::max<int64_t>(2 * ftot, 1), &f));
  if (ftot) {
    fill_kernel<<<cdiv(2 * ftot, 256), 256, 0, st>>>(f, 2 * ftot, 0.0);
    HIPCHK(hipGetLastError());
  }
  ReduceGroup *dg;
  CHK(get_buf(h, "groups", std::max<size_t>(groups.size(), 1), &dg));
  if (!groups.empty()) {
    HIPCHK(hipMemcpyAsync(dg, groups.data(), sizeof(ReduceGroup) * groups.size(), hipMemcpyHostToDevice, st));
    int chunks = 1;
    for (const ReduceGroup &g : groups)
      chunks = std::max<int>(chunks, (int)cdiv((int64_t)g.nrows * 2 * (g.l + 1), 256));
    reduce_partials_kernel<<<dim3((unsigned)groups.size(), (unsigned)std::min(chunks, 64)), 256, 0, st>>>(dg, partial,
                                                                                                      dk, f);
    HIPCHK(hipGetLastError());
  }
  for (int c = 0; c < ncat; ++c) {
    const int Lc = cats[c].lmax, Nc = cats[c].nmax;
    if (Lc == 0) continue;
    cln_kernel<<<cdiv((int64_t)Lc * Nc, 256), 256, 0, st>>>(Lc, Nc, dnl + nlbase[c], dfoff + foffbase[c],
                                                           f + 2 * fbase[c], cats[c].d_c_ln);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cats[c].d_k_ln, dk + kbase[c], sizeof(double) * Lc * Nc, hipMemcpyDeviceToDevice, st));
    if (cats[c].d_f_lmn)
      HIPCHK(hipMemcpyAsync(cats[c].d_f_lmn, f + 2 * fbase[c], sizeof(double) * 2 * sfbhip_flmn_count(Lc, Nc),
                            hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipEventRecord(h->ev[3], st));
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  CHK(read_timings(h, mixed && lsplit < L ? 1 : 0));
  if (e & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  if (e & ERR_DOMAIN) return fail(SFBHIP_DOMAIN_ERROR, "argument (z) outside of defined range.");
  return SFBHIP_OK;
}

sfbhip_status sfbhip_pixelize(sfbhip_handle *h, const double *d_points, int64_t n_points, int32_t nside,
                              int64_t *d_offsets, double *d_radii, void *stream) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  if (nside < 1 || (nside & (nside - 1)) || nside > 8192)
    return fail(SFBHIP_INVALID_ARGUMENT, "argument (nside) must be a power of 2");  // PixelizedPointContainer.cpp:13-15
  if (n_points < 0 || (n_points > 0 && (!d_points || !d_radii)) || !d_offsets)
    return fail(SFBHIP_INVALID_ARGUMENT, "NULL buffer");
  if (n_points > 0x7fffffffLL) return fail(SFBHIP_NOT_SUPPORTED, "more than 2^31 points");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  const int64_t npix = 12LL * nside * nside;
  int *derr;
  CHK(zero_err(h, st, &derr));
  const int64_t n = std::max<int64_t>(n_points, 1);
  uint32_t *k0, *k1;
  double *v0;
  CHK(get_buf(h, "pix_k0", (size_t)n, &k0));
  CHK(get_buf(h, "pix_k1", (size_t)n, &k1));
  CHK(get_buf(h, "pix_v0", (size_t)n, &v0));
  if (n_points > 0) {
    ang2pix_kernel<<<cdiv(n_points, 256), 256, 0, st>>>(d_points, n_points, nside, k0, v0, derr);
    HIPCHK(hipGetLastError());
    int end_bit = 1;
    while ((1LL << end_bit) < npix) ++end_bit;
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, v0, d_radii, (int)n_points, 0, end_bit, st));
    char *tmp;
    CHK(get_buf(h, "pix_tmp", tmp_bytes, &tmp));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, v0, d_radii, (int)n_points, 0, end_bit, st));
  }
  csr_offsets_kernel<<<cdiv(npix + 1, 256), 256, 0, st>>>(k1, n_points, npix, d_offsets);
  HIPCHK(hipGetLastError());
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  if (e & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  return SFBHIP_OK;
}

sfbhip_status sfbhip_ang2pix_ring(int32_t nside, double theta, double phi, int64_t *pix) {
  g_last_error.clear();
  if (!pix) return fail(SFBHIP_INVALID_ARGUMENT, "sfbhip_ang2pix_ring: pix is NULL");
  if (nside <= 0 || nside > 8192 || (nside & (nside - 1)))
    return fail(SFBHIP_INVALID_ARGUMENT, "argument (nside) must be a power of 2");
  if (!std::isfinite(theta) || !std::isfinite(phi))
    return fail(SFBHIP_INVALID_ARGUMENT, "sfbhip_ang2pix_ring: non-finite angle");
  *pix = hp_ang2pix(nside, theta, phi);
  return SFBHIP_OK;
}

sfbhip_status sfbhip_healpix(sfbhip_handle *h, const int64_t *d_offsets, const double *d_radii, int32_t nside,
                             const sfbhip_params *p, double *d_k_ln, double *d_c_ln, double *d_f_lmn,
                             void *stream) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  clear_timings(h);  // a call refused before its contraction reports zero phase times
  CHK(validate(p));
  if (nside < 1 || (nside & (nside - 1)) || nside > 8192)
    return fail(SFBHIP_INVALID_ARGUMENT, "argument (nside) must be a power of 2");
  if (!d_offsets || !d_k_ln || !d_c_ln) return fail(SFBHIP_INVALID_ARGUMENT, "NULL buffer");
  if (p->precision == SFBHIP_PREC_MIXED && p->mixed_operands != SFBHIP_OPERANDS_SPLIT_FP16)
    return fail(SFBHIP_NOT_SUPPORTED, "the ring path (precision MIXED) uses split-FP16 operands");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  const int64_t npix = 12LL * nside * nside;
  int *derr;
  CHK(zero_err(h, st, &derr));
  // total number of points (the normalisation's N, sfb_reverse_fft.cpp:49-50)
  int64_t npts = 0;
  HIPCHK(hipMemcpyAsync(h->h_err, d_offsets + npix, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(&npts, h->h_err, sizeof(int64_t));
  if (npts <= 0 || !d_radii)
    return fail(SFBHIP_INVALID_ARGUMENT, "empty catalogue: the normalisation divides by the number of points");
  if (npts > 0x7fffffffLL) return fail(SFBHIP_NOT_SUPPORTED, "more than 2^31 points");
  // the CSR and its radii are checked before any kernel reads through them
  csr_check_kernel<<<(unsigned)std::min<int64_t>(cdiv(std::max(npix, npts), 256), 8192), 256, 0, st>>>(
      d_offsets, npix, d_radii, npts, derr);
  HIPCHK(hipGetLastError());
  int maxcount = 0;
  double rdata = 0.0;
  {
    HIPCHK(hipMemcpyAsync(h->h_err, derr, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int e0 = h->h_err[0];
    if (e0 & ERR_BADCSR)
      return fail(SFBHIP_INVALID_ARGUMENT, "pixel offsets are not a CSR of the radii (off[0] = 0, non-decreasing)");
    if (e0 & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
    maxcount = h->h_err[1];
    std::memcpy(&rdata, h->h_err + 2, sizeof(double));
    fill_int_kernel<<<1, 64, 0, st>>>(derr, 4, 0);
    HIPCHK(hipGetLastError());
  }
  if (p->verbose) {
    std::printf("SFB decomposition (reverse method): lmax=%d nmax=%d NSide=%d\n", p->lmax, p->nmax, nside);
    std::printf("Catalog consists of %lld points.\n", (long long)npts);
  }
  CHK(ensure_kln(h, p, st));
  LegendreTables lt;
  CHK(ensure_legendre(h, std::max(p->lmax, 1), st, &lt));
  // pixel tables
  double *pb;
  int32_t *pnp;
  CHK(get_buf(h, "pix_tab", (size_t)8 * npix, &pb));
  CHK(get_buf(h, "pix_np", (size_t)npix, &pnp));
  if (h->pix_nside != nside) {
    pixel_setup_kernel<<<cdiv(npix, 256), 256, 0, st>>>(nside, pb, pb + npix, pb + 2 * npix, pb + 3 * npix,
                                                          pb + 4 * npix, pb + 5 * npix, pb + 6 * npix, pb + 7 * npix,
                                                          pnp);
    HIPCHK(hipGetLastError());
    h->pix_nside = nside;
  }
  PixelsSoA pix;
  pix.x = pb;
  pix.s = pb + npix;
  pix.phi = pb + 2 * npix;
  pix.cphi = pb + 3 * npix;
  pix.sphi = pb + 4 * npix;
  pix.phi2 = pb + 5 * npix;
  pix.cphi2 = pb + 6 * npix;
  pix.sphi2 = pb + 7 * npix;
  pix.np = pnp;
  pix.off = d_offsets;
  pix.radii = d_radii;
  pix.n = npix;
  {
    // contraction order of the pixels: by point count (stable sort), so that
    // the lanes of a wave, one pixel each, loop over about equally many points
    uint32_t *ck0, *ck1;
    int32_t *pv0, *pv1;
    CHK(get_buf(h, "pix_ck0", (size_t)npix, &ck0));
    CHK(get_buf(h, "pix_ck1", (size_t)npix, &ck1));
    CHK(get_buf(h, "pix_pv0", (size_t)npix, &pv0));
    CHK(get_buf(h, "pix_pv1", (size_t)npix, &pv1));
    pixel_count_keys_kernel<<<cdiv(npix, 256), 256, 0, st>>>(d_offsets, npix, ck0, pv0);
    HIPCHK(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && (1LL << end_bit) <= npts) ++end_bit;
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ck0, ck1, pv0, pv1, (int)npix, 0, end_bit, st));
    char *tmp;
    CHK(get_buf(h, "pix_ptmp", std::max<size_t>(tmp_bytes, 1), &tmp));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, ck0, ck1, pv0, pv1, (int)npix, 0, end_bit, st));
    pix.perm = pv1;
  }
  pix.alias_compat = p->alias_compat;
  const double norm = norm_factor(p, (double)npts);
  double *dk;
  CHK(get_buf(h, "kln", (size_t)std::max(p->lmax, 1) * p->nmax, &dk));
  if (p->precision == SFBHIP_PREC_MIXED)
    return healpix_ring(h, d_offsets, d_radii, nside, npts, maxcount, rdata, p, pix, lt, norm, dk, d_k_ln, d_c_ln,
                        d_f_lmn, derr, st);
  const bool use_table = p->bessel == SFBHIP_BESSEL_REFERENCE || (p->bessel == SFBHIP_BESSEL_AUTO && npts > 1000);
  BesselTableRef tab{nullptr, nullptr, nullptr, 0};
  if (use_table) CHK(ensure_table(h, p, st, &tab));
  DirectPlan pl;
  PlanCtx cx;
  cx.nmax = p->nmax;
  cx.norm = norm;
  plan_direct(h->kc.nl, 0, p->lmax, npix, F64_NT, F64_TP, F64_MC, 16, f64_target_items(npix, true), cx, &pl);
  double *partial;
  CHK(get_buf(h, "partial", (size_t)std::max<int64_t>(pl.partial_size, 1), &partial));
  std::vector<DirectItem> all;
  for (auto &v : pl.items) all.insert(all.end(), v.begin(), v.end());
  DirectItem *ditems;
  CHK(get_buf(h, "items", std::max<size_t>(all.size(), 1), &ditems));
  if (!all.empty())
    HIPCHK(hipMemcpyAsync(ditems, all.data(), sizeof(DirectItem) * all.size(), hipMemcpyHostToDevice, st));
  clear_timings(h);
  HIPCHK(hipEventRecord(h->ev[0], st));
  size_t off = 0;
  PointsSoA none{};
  CHK(launch_plan_f64(h, pl, ditems, &off, st, none, &pix, dk, p->nmax, lt, use_table, tab, partial, derr));
  HIPCHK(hipEventRecord(h->ev[1], st));
  HIPCHK(hipEventRecord(h->ev[2], st));
  CHK(finalize(h, p, pl.groups, partial, norm, d_k_ln, d_c_ln, d_f_lmn, st));
  HIPCHK(hipEventRecord(h->ev[3], st));
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  CHK(read_timings(h, 0));
  if (e & ERR_DOMAIN) return fail(SFBHIP_DOMAIN_ERROR, "argument (z) outside of defined range.");
  return SFBHIP_OK;
}

sfbhip_status sfbhip_bessel_zeros(int32_t l, int32_t n, int32_t mode, double *z_host) {
  g_last_error.clear();
  if (l < 0 || n < 0 || (n > 0 && !z_host)) return fail(SFBHIP_INVALID_ARGUMENT, "bessel_zeros: l, n >= 0 required");
  if (mode != SFBHIP_ZEROS_REFERENCE && mode != SFBHIP_ZEROS_CONVERGED)
    return fail(SFBHIP_INVALID_ARGUMENT, "bessel_zeros: unknown mode");
  bessel_zeros(l, n, mode, z_host);
  return SFBHIP_OK;
}

}  // extern "C"

namespace {

// 16-point Gauss-Legendre nodes and weights on [-1, 1]: Newton on P_16 from
// the estimate cos(pi (i + 3/4) / (n + 1/2)) (textbook), weights
// 2 / ((1 - x^2) P_16'(x)^2).
void gauss_legendre16(double *x, double *w) {
  const int n = IG_NODES;
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double d = p1 / dp;
      t -= d;
      if (std::fabs(d) < 1e-16) break;
    }
    x[i] = t;
    w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

struct IntegralCall {
  IntegralPlan *plan = nullptr;
  const double *zeros = nullptr;
  const double *gl = nullptr;
  int nz = 0;
};

// Validates the triples, sizes and computes the zero rows (cached per
// handle), uploads the triples and runs the plan kernel.
sfbhip_status prep_integrals(sfbhip_handle *h, int64_t n, const int32_t *l, const double *k1, const double *k2,
                             double rmax, hipStream_t st, IntegralCall *c) {
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  if (n <= 0 || !l || !k1 || !k2) return fail(SFBHIP_INVALID_ARGUMENT, "sbessel integrals: no triples");
  if (!(rmax > 0.0) || !std::isfinite(rmax)) return fail(SFBHIP_INVALID_ARGUMENT, "r_max must be positive");
  int L = 0;
  double X = 0.0;
  for (int64_t t = 0; t < n; ++t) {
    if (l[t] < 0 || l[t] > 100000) return fail(SFBHIP_INVALID_ARGUMENT, "multipole l out of range");
    if (!(k1[t] >= 0.0) || !(k2[t] >= 0.0) || !std::isfinite(k1[t]) || !std::isfinite(k2[t]))
      return fail(SFBHIP_INVALID_ARGUMENT, "k1, k2 must be finite and >= 0");
    L = std::max(L, l[t] + 1);
    X = std::max(X, std::max(k1[t], k2[t]) * rmax);
  }
  // zeros of j_l are more than pi apart, so fewer than X / pi + 1 lie below X
  const double nzd = std::floor(X / M_PI) + 3.0;
  if (nzd * L > (double)(1 << 27))
    return fail(SFBHIP_NOT_SUPPORTED, "sbessel integrals: k r_max too large for the zero table");
  const int nz = (int)nzd;
  HIPCHK(hipSetDevice(h->device));
  double *dz, *dgl;
  CHK(get_buf(h, "ig_zeros", (size_t)L * nz, &dz));
  CHK(get_buf(h, "ig_gl", 2 * IG_NODES, &dgl));
  static thread_local std::vector<double> glh;
  if (glh.empty()) {
    glh.resize(2 * IG_NODES);
    gauss_legendre16(glh.data(), glh.data() + IG_NODES);
  }
  HIPCHK(hipMemcpyAsync(dgl, glh.data(), sizeof(double) * 2 * IG_NODES, hipMemcpyHostToDevice, st));
  if (h->ig_L != L || h->ig_nz != nz) {
    bessel_zeros_kernel<<<(unsigned)cdiv(L, 64), 64, 0, st>>>(L, nz, SFBHIP_ZEROS_REFERENCE, dz);
    HIPCHK(hipGetLastError());
    h->ig_L = L;
    h->ig_nz = nz;
  }
  int32_t *dl;
  double *dk1, *dk2;
  IntegralPlan *dplan;
  CHK(get_buf(h, "ig_l", (size_t)n, &dl));
  CHK(get_buf(h, "ig_k1", (size_t)n, &dk1));
  CHK(get_buf(h, "ig_k2", (size_t)n, &dk2));
  CHK(get_buf(h, "ig_plan", (size_t)n, &dplan));
  HIPCHK(hipMemcpyAsync(dl, l, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dk1, k1, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dk2, k2, sizeof(double) * n, hipMemcpyHostToDevice, st));
  int *derr;
  CHK(zero_err(h, st, &derr));
  sbessel_plan_kernel<<<(unsigned)cdiv(n, 256), 256, 0, st>>>(n, dl, dk1, dk2, rmax, dz, nz, dplan, derr);
  HIPCHK(hipGetLastError());
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  if (e) return fail(SFBHIP_HIP_ERROR, "sbessel integrals: zero table too short");
  c->plan = dplan;
  c->zeros = dz;
  c->gl = dgl;
  c->nz = nz;
  return SFBHIP_OK;
}

// Node offsets of the triples (host prefix sum of 16 x pieces), uploaded.
sfbhip_status integral_offsets(sfbhip_handle *h, int64_t n, const IntegralCall &c, hipStream_t st,
                               int64_t **d_off, int64_t *total) {
  std::vector<IntegralPlan> hp((size_t)n);
  HIPCHK(hipMemcpyAsync(hp.data(), c.plan, sizeof(IntegralPlan) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int64_t> off((size_t)n);
  int64_t s = 0;
  for (int64_t t = 0; t < n; ++t) {
    off[t] = s;
    s += IG_NODES * hp[t].npieces;
  }
  CHK(get_buf(h, "ig_off", (size_t)n, d_off));
  HIPCHK(hipMemcpyAsync(*d_off, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host vector goes out of scope
  *total = s;
  return SFBHIP_OK;
}

}  // namespace

extern "C" {

sfbhip_status sfbhip_sbessel_nodes(sfbhip_handle *h, int64_t n, const int32_t *l, const double *k1,
                                   const double *k2, double rmax, int64_t *n_nodes, double *r_nodes) {
  g_last_error.clear();
  if (!n_nodes) return fail(SFBHIP_INVALID_ARGUMENT, "n_nodes is NULL");
  hipStream_t st = nullptr;
  IntegralCall c;
  CHK(prep_integrals(h, n, l, k1, k2, rmax, st, &c));
  int64_t *doff, total = 0;
  CHK(integral_offsets(h, n, c, st, &doff, &total));
  *n_nodes = total;
  if (!r_nodes) return SFBHIP_OK;
  double *dr;
  CHK(get_buf(h, "ig_nodes", (size_t)std::max<int64_t>(total, 1), &dr));
  sbessel_nodes_kernel<<<(unsigned)cdiv(n, 4), 256, 0, st>>>(n, c.plan, c.zeros, c.nz, rmax, c.gl, doff, dr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(r_nodes, dr, sizeof(double) * total, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SFBHIP_OK;
}

sfbhip_status sfbhip_sbessel_integrals(sfbhip_handle *h, int64_t n, const int32_t *l, const double *k1,
                                       const double *k2, double rmax, const sfbhip_weight *w, double *out) {
  g_last_error.clear();
  if (!w || !out) return fail(SFBHIP_INVALID_ARGUMENT, "weight or output is NULL");
  if (w->kind != SFBHIP_WEIGHT_NODES && w->kind != SFBHIP_WEIGHT_TABLE)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown weight kind");
  if (w->n > 0 && !w->values) return fail(SFBHIP_INVALID_ARGUMENT, "weight values are NULL");
  if (w->r_power < 0 || w->r_power > 8) return fail(SFBHIP_INVALID_ARGUMENT, "weight r_power must be in 0..8");
  if (w->kind == SFBHIP_WEIGHT_TABLE) {
    if (w->n < 2 || w->n > INT32_MAX) return fail(SFBHIP_INVALID_ARGUMENT, "weight table needs >= 2 samples");
    if (!(w->x_max > w->x_min) || !std::isfinite(w->x_min) || !std::isfinite(w->x_max))
      return fail(SFBHIP_INVALID_ARGUMENT, "weight table range must satisfy x_min < x_max");
  }
  hipStream_t st = nullptr;
  IntegralCall c;
  CHK(prep_integrals(h, n, l, k1, k2, rmax, st, &c));
  WeightRef W{};
  W.kind = w->kind == SFBHIP_WEIGHT_NODES ? W_NODES : W_TABLE;
  W.r_power = w->r_power;
  double *dv;
  if (w->kind == SFBHIP_WEIGHT_NODES) {
    int64_t *doff, total = 0;
    CHK(integral_offsets(h, n, c, st, &doff, &total));
    if (w->n != total)
      return fail(SFBHIP_INVALID_ARGUMENT, "weight: " + std::to_string(w->n) + " node values given, the triples have " +
                                               std::to_string(total) + " nodes");
    W.node_off = doff;
  } else {
    W.ntab = (int32_t)w->n;
    W.x_min = w->x_min;
    W.x_max = w->x_max;
    W.dx_inv = (double)(w->n - 1) / (w->x_max - w->x_min);
  }
  CHK(get_buf(h, "ig_w", (size_t)std::max<int64_t>(w->n, 1), &dv));
  HIPCHK(hipMemcpyAsync(dv, w->values, sizeof(double) * w->n, hipMemcpyHostToDevice, st));
  W.values = dv;
  double *dout;
  CHK(get_buf(h, "ig_out", (size_t)n, &dout));
  int *derr;
  CHK(zero_err(h, st, &derr));
  clear_timings(h);
  // the triples in order of decreasing piece count (stable radix sort)
  int32_t *order = nullptr;
  if (n > 1 && n <= INT32_MAX) {
    uint32_t *k0, *k1;
    int32_t *i0;
    CHK(get_buf(h, "ig_ok0", (size_t)n, &k0));
    CHK(get_buf(h, "ig_ok1", (size_t)n, &k1));
    CHK(get_buf(h, "ig_oi0", (size_t)n, &i0));
    CHK(get_buf(h, "ig_oi1", (size_t)n, &order));
    sbessel_order_keys_kernel<<<(unsigned)cdiv(n, 256), 256, 0, st>>>(n, c.plan, k0, i0);
    HIPCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, i0, order, (int)n, 0, 32, st));
    char *tmp;
    CHK(get_buf(h, "ig_otmp", std::max<size_t>(tmp_bytes, 1), &tmp));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i0, order, (int)n, 0, 32, st));
  }
  HIPCHK(hipEventRecord(h->ev[0], st));
  sbessel_integral_kernel<<<(unsigned)cdiv(n, 4), 256, 0, st>>>(n, c.plan, c.zeros, c.nz, rmax, c.gl, W, order, dout,
                                                               derr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev[1], st));
  HIPCHK(hipEventRecord(h->ev[2], st));
  HIPCHK(hipEventRecord(h->ev[3], st));
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  int e = 0;
  CHK(read_err(h, st, derr, &e));
  CHK(read_timings(h, 0));
  if (e & ERR_DOMAIN)
    return fail(SFBHIP_DOMAIN_ERROR, "argument outside range for which this FunctionInterpolator is defined");
  return SFBHIP_OK;
}

}  // extern "C"

extern "C" {

sfbhip_status sfbhip_filter_points(sfbhip_handle *h, const void *d_records, int64_t n, int32_t record_kind,
                                   double shift, const sfbhip_filter *filters, int32_t nfilters, uint64_t seed,
                                   uint64_t index_base, double *d_points_out, int64_t *n_out, void *stream) {
  g_last_error.clear();
  if (!h || !n_out) return fail(SFBHIP_INVALID_ARGUMENT, "handle or n_out is NULL");
  if (n < 0) return fail(SFBHIP_INVALID_ARGUMENT, "negative record count");
  if (n > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "filter: more than 2^31 - 1 records per call");
  if (record_kind < SFBHIP_RECORD_POINT || record_kind > SFBHIP_RECORD_SPHERICAL_F64)
    return fail(SFBHIP_INVALID_ARGUMENT, "unknown record kind");
  if (nfilters < 0 || nfilters > IN_MAX_FILTERS || (nfilters > 0 && !filters))
    return fail(SFBHIP_INVALID_ARGUMENT, "between 0 and " + std::to_string(IN_MAX_FILTERS) + " filters");
  *n_out = 0;
  if (n == 0) return SFBHIP_OK;
  if (!d_records || !d_points_out) return fail(SFBHIP_INVALID_ARGUMENT, "records or output is NULL");
  FilterSet fs{};
  fs.count = nfilters;
  for (int k = 0; k < nfilters; ++k) {
    const sfbhip_filter &F = filters[k];
    FilterDesc &D = fs.f[k];
    D.kind = F.kind;
    D.nside = F.nside;
    D.a = F.a;
    D.x_min = F.x_min;
    D.x_max = F.x_max;
    D.n = F.n;
    D.values = F.d_values;
    switch (F.kind) {
      case SFBHIP_FILTER_TOPHAT:
        break;
      case SFBHIP_FILTER_GAUSSIAN:
        if (!(F.a > 0.0)) return fail(SFBHIP_INVALID_ARGUMENT, "Gaussian filter: R0 must be positive");
        break;
      case SFBHIP_FILTER_WINDOW_TABLE:
        if (F.n < 2 || F.n > INT32_MAX || !F.d_values || !(F.x_max > F.x_min))
          return fail(SFBHIP_INVALID_ARGUMENT, "window table: >= 2 samples on x_min < x_max required");
        D.dx_inv = (double)(F.n - 1) / (F.x_max - F.x_min);
        break;
      case SFBHIP_FILTER_WINDOW_VALUES:
        if (F.n != n || !F.d_values)
          return fail(SFBHIP_INVALID_ARGUMENT, "window values: one probability per record required");
        break;
      case SFBHIP_FILTER_ANGULAR_MASK:
        if (F.nside < 1 || (F.nside & (F.nside - 1)) || F.n != 12 * (int64_t)F.nside * F.nside || !F.d_values)
          return fail(SFBHIP_INVALID_ARGUMENT, "healpix_map is not a valid map (check size!)");
        break;
      default:
        return fail(SFBHIP_INVALID_ARGUMENT, "unknown filter kind");
    }
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  int32_t *flag;
  int64_t *pos;
  CHK(get_buf(h, "in_flag", (size_t)n, &flag));
  CHK(get_buf(h, "in_pos", (size_t)n, &pos));
  int *derr;
  CHK(zero_err(h, st, &derr));
  const unsigned nb = (unsigned)cdiv(n, 256);
  ingest_flag_kernel<<<nb, 256, 0, st>>>(d_records, n, record_kind, shift, fs, seed, index_base, flag, derr);
  HIPCHK(hipGetLastError());
  size_t tmp_bytes = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, flag, pos, (int)n, st));
  unsigned char *tmp;
  CHK(get_buf(h, "in_scan_tmp", std::max<size_t>(tmp_bytes, 1), &tmp));
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flag, pos, (int)n, st));
  compact_convert_kernel<<<nb, 256, 0, st>>>(d_records, flag, pos, n, record_kind, shift, d_points_out);
  HIPCHK(hipGetLastError());
  int64_t last_pos = 0;
  int32_t last_flag = 0;
  HIPCHK(hipMemcpyAsync(&last_pos, pos + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&last_flag, flag + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  int e = 0;
  CHK(read_err(h, st, derr, &e));  // synchronises the stream
  if (e & ERR_DOMAIN)
    return fail(SFBHIP_DOMAIN_ERROR, "argument outside range for which this FunctionInterpolator is defined");
  *n_out = last_pos + last_flag;  // exclusive scan: the count is the last position plus the last flag
  return SFBHIP_OK;
}

sfbhip_status sfbhip_random_subset(sfbhip_handle *h, const double *d_points, int64_t n, int64_t k, uint64_t seed,
                                   double *d_out, void *stream) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  if (n < 0 || k < 0) return fail(SFBHIP_INVALID_ARGUMENT, "negative size");
  if (n > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "random subset: more than 2^31 - 1 points");
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(h->device));
  if (k > n) k = n;
  if (k == 0) return SFBHIP_OK;
  if (!d_points || !d_out) return fail(SFBHIP_INVALID_ARGUMENT, "points or output is NULL");
  if (k == n) {
    HIPCHK(hipMemcpyAsync(d_out, d_points, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return SFBHIP_OK;
  }
  uint64_t *k0, *k1;
  int64_t *i0, *i1;
  CHK(get_buf(h, "ss_k0", (size_t)n, &k0));
  CHK(get_buf(h, "ss_k1", (size_t)n, &k1));
  CHK(get_buf(h, "ss_i0", (size_t)n, &i0));
  CHK(get_buf(h, "ss_i1", (size_t)n, &i1));
  const unsigned nb = (unsigned)cdiv(n, 256);
  subset_keys_kernel<<<nb, 256, 0, st>>>(n, seed, k0, i0);
  HIPCHK(hipGetLastError());
  // the k smallest keys are the subset; their indices sorted back into input order
  size_t b1 = 0, b2 = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, k0, k1, i0, i1, (int)n, 0, 64, st));
  HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, b2, i1, i0, (int)k, 0, 64, st));
  unsigned char *tmp;
  CHK(get_buf(h, "ss_tmp", std::max<size_t>(std::max(b1, b2), 1), &tmp));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, b1, k0, k1, i0, i1, (int)n, 0, 64, st));
  HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp, b2, i1, i0, (int)k, 0, 64, st));
  gather_aos_kernel<<<(unsigned)cdiv(k, 256), 256, 0, st>>>(d_points, i0, k, d_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return SFBHIP_OK;
}

}  // extern "C"

// ---- Batch plans (catana_hip.h): every host step of a batched call done
// once at creation; an enqueue launches only the library's kernels on one
// stream, so it can be captured in a hipGraph.

struct sfbhip_batch_plan {
  int device = 0;
  int ncat = 0;
  std::vector<void *> mem;         // device allocations owned by the plan
  hipStream_t own = nullptr;       // NULL stream argument: enqueue / launch here (synchronises with stream 0)
  hipStream_t cap = nullptr;       // NULL stream argument: capture here
  // the grouped launches of point splits 2 and 4 run on these (default
  // priority), forked from and joined back to the enqueue's stream by events
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int *h_err = nullptr;            // pinned copy of the error flags
  int timed = 0;                   // last run: 0 none, 1 eager enqueue (phases), 2 graph launch (total)
  // device data
  int *derr = nullptr;             // [0] error bits, [8 + 8 k + x] counter of lane x of grouped launch k
  double *soa = nullptr;           // point table in radius order (what the kernels read)
  double *soa_in = nullptr;        // point table in input order
  int64_t ntot = 0;
  uint16_t *keys[2] = {nullptr, nullptr};  // radius buckets, input order / sort ping-pong
  int32_t *idx[2] = {nullptr, nullptr};
  uint32_t *hist = nullptr;
  SortSeg *dseg = nullptr;
  int sort_nb = 1;                 // blocks of the largest catalogue's sort
  CatPrep *dprep = nullptr;
  int prep_bx = 1;
  double *dk = nullptr;
  int32_t *dnl = nullptr;
  int64_t *dfoff = nullptr;
  LegendreTables lt{};
  LegendreScaled ls{};
  CubicTables ct{};
  BesselTableRef q5{};
  DirectItem *ditems = nullptr;
  int nlow = 0;
  int n64[5] = {0, 0, 0, 0, 0};
  int nmx = 0;
  int nmx_ks[3] = {0, 0, 0};   // grouped items with point split 1, 2, 4 (in this order after the FP64 items)
  GroupedLanes lanes_ks[3] = {};  // their XCD lanes
  int grid_ks[3] = {1, 1, 1};
  ReduceGroup *dgroups = nullptr;
  int ngroups = 0, red_chunks = 1;
  double *partial = nullptr;
  double *f = nullptr;
  int64_t ftot = 0;
  CatOut *dout = nullptr;
  int out_bx = 1;
};

namespace {

template <class T>
sfbhip_status plan_alloc(sfbhip_batch_plan *pl, size_t count, T **out) {
  void *ptr = nullptr;
  const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  const hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess)
    return fail(SFBHIP_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + " bytes) for a batch plan: " +
                                          hipGetErrorString(e));
  pl->mem.push_back(ptr);
  *out = static_cast<T *>(ptr);
  return SFBHIP_OK;
}

template <class T>
sfbhip_status plan_upload(sfbhip_batch_plan *pl, const std::vector<T> &v, T **out) {
  CHK(plan_alloc(pl, v.size(), out));
  if (!v.empty()) HIPCHK(hipMemcpy(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return SFBHIP_OK;
}

// Items of the grouped split-FP16 kernel for one catalogue: per (l, 64-row
// tile) group, S point ranges of one length, each cut into passes of the
// band's chunk count; an item of <= 16 (32) rows splits its tiles' points over
// 4 (2) waves where the band allows.  The points are ordered by 16-bit radius
// bucket (R / 2^16 wide), so an item's radii lie within one bucket of
// [r_first, r_last]: its Bessel window, which has a node of margin on either
// side, covers them while a bucket spans less than a node (table reach below
// 2^15 intervals); past that the window starts at x = 0 (ITEM_FULL_WINDOW).
void plan_grouped(const std::vector<int32_t> &nl, int lmin, int lmax, int64_t npts, const sfbhip_tile_table &tt,
                  int64_t target_items, const PlanCtx &cx, std::vector<DirectItem> *items,
                  std::vector<ReduceGroup> *groups, int64_t *partial_size) {
  int64_t passes_total = 0;
  for (int l = lmin; l < lmax; ++l)
    passes_total += cdiv(nl[l], MX_NT) * cdiv(l / MX_MC + 1, tt.chunks[tile_band(l)]);
  if (passes_total == 0) return;
  int64_t S = std::max<int64_t>(1, cdiv(target_items, passes_total));
  S = std::min<int64_t>(S, cdiv(npts, MX_TP));
  const int64_t chunk = cdiv(cdiv(npts, S), MX_TP) * MX_TP;
  S = cdiv(npts, chunk);
  for (int l = lmax - 1; l >= lmin; --l) {
    const int b = tile_band(l);
    const int qcap = tt.chunks[b], kscap = tt.point_split[b];
    const int Q = l / MX_MC + 1;
    const int ncol = 2 * (l + 1);
    for (int n0 = 0; n0 < nl[l]; n0 += MX_NT) {
      ReduceGroup g;
      g.l = l;
      g.n0 = n0;
      g.nrows = std::min(MX_NT, nl[l] - n0);
      g.nsplit = (int32_t)S;
      g.base = *partial_size;
      g.stride = (int64_t)MX_NT * ncol;
      g.fofs = cx.fbase + (int64_t)l * (l + 1) / 2 * cx.nmax;
      g.kofs = cx.kbase + (int64_t)l * cx.nmax + n0;
      g.norm = cx.norm;
      g.nmax = cx.nmax;
      g.pad = 0;
      *partial_size += S * g.stride;
      groups->push_back(g);
      int32_t ks = 0;
      if (kscap >= 4 && g.nrows <= 16)
        ks = ITEM_SPLIT_KS4;
      else if (kscap >= 2 && g.nrows <= 32)
        ks = ITEM_SPLIT_KS2;
      const int32_t tab = (int32_t)std::min<double>(
          std::floor(cx.kln[(size_t)l * cx.nmax + n0 + g.nrows - 1] * cx.R / MX_H) + 2, 1e9);
      for (int64_t s = 0; s < S; ++s) {
        for (const auto &qr : chunk_ranges(Q, qcap, true)) {
          DirectItem it;
          it.l = l;
          it.n0 = n0;
          it.nrows = g.nrows;
          it.qlo = qr.first;
          it.qhi = qr.second;
          it.cat = cx.cat;
          it.tab = tab;
          it.flags = ks;
          it.p0 = cx.pbase + s * chunk;
          it.p1 = cx.pbase + std::min(npts, (s + 1) * chunk);
          it.out = g.base + s * g.stride;
          it.kofs = g.kofs;
          items->push_back(it);
        }
      }
    }
  }
}

// Relative cost of a grouped item, for ordering the list longest first: the
// Y* rows of its chunks and the Legendre steps down to its first chunk, per
// point.  (Ordering by kernel variant first, so that the blocks of a compute
// unit share one variant's code, measured 1 % slower on 64 catalogues.)
double item_cost(const DirectItem &it) {
  return (double)(it.p1 - it.p0) * ((double)(it.qhi - it.qlo) + 0.1 * it.qlo + 0.3);
}

sfbhip_status plan_enqueue(sfbhip_batch_plan *pl, hipStream_t st, bool timed) {
  if (timed) HIPCHK(hipEventRecord(pl->ev[0], st));
  fill_int_kernel<<<1, 64, 0, st>>>(pl->derr, 8 + 3 * GROUPED_LANES, 0);
  HIPCHK(hipGetLastError());
  const dim3 gcat((unsigned)pl->prep_bx, (unsigned)pl->ncat);
  prep_points_batch_kernel<<<gcat, 256, 0, st>>>(pl->dprep, pl->soa_in, pl->ntot, pl->keys[0], pl->derr);
  HIPCHK(hipGetLastError());
  // radius order within each catalogue: two 8-bit passes (keys 0 -> 1 -> 0), then the gather
  const dim3 gsort((unsigned)pl->sort_nb, (unsigned)pl->ncat);
  for (int pass = 0; pass < 2; ++pass) {
    sort_hist_kernel<<<gsort, 256, 0, st>>>(pl->dseg, pl->keys[pass], 8 * pass, pl->hist);
    HIPCHK(hipGetLastError());
    sort_scan_kernel<<<(unsigned)pl->ncat, 1024, 0, st>>>(pl->dseg, pl->hist);
    HIPCHK(hipGetLastError());
    sort_scatter_kernel<<<gsort, 256, 0, st>>>(pl->dseg, pl->keys[pass], pass ? pl->idx[0] : nullptr, 8 * pass,
                                               pl->hist, pl->keys[1 - pass], pl->idx[pass]);
    HIPCHK(hipGetLastError());
  }
  sort_gather_kernel<<<gcat, 256, 0, st>>>(pl->dseg, pl->idx[1], pl->soa_in, pl->soa, pl->ntot);
  HIPCHK(hipGetLastError());
  const int64_t n = pl->ntot;
  const PointsSoA pts{pl->soa, pl->soa + n, pl->soa + 2 * n, pl->soa + 3 * n, pl->soa + 4 * n, pl->soa + 5 * n, n};
  if (timed) HIPCHK(hipEventRecord(pl->ev[1], st));
  size_t off = 0;
  if (pl->nlow) {
    sfb_lowl_f64_kernel<<<pl->nlow, 256, 0, st>>>(pts, pl->ditems, pl->dk, pl->q5, pl->partial);
    HIPCHK(hipGetLastError());
    off += pl->nlow;
  }
  for (int ci = 0; ci < 5; ++ci) {
    const int k = pl->n64[ci];
    if (!k) continue;
    const DirectItem *di = pl->ditems + off;
    switch (ci) {
      case 0: launch_f64<1>(k, di, st, pts, nullptr, pl->dk, 0, pl->lt, JL_QUINTIC, pl->q5, pl->partial, pl->derr); break;
      case 1: launch_f64<2>(k, di, st, pts, nullptr, pl->dk, 0, pl->lt, JL_QUINTIC, pl->q5, pl->partial, pl->derr); break;
      case 2: launch_f64<4>(k, di, st, pts, nullptr, pl->dk, 0, pl->lt, JL_QUINTIC, pl->q5, pl->partial, pl->derr); break;
      case 3: launch_f64<8>(k, di, st, pts, nullptr, pl->dk, 0, pl->lt, JL_QUINTIC, pl->q5, pl->partial, pl->derr); break;
      default: launch_f64<16>(k, di, st, pts, nullptr, pl->dk, 0, pl->lt, JL_QUINTIC, pl->q5, pl->partial, pl->derr); break;
    }
    HIPCHK(hipGetLastError());
    off += k;
  }
  if (timed) HIPCHK(hipEventRecord(pl->ev[2], st));
  // one persistent launch per point split, each with its own item counter:
  // point split 1 on st, 2 and 4 on the plan's side streams, so that each
  // launch's blocks take over the compute units the others' blocks leave
  // (running the FP64 launches there too, beside the grouped ones, measured
  // no faster: 166.3 ms per 64-catalogue replay either way)
  HIPCHK(hipEventRecord(pl->fork, st));
  for (int k = 0; k < 2; ++k) HIPCHK(hipStreamWaitEvent(pl->side[k], pl->fork, 0));
  for (int k = 0; k < 3; ++k) {
    const int n = pl->nmx_ks[k];
    const DirectItem *di = pl->ditems + off;
    int *ctr = pl->derr + 8 + GROUPED_LANES * k;
    const GroupedLanes &ln = pl->lanes_ks[k];
    const unsigned gx = (unsigned)pl->grid_ks[k];
    const hipStream_t s = k == 0 ? st : pl->side[k - 1];
    if (n && k == 0) sfb_direct_split16_grouped_kernel<1><<<gx, 256, 0, s>>>(pts, di, ln, ctr, pl->dk, pl->ls, pl->ct, pl->partial);
    if (n && k == 1) sfb_direct_split16_grouped_kernel<2><<<gx, 256, 0, s>>>(pts, di, ln, ctr, pl->dk, pl->ls, pl->ct, pl->partial);
    if (n && k == 2) sfb_direct_split16_grouped_kernel<4><<<gx, 256, 0, s>>>(pts, di, ln, ctr, pl->dk, pl->ls, pl->ct, pl->partial);
    HIPCHK(hipGetLastError());
    off += n;
  }
  for (int k = 0; k < 2; ++k) {
    HIPCHK(hipEventRecord(pl->join[k], pl->side[k]));
    HIPCHK(hipStreamWaitEvent(st, pl->join[k], 0));
  }
  if (timed) HIPCHK(hipEventRecord(pl->ev[3], st));
  if (pl->ftot) {
    fill_kernel<<<(unsigned)cdiv(2 * pl->ftot, 256), 256, 0, st>>>(pl->f, 2 * pl->ftot, 0.0);
    HIPCHK(hipGetLastError());
  }
  if (pl->ngroups) {
    reduce_partials_kernel<<<dim3((unsigned)pl->ngroups, (unsigned)pl->red_chunks), 256, 0, st>>>(pl->dgroups, pl->partial,
                                                                                                pl->dk, pl->f);
    HIPCHK(hipGetLastError());
  }
  batch_outputs_kernel<<<dim3((unsigned)pl->out_bx, (unsigned)pl->ncat), 256, 0, st>>>(pl->dout, pl->f, pl->dk, pl->dnl,
                                                                                     pl->dfoff);
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(pl->ev[4], st));
  return SFBHIP_OK;
}

sfbhip_status plan_create(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat, const sfbhip_params *p,
                          const double *r_bound, sfbhip_batch_plan *pl) {
  if (!cats || ncat < 1 || ncat > 128) return fail(SFBHIP_INVALID_ARGUMENT, "ncat must be in [1, 128]");
  if (!p) return fail(SFBHIP_INVALID_ARGUMENT, "params is NULL");
  if (p->precision != SFBHIP_PREC_MIXED || p->mixed_operands != SFBHIP_OPERANDS_SPLIT_FP16)
    return fail(SFBHIP_NOT_SUPPORTED,
                "batch plans run precision MIXED with SPLIT_FP16 operands; use sfbhip_direct_batched otherwise");
  if (p->mixed_kernel == SFBHIP_MIXED_KERNEL_512)
    return fail(SFBHIP_NOT_SUPPORTED, "batch plans run the grouped split-FP16 kernel (256 threads); MIXED_KERNEL_512 is not one");
  std::vector<sfbhip_params> cp(ncat, *p);
  std::vector<double> R(ncat);
  int64_t ntot = 0;
  int L = 0, NZ = 1;
  for (int c = 0; c < ncat; ++c) {
    cp[c].lmax = cats[c].lmax;
    cp[c].nmax = cats[c].nmax;
    cp[c].rmax = cats[c].rmax;
    cp[c].k_max = cats[c].k_max;
    cp[c].window_volume = cats[c].window_volume;
    CHK(validate(&cp[c]));
    if (cats[c].n_points <= 0 || !cats[c].d_points)
      return fail(SFBHIP_INVALID_ARGUMENT, "catalogue " + std::to_string(c) +
                                               ": empty catalogue: the normalisation divides by the number of points");
    if (!cats[c].d_k_ln || !cats[c].d_c_ln)
      return fail(SFBHIP_INVALID_ARGUMENT, "catalogue " + std::to_string(c) + ": k_ln / c_ln output is NULL");
    double rb = cats[c].rmax;
    if (r_bound) {
      if (!(r_bound[c] >= 0.0) || !std::isfinite(r_bound[c]))
        return fail(SFBHIP_INVALID_ARGUMENT, "catalogue " + std::to_string(c) + ": r_bound must be finite and >= 0");
      rb = std::max(rb, r_bound[c]);
    }
    R[c] = rb;
    ntot += cats[c].n_points;
    L = std::max(L, cats[c].lmax);
    NZ = std::max(NZ, cats[c].nmax + 1);
  }
  if (ntot > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "more than 2^31 points in one batch");
  if (L > MX_LMAX)
    return fail(SFBHIP_NOT_SUPPORTED, "mixed path supports lmax <= " + std::to_string(MX_LMAX) + "; use precision FP64");
  pl->ncat = ncat;
  pl->ntot = ntot;

  // zeros of j_l, l < L (the first nmax_c + 1 of them are those of a single call with nmax_c)
  std::vector<double> zeros((size_t)std::max(L, 1) * NZ, 0.0);
  if (L > 0) {
    double *dz;
    CHK(plan_alloc(pl, (size_t)L * NZ, &dz));
    bessel_zeros_kernel<<<cdiv(L, 64), 64, 0, nullptr>>>(L, NZ, p->zeros, dz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(zeros.data(), dz, sizeof(double) * L * NZ, hipMemcpyDeviceToHost));
  }
  // k_ln, n_l, f_lmn offsets per catalogue
  std::vector<double> kcat;
  std::vector<int32_t> nlcat;
  std::vector<int64_t> foffcat, kbase(ncat), nlbase(ncat), fbase(ncat), foffbase(ncat), pbase(ncat);
  std::vector<std::vector<int32_t>> nl(ncat);
  int64_t ftot = 0, o = 0;
  for (int c = 0; c < ncat; ++c) {
    const int Lc = cats[c].lmax, Nc = cats[c].nmax;
    const double kmax = cats[c].k_max > 0 ? cats[c].k_max : 0.0;
    pbase[c] = o;
    o += cats[c].n_points;
    kbase[c] = (int64_t)kcat.size();
    nlbase[c] = (int64_t)nlcat.size();
    foffbase[c] = (int64_t)foffcat.size();
    fbase[c] = ftot;
    nl[c].assign(std::max(Lc, 1), 0);
    for (int l = 0; l < Lc; ++l) {
      int cnt = 0;
      for (int n = 0; n < Nc; ++n) {
        const double z = zeros[(size_t)l * NZ + n];
        kcat.push_back(z / cats[c].rmax);
        if (kmax <= 0.0 || z < kmax * cats[c].rmax) cnt = n + 1;
      }
      nl[c][l] = cnt;
    }
    nlcat.insert(nlcat.end(), nl[c].begin(), nl[c].end());
    const std::vector<int64_t> fo = flmn_offsets(Lc, Nc);
    foffcat.insert(foffcat.end(), fo.begin(), fo.end());
    ftot += (int64_t)sfbhip_flmn_count(Lc, Nc);
  }
  CHK(plan_upload(pl, kcat, &pl->dk));
  CHK(plan_upload(pl, nlcat, &pl->dnl));
  CHK(plan_upload(pl, foffcat, &pl->dfoff));
  pl->ftot = ftot;

  // tables shared by all catalogues
  {
    std::vector<double2> hs, hc;
    std::vector<int32_t> ho;
    legendre_host(std::max(L, 1), hs, hc, ho);
    double2 *ds, *dc;
    int32_t *dof;
    CHK(plan_upload(pl, hs, &ds));
    CHK(plan_upload(pl, hc, &dc));
    CHK(plan_upload(pl, ho, &dof));
    pl->lt = LegendreTables{ds, dc, dof};
  }
  const int lsplit = std::min(L, std::max(0, p->mixed_fp64_below));
  std::vector<double> xmax(std::max(L, 1), 0.0);
  for (int c = 0; c < ncat; ++c)
    for (int l = 0; l < cats[c].lmax; ++l)
      if (nl[c][l] > 0)
        xmax[l] = std::max(xmax[l], kcat[kbase[c] + (int64_t)l * cats[c].nmax + nl[c][l] - 1] * R[c]);
  if (lsplit > 0) {
    std::vector<int32_t> qoff(lsplit), qni(lsplit);
    int64_t qo = 0;
    int32_t mx = 1;
    for (int l = 0; l < lsplit; ++l) {
      qni[l] = (int32_t)std::floor(xmax[l] * Q5_HINV) + 2;
      qoff[l] = (int32_t)qo;
      qo += qni[l];
      mx = std::max(mx, qni[l]);
    }
    double *q5;
    int32_t *doff, *dni;
    CHK(plan_alloc(pl, (size_t)6 * qo, &q5));
    CHK(plan_upload(pl, qoff, &doff));
    CHK(plan_upload(pl, qni, &dni));
    quintic_table_kernel<<<dim3((unsigned)cdiv(mx, 256), (unsigned)lsplit), 256, 0, nullptr>>>(lsplit, doff, dni, q5);
    HIPCHK(hipGetLastError());
    pl->q5.q5 = q5;
    pl->q5.q5_off = doff;
    pl->q5.q5_nint = dni;
  }
  if (lsplit < L) {
    // the split-FP16 kernel's tables, owned by the plan: cubic j_l coefficients
    // over each l's largest argument and the scaled Legendre constants
    std::vector<int32_t> coff(L, 0), cni(L, 1);
    int64_t co = 0;
    int32_t mx = 1;
    for (int l = 0; l < L; ++l) {
      const int64_t k = (int64_t)std::floor(xmax[l] / MX_H) + 2;
      if (k > MX_TAB_MAX)
        return fail(SFBHIP_NOT_SUPPORTED, "mixed path: largest Bessel argument k r exceeds " +
                                              std::to_string(MX_TAB_MAX * MX_H) + "; use precision FP64");
      cni[l] = (int32_t)k;
      coff[l] = (int32_t)co;
      co += k;
      mx = std::max<int32_t>(mx, (int32_t)k);
      if (co > INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "mixed path: Bessel tables too large");
    }
    float4 *tb, *tb16;
    int32_t *doff, *dni;
    CHK(plan_alloc(pl, (size_t)co, &tb));
    CHK(plan_alloc(pl, (size_t)co, &tb16));
    CHK(plan_upload(pl, coff, &doff));
    CHK(plan_upload(pl, cni, &dni));
    cubic_table_kernel<<<dim3((unsigned)cdiv(mx, 256), (unsigned)L), 256, 0, nullptr>>>(L, doff, dni, tb, tb16);
    HIPCHK(hipGetLastError());
    pl->ct = CubicTables{tb, doff, dni, tb16};
    std::vector<double2> hs;
    std::vector<double> hg, hd;
    std::vector<int32_t> ho;
    legendre_scaled_host(L, hs, hg, hd, ho);
    double2 *ds2;
    double *dg, *dd;
    int32_t *dso;
    CHK(plan_upload(pl, hs, &ds2));
    CHK(plan_upload(pl, hg, &dg));
    CHK(plan_upload(pl, hd, &dd));
    CHK(plan_upload(pl, ho, &dso));
    pl->ls = LegendreScaled{ds2, dg, dd, dso};
  }

  // work items: FP64 low l, FP64 l < lsplit beyond the low-l kernel's range, grouped mixed
  double wtot = 0.0;
  std::vector<double> work(ncat, 0.0);
  for (int c = 0; c < ncat; ++c) {
    for (int l = 0; l < cats[c].lmax; ++l) work[c] += (double)(l + 1) * nl[c][l];
    work[c] *= (double)cats[c].n_points;
    wtot += work[c];
  }
  auto ctx = [&](int c) {
    PlanCtx cx;
    cx.nmax = cats[c].nmax;
    cx.norm = norm_factor(&cp[c], (double)cats[c].n_points);
    cx.kbase = kbase[c];
    cx.fbase = fbase[c];
    cx.pbase = pbase[c];
    cx.cat = c;
    cx.kln = kcat.data() + kbase[c];
    cx.R = R[c];
    return cx;
  };
  const sfbhip_tile_table &tt = h->tiles;
  const double ttot_f64 = (double)mixed_target_items(ntot);
  const double ttot_mx = (double)tt.items_per_cu * h->ncu;
  auto share = [&](int c, double t) { return std::max<int64_t>(16, (int64_t)(t * work[c] / std::max(wtot, 1.0))); };
  const int llow = std::min(lsplit, LW_LMAX);
  DirectPlan pllw, pl64;
  for (int c = 0; c < ncat; ++c)
    if (llow > 0)
      plan_direct(nl[c], 0, std::min(llow, cats[c].lmax), cats[c].n_points, LW_ROWS, LW_TILE, F64_MC, 1,
                  share(c, ttot_f64), ctx(c), &pllw);
  pl64.partial_size = pllw.partial_size;
  for (int c = 0; c < ncat; ++c)
    plan_direct(nl[c], std::min(llow, cats[c].lmax), std::min(lsplit, cats[c].lmax), cats[c].n_points, F64_NT, F64_TP,
                F64_MC, 16, share(c, ttot_f64), ctx(c), &pl64);
  std::vector<DirectItem> mx;
  std::vector<ReduceGroup> groups = pllw.groups;
  groups.insert(groups.end(), pl64.groups.begin(), pl64.groups.end());
  int64_t psize = pl64.partial_size;
  for (int c = 0; c < ncat; ++c)
    plan_grouped(nl[c], std::min(lsplit, cats[c].lmax), cats[c].lmax, cats[c].n_points, tt, share(c, ttot_mx), ctx(c),
                 &mx, &groups, &psize);
  std::stable_sort(mx.begin(), mx.end(),
                   [](const DirectItem &a, const DirectItem &b) { return item_cost(a) > item_cost(b); });
  // by point split (1, 2, 4), each part still by decreasing cost
  {
    std::vector<DirectItem> byks;
    const int32_t ksf[3] = {0, ITEM_SPLIT_KS2, ITEM_SPLIT_KS4};
    // ... and each part in XCD lanes: lane x the items of the point ranges of
    // rank x mod 8 (ranks of the ranges' starts over all catalogues)
    std::vector<int64_t> starts;
    for (const DirectItem &it : mx) starts.push_back(it.p0);
    std::sort(starts.begin(), starts.end());
    starts.erase(std::unique(starts.begin(), starts.end()), starts.end());
    for (int k = 0; k < 3; ++k) {
      GroupedLanes &ln = pl->lanes_ks[k];
      const size_t before = byks.size();
      for (int x = 0; x < GROUPED_LANES; ++x) {
        ln.off[x] = (int32_t)(byks.size() - before);
        for (const DirectItem &it : mx) {
          if ((it.flags & (ITEM_SPLIT_KS2 | ITEM_SPLIT_KS4)) != ksf[k]) continue;
          const int64_t r = std::lower_bound(starts.begin(), starts.end(), it.p0) - starts.begin();
          if (r % GROUPED_LANES == x) byks.push_back(it);
        }
      }
      ln.off[GROUPED_LANES] = (int32_t)(byks.size() - before);
      pl->nmx_ks[k] = (int)(byks.size() - before);
    }
    mx.swap(byks);
  }
  std::vector<DirectItem> all;
  for (auto &v : pllw.items) all.insert(all.end(), v.begin(), v.end());
  pl->nlow = (int)all.size();
  for (int ci = 0; ci < 5; ++ci) {
    all.insert(all.end(), pl64.items[ci].begin(), pl64.items[ci].end());
    pl->n64[ci] = (int)pl64.items[ci].size();
  }
  if (!pl64.items[5].empty()) return fail(SFBHIP_NOT_SUPPORTED, "FP64 plan class beyond 16 chunks");
  all.insert(all.end(), mx.begin(), mx.end());
  if (all.size() > (size_t)INT32_MAX) return fail(SFBHIP_NOT_SUPPORTED, "too many work items");
  pl->nmx = (int)mx.size();
  CHK(plan_upload(pl, all, &pl->ditems));
  CHK(plan_upload(pl, groups, &pl->dgroups));
  pl->ngroups = (int)groups.size();
  {
    int chunks = 1;
    for (const ReduceGroup &g : groups)
      chunks = std::max<int>(chunks, (int)cdiv((int64_t)g.nrows * 2 * (g.l + 1), 256));
    pl->red_chunks = std::min(chunks, 64);
  }
  CHK(plan_alloc(pl, (size_t)std::max<int64_t>(psize, 1), &pl->partial));
  CHK(plan_alloc(pl, (size_t)std::max<int64_t>(2 * ftot, 1), &pl->f));
  CHK(plan_alloc(pl, (size_t)6 * ntot, &pl->soa));
  CHK(plan_alloc(pl, (size_t)6 * ntot, &pl->soa_in));
  for (int k = 0; k < 2; ++k) {
    CHK(plan_alloc(pl, (size_t)ntot, &pl->keys[k]));
    CHK(plan_alloc(pl, (size_t)ntot, &pl->idx[k]));
  }
  CHK(plan_alloc(pl, 8 + 3 * GROUPED_LANES, &pl->derr));
  {
    std::vector<SortSeg> seg(ncat);
    int64_t hb = 0;
    int nbmax = 1;
    for (int c = 0; c < ncat; ++c) {
      const int nb = (int)cdiv(cats[c].n_points, SORT_CHUNK);
      seg[c] = SortSeg{pbase[c], cats[c].n_points, hb, nb, 0};
      hb += 256LL * nb;
      nbmax = std::max(nbmax, nb);
    }
    CHK(plan_upload(pl, seg, &pl->dseg));
    CHK(plan_alloc(pl, (size_t)hb, &pl->hist));
    pl->sort_nb = nbmax;
  }
  for (int k = 0; k < 3; ++k) pl->grid_ks[k] = (int)std::max<int64_t>(1, std::min<int64_t>(pl->nmx_ks[k], 2LL * h->ncu));
  // catalogue descriptors
  std::vector<CatPrep> prep(ncat);
  std::vector<CatOut> outs(ncat);
  int64_t nmaxp = 1, omax = 1;
  for (int c = 0; c < ncat; ++c) {
    prep[c] = CatPrep{cats[c].d_points, cats[c].n_points, pbase[c], R[c]};
    outs[c] = CatOut{cats[c].d_k_ln, cats[c].d_c_ln, cats[c].d_f_lmn, cats[c].lmax, cats[c].nmax,
                     kbase[c], nlbase[c], foffbase[c], fbase[c]};
    nmaxp = std::max<int64_t>(nmaxp, cats[c].n_points);
    omax = std::max<int64_t>(omax, (int64_t)cats[c].lmax * cats[c].nmax);
    if (cats[c].d_f_lmn) omax = std::max<int64_t>(omax, 2 * (int64_t)sfbhip_flmn_count(cats[c].lmax, cats[c].nmax));
  }
  CHK(plan_upload(pl, prep, &pl->dprep));
  CHK(plan_upload(pl, outs, &pl->dout));
  pl->prep_bx = (int)std::min<int64_t>(2048, cdiv(nmaxp, 1024));
  pl->out_bx = (int)std::min<int64_t>(256, cdiv(omax, 1024));
  HIPCHK(hipDeviceSynchronize());  // tables built, host vectors go out of scope
  return SFBHIP_OK;
}

void plan_release(sfbhip_batch_plan *pl) {
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  for (void *ptr : pl->mem) (void)hipFree(ptr);
  if (pl->h_err) (void)hipHostFree(pl->h_err);
  for (auto &e : pl->ev)
    if (e) (void)hipEventDestroy(e);
  if (pl->own) (void)hipStreamDestroy(pl->own);
  if (pl->cap) (void)hipStreamDestroy(pl->cap);
  if (pl->fork) (void)hipEventDestroy(pl->fork);
  for (int k = 0; k < 2; ++k) {
    if (pl->join[k]) (void)hipEventDestroy(pl->join[k]);
    if (pl->side[k]) (void)hipStreamDestroy(pl->side[k]);
  }
  delete pl;
}

sfbhip_status check_tiles(const sfbhip_tile_table *t) {
  for (int b = 0; b < SFBHIP_TILE_BANDS; ++b) {
    if (t->chunks[b] != 1 && t->chunks[b] != 2 && t->chunks[b] != 4)
      return fail(SFBHIP_INVALID_ARGUMENT, "tile table: chunks must be 1, 2 or 4");
    if (t->point_split[b] != 1 && t->point_split[b] != 2 && t->point_split[b] != 4)
      return fail(SFBHIP_INVALID_ARGUMENT, "tile table: point_split must be 1, 2 or 4");
  }
  if (t->items_per_cu < 1 || t->items_per_cu > 4096)
    return fail(SFBHIP_INVALID_ARGUMENT, "tile table: items_per_cu must be in [1, 4096]");
  return SFBHIP_OK;
}

}  // namespace

extern "C" {

sfbhip_status sfbhip_set_tile_table(sfbhip_handle *h, const sfbhip_tile_table *t) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  if (!t) {
    default_tiles(&h->tiles);
    return SFBHIP_OK;
  }
  CHK(check_tiles(t));
  h->tiles = *t;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_get_tile_table(sfbhip_handle *h, sfbhip_tile_table *t) {
  g_last_error.clear();
  if (!h || !t) return fail(SFBHIP_INVALID_ARGUMENT, "NULL argument");
  *t = h->tiles;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_create(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                       const sfbhip_params *p, const double *r_bound, sfbhip_batch_plan **out) {
  g_last_error.clear();
  if (!h || !out) return fail(SFBHIP_INVALID_ARGUMENT, "handle or out is NULL");
  *out = nullptr;
  HIPCHK(hipSetDevice(h->device));
  sfbhip_batch_plan *pl = new sfbhip_batch_plan();
  pl->device = h->device;
  sfbhip_status rc = SFBHIP_OK;
  auto step = [&](hipError_t e, const char *what) {
    if (rc == SFBHIP_OK && e != hipSuccess) rc = fail(SFBHIP_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
  };
  step(hipHostMalloc((void **)&pl->h_err, sizeof(int) * 8, hipHostMallocDefault), "hipHostMalloc");
  for (auto &e : pl->ev) step(hipEventCreate(&e), "hipEventCreate");
  step(hipStreamCreateWithFlags(&pl->own, hipStreamDefault), "hipStreamCreate");
  step(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking), "hipStreamCreate");
  step(hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming), "hipEventCreate");
  for (int k = 0; k < 2; ++k) {
    step(hipEventCreateWithFlags(&pl->join[k], hipEventDisableTiming), "hipEventCreate");
    step(hipStreamCreateWithFlags(&pl->side[k], hipStreamNonBlocking), "hipStreamCreate");
  }
  if (rc == SFBHIP_OK) rc = plan_create(h, cats, ncat, p, r_bound, pl);
  if (rc != SFBHIP_OK) {
    (void)hipDeviceSynchronize();
    plan_release(pl);
    return rc;
  }
  *out = pl;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_destroy(sfbhip_batch_plan *plan) {
  if (!plan) return SFBHIP_OK;
  (void)hipSetDevice(plan->device);
  (void)hipDeviceSynchronize();
  plan_release(plan);
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_enqueue(sfbhip_batch_plan *plan, void *stream) {
  g_last_error.clear();
  if (!plan) return fail(SFBHIP_INVALID_ARGUMENT, "plan is NULL");
  HIPCHK(hipSetDevice(plan->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->own;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(st, &cs));
  const bool timed = cs == hipStreamCaptureStatusNone;
  CHK(plan_enqueue(plan, st, timed));
  if (timed) plan->timed = 1;
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_capture(sfbhip_batch_plan *plan, void *stream) {
  g_last_error.clear();
  if (!plan) return fail(SFBHIP_INVALID_ARGUMENT, "plan is NULL");
  HIPCHK(hipSetDevice(plan->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->cap;
  if (plan->exec) {
    HIPCHK(hipGraphExecDestroy(plan->exec));
    plan->exec = nullptr;
  }
  if (plan->graph) {
    HIPCHK(hipGraphDestroy(plan->graph));
    plan->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const sfbhip_status rc = plan_enqueue(plan, st, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != SFBHIP_OK) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIPCHK(e);
  plan->graph = g;
  HIPCHK(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_launch(sfbhip_batch_plan *plan, void *stream) {
  g_last_error.clear();
  if (!plan) return fail(SFBHIP_INVALID_ARGUMENT, "plan is NULL");
  if (!plan->exec) return fail(SFBHIP_INVALID_ARGUMENT, "plan has no captured graph (sfbhip_batch_plan_capture)");
  HIPCHK(hipSetDevice(plan->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->own;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(st, &cs));
  const bool timed = cs == hipStreamCaptureStatusNone;
  if (timed) HIPCHK(hipEventRecord(plan->ev[0], st));
  HIPCHK(hipGraphLaunch(plan->exec, st));
  if (timed) {
    HIPCHK(hipEventRecord(plan->ev[4], st));
    plan->timed = 2;
  }
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_plan_sync(sfbhip_batch_plan *plan, void *stream, float *ms4) {
  g_last_error.clear();
  if (!plan) return fail(SFBHIP_INVALID_ARGUMENT, "plan is NULL");
  HIPCHK(hipSetDevice(plan->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->own;
  HIPCHK(hipMemcpyAsync(plan->h_err, plan->derr, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ms4) {
    for (int i = 0; i < 4; ++i) ms4[i] = 0.f;
    if (plan->timed == 1)
      for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&ms4[i], plan->ev[i + 1], plan->ev[i + 2]));
    if (plan->timed) HIPCHK(hipEventElapsedTime(&ms4[3], plan->ev[0], plan->ev[4]));
  }
  const int e = plan->h_err[0];
  if (e & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  if (e & ERR_DOMAIN) return fail(SFBHIP_DOMAIN_ERROR, "point beyond its catalogue's radius bound");
  return SFBHIP_OK;
}

sfbhip_status sfbhip_batch_autotune(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                    const sfbhip_params *p, const double *r_bound, int32_t reps,
                                    sfbhip_tile_table *out) {
  g_last_error.clear();
  if (!h) return fail(SFBHIP_INVALID_ARGUMENT, "handle is NULL");
  if (!cats || ncat < 1 || ncat > 128) return fail(SFBHIP_INVALID_ARGUMENT, "ncat must be in [1, 128]");
  if (!p) return fail(SFBHIP_INVALID_ARGUMENT, "params is NULL");
  reps = std::max(1, std::min(reps, 100));
  const sfbhip_tile_table start = h->tiles;
  // time of a table: the fastest of `reps` eager runs of the contraction and reduction
  auto measure = [&](const sfbhip_tile_table &t, float *ms) -> sfbhip_status {
    h->tiles = t;
    sfbhip_batch_plan *pl = nullptr;
    sfbhip_status rc = sfbhip_batch_plan_create(h, cats, ncat, p, r_bound, &pl);
    if (rc != SFBHIP_OK) return rc;
    float best = 3.0e38f;
    for (int r = 0; r <= reps && rc == SFBHIP_OK; ++r) {
      float m4[4];
      rc = sfbhip_batch_plan_enqueue(pl, nullptr);
      if (rc == SFBHIP_OK) rc = sfbhip_batch_plan_sync(pl, nullptr, m4);
      if (rc == SFBHIP_OK && r > 0) best = std::min(best, m4[1] + m4[2]);  // run 0 warms up
    }
    sfbhip_batch_plan_destroy(pl);
    *ms = best;
    return rc;
  };
  int L = 0;
  for (int c = 0; c < ncat; ++c) L = std::max(L, cats[c].lmax);
  const int lsplit = std::min(L, std::max(0, p->mixed_fp64_below));
  sfbhip_tile_table best = start;
  float best_ms = 0.f;
  sfbhip_status rc = measure(best, &best_ms);
  // bands from the top (the largest l carry most of the work)
  for (int b = SFBHIP_TILE_BANDS - 1; b >= 0 && rc == SFBHIP_OK; --b) {
    const int blo = b == 0 ? 0 : (16 << (b - 1)), bhi = 16 << b;
    if (blo >= L || bhi <= lsplit) continue;
    const sfbhip_tile_table base = best;
    for (int ch : {1, 2, 4})
      for (int ks : {1, 2, 4}) {
        if (ch == base.chunks[b] && ks == base.point_split[b]) continue;
        sfbhip_tile_table t = base;
        t.chunks[b] = ch;
        t.point_split[b] = ks;
        float ms = 0.f;
        rc = measure(t, &ms);
        if (rc != SFBHIP_OK) break;
        if (ms < best_ms) {
          best = t;
          best_ms = ms;
        }
      }
  }
  if (rc == SFBHIP_OK) {
    const sfbhip_tile_table base = best;
    for (int ipc : {16, 32, 64, 128, 256}) {
      if (ipc == base.items_per_cu) continue;
      sfbhip_tile_table t = base;
      t.items_per_cu = ipc;
      float ms = 0.f;
      rc = measure(t, &ms);
      if (rc != SFBHIP_OK) break;
      if (ms < best_ms) {
        best = t;
        best_ms = ms;
      }
    }
  }
  if (rc != SFBHIP_OK) {
    h->tiles = start;
    return rc;
  }
  h->tiles = best;
  if (out) *out = best;
  return SFBHIP_OK;
}

}  // extern "C"

// ---- sfbhip_direct_noalloc: sfbhip_direct (FP64, exact j_l) for callers
// that cannot allocate device memory.  Every device buffer beyond the
// caller's points and outputs -- error flags, Bessel zeros, n_l, Legendre
// constants, the point table, work items, reduction groups, partial blocks
// and, without a caller f_lmn, the f_lmn workspace -- is carved from one
// static device array per GPU.  Calls that overlap on one GPU race for it.

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch` on each GPU

// The device address of this GPU's copy of `scratch`, looked up once per
// device with the device current (host code never uses the array's name).
sfbhip_status scratch_of_device(int dev, unsigned char **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(SFBHIP_NOT_SUPPORTED, "no-allocation entry point: device index beyond " +
                                          std::to_string(NOALLOC_MAX_DEVICES - 1));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(SFBHIP_HIP_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return SFBHIP_OK;
}

// Consecutive 256-byte aligned pieces of the scratch array; `used` may pass
// the capacity, which the caller checks before touching any piece.
struct Carve {
  unsigned char *base;
  size_t used = 0;
  template <class T>
  T *take(size_t count) {
    T *p = reinterpret_cast<T *>(base + used);
    used += (count * sizeof(T) + 255) & ~size_t(255);
    return p;
  }
};

sfbhip_status noalloc_too_big(size_t need) {
  return fail(SFBHIP_NOT_SUPPORTED, "no-allocation entry point: this call needs " + std::to_string(need) +
                                        " bytes of device workspace, the library holds " +
                                        std::to_string(NOALLOC_SCRATCH_BYTES) + " per GPU; use sfbhip_direct");
}

}  // namespace

extern "C" {

sfbhip_status sfbhip_direct_noalloc(const double *d_points, int64_t n_points, const sfbhip_params *p,
                                    double *d_k_ln, double *d_c_ln, double *d_f_lmn, void *stream) {
  g_last_error.clear();
  CHK(validate(p));
  if (p->precision != SFBHIP_PREC_FP64)
    return fail(SFBHIP_NOT_SUPPORTED, "no-allocation entry point: precision FP64 only");
  if (p->bessel == SFBHIP_BESSEL_REFERENCE || (p->bessel == SFBHIP_BESSEL_AUTO && n_points > 1000))
    return fail(SFBHIP_NOT_SUPPORTED,
                "no-allocation entry point: exact j_l only (bessel EXACT, or AUTO with at most 1000 points)");
  if (n_points <= 0 || !d_points)
    return fail(SFBHIP_INVALID_ARGUMENT, "empty catalogue: the normalisation divides by the number of points");
  if (!d_k_ln || !d_c_ln) return fail(SFBHIP_INVALID_ARGUMENT, "k_ln / c_ln output is NULL");
  if ((uint64_t)n_points > NOALLOC_SCRATCH_BYTES / (6 * sizeof(double))) return noalloc_too_big(48 * (size_t)n_points);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  CHK(scratch_of_device(dev, &base));
  if (p->verbose) {
    std::printf("SFB decomposition (raw method): lmax=%d nmax=%d Rmax=%g\n", p->lmax, p->nmax, p->rmax);
    std::printf("Catalog consists of %lld points.\n", (long long)n_points);
  }
  const int L = p->lmax, N = p->nmax, NZ = N + 1;
  Carve cv{base};
  int *derr = cv.take<int>(4);
  double *dz = cv.take<double>((size_t)std::max(L, 1) * NZ);
  if (cv.used > NOALLOC_SCRATCH_BYTES) return noalloc_too_big(cv.used);
  fill_int_kernel<<<1, 64, 0, st>>>(derr, 4, 0);
  HIPCHK(hipGetLastError());
  // zeros of j_l on the device, k_ln and n_l on the host (as sfbhip_direct)
  std::vector<double> zeros((size_t)std::max(L, 1) * NZ, 0.0), kln((size_t)std::max(L, 1) * N, 0.0);
  std::vector<int32_t> nl(std::max(L, 1), 0);
  if (L > 0) {
    bessel_zeros_kernel<<<cdiv(L, 64), 64, 0, st>>>(L, NZ, p->zeros, dz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(zeros.data(), dz, sizeof(double) * L * NZ, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  const double kmax = p->k_max > 0 ? p->k_max : 0.0;
  for (int l = 0; l < L; ++l) {
    int cnt = 0;
    for (int n = 0; n < N; ++n) {
      const double z = zeros[(size_t)l * NZ + n];
      kln[(size_t)l * N + n] = z / p->rmax;
      if (kmax <= 0.0 || z < kmax * p->rmax) cnt = n + 1;
    }
    nl[l] = cnt;
  }
  // the work items of sfbhip_direct's FP64 path
  const double norm = norm_factor(p, (double)n_points);
  PlanCtx cx;
  cx.nmax = N;
  cx.norm = norm;
  cx.kln = kln.data();
  cx.R = p->rmax;
  DirectPlan pl64;
  plan_direct(nl, 0, L, n_points, F64_NT, F64_TP, F64_MC, 16, f64_target_items(n_points, false), cx, &pl64);
  if (!pl64.items[5].empty()) return fail(SFBHIP_NOT_SUPPORTED, "FP64 plan class beyond 16 chunks");
  std::vector<DirectItem> all;
  for (auto &v : pl64.items) all.insert(all.end(), v.begin(), v.end());
  const std::vector<ReduceGroup> &groups = pl64.groups;
  const std::vector<int64_t> foff = flmn_offsets(L, N);
  std::vector<double2> hs, hc;
  std::vector<int32_t> ho;
  legendre_host(std::max(L, 1), hs, hc, ho);
  const size_t nf = sfbhip_flmn_count(L, N);
  // the rest of the workspace
  int32_t *dnl = cv.take<int32_t>(nl.size());
  int64_t *dfoff = cv.take<int64_t>(foff.size());
  double2 *dsect = cv.take<double2>(hs.size());
  double2 *dcoef = cv.take<double2>(hc.size());
  int32_t *dloff = cv.take<int32_t>(ho.size());
  double *soa = cv.take<double>((size_t)6 * n_points);
  DirectItem *ditems = cv.take<DirectItem>(std::max<size_t>(all.size(), 1));
  ReduceGroup *dg = cv.take<ReduceGroup>(std::max<size_t>(groups.size(), 1));
  double *partial = cv.take<double>((size_t)std::max<int64_t>(pl64.partial_size, 1));
  double *f = d_f_lmn ? d_f_lmn : cv.take<double>(std::max<size_t>(2 * nf, 1));
  if (cv.used > NOALLOC_SCRATCH_BYTES) return noalloc_too_big(cv.used);
  if (L > 0)
    HIPCHK(hipMemcpyAsync(d_k_ln, kln.data(), sizeof(double) * L * N, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dnl, nl.data(), sizeof(int32_t) * nl.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dfoff, foff.data(), sizeof(int64_t) * foff.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dsect, hs.data(), sizeof(double2) * hs.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dcoef, hc.data(), sizeof(double2) * hc.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dloff, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice, st));
  if (!all.empty())
    HIPCHK(hipMemcpyAsync(ditems, all.data(), sizeof(DirectItem) * all.size(), hipMemcpyHostToDevice, st));
  if (!groups.empty())
    HIPCHK(hipMemcpyAsync(dg, groups.data(), sizeof(ReduceGroup) * groups.size(), hipMemcpyHostToDevice, st));
  // point table; a bad point is refused before any kernel evaluates a Bessel function of it
  const int64_t n = n_points;
  double *r = soa, *x = soa + n, *s = soa + 2 * n, *ph = soa + 3 * n, *cp = soa + 4 * n, *sp = soa + 5 * n;
  prep_points_kernel<<<cdiv(n, 256), 256, 0, st>>>(d_points, n, r, x, s, ph, cp, sp, derr,
                                                    reinterpret_cast<unsigned long long *>(derr + 2));
  HIPCHK(hipGetLastError());
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (herr & ERR_BADPOINT) return fail(SFBHIP_DOMAIN_ERROR, "point with r < 0 or a non-finite coordinate");
  const PointsSoA pts{r, x, s, ph, cp, sp, n};
  const LegendreTables lt{dsect, dcoef, dloff};
  size_t off = 0;
  for (int ci = 0; ci < 5; ++ci) {
    const int k = (int)pl64.items[ci].size();
    if (!k) continue;
    const DirectItem *di = ditems + off;
    switch (ci) {
      case 0: launch_f64<1>(k, di, st, pts, nullptr, d_k_ln, N, lt, JL_EXACT, BesselTableRef{}, partial, derr); break;
      case 1: launch_f64<2>(k, di, st, pts, nullptr, d_k_ln, N, lt, JL_EXACT, BesselTableRef{}, partial, derr); break;
      case 2: launch_f64<4>(k, di, st, pts, nullptr, d_k_ln, N, lt, JL_EXACT, BesselTableRef{}, partial, derr); break;
      case 3: launch_f64<8>(k, di, st, pts, nullptr, d_k_ln, N, lt, JL_EXACT, BesselTableRef{}, partial, derr); break;
      default: launch_f64<16>(k, di, st, pts, nullptr, d_k_ln, N, lt, JL_EXACT, BesselTableRef{}, partial, derr); break;
    }
    HIPCHK(hipGetLastError());
    off += k;
  }
  // f_lmn (zeroed, then every group's sum), c_ln
  if (nf) {
    fill_kernel<<<cdiv(2 * nf, 256), 256, 0, st>>>(f, (int64_t)(2 * nf), 0.0);
    HIPCHK(hipGetLastError());
  }
  if (!groups.empty()) {
    int chunks = 1;
    for (const ReduceGroup &g : groups)
      chunks = std::max<int>(chunks, (int)cdiv((int64_t)g.nrows * 2 * (g.l + 1), 256));
    reduce_partials_kernel<<<dim3((unsigned)groups.size(), (unsigned)std::min(chunks, 64)), 256, 0, st>>>(
        dg, partial, d_k_ln, f);
    HIPCHK(hipGetLastError());
  }
  if (L > 0) {
    cln_kernel<<<cdiv((int64_t)L * N, 256), 256, 0, st>>>(L, N, dnl, dfoff, f, d_c_ln);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));  // the host vectors the copies read go out of scope
  return SFBHIP_OK;
}

}  // extern "C"
