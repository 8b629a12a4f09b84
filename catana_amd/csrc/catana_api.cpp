// catana_api.cpp -- host C++ API of catana_api.hpp over the C-ABI.
#include "catana_api.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>

#include "catana_hip.h"

namespace catana {

namespace {

constexpr double kTwoPi = 6.283185307179586476925;

void throw_status(sfbhip_status st, const char *what) {
  if (st == SFBHIP_OK) return;
  const std::string msg = std::string(what) + ": " + sfbhip_last_error();
  if (st == SFBHIP_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (st == SFBHIP_DOMAIN_ERROR) throw std::domain_error(msg);
  throw std::runtime_error(msg);
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer owned by the host API (the C-ABI itself allocates only its workspace).
struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
  }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

// Process-wide context on device 0: one handle and reusable staging buffers.
struct Context {
  sfbhip_handle *h = nullptr;
  DevMem pts, off, rad, kln, cln, flmn;
  std::mutex mu;
  Context() { throw_status(sfbhip_create(&h, 0), "sfbhip_create"); }
  ~Context() { sfbhip_destroy(h); }
};

// Never destroyed: at process exit the HIP runtime may already be gone, and
// the operating system's GPU driver releases the device memory anyway.
Context &ctx() {
  static Context *c = new Context;
  return *c;
}

sfbhip_params make_params(unsigned short lmax, unsigned short nmax, double rmax, bool verbose, const Options &o) {
  sfbhip_params p;
  sfbhip_default_params(&p);
  p.lmax = lmax;
  p.nmax = nmax;
  p.rmax = rmax;
  p.k_max = o.k_max;
  p.window_volume = o.window_volume;
  p.precision = o.precision ? SFBHIP_PREC_MIXED : SFBHIP_PREC_FP64;
  p.bessel = o.interpolated < 0 ? SFBHIP_BESSEL_AUTO : (o.interpolated ? SFBHIP_BESSEL_REFERENCE : SFBHIP_BESSEL_EXACT);
  p.zeros = o.zeros ? SFBHIP_ZEROS_CONVERGED : SFBHIP_ZEROS_REFERENCE;
  p.alias_compat = o.alias_compat ? 1 : 0;
  p.verbose = verbose ? 1 : 0;
  return p;
}

KClkk download(Context &c, unsigned short lmax, unsigned short nmax, bool store_flmn) {
  KClkk out;
  out.k_ln = Array2D(lmax, nmax);
  out.c_ln = Array2D(lmax, nmax);
  const size_t n2 = (size_t)lmax * nmax;
  if (n2) {
    hip_check(hipMemcpy(out.k_ln.data.data(), c.kln.p, n2 * sizeof(double), hipMemcpyDeviceToHost), "copy k_ln");
    hip_check(hipMemcpy(out.c_ln.data.data(), c.cln.p, n2 * sizeof(double), hipMemcpyDeviceToHost), "copy c_ln");
  }
  if (store_flmn) {
    const size_t nf = sfbhip_flmn_count(lmax, nmax);
    std::vector<std::complex<double>> flat(nf);
    if (nf)
      hip_check(hipMemcpy(flat.data(), c.flmn.p, nf * sizeof(std::complex<double>), hipMemcpyDeviceToHost),
                "copy f_lmn");
    size_t off = 0;
    for (unsigned l = 0; l < lmax; ++l) {
      ComplexBlock b;
      b.rows = l + 1;
      b.cols = nmax;
      b.data.assign(flat.begin() + off, flat.begin() + off + (size_t)(l + 1) * nmax);
      off += (size_t)(l + 1) * nmax;
      out.f_lmn.push_back(std::move(b));
    }
  }
  return out;
}

void reserve_outputs(Context &c, unsigned short lmax, unsigned short nmax, bool store_flmn) {
  c.kln.reserve(std::max<size_t>((size_t)lmax * nmax * sizeof(double), 8));
  c.cln.reserve(std::max<size_t>((size_t)lmax * nmax * sizeof(double), 8));
  if (store_flmn) c.flmn.reserve(std::max<size_t>(sfbhip_flmn_count(lmax, nmax) * 16, 16));
}

KClkk run_direct(const PointContainer &points, unsigned short lmax, unsigned short nmax, double rmax, bool store,
                 bool verbose, const Options &o) {
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const size_t n = points.size();
  c.pts.reserve(std::max<size_t>(n * sizeof(Point), 24));
  if (n) hip_check(hipMemcpy(c.pts.p, points.data(), n * sizeof(Point), hipMemcpyHostToDevice), "copy points");
  reserve_outputs(c, lmax, nmax, store);
  sfbhip_params p = make_params(lmax, nmax, rmax, verbose, o);
  throw_status(sfbhip_direct(c.h, static_cast<const double *>(c.pts.p), (int64_t)n, &p,
                             static_cast<double *>(c.kln.p), static_cast<double *>(c.cln.p),
                             store ? static_cast<double *>(c.flmn.p) : nullptr, nullptr),
               "sfb_decomposition");
  return download(c, lmax, nmax, store);
}

KClkk run_healpix(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                  bool store, bool verbose, const Options &o) {
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  std::vector<int64_t> off;
  std::vector<double> rad;
  pix.to_csr(off, rad);
  c.off.reserve(off.size() * sizeof(int64_t));
  c.rad.reserve(std::max<size_t>(rad.size() * sizeof(double), 8));
  hip_check(hipMemcpy(c.off.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice), "copy offsets");
  if (!rad.empty())
    hip_check(hipMemcpy(c.rad.p, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice), "copy radii");
  reserve_outputs(c, lmax, nmax, store);
  sfbhip_params p = make_params(lmax, nmax, rmax, verbose, o);
  throw_status(sfbhip_healpix(c.h, static_cast<const int64_t *>(c.off.p), static_cast<const double *>(c.rad.p),
                              (int32_t)pix.get_nside(), &p, static_cast<double *>(c.kln.p),
                              static_cast<double *>(c.cln.p), store ? static_cast<double *>(c.flmn.p) : nullptr,
                              nullptr),
               "sfb_decomposition");
  return download(c, lmax, nmax, store);
}

bool is_power_of_2(unsigned int n) { return n && ((n & (n - 1)) == 0); }

}  // namespace

// ---------------------------------------------------------------- points
Point::Point(double x, double y, double z) {
  r = std::sqrt(x * x + y * y + z * z);
  p.theta = std::atan2(std::sqrt(x * x + y * y), z);
  double ph = (x == 0.0 && y == 0.0) ? 0.0 : std::atan2(y, x);
  if (ph < 0.0) ph += kTwoPi;
  if (ph >= kTwoPi) ph -= kTwoPi;
  p.phi = ph;
}

Point point_from_spherical_position(double r, double theta, double phi) {
  Point o;
  o.r = r;
  o.p.theta = theta;
  o.p.phi = phi;
  return o;
}

Point point_from_box_position(double x, double y, double z, double shift) {
  return Point(x + shift, y + shift, z + shift);
}

// ---------------------------------------------------------------- KClkk
Array2D compute_kln(unsigned short lmax, unsigned short nmax, double rmax, int zeros) {
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  Options o;
  o.zeros = zeros;
  sfbhip_params p = make_params(lmax, nmax, rmax, false, o);
  Array2D k(lmax, nmax);
  std::vector<int32_t> nl(std::max<int>(lmax, 1));
  throw_status(sfbhip_kln(c.h, &p, k.data.data(), nl.data()), "compute_kln");
  return k;
}

// ---- FunctionInterpolator (FunctionInterpolator.cpp:7-45) ----

FunctionInterpolator::FunctionInterpolator(std::function<double(double)> fct, unsigned int interpolation_points,
                                           double x_min, double x_max, bool /*parallel*/)
    : x_min_(x_min), x_max_(x_max) {
  if (interpolation_points < 2 || !(x_max > x_min))
    throw std::invalid_argument("FunctionInterpolator: needs >= 2 points and x_min < x_max");
  const double dx = (x_max - x_min) / (interpolation_points - 1);
  dx_inv_ = 1.0 / dx;
  y_.resize(interpolation_points);
  for (unsigned int i = 0; i < interpolation_points; ++i) y_[i] = fct(x_min + i * dx);
}

unsigned int FunctionInterpolator::get_index(double x) const {
  if (x < x_min_ || x >= x_max_)
    throw std::domain_error("argument outside range for which this FunctionInterpolator is defined");
  return static_cast<unsigned int>((x - x_min_) * dx_inv_);
}

double FunctionInterpolator::operator()(double x) const {
  if (x < x_min_ || x >= x_max_)
    throw std::domain_error("argument outside range for which this FunctionInterpolator is defined");
  const double u = (x - x_min_) * dx_inv_;
  unsigned int i = static_cast<unsigned int>(u);
  if (i > y_.size() - 2) i = (unsigned int)y_.size() - 2;
  return y_[i] + (y_[i + 1] - y_[i]) * (u - i);
}

namespace besseltools {

// ---- SphericalBesselZeros (SphericalBesselZeros.cpp:11-101), zeros from the C-ABI ----

SphericalBesselZeros::SphericalBesselZeros(unsigned int l) : l_(l) { grow(2); }

void SphericalBesselZeros::grow(size_t n) {
  if (zeros_.size() >= n) return;
  // the zero sequence is deterministic, so a longer prefix extends the shorter
  size_t m = std::max<size_t>(n, 2 * zeros_.size());
  std::vector<double> z(m);
  throw_status(sfbhip_bessel_zeros((int32_t)l_, (int32_t)m, SFBHIP_ZEROS_REFERENCE, z.data()), "SphericalBesselZeros");
  zeros_.swap(z);
}

double SphericalBesselZeros::operator[](unsigned int n) {
  grow((size_t)n + 1);
  return zeros_[n];
}

void SphericalBesselZeros::compute_up_to(double z_max) {
  while (zeros_.back() < z_max) grow(zeros_.size() + 1);
}

unsigned int SphericalBesselZeros::next_zero(double z) {
  compute_up_to(z);
  return (unsigned int)(std::upper_bound(zeros_.begin(), zeros_.end(), z) - zeros_.begin());
}

namespace {

void check_triples(const std::vector<int32_t> &l, const std::vector<double> &k1, const std::vector<double> &k2) {
  if (l.size() != k1.size() || l.size() != k2.size())
    throw std::invalid_argument("double_sbessel_integrals: l, k1, k2 must have the same length");
}

std::vector<double> run_integrals(const std::vector<int32_t> &l, const std::vector<double> &k1,
                                  const std::vector<double> &k2, double Rmax, const sfbhip_weight &w) {
  check_triples(l, k1, k2);
  std::vector<double> out(l.size());
  if (l.empty()) return out;
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  throw_status(sfbhip_sbessel_integrals(c.h, (int64_t)l.size(), l.data(), k1.data(), k2.data(), Rmax, &w, out.data()),
               "double_sbessel_integrator");
  return out;
}

std::vector<double> integrals_callable(std::function<double(double)> f, int r_power, const std::vector<int32_t> &l,
                                       const std::vector<double> &k1, const std::vector<double> &k2, double Rmax) {
  std::vector<double> r = sbessel_nodes(l, k1, k2, Rmax);
  for (double &x : r) x = f(x);
  sfbhip_weight w{SFBHIP_WEIGHT_NODES, r_power, (int64_t)r.size(), r.data(), 0.0, 0.0};
  return run_integrals(l, k1, k2, Rmax, w);
}

sfbhip_weight table_weight(const FunctionInterpolator &fi, int r_power) {
  return sfbhip_weight{SFBHIP_WEIGHT_TABLE, r_power, (int64_t)fi.samples().size(), fi.samples().data(), fi.x_min(),
                       fi.x_max()};
}

}  // namespace

std::vector<double> sbessel_nodes(const std::vector<int32_t> &l, const std::vector<double> &k1,
                                  const std::vector<double> &k2, double Rmax) {
  check_triples(l, k1, k2);
  if (l.empty()) return {};
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int64_t n = 0;
  throw_status(sfbhip_sbessel_nodes(c.h, (int64_t)l.size(), l.data(), k1.data(), k2.data(), Rmax, &n, nullptr),
               "sbessel_nodes");
  std::vector<double> r((size_t)n);
  throw_status(sfbhip_sbessel_nodes(c.h, (int64_t)l.size(), l.data(), k1.data(), k2.data(), Rmax, &n, r.data()),
               "sbessel_nodes");
  return r;
}

std::vector<double> sbessel_integrals_at_nodes(const std::vector<int32_t> &l, const std::vector<double> &k1,
                                               const std::vector<double> &k2, double Rmax,
                                               const std::vector<double> &f_at_nodes) {
  sfbhip_weight w{SFBHIP_WEIGHT_NODES, 0, (int64_t)f_at_nodes.size(), f_at_nodes.data(), 0.0, 0.0};
  return run_integrals(l, k1, k2, Rmax, w);
}

std::vector<double> double_sbessel_integrals(std::function<double(double)> f, const std::vector<int32_t> &l,
                                             const std::vector<double> &k1, const std::vector<double> &k2,
                                             double Rmax) {
  return integrals_callable(std::move(f), 0, l, k1, k2, Rmax);
}

std::vector<double> double_sbessel_integrals(const FunctionInterpolator &fi, const std::vector<int32_t> &l,
                                             const std::vector<double> &k1, const std::vector<double> &k2,
                                             double Rmax) {
  return run_integrals(l, k1, k2, Rmax, table_weight(fi, 0));
}

double double_sbessel_integrator(std::function<double(double)> f, const unsigned int &l, const double &Rmax, double k1,
                                 double k2) {
  return double_sbessel_integrals(std::move(f), {(int32_t)l}, {k1}, {k2}, Rmax)[0];
}

double double_sbessel_integrator_bz(std::function<double(double)> f, const unsigned int &l, SphericalBesselZeros &bz,
                                    const double &Rmax, double k1, double k2) {
  if (bz.l() != l) throw std::invalid_argument("double_sbessel_integrator_bz: zeros of another multipole");
  return double_sbessel_integrator(std::move(f), l, Rmax, k1, k2);
}

double double_sbessel_integrator(const FunctionInterpolator &fi, const unsigned int &l, const double &Rmax, double k1,
                                 double k2) {
  return double_sbessel_integrals(fi, {(int32_t)l}, {k1}, {k2}, Rmax)[0];
}

}  // namespace besseltools

void kept_modes(unsigned short lmax, double k_max, double rmax, unsigned short nmax, std::vector<int32_t> *l,
                std::vector<int32_t> *n, std::vector<double> *k) {
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  Options o;
  o.k_max = k_max;
  sfbhip_params p = make_params(lmax, nmax, rmax, false, o);
  Array2D kln(lmax, nmax);
  std::vector<int32_t> nl(std::max<int>(lmax, 1));
  throw_status(sfbhip_kln(c.h, &p, kln.data.data(), nl.data()), "kept_modes");
  for (int ll = 0; ll < lmax; ++ll)
    for (int nn = 0; nn < nl[ll]; ++nn) {
      l->push_back(ll);
      n->push_back(nn);
      k->push_back(kln(ll, nn));
    }
}

namespace {

Array2D shot_noise_finish(const std::vector<double> &I, const std::vector<int32_t> &l, const std::vector<int32_t> &n,
                          const std::vector<double> &k, unsigned short lmax, unsigned short nmax, double rmax,
                          double n_points, double window_volume) {
  const double V = window_volume > 0 ? window_volume : 4.0 / 3.0 * M_PI * rmax * rmax * rmax;
  Array2D out(lmax, nmax);
  for (size_t i = 0; i < I.size(); ++i) out(l[i], n[i]) = V / n_points * (2.0 / M_PI) * k[i] * k[i] * I[i];
  return out;
}

}  // namespace

Array2D shot_noise(std::function<double(double)> w, unsigned short lmax, double k_max, double rmax,
                   unsigned short nmax, double n_points, double window_volume) {
  if (!(n_points > 0)) throw std::invalid_argument("shot_noise: n_points must be positive");
  std::vector<int32_t> l, n;
  std::vector<double> k;
  kept_modes(lmax, k_max, rmax, nmax, &l, &n, &k);
  if (l.empty()) return Array2D(lmax, nmax);
  std::vector<double> r = besseltools::sbessel_nodes(l, k, k, rmax);
  for (double &x : r) x = w(x);
  sfbhip_weight wt{SFBHIP_WEIGHT_NODES, 2, (int64_t)r.size(), r.data(), 0.0, 0.0};
  std::vector<double> I = besseltools::run_integrals(l, k, k, rmax, wt);
  return shot_noise_finish(I, l, n, k, lmax, nmax, rmax, n_points, window_volume);
}

Array2D shot_noise(const FunctionInterpolator &w, unsigned short lmax, double k_max, double rmax,
                   unsigned short nmax, double n_points, double window_volume) {
  if (!(n_points > 0)) throw std::invalid_argument("shot_noise: n_points must be positive");
  std::vector<int32_t> l, n;
  std::vector<double> k;
  kept_modes(lmax, k_max, rmax, nmax, &l, &n, &k);
  if (l.empty()) return Array2D(lmax, nmax);
  std::vector<double> I = besseltools::run_integrals(l, k, k, rmax, besseltools::table_weight(w, 2));
  return shot_noise_finish(I, l, n, k, lmax, nmax, rmax, n_points, window_volume);
}

KClkk::KClkk(unsigned short lmax, unsigned short nmax, double rmax)
    : k_ln(compute_kln(lmax, nmax, rmax)), c_ln(lmax, nmax) {}

void KClkk::savetxt(const std::string &base) const {
  auto dump = [](const std::string &path, const Array2D &a) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "Error creating output file: %s\n", path.c_str());
      return;
    }
    for (size_t i = 0; i < a.rows; ++i) {
      for (size_t j = 0; j < a.cols; ++j) {
        std::fprintf(f, "%.18e", a(i, j));
        if (j != a.cols - 1) std::fputc(' ', f);
      }
      std::fputc('\n', f);
    }
    std::fclose(f);
  };
  dump(base + ".k_ln", k_ln);
  dump(base + ".c_ln", c_ln);
}

// ---------------------------------------------------------------- pixelised container
PixelizedPointContainer::PixelizedPointContainer(unsigned int nside) : nside_(nside) {
  if (!is_power_of_2(nside)) throw std::invalid_argument("argument (nside) must be a power of 2");
  pixels_.resize((size_t)12 * nside * nside);
}

PixelizedPointContainer::PixelizedPointContainer(unsigned int nside, const PointContainer &pc)
    : PixelizedPointContainer(nside) {
  if (pc.empty()) return;
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const size_t n = pc.size(), npix = pixels_.size();
  c.pts.reserve(n * sizeof(Point));
  c.off.reserve((npix + 1) * sizeof(int64_t));
  c.rad.reserve(n * sizeof(double));
  hip_check(hipMemcpy(c.pts.p, pc.data(), n * sizeof(Point), hipMemcpyHostToDevice), "copy points");
  throw_status(sfbhip_pixelize(c.h, static_cast<const double *>(c.pts.p), (int64_t)n, (int32_t)nside,
                               static_cast<int64_t *>(c.off.p), static_cast<double *>(c.rad.p), nullptr),
               "PixelizedPointContainer");
  std::vector<int64_t> off(npix + 1);
  std::vector<double> rad(n);
  hip_check(hipMemcpy(off.data(), c.off.p, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost), "copy offsets");
  hip_check(hipMemcpy(rad.data(), c.rad.p, n * sizeof(double), hipMemcpyDeviceToHost), "copy radii");
  for (size_t q = 0; q < npix; ++q) pixels_[q].assign(rad.begin() + off[q], rad.begin() + off[q + 1]);
}

void PixelizedPointContainer::add_point(const Point &pt) {
  int64_t q = 0;
  throw_status(sfbhip_ang2pix_ring((int32_t)nside_, pt.p.theta, pt.p.phi, &q), "add_point");
  pixels_[(size_t)q].push_back(pt.r);
}

size_t PixelizedPointContainer::get_npoints() const {
  size_t s = 0;
  for (const auto &v : pixels_) s += v.size();
  return s;
}

std::vector<int> PixelizedPointContainer::get_countmap() const {
  std::vector<int> m(pixels_.size());
  for (size_t i = 0; i < pixels_.size(); ++i) m[i] = (int)pixels_[i].size();
  return m;
}

void PixelizedPointContainer::to_csr(std::vector<int64_t> &offsets, std::vector<double> &radii) const {
  offsets.assign(pixels_.size() + 1, 0);
  radii.clear();
  radii.reserve(get_npoints());
  for (size_t i = 0; i < pixels_.size(); ++i) {
    radii.insert(radii.end(), pixels_[i].begin(), pixels_[i].end());
    offsets[i + 1] = (int64_t)radii.size();
  }
}

// ---------------------------------------------------------------- decompositions
KClkk sfb_decomposition(const PointContainer &points, unsigned short lmax, unsigned short nmax, double rmax,
                        bool store_flmn, bool verbose, const Options &opt) {
  return run_direct(points, lmax, nmax, rmax, store_flmn, verbose, opt);
}

KClkk sfb_decomposition(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                        bool store_flmn, bool verbose, const Options &opt) {
  return run_healpix(pix, lmax, nmax, rmax, store_flmn, verbose, opt);
}

KClkk _sfb_raw(const PointContainer &points, unsigned short lmax, unsigned short nmax, double rmax, bool store_flmn,
               bool verbose, bool /*parallel*/, bool interpolated) {
  Options o;
  o.interpolated = interpolated ? 1 : 0;
  return run_direct(points, lmax, nmax, rmax, store_flmn, verbose, o);
}

KClkk _sfb_reverse_fft(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                       bool store_flmn, bool verbose, bool /*parallel*/, bool interpolated) {
  Options o;
  o.interpolated = interpolated ? 1 : 0;
  o.alias_compat = true;
  return run_healpix(pix, lmax, nmax, rmax, store_flmn, verbose, o);
}

KClkk _sfb_reverse(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                   bool store_flmn, bool verbose, bool /*parallel*/, bool interpolated) {
  Options o;
  o.interpolated = interpolated ? 1 : 0;
  o.alias_compat = false;
  return run_healpix(pix, lmax, nmax, rmax, store_flmn, verbose, o);
}

KClkk sfb_decomposition_reverse(const PointContainer &points, unsigned short lmax, double k_max, double rmax,
                                unsigned short nmax, double window_volume, bool store_flmn, bool verbose,
                                int precision) {
  Options o;
  o.precision = precision;
  o.interpolated = 0;
  o.k_max = k_max;
  o.window_volume = window_volume;
  return run_direct(points, lmax, nmax, rmax, store_flmn, verbose, o);
}

std::vector<KClkk> sfb_decomposition_reverse_batched(const std::vector<CatalogSpec> &cats, bool store_flmn,
                                                     bool verbose, int precision) {
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  // one device region for all points, one for all outputs
  size_t npts = 0, nout = 0;
  std::vector<size_t> poff(cats.size()), koff(cats.size()), foff(cats.size());
  for (size_t i = 0; i < cats.size(); ++i) {
    if (!cats[i].points) throw std::invalid_argument("catalogue " + std::to_string(i) + ": no points");
    poff[i] = npts;
    npts += cats[i].points->size();
    koff[i] = nout;
    nout += 2 * (size_t)cats[i].lmax * cats[i].nmax;  // k_ln, c_ln
    foff[i] = nout;
    if (store_flmn) nout += 2 * sfbhip_flmn_count(cats[i].lmax, cats[i].nmax);
  }
  c.pts.reserve(std::max<size_t>(npts * sizeof(Point), 24));
  c.flmn.reserve(std::max<size_t>(nout * sizeof(double), 8));
  std::vector<sfbhip_catalog> cc(cats.size());
  double *obuf = static_cast<double *>(c.flmn.p);
  for (size_t i = 0; i < cats.size(); ++i) {
    const size_t n = cats[i].points->size();
    if (n)
      hip_check(hipMemcpy(static_cast<char *>(c.pts.p) + poff[i] * sizeof(Point), cats[i].points->data(),
                          n * sizeof(Point), hipMemcpyHostToDevice),
                "copy points");
    const size_t k2 = (size_t)cats[i].lmax * cats[i].nmax;
    cc[i].d_points = reinterpret_cast<const double *>(static_cast<char *>(c.pts.p) + poff[i] * sizeof(Point));
    cc[i].n_points = (int64_t)n;
    cc[i].lmax = cats[i].lmax;
    cc[i].nmax = cats[i].nmax;
    cc[i].rmax = cats[i].rmax;
    cc[i].k_max = cats[i].k_max;
    cc[i].window_volume = cats[i].window_volume;
    cc[i].d_k_ln = obuf + koff[i];
    cc[i].d_c_ln = obuf + koff[i] + k2;
    cc[i].d_f_lmn = store_flmn ? obuf + foff[i] : nullptr;
  }
  Options o;
  o.precision = precision;
  sfbhip_params p = make_params(1, 1, 1.0, verbose, o);
  throw_status(sfbhip_direct_batched(c.h, cc.data(), (int32_t)cc.size(), &p, nullptr),
               "sfb_decomposition_reverse_batched");
  std::vector<double> host(nout);
  if (nout) hip_check(hipMemcpy(host.data(), obuf, nout * sizeof(double), hipMemcpyDeviceToHost), "copy results");
  std::vector<KClkk> res(cats.size());
  for (size_t i = 0; i < cats.size(); ++i) {
    const unsigned short L = cats[i].lmax, N = cats[i].nmax;
    const size_t k2 = (size_t)L * N;
    res[i].k_ln = Array2D(L, N);
    res[i].c_ln = Array2D(L, N);
    std::copy(host.begin() + koff[i], host.begin() + koff[i] + k2, res[i].k_ln.data.begin());
    std::copy(host.begin() + koff[i] + k2, host.begin() + koff[i] + 2 * k2, res[i].c_ln.data.begin());
    if (store_flmn) {
      size_t o2 = foff[i];
      for (unsigned l = 0; l < L; ++l) {
        ComplexBlock b;
        b.rows = l + 1;
        b.cols = N;
        b.data.resize((size_t)(l + 1) * N);
        for (size_t e = 0; e < b.data.size(); ++e) b.data[e] = {host[o2 + 2 * e], host[o2 + 2 * e + 1]};
        o2 += 2 * b.data.size();
        res[i].f_lmn.push_back(std::move(b));
      }
    }
  }
  return res;
}

KClkk sfb_decomposition_healpix(const PixelizedPointContainer &pix, unsigned short lmax, double k_max, double rmax,
                                unsigned short nmax, double window_volume, bool store_flmn, bool verbose,
                                int precision) {
  Options o;
  o.precision = precision;
  o.interpolated = 0;
  o.k_max = k_max;
  o.window_volume = window_volume;
  return run_healpix(pix, lmax, nmax, rmax, store_flmn, verbose, o);
}

KClkk sfb_decomposition_healpix(const PointContainer &points, unsigned int nside, unsigned short lmax, double k_max,
                                double rmax, unsigned short nmax, double window_volume, bool store_flmn,
                                bool verbose, int precision) {
  return sfb_decomposition_healpix(PixelizedPointContainer(nside, points), lmax, k_max, rmax, nmax, window_volume,
                                   store_flmn, verbose, precision);
}

}  // namespace catana

// ================================================================ ingestion (io)
namespace catana {

namespace {

std::atomic<uint64_t> g_seed{0};
std::atomic<uint64_t> g_draws{0};  // records filtered so far: the index base of the next draws

constexpr double kPi = 3.14159265358979323846;

struct Unit {
  double x, y, z;
};
Unit unit_vector(const Point &p) {
  const double st = std::sin(p.p.theta);
  return {st * std::cos(p.p.phi), st * std::sin(p.p.phi), std::cos(p.p.theta)};
}

FILE *fh(void *p) { return static_cast<FILE *>(p); }

sfbhip_filter to_c(const io::FilterSpec &s) {
  sfbhip_filter f;
  f.kind = s.kind;
  f.nside = s.nside;
  f.a = s.a;
  f.x_min = s.x_min;
  f.x_max = s.x_max;
  f.n = s.n;
  f.d_values = s.d_values;
  return f;
}

// Device buffers of the streaming filter (one set per thread).
struct StreamBufs {
  DevMem chunk, out, all, sub;
  size_t all_count = 0;
};
StreamBufs &stream_bufs() {
  static thread_local StreamBufs *b = new StreamBufs;
  return *b;
}

// Filters n host points with the given filters on the device; appends the
// survivors to bufs.all (device) and returns their count.  The filters'
// host work (a Python window function, say) runs before the handle's lock is
// taken, so such a callback may itself call into this library.
size_t filter_chunk(Context &c, StreamBufs &b, const Point *pts, size_t n, const std::vector<io::Filter *> &filters) {
  if (n == 0) return 0;
  std::vector<sfbhip_filter> fs;
  for (io::Filter *f : filters) {
    f->prepare(pts, n);
    fs.push_back(to_c(f->spec()));
  }
  std::lock_guard<std::mutex> lk(c.mu);
  b.chunk.reserve(n * sizeof(Point));
  b.out.reserve(n * sizeof(Point));
  hip_check(hipMemcpy(b.chunk.p, pts, n * sizeof(Point), hipMemcpyHostToDevice), "copy points");
  int64_t kept = 0;
  const uint64_t base = g_draws.fetch_add(n);
  throw_status(sfbhip_filter_points(c.h, b.chunk.p, (int64_t)n, SFBHIP_RECORD_POINT, 0.0, fs.data(),
                                    (int32_t)fs.size(), g_seed.load(), base, static_cast<double *>(b.out.p), &kept,
                                    nullptr),
               "filter");
  if (kept > 0) {
    const size_t need = (b.all_count + (size_t)kept) * sizeof(Point);
    if (need > b.all.bytes) {  // grow, keeping what was collected
      DevMem bigger;
      bigger.reserve(std::max(need, 2 * b.all.bytes));
      if (b.all_count)
        hip_check(hipMemcpy(bigger.p, b.all.p, b.all_count * sizeof(Point), hipMemcpyDeviceToDevice), "grow");
      std::swap(b.all.p, bigger.p);
      std::swap(b.all.bytes, bigger.bytes);
    }
    hip_check(hipMemcpy(static_cast<char *>(b.all.p) + b.all_count * sizeof(Point), b.out.p,
                        (size_t)kept * sizeof(Point), hipMemcpyDeviceToDevice),
              "append");
    b.all_count += (size_t)kept;
  }
  return (size_t)kept;
}

}  // namespace

void init_random(unsigned int seed) {
  g_seed = seed;
  g_draws = 0;
}

void init_random() { init_random(std::random_device{}()); }

namespace io {

// ---- sources ----

PointContainer Source::get_point_container() {
  reset();
  PointContainer pc;
  std::vector<Point> buf(100000);
  long long k;
  while ((k = read(buf.data(), buf.size())) != -1) pc.insert(pc.end(), buf.begin(), buf.begin() + k);
  return pc;
}

PixelizedPointContainer Source::get_pixelized_point_container(unsigned int nside) {
  return PixelizedPointContainer(nside, get_point_container());
}

long long PointContainerSource::read(Point *out, size_t n) {
  if (pos_ >= pc_.size()) return -1;
  const size_t k = std::min(n, pc_.size() - pos_);
  std::copy(pc_.begin() + pos_, pc_.begin() + pos_ + k, out);
  pos_ += k;
  return (long long)k;
}

TextSource::TextSource(std::string filename, TextFormat format, double shift)
    : filename_(std::move(filename)), format_(format), shift_(shift) {
  fp_ = std::fopen(filename_.c_str(), "r");
  if (!fp_) std::printf("WARNING: Could not open file %s\n", filename_.c_str());
}

TextSource::~TextSource() {
  if (fp_) std::fclose(fh(fp_));
}

long long TextSource::read(Point *out, size_t n) {
  if (!fp_) return -1;
  size_t i = 0;
  std::string line;
  char buf[4096];
  while (i < n) {
    line.clear();
    bool got = false;
    while (std::fgets(buf, sizeof(buf), fh(fp_))) {
      got = true;
      line += buf;
      if (!line.empty() && line.back() == '\n') break;
    }
    if (!got) break;
    // three numbers; a missing one reads as 0 (std::istream >> on failure)
    double a[3] = {0.0, 0.0, 0.0};
    const char *s = line.c_str();
    for (int k = 0; k < 3; ++k) {
      char *end;
      const double v = std::strtod(s, &end);
      if (end == s) break;
      a[k] = v;
      s = end;
    }
    if (format_ == TextFormat::CARTESIAN)
      out[i] = point_from_box_position(a[0], a[1], a[2], shift_);
    else if (format_ == TextFormat::SPHERICAL)
      out[i] = point_from_spherical_position(a[0], a[1], a[2]);
    else
      out[i] = point_from_spherical_position(a[2], a[1], a[0]);  // phi theta r
    ++i;
  }
  return i == 0 ? -1 : (long long)i;
}

void TextSource::reset() {
  if (fp_) std::rewind(fh(fp_));
}

namespace {
size_t record_bytes(RecordType t) {
  return (t == RecordType::CARTESIAN_F32 || t == RecordType::SPHERICAL_F32) ? 12 : 24;
}
Point record_to_point(const unsigned char *rec, RecordType t) {
  double a[3];
  if (record_bytes(t) == 12) {
    float f[3];
    std::memcpy(f, rec, 12);
    a[0] = f[0];
    a[1] = f[1];
    a[2] = f[2];
  } else {
    std::memcpy(a, rec, 24);
  }
  if (t == RecordType::CARTESIAN_F32 || t == RecordType::CARTESIAN_F64) return point_from_box_position(a[0], a[1], a[2], 0.0);
  return point_from_spherical_position(a[0], a[1], a[2]);
}
void point_to_record(const Point &p, RecordType t, unsigned char *rec) {
  double a[3];
  if (t == RecordType::CARTESIAN_F32 || t == RecordType::CARTESIAN_F64) {
    const Unit u = unit_vector(p);
    a[0] = p.r * u.x;
    a[1] = p.r * u.y;
    a[2] = p.r * u.z;
  } else {
    a[0] = p.r;
    a[1] = p.p.theta;
    a[2] = p.p.phi;
  }
  if (record_bytes(t) == 12) {
    const float f[3] = {(float)a[0], (float)a[1], (float)a[2]};
    std::memcpy(rec, f, 12);
  } else {
    std::memcpy(rec, a, 24);
  }
}
}  // namespace

RawBinarySource::RawBinarySource(std::string filename, bool verbose, RecordType type)
    : filename_(std::move(filename)), type_(type) {
  fp_ = std::fopen(filename_.c_str(), "rb");
  if (!fp_) {
    std::printf("WARNING: Could not open file %s\n", filename_.c_str());
    return;
  }
  if (verbose) std::printf("Opened file %s\n", filename_.c_str());
  std::fseek(fh(fp_), 0, SEEK_END);
  ntot_ = (size_t)std::ftell(fh(fp_)) / record_bytes(type_);
  std::fseek(fh(fp_), 0, SEEK_SET);
}

RawBinarySource::~RawBinarySource() {
  if (fp_) std::fclose(fh(fp_));
}

long long RawBinarySource::read(Point *out, size_t n) {
  if (!fp_) return -1;
  const size_t k = std::min(n, ntot_ - pos_);
  if (k == 0) return -1;
  const size_t rb = record_bytes(type_);
  std::vector<unsigned char> buf(k * rb);
  const size_t got = std::fread(buf.data(), rb, k, fh(fp_));
  for (size_t i = 0; i < got; ++i) out[i] = record_to_point(buf.data() + i * rb, type_);
  pos_ += got;
  return got == 0 ? -1 : (long long)got;
}

void RawBinarySource::reset() {
  pos_ = 0;
  if (fp_) std::fseek(fh(fp_), 0, SEEK_SET);
}

// Gadget format 1 (the header block of read_snapshot.c, as GadgetSource.cpp:10-24 reads it)
namespace {
#pragma pack(push, 1)
struct GadgetHeader {
  uint32_t npart[6];
  double mass[6];
  double time, redshift;
  int32_t flag_sfr, flag_feedback;
  int32_t npartTotal[6];
  int32_t flag_cooling, num_files;
  double BoxSize, Omega0, OmegaLambda, HubbleParam;
  char fill[256 - 6 * 4 - 6 * 8 - 2 * 8 - 2 * 4 - 6 * 4 - 2 * 4 - 4 * 8];
};
#pragma pack(pop)
static_assert(sizeof(GadgetHeader) == 256, "Gadget header is 256 bytes");
}  // namespace

GadgetSource::GadgetSource(std::string filename, bool verbose) : base_(filename), verbose_(verbose) {
  load_file(filename, true);
  if (files_ > 1) {
    base_ = filename.substr(0, filename.find_last_of('.'));
    if (verbose_) std::printf("Detected %u files.\n", files_);
    if (current_file_ != 0) load_file(base_ + ".0", false);
  }
}

GadgetSource::~GadgetSource() {
  if (fp_) std::fclose(fh(fp_));
}

void GadgetSource::load_file(const std::string &fname, bool initial) {
  if (fp_) std::fclose(fh(fp_));
  fp_ = std::fopen(fname.c_str(), "rb");
  if (!fp_) {
    std::printf("WARNING: Could not open file %s\n", fname.c_str());
    n_points_file_ = 0;
    return;
  }
  if (verbose_) std::printf("Opened file %s\n", fname.c_str());
  GadgetHeader hd;
  std::fseek(fh(fp_), 4, SEEK_SET);  // block marker
  if (std::fread(&hd, sizeof(hd), 1, fh(fp_)) != 1) throw std::runtime_error("GadgetSource: short header in " + fname);
  // skip the header's end marker, the position block's start marker and the gas particles
  std::fseek(fh(fp_), 4 + (long)hd.npart[0] * 12 + 4, SEEK_CUR);
  current_point_file_ = 0;
  n_points_file_ = hd.npart[1];
  const size_t dot = fname.find_last_of('.');
  current_file_ = dot == std::string::npos ? 0 : (unsigned)std::strtoul(fname.c_str() + dot + 1, nullptr, 10);
  if (initial) {
    files_ = (unsigned)std::max(1, hd.num_files);
    n_points_ = (size_t)hd.npartTotal[1];
    hubble_ = hd.HubbleParam;
    box_size_ = hd.BoxSize;
  }
}

long long GadgetSource::read(Point *out, size_t n) {
  const size_t to_load = std::min(n, n_points_ - current_point_);
  if (to_load == 0 || !fp_) return -1;
  const double shift = -box_size_ / 2.0;
  size_t loaded = 0;
  while (loaded < to_load) {
    float rec[3];
    if (std::fread(rec, sizeof(rec), 1, fh(fp_)) != 1) break;
    out[loaded++] = point_from_box_position(rec[0], rec[1], rec[2], shift);
    ++current_point_;
    ++current_point_file_;
    if (current_point_file_ == n_points_file_) {
      if (current_file_ + 1 >= files_) {
        if (loaded < to_load)
          std::printf("WARNING: Did not find as many points as expected (%zu instead of %zu).\n", current_point_,
                      n_points_);
        break;
      }
      load_file(base_ + "." + std::to_string(current_file_ + 1), false);
    }
  }
  return loaded == 0 ? -1 : (long long)loaded;
}

void GadgetSource::reset() {
  load_file(files_ > 1 ? base_ + ".0" : base_, false);
  current_point_ = 0;
}

// ---- sinks ----

long long PointContainerSink::write(const Point *in, size_t n) {
  pc_.insert(pc_.end(), in, in + n);
  return (long long)n;
}

long long PixelizedPointContainerSink::write(const Point *in, size_t n) {
  for (size_t i = 0; i < n; ++i) pix_.add_point(in[i]);
  return (long long)n;
}

TextSink::TextSink(std::string filename, TextFormat format, bool verbose) : format_(format) {
  fp_ = std::fopen(filename.c_str(), "w");
  if (!fp_)
    std::printf("WARNING: Could not create file %s\n", filename.c_str());
  else if (verbose)
    std::printf("Created file %s\n", filename.c_str());
}

TextSink::~TextSink() { close(); }

void TextSink::close() {
  if (fp_) std::fclose(fh(fp_));
  fp_ = nullptr;
}

long long TextSink::write(const Point *in, size_t n) {
  if (!fp_) return -1;
  for (size_t i = 0; i < n; ++i) {
    const Point &p = in[i];
    if (format_ == TextFormat::CARTESIAN) {
      const Unit u = unit_vector(p);
      std::fprintf(fh(fp_), "%.16e %.16e %.16e\n", p.r * u.x, p.r * u.y, p.r * u.z);
    } else if (format_ == TextFormat::SPHERICAL) {
      std::fprintf(fh(fp_), "%.16e %.16e %.16e\n", p.r, p.p.theta, p.p.phi);
    } else {
      std::fprintf(fh(fp_), " %18.14F %18.14F %18.12F\n", p.p.phi, p.p.theta, p.r);
    }
  }
  std::fflush(fh(fp_));
  return (long long)n;
}

RawBinarySink::RawBinarySink(std::string filename, bool verbose, bool append, RecordType type) : type_(type) {
  fp_ = std::fopen(filename.c_str(), append ? "ab" : "wb");
  if (!fp_)
    std::printf("WARNING: Could not create file %s\n", filename.c_str());
  else if (verbose)
    std::printf("Opened RawBinarySink %s in %s-mode\n", filename.c_str(), append ? "append" : "overwrite");
}

RawBinarySink::~RawBinarySink() { close(); }

void RawBinarySink::close() {
  if (fp_) std::fclose(fh(fp_));
  fp_ = nullptr;
}

long long RawBinarySink::write(const Point *in, size_t n) {
  if (!fp_) return -1;
  const size_t rb = record_bytes(type_);
  std::vector<unsigned char> buf(n * rb);
  for (size_t i = 0; i < n; ++i) point_to_record(in[i], type_, buf.data() + i * rb);
  std::fwrite(buf.data(), rb, n, fh(fp_));
  std::fflush(fh(fp_));
  return (long long)n;
}

// ---- filters ----

const void *Filter::upload(const void *host, size_t bytes) const {
  if (bytes == 0) return nullptr;
  if (bytes > dev_bytes_) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc (filter)");
    dev_ = std::shared_ptr<void>(p, [](void *q) { (void)hipFree(q); });
    dev_bytes_ = bytes;
  }
  hip_check(hipMemcpy(dev_.get(), host, bytes, hipMemcpyHostToDevice), "copy filter data");
  return dev_.get();
}

size_t Filter::operator()(PointContainer &pc, bool resize) {
  Context &c = ctx();
  StreamBufs &b = stream_bufs();
  b.all_count = 0;
  const size_t kept = filter_chunk(c, b, pc.data(), pc.size(), {this});
  if (kept) hip_check(hipMemcpy(pc.data(), b.all.p, kept * sizeof(Point), hipMemcpyDeviceToHost), "copy points");
  if (resize) pc.resize(kept);
  return kept;
}

FilterSpec TophatRadialWindowFunctionFilter::spec() const {
  FilterSpec s;
  s.kind = SFBHIP_FILTER_TOPHAT;
  s.a = R0_;
  return s;
}

GenericRadialWindowFunctionFilter::GenericRadialWindowFunctionFilter(std::function<double(double)> window_function)
    : w_(std::move(window_function)) {}

GenericRadialWindowFunctionFilter::GenericRadialWindowFunctionFilter(std::function<double(double)> window_function,
                                                                     size_t interpolation_points, double min,
                                                                     double max, bool /*parallel_init*/)
    : x_min_(min), x_max_(max) {
  FunctionInterpolator fi(std::move(window_function), (unsigned)interpolation_points, min, max);
  table_ = fi.samples();
}

void GenericRadialWindowFunctionFilter::prepare(const Point *pts, size_t n) {
  if (!table_.empty()) return;
  values_.resize(n);
  for (size_t i = 0; i < n; ++i) values_[i] = w_(pts[i].r);
}

FilterSpec GenericRadialWindowFunctionFilter::spec() const {
  FilterSpec s;
  if (!table_.empty()) {
    s.kind = SFBHIP_FILTER_WINDOW_TABLE;
    s.x_min = x_min_;
    s.x_max = x_max_;
    s.n = (int64_t)table_.size();
    s.d_values = upload(table_.data(), table_.size() * sizeof(double));
  } else {
    s.kind = SFBHIP_FILTER_WINDOW_VALUES;
    s.n = (int64_t)values_.size();
    s.d_values = upload(values_.data(), values_.size() * sizeof(double));
  }
  return s;
}

FilterSpec GaussianRadialWindowFunctionFilter::spec() const {
  FilterSpec s;
  s.kind = SFBHIP_FILTER_GAUSSIAN;
  s.a = R0_;
  return s;
}

AngularMaskFilter::AngularMaskFilter(const std::vector<float> &healpix_map) : map_(healpix_map) {
  const unsigned nside = (unsigned)std::lround(std::sqrt(map_.size() / 12.0));
  if ((size_t)12 * nside * nside != map_.size() || !is_power_of_2(nside))
    throw std::runtime_error("healpix_map is not a valid map (check size!)");
  nside_ = nside;
}

AngularMaskFilter::AngularMaskFilter(const std::string &healpix_mask_fits) {
  throw std::runtime_error("AngularMaskFilter(" + healpix_mask_fits +
                           "): FITS maps are not read by this build (no cfitsio); pass the RING map as an array");
}

FilterSpec AngularMaskFilter::spec() const {
  FilterSpec s;
  s.kind = SFBHIP_FILTER_ANGULAR_MASK;
  s.nside = (int)nside_;
  s.n = (int64_t)map_.size();
  s.d_values = upload(map_.data(), map_.size() * sizeof(float));
  return s;
}

// ---- FilterStream ----

FilterStream::FilterStream(Source *source, Sink *sink, size_t buffer_size, bool verbose)
    : source_(source), sink_(sink), buffer_size_(std::max<size_t>(buffer_size, 1)), verbose_(verbose) {}

size_t FilterStream::run_device(size_t subsample_size, const double **d_points) {
  if (!source_) throw std::invalid_argument("FilterStream: no source");
  Context &c = ctx();
  StreamBufs &b = stream_bufs();  // per thread: only the handle needs the lock
  b.all_count = 0;
  std::vector<Point> buf(buffer_size_);
  long long k;
  size_t n_read = 0;
  while ((k = source_->read(buf.data(), buf.size())) != -1) {
    n_read += (size_t)k;
    const size_t kept = filter_chunk(c, b, buf.data(), (size_t)k, filters_);
    if (verbose_)
      std::printf("\tPoints read: %10lld (%10zu)\twritten: %10zu ( %10zu)\n", k, n_read, kept, b.all_count);
  }
  size_t n = b.all_count;
  const double *out = static_cast<const double *>(b.all.p);
  if (subsample_size != 0 && subsample_size < n) {
    std::lock_guard<std::mutex> lk(c.mu);
    b.sub.reserve(subsample_size * sizeof(Point));
    const uint64_t seed = g_seed.load() ^ (0xd1b54a32d192ed03ULL * (g_draws.fetch_add(1) + 1));
    throw_status(sfbhip_random_subset(c.h, out, (int64_t)n, (int64_t)subsample_size, seed,
                                      static_cast<double *>(b.sub.p), nullptr),
                 "random subset");
    n = subsample_size;
    out = static_cast<const double *>(b.sub.p);
  }
  if (d_points) *d_points = out;
  return n;
}

size_t FilterStream::run(size_t subsample_size, std::string /*temp_filename*/) {
  const double *d = nullptr;
  const size_t n = run_device(subsample_size, &d);
  if (!sink_) return n;
  std::vector<Point> host(n);
  if (n) hip_check(hipMemcpy(host.data(), d, n * sizeof(Point), hipMemcpyDeviceToHost), "copy points");
  const size_t chunk = buffer_size_;
  for (size_t i = 0; i < n; i += chunk) sink_->write(host.data() + i, std::min(chunk, n - i));
  return n;
}

size_t FilterStream::run_totemp(std::string temp_filename, bool append) {
  const double *d = nullptr;
  const size_t n = run_device(0, &d);
  std::vector<Point> host(n);
  if (n) hip_check(hipMemcpy(host.data(), d, n * sizeof(Point), hipMemcpyDeviceToHost), "copy points");
  RawBinarySink tmp(temp_filename, verbose_, append, RecordType::SPHERICAL_F64);
  tmp.write(host.data(), n);
  return n;
}

size_t FilterStream::run_fromtemp(std::string temp_filename, size_t subsample_size, bool remove_temp) {
  RawBinarySource tmp(temp_filename, verbose_, RecordType::SPHERICAL_F64);
  if (verbose_) std::printf("Temporary file %s contains %zu points.\n", temp_filename.c_str(), tmp.get_npoints());
  Source *keep_source = source_;
  std::vector<Filter *> keep_filters;
  keep_filters.swap(filters_);
  source_ = &tmp;
  size_t n = 0;
  try {
    n = run(subsample_size);
  } catch (...) {
    source_ = keep_source;
    filters_.swap(keep_filters);
    throw;
  }
  source_ = keep_source;
  filters_.swap(keep_filters);
  if (remove_temp) std::remove(temp_filename.c_str());
  return n;
}

}  // namespace io

// ---- Analyzer ----

namespace {
size_t analyzer_points(io::Source *source, const std::vector<io::Filter *> &filters, double rmax, unsigned subsample,
                       bool verbose, const double **d) {
  if (!source) throw std::invalid_argument("Analyzer: no source");
  source->reset();
  io::FilterStream fs(source, nullptr, 1000000, verbose);
  io::TophatRadialWindowFunctionFilter tophat(rmax);
  fs.add_filter(&tophat);
  for (io::Filter *f : filters) fs.add_filter(f);
  return fs.run_device(subsample, d);
}
}  // namespace

KClkk Analyzer::compute_sfb(unsigned short lmax, unsigned short nmax, double rmax, bool store_flmn, bool verbose) {
  const double *d = nullptr;
  const size_t n = analyzer_points(source_, filters_, rmax, subsample_, verbose, &d);
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  reserve_outputs(c, lmax, nmax, store_flmn);
  Options o;  // sfb_decomposition: FP64, the reference's table when N > 1000
  sfbhip_params p = make_params(lmax, nmax, rmax, verbose, o);
  throw_status(sfbhip_direct(c.h, d, (int64_t)n, &p, static_cast<double *>(c.kln.p), static_cast<double *>(c.cln.p),
                             store_flmn ? static_cast<double *>(c.flmn.p) : nullptr, nullptr),
               "Analyzer::compute_sfb");
  return download(c, lmax, nmax, store_flmn);
}

KClkk Analyzer::compute_sfb_pixelized(unsigned short lmax, unsigned short nmax, double rmax, unsigned int nside,
                                      bool store_flmn, bool verbose) {
  if (!is_power_of_2(nside)) throw std::invalid_argument("nside must be a power of 2");
  const double *d = nullptr;
  const size_t n = analyzer_points(source_, filters_, rmax, subsample_, verbose, &d);
  Context &c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const size_t npix = (size_t)12 * nside * nside;
  c.off.reserve((npix + 1) * sizeof(int64_t));
  c.rad.reserve(std::max<size_t>(n * sizeof(double), 8));
  throw_status(sfbhip_pixelize(c.h, d, (int64_t)n, (int32_t)nside, static_cast<int64_t *>(c.off.p),
                               static_cast<double *>(c.rad.p), nullptr),
               "Analyzer::compute_sfb_pixelized");
  reserve_outputs(c, lmax, nmax, store_flmn);
  Options o;
  sfbhip_params p = make_params(lmax, nmax, rmax, verbose, o);
  throw_status(sfbhip_healpix(c.h, static_cast<const int64_t *>(c.off.p), static_cast<const double *>(c.rad.p),
                              (int32_t)nside, &p, static_cast<double *>(c.kln.p), static_cast<double *>(c.cln.p),
                              store_flmn ? static_cast<double *>(c.flmn.p) : nullptr, nullptr),
               "Analyzer::compute_sfb_pixelized");
  return download(c, lmax, nmax, store_flmn);
}

}  // namespace catana
