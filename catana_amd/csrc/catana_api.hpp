// catana_api.hpp -- host C++ API with the names and argument meanings of
// CatAna's decomposition interface (catana/include/catana/decomposition/,
// catana/include/catana/points/), implemented over the C-ABI of
// libcatana_hip.so.  Errors are the reference's exception types:
// std::invalid_argument (nside, sizes), std::domain_error (Bessel argument out
// of the interpolation range), std::runtime_error (device failures).
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace catana {

// Angular position (theta from the z axis, phi in [0, 2 pi)); the layout of HEALPix's pointing.
struct pointing {
  double theta = 0.0, phi = 0.0;
};

// Point.hpp:10-19: {r, pointing} -- 24 bytes, the AoS the kernels read.
struct Point {
  Point() = default;
  Point(double x, double y, double z);  // from cartesian coordinates (Point.cpp:5-10)
  double r = 0.0;
  pointing p;
};
static_assert(sizeof(Point) == 3 * sizeof(double), "Point must be the 24-byte (r, theta, phi) record");

Point point_from_spherical_position(double r, double theta, double phi);
Point point_from_box_position(double x, double y, double z, double shift = 0.0);

// PointContainer.hpp:11-19
class PointContainer : public std::vector<Point> {
 public:
  using std::vector<Point>::vector;
  void add_point(const Point &p) { push_back(p); }
};

// Row-major 2-D array of doubles (the reference uses Eigen arrays).
struct Array2D {
  Array2D() = default;
  Array2D(std::size_t r, std::size_t c) : rows(r), cols(c), data(r * c, 0.0) {}
  double &operator()(std::size_t i, std::size_t j) { return data[i * cols + j]; }
  double operator()(std::size_t i, std::size_t j) const { return data[i * cols + j]; }
  std::size_t rows = 0, cols = 0;
  std::vector<double> data;
};

// Complex (l+1) x nmax block, row-major (m, n).
struct ComplexBlock {
  std::size_t rows = 0, cols = 0;
  std::vector<std::complex<double>> data;
};

// return_types.hpp:19-38
struct KClkk {
  KClkk(unsigned short lmax, unsigned short nmax, double rmax);  // k_ln from the zeros of j_l, c_ln = 0
  KClkk() = default;
  Array2D k_ln;                     // k at which C_l(k, k) is evaluated, (lmax, nmax)
  Array2D c_ln;                     // C_l(k_ln, k_ln), (lmax, nmax)
  std::vector<ComplexBlock> f_lmn;  // per l: (l+1) x nmax, m >= 0 (when store_flmn)
  // "<base>.k_ln" and "<base>.c_ln": rows = l, precision 18, scientific (return_types.cpp:25-64)
  void savetxt(const std::string &filename_base) const;
};

// PixelizedPointContainer.hpp: HEALPix RING map of radii per pixel.
class PixelizedPointContainer {
 public:
  explicit PixelizedPointContainer(unsigned int nside);  // throws std::invalid_argument if not a power of 2
  PixelizedPointContainer(unsigned int nside, const PointContainer &pc);  // binned on the GPU
  void add_point(const Point &p);
  unsigned int get_nside() const { return nside_; }
  std::size_t get_npoints() const;
  std::size_t size() const { return pixels_.size(); }
  std::vector<int> get_countmap() const;
  const std::vector<double> &pixel(std::size_t i) const { return pixels_[i]; }
  // CSR view: offsets (npix + 1) and radii in pixel order
  void to_csr(std::vector<int64_t> &offsets, std::vector<double> &radii) const;

 private:
  unsigned int nside_;
  std::vector<std::vector<double>> pixels_;
};

// How the call is evaluated on the GPU (extensions; defaults reproduce the reference).
struct Options {
  int precision = 0;         // 0 FP64 (drop-in), 1 mixed
  int interpolated = -1;     // -1: reference rule N > 1000; 0 exact j_l; 1 reference linear table
  int zeros = 0;             // 0 reference Newton loop, 1 converged
  double k_max = 0.0;        // > 0: per-l truncation n_l = #{z_ln < k_max rmax}
  double window_volume = 0;  // > 0 replaces (4/3) pi rmax^3
  bool alias_compat = true;  // pixelised path: keep the reference's unconjugated fold
};

// sfb_decomposition.hpp:29-47 (public wrappers: interpolated = N > 1000)
KClkk sfb_decomposition(const PointContainer &points, unsigned short lmax, unsigned short nmax, double rmax,
                        bool store_flmn, bool verbose, const Options &opt = Options());
KClkk sfb_decomposition(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                        bool store_flmn, bool verbose, const Options &opt = Options());

// sfb_raw.hpp / sfb_reverse.hpp / sfb_reverse_fft.hpp (parallel has no effect: the GPU is always parallel)
KClkk _sfb_raw(const PointContainer &points, unsigned short lmax, unsigned short nmax, double rmax,
               bool store_flmn, bool verbose, bool parallel, bool interpolated);
KClkk _sfb_reverse_fft(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                       bool store_flmn, bool verbose, bool parallel, bool interpolated);
KClkk _sfb_reverse(const PixelizedPointContainer &pix, unsigned short lmax, unsigned short nmax, double rmax,
                   bool store_flmn, bool verbose, bool parallel, bool interpolated);

// Variants with a k_max cut and a window volume (n_l per l; nmax bounds the columns);
// j_l is evaluated exactly (no interpolation table) whatever N.
KClkk sfb_decomposition_reverse(const PointContainer &points, unsigned short lmax, double k_max, double rmax,
                                unsigned short nmax, double window_volume, bool store_flmn, bool verbose,
                                int precision = 1);
// precision 0: the FP64 pixel path; 1: the ring path (FP32 maps, split-FP16
// ring contraction; C_l within 1e-6 of FP64 for l >= 8, 5e-6 below).
KClkk sfb_decomposition_healpix(const PixelizedPointContainer &pix, unsigned short lmax, double k_max, double rmax,
                                unsigned short nmax, double window_volume, bool store_flmn, bool verbose,
                                int precision = 0);
KClkk sfb_decomposition_healpix(const PointContainer &points, unsigned int nside, unsigned short lmax,
                                double k_max, double rmax, unsigned short nmax, double window_volume,
                                bool store_flmn, bool verbose, int precision = 0);

// One catalogue of a batched decomposition (sfb_decomposition_reverse's arguments).
struct CatalogSpec {
  const PointContainer *points = nullptr;
  unsigned short lmax = 0;
  double k_max = 0.0;
  double rmax = 1.0;
  unsigned short nmax = 1;
  double window_volume = 0.0;
};

// sfb_decomposition_reverse of several catalogues in one batched device call
// (one grouped launch per kernel); result i belongs to cats[i].
std::vector<KClkk> sfb_decomposition_reverse_batched(const std::vector<CatalogSpec> &cats, bool store_flmn,
                                                     bool verbose, int precision = 1);

// Zeros-based k_ln table (return_types.cpp:10-19).
Array2D compute_kln(unsigned short lmax, unsigned short nmax, double rmax, int zeros = 0);

// A function sampled at interpolation_points equidistant points on
// [x_min, x_max] and interpolated linearly; arguments outside [x_min, x_max)
// throw std::domain_error (FunctionInterpolator.cpp:7-45).
class FunctionInterpolator {
 public:
  FunctionInterpolator(std::function<double(double)> fct, unsigned int interpolation_points, double x_min,
                       double x_max, bool parallel = false);
  double operator()(double x) const;
  double operator[](unsigned int index) const { return y_[index]; }
  unsigned int get_index(double x) const;
  const std::vector<double> &samples() const { return y_; }
  double x_min() const { return x_min_; }
  double x_max() const { return x_max_; }

 private:
  std::vector<double> y_;
  double x_min_, x_max_, dx_inv_;
};

namespace besseltools {

// Zeros of j_l, computed on demand and kept (SphericalBesselZeros.cpp:11-101).
class SphericalBesselZeros {
 public:
  explicit SphericalBesselZeros(unsigned int l);
  double operator[](unsigned int n);
  void compute_up_to(double z_max);     // until the last zero is >= z_max
  unsigned int next_zero(double z);     // index of the first zero > z
  unsigned int l() const { return l_; }

 private:
  void grow(size_t n);
  unsigned int l_;
  std::vector<double> zeros_;
};

// int_0^Rmax f(r) j_l(k1 r) j_l(k2 r) dr (sbessel_integrator.cpp:45-110), on the device.
double double_sbessel_integrator(std::function<double(double)> f, const unsigned int &l, const double &Rmax, double k1,
                                 double k2);
double double_sbessel_integrator_bz(std::function<double(double)> f, const unsigned int &l, SphericalBesselZeros &bz,
                                    const double &Rmax, double k1, double k2);
double double_sbessel_integrator(const FunctionInterpolator &fi, const unsigned int &l, const double &Rmax, double k1,
                                 double k2);

// Many triples (l[i], k1[i], k2[i]) in one device call.  f is evaluated on the
// host at every quadrature node (sbessel_nodes); the interpolator overload
// runs entirely on the device.
std::vector<double> double_sbessel_integrals(std::function<double(double)> f, const std::vector<int32_t> &l,
                                             const std::vector<double> &k1, const std::vector<double> &k2,
                                             double Rmax);
std::vector<double> double_sbessel_integrals(const FunctionInterpolator &fi, const std::vector<int32_t> &l,
                                             const std::vector<double> &k1, const std::vector<double> &k2,
                                             double Rmax);
// The two halves of the callable path, for callers that evaluate f in bulk.
std::vector<double> sbessel_nodes(const std::vector<int32_t> &l, const std::vector<double> &k1,
                                  const std::vector<double> &k2, double Rmax);
std::vector<double> sbessel_integrals_at_nodes(const std::vector<int32_t> &l, const std::vector<double> &k1,
                                               const std::vector<double> &k2, double Rmax,
                                               const std::vector<double> &f_at_nodes);

}  // namespace besseltools

// (l, n, k_ln) of every mode sfb_decomposition_reverse keeps (n < n_l with
// z_ln < k_max rmax; k_max <= 0: all nmax), in (l, n) order.
void kept_modes(unsigned short lmax, double k_max, double rmax, unsigned short nmax, std::vector<int32_t> *l,
                std::vector<int32_t> *n, std::vector<double> *k);

// Shot-noise term of c_ln (docs/catana_usage.rst:126-134):
//   C_l^sn(k_ln) = (V/N) (2/pi) k_ln^2 int_0^rmax r^2 w(r) j_l(k_ln r)^2 dr,
// V = window_volume or (4/3) pi rmax^3; n modes as sfb_decomposition_reverse
// keeps them (k_max <= 0: all nmax), the rest 0.  (lmax x nmax), row = l.
Array2D shot_noise(std::function<double(double)> w, unsigned short lmax, double k_max, double rmax,
                   unsigned short nmax, double n_points, double window_volume = 0.0);
Array2D shot_noise(const FunctionInterpolator &w, unsigned short lmax, double k_max, double rmax,
                   unsigned short nmax, double n_points, double window_volume = 0.0);

// Seed of the random point filters and subsets (tools/random.cpp init_random):
// every draw is a hash of (seed, filter, running record count), so a seeded
// sequence of runs is reproducible.  Without a call the seed is 0.
void init_random(unsigned int seed);
void init_random();  // a seed from std::random_device

namespace io {

// ---- sources (io/Source.hpp and io/sources/*): read() fills up to n points
// and returns how many, or -1 when nothing is left ----
class Source {
 public:
  virtual ~Source() = default;
  virtual long long read(Point *out, size_t n) = 0;
  virtual void reset() = 0;
  virtual size_t get_npoints() { return 0; }
  PointContainer get_point_container();
  PixelizedPointContainer get_pixelized_point_container(unsigned int nside);
};

class PointContainerSource : public Source {
 public:
  explicit PointContainerSource(const PointContainer &pc) : pc_(pc) {}
  long long read(Point *out, size_t n) override;
  void reset() override { pos_ = 0; }
  size_t get_npoints() override { return pc_.size(); }

 private:
  const PointContainer &pc_;
  size_t pos_ = 0;
};

// Text records, one per line (TextSource.cpp:10-60): "x y z" (+ shift on
// every axis), "r theta phi", or the 3dex order "phi theta r".
enum class TextFormat { CARTESIAN, SPHERICAL, SPHERICAL_3DEX };
class TextSource : public Source {
 public:
  TextSource(std::string filename, TextFormat format, double shift = 0.0);
  ~TextSource() override;
  long long read(Point *out, size_t n) override;
  void reset() override;

 private:
  std::string filename_;
  TextFormat format_;
  double shift_;
  void *fp_ = nullptr;
};

// Raw binary records (RawBinarySource.cpp:9-90): 3 floats or doubles each,
// cartesian (x, y, z) or spherical (r, theta, phi).
enum class RecordType { CARTESIAN_F32, CARTESIAN_F64, SPHERICAL_F32, SPHERICAL_F64 };
class RawBinarySource : public Source {
 public:
  RawBinarySource(std::string filename, bool verbose = false, RecordType type = RecordType::SPHERICAL_F32);
  ~RawBinarySource() override;
  long long read(Point *out, size_t n) override;
  void reset() override;
  size_t get_npoints() override { return ntot_; }

 private:
  std::string filename_;
  RecordType type_;
  void *fp_ = nullptr;
  size_t ntot_ = 0, pos_ = 0;
};

// Gadget (format 1) snapshots, one file or name.0 .. name.(N-1): the halo
// particles' float positions shifted by -BoxSize/2 (GadgetSource.cpp:38-177).
class GadgetSource : public Source {
 public:
  GadgetSource(std::string filename, bool verbose = false);
  ~GadgetSource() override;
  long long read(Point *out, size_t n) override;
  void reset() override;
  size_t get_npoints() override { return n_points_; }
  double box_size() const { return box_size_; }
  double hubble_parameter() const { return hubble_; }

 private:
  void load_file(const std::string &fname, bool initial);
  std::string base_;
  bool verbose_;
  void *fp_ = nullptr;
  unsigned files_ = 1, current_file_ = 0;
  size_t n_points_ = 0, current_point_ = 0, n_points_file_ = 0, current_point_file_ = 0;
  double box_size_ = 0.0, hubble_ = 0.0;
};

// ---- sinks (io/Sink.hpp and io/sinks/*): write() returns the count written, or -1 when closed ----
class Sink {
 public:
  virtual ~Sink() = default;
  virtual long long write(const Point *in, size_t n) = 0;
  long long write(const PointContainer &pc) { return write(pc.data(), pc.size()); }
};

class PointContainerSink : public Sink {
 public:
  explicit PointContainerSink(PointContainer &pc) : pc_(pc) {}
  long long write(const Point *in, size_t n) override;

 private:
  PointContainer &pc_;
};

class PixelizedPointContainerSink : public Sink {
 public:
  explicit PixelizedPointContainerSink(PixelizedPointContainer &pix) : pix_(pix) {}
  long long write(const Point *in, size_t n) override;

 private:
  PixelizedPointContainer &pix_;
};

// TextSink.cpp: "%.18e"-style columns in the source's order (the 3dex order
// with the reference's " %18.14F %18.14F %18.12F" format, record_type.cpp:82-86).
class TextSink : public Sink {
 public:
  TextSink(std::string filename, TextFormat format, bool verbose = false);
  ~TextSink() override;
  long long write(const Point *in, size_t n) override;
  void close();

 private:
  TextFormat format_;
  void *fp_ = nullptr;
};

class RawBinarySink : public Sink {
 public:
  RawBinarySink(std::string filename, bool verbose = false, bool append = false,
                RecordType type = RecordType::SPHERICAL_F32);
  ~RawBinarySink() override;
  long long write(const Point *in, size_t n) override;
  void close();

 private:
  RecordType type_;
  void *fp_ = nullptr;
};

// ---- filters (io/Filter.hpp, io/filters/*), evaluated on the device ----
// The device description of a filter: the fields of sfbhip_filter (kind as
// SFBHIP_FILTER_*, d_values in device memory the filter owns).
struct FilterSpec {
  int kind = 0, nside = 0;
  double a = 0.0, x_min = 0.0, x_max = 0.0;
  int64_t n = 0;
  const void *d_values = nullptr;
};

class Filter {
 public:
  virtual ~Filter() = default;
  // Filters the container in place on the device; the survivors keep their
  // input order (the reference swaps rejected points to the end).  Returns the
  // new size (Filter.cpp:7-23).
  size_t operator()(PointContainer &pc, bool resize = true);
  // host evaluation needed before the device pass (callables), per chunk
  virtual void prepare(const Point *pts, size_t n) {}
  virtual FilterSpec spec() const = 0;

 protected:
  // uploads host data into the filter's device buffer (grown as needed)
  const void *upload(const void *host, size_t bytes) const;

 private:
  mutable std::shared_ptr<void> dev_;
  mutable size_t dev_bytes_ = 0;
};

class TophatRadialWindowFunctionFilter : public Filter {
 public:
  explicit TophatRadialWindowFunctionFilter(double R0) : R0_(R0) {}
  FilterSpec spec() const override;

 private:
  double R0_;
};

class GenericRadialWindowFunctionFilter : public Filter {
 public:
  // acceptance probability w(r), evaluated on the host at every point
  explicit GenericRadialWindowFunctionFilter(std::function<double(double)> window_function);
  // w sampled at interpolation_points on [min, max] and interpolated on the device
  GenericRadialWindowFunctionFilter(std::function<double(double)> window_function, size_t interpolation_points,
                                    double min, double max, bool parallel_init = false);
  void prepare(const Point *pts, size_t n) override;
  FilterSpec spec() const override;

 protected:
  GenericRadialWindowFunctionFilter() = default;
  std::function<double(double)> w_;
  std::vector<double> table_, values_;
  double x_min_ = 0.0, x_max_ = 0.0;
};

class GaussianRadialWindowFunctionFilter : public GenericRadialWindowFunctionFilter {
 public:
  explicit GaussianRadialWindowFunctionFilter(double R0) : R0_(R0) {}
  void prepare(const Point *, size_t) override {}
  FilterSpec spec() const override;

 private:
  double R0_;
};

class AngularMaskFilter : public Filter {
 public:
  // HEALPix RING float map, 12 nside^2 pixels: 1 keeps a point, 0 removes it
  explicit AngularMaskFilter(const std::vector<float> &healpix_map);
  // FITS maps need cfitsio, which this build does not link: throws std::runtime_error
  explicit AngularMaskFilter(const std::string &healpix_mask_fits);
  FilterSpec spec() const override;

 private:
  std::vector<float> map_;
  unsigned nside_ = 0;
};

// Reads the source in chunks of buffer_size points, filters each chunk on the
// device and collects the survivors in device memory; a subsample of the
// survivors is drawn on the device; the result goes to the sink (host) or
// stays on the device (run_device).  The reference streams through a
// temporary file for the subsample (FilterStream.cpp:57-101); HBM holds the
// filtered catalogue instead.
class FilterStream {
 public:
  FilterStream(Source *source, Sink *sink, size_t buffer_size = 1000000, bool verbose = false);
  void set_source(Source *source) { source_ = source; }
  void add_filter(Filter *filter) { filters_.push_back(filter); }
  size_t run(size_t subsample_size = 0, std::string temp_filename = "");
  // the reference's two halves of a subsampled run (FilterStream.cpp:57-91):
  // filter into a raw file of (r, theta, phi) doubles, then read it back
  // through the random subset into the sink
  size_t run_totemp(std::string temp_filename, bool append);
  size_t run_fromtemp(std::string temp_filename, size_t subsample_size, bool remove_temp);
  // the filtered (and subsampled) points stay on the device: returns their
  // count; *d_points is valid until the next run on this thread
  size_t run_device(size_t subsample_size, const double **d_points);

 private:
  Source *source_;
  Sink *sink_;
  size_t buffer_size_;
  bool verbose_;
  std::vector<Filter *> filters_;
};

}  // namespace io

// Source + filters (+ tophat at rmax) + subsample + decomposition on the
// device without a host round trip of the filtered catalogue (Analyzer.cpp:10-62).
class Analyzer {
 public:
  Analyzer() = default;
  explicit Analyzer(io::Source *source) : source_(source) {}
  void set_source(io::Source *source) { source_ = source; }
  void set_subsample_size(unsigned int n) { subsample_ = n; }
  void add_filter(io::Filter *filter) { filters_.push_back(filter); }
  KClkk compute_sfb(unsigned short lmax, unsigned short nmax, double rmax, bool store_flmn, bool verbose);
  KClkk compute_sfb_pixelized(unsigned short lmax, unsigned short nmax, double rmax, unsigned int nside,
                              bool store_flmn, bool verbose);

 private:
  io::Source *source_ = nullptr;
  std::vector<io::Filter *> filters_;
  unsigned int subsample_ = 0;
};

}  // namespace catana
