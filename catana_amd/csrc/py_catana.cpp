// py_catana.cpp -- pybind11 module catana_amd._catana: the Python names of
// CatAna's basictypes and decomposition modules for the SFB path
// (python/py_basictypes.cpp:22-154, python/py_decomposition.cpp:27-153), plus
// the k_max / window-volume entry points sfb_decomposition_reverse and
// sfb_decomposition_healpix.  The GIL is released around device work.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstring>
#include <stdexcept>

#include "catana_api.hpp"

namespace py = pybind11;
using namespace catana;

namespace {

py::array_t<double> to_numpy(const Array2D &a) {
  py::array_t<double> out({(py::ssize_t)a.rows, (py::ssize_t)a.cols});
  if (!a.data.empty()) std::memcpy(out.mutable_data(), a.data.data(), a.data.size() * sizeof(double));
  return out;
}

py::list flmn_to_list(const KClkk &k) {
  py::list out;
  for (const auto &b : k.f_lmn) {
    py::array_t<std::complex<double>> a({(py::ssize_t)b.rows, (py::ssize_t)b.cols});
    if (!b.data.empty()) std::memcpy(a.mutable_data(), b.data.data(), b.data.size() * sizeof(std::complex<double>));
    out.append(a);
  }
  return out;
}

// (N, 3) float64 row-major -> PointContainer; same checks and messages as py_basictypes.cpp:82-110.
PointContainer container_from_array(py::array_t<double> array, const std::string &coord_type) {
  bool spherical;
  if (coord_type == "cartesian")
    spherical = false;
  else if (coord_type == "spherical")
    spherical = true;
  else
    throw std::runtime_error("coord_type either 'cartesian' or 'spherical'");
  py::buffer_info info = array.request();
  if (info.ndim != 2) throw std::runtime_error("Incompatible number of dimensions of array (needs 2)");
  if (info.shape[1] != 3) throw std::runtime_error("Incompatible number of columns of array (needs 3)");
  const size_t n = (size_t)info.shape[0];
  PointContainer pc(n);
  if (n == 0) return pc;
  if (info.strides[1] != (py::ssize_t)sizeof(double))
    throw std::runtime_error("Data ordering of array is ColumnMajor, but RowMajor needed.");
  const char *base = static_cast<const char *>(info.ptr);
  const py::ssize_t rs = info.strides[0];
  for (size_t i = 0; i < n; ++i) {
    const double *c = reinterpret_cast<const double *>(base + (py::ssize_t)i * rs);
    pc[i] = spherical ? point_from_spherical_position(c[0], c[1], c[2]) : point_from_box_position(c[0], c[1], c[2]);
  }
  return pc;
}

// w at the quadrature nodes: one vectorised call when w accepts an array
// (a scalar result is broadcast), else one call per node.
std::vector<double> eval_on_nodes(const py::function &w, const std::vector<double> &r) {
  if (r.empty()) return {};
  try {
    py::array_t<double> ra((py::ssize_t)r.size(), r.data());
    py::object res = w(ra);
    auto a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(res);
    if (a && a.size() == (py::ssize_t)r.size()) return std::vector<double>(a.data(), a.data() + a.size());
    if (a && a.size() == 1) return std::vector<double>(r.size(), *a.data());
  } catch (py::error_already_set &) {
  }
  std::vector<double> out(r.size());
  for (size_t i = 0; i < r.size(); ++i) out[i] = w(r[i]).cast<double>();
  return out;
}

// numpy broadcasting of (k1, k2) as py::vectorize does (py_besseltools.cpp:45-78);
// a float for scalar inputs
template <class Run>
py::object vectorized_k(py::array_t<double> k1, py::array_t<double> k2, unsigned int l, Run run) {
  py::module_ np = py::module_::import("numpy");
  py::tuple b = np.attr("broadcast_arrays")(k1, k2);
  auto a1 = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(b[0]);
  auto a2 = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(b[1]);
  std::vector<double> v1(a1.data(), a1.data() + a1.size()), v2(a2.data(), a2.data() + a2.size());
  std::vector<int32_t> ls(v1.size(), (int32_t)l);
  std::vector<double> res = run(ls, v1, v2);
  if (a1.ndim() == 0) return py::float_(res.empty() ? 0.0 : res[0]);
  py::array_t<double> out(std::vector<py::ssize_t>(a1.shape(), a1.shape() + a1.ndim()));
  if (!res.empty()) std::memcpy(out.mutable_data(), res.data(), res.size() * sizeof(double));
  return std::move(out);
}

std::vector<double> integrals_py(const py::function &w, const std::vector<int32_t> &l, const std::vector<double> &k1,
                                 const std::vector<double> &k2, double rmax) {
  std::vector<double> r;
  {
    py::gil_scoped_release nogil;
    r = besseltools::sbessel_nodes(l, k1, k2, rmax);
  }
  std::vector<double> wv = eval_on_nodes(w, r);
  py::gil_scoped_release nogil;
  return besseltools::sbessel_integrals_at_nodes(l, k1, k2, rmax, wv);
}

// the reference's Python-visible record classes (python/py_io.cpp:104-250)
struct CartesianTextSource : io::TextSource {
  CartesianTextSource(const std::string &f, double shift) : io::TextSource(f, io::TextFormat::CARTESIAN, shift) {}
};
struct SphericalTextSource : io::TextSource {
  explicit SphericalTextSource(const std::string &f) : io::TextSource(f, io::TextFormat::SPHERICAL) {}
};
struct SphericalTextSource3dex : io::TextSource {
  explicit SphericalTextSource3dex(const std::string &f) : io::TextSource(f, io::TextFormat::SPHERICAL_3DEX) {}
};
struct CartesianTextSink : io::TextSink {
  CartesianTextSink(const std::string &f, bool v) : io::TextSink(f, io::TextFormat::CARTESIAN, v) {}
};
struct SphericalTextSink : io::TextSink {
  SphericalTextSink(const std::string &f, bool v) : io::TextSink(f, io::TextFormat::SPHERICAL, v) {}
};
struct SphericalTextSink3dex : io::TextSink {
  SphericalTextSink3dex(const std::string &f, bool v) : io::TextSink(f, io::TextFormat::SPHERICAL_3DEX, v) {}
};

int precision_code(const std::string &s) {
  if (s == "fp64") return 0;
  if (s == "mixed") return 1;
  throw std::invalid_argument("precision must be 'fp64' or 'mixed'");
}

}  // namespace

PYBIND11_MODULE(_catana, m) {
  m.doc() = "SFB decomposition of point catalogues on MI355X (CatAna's basictypes and decomposition API)";

  py::class_<Point>(m, "Point", py::buffer_protocol(),
                    "A point in spherical coordinates (r, theta, phi); np.array(point) gives (r, theta, phi).")
      .def(py::init<>())
      .def(py::init<double, double, double>(), py::arg("x"), py::arg("y"), py::arg("z"))
      .def("spherical", [](const Point &p) { return py::make_tuple(p.r, p.p.theta, p.p.phi); })
      .def("cartesian",
           [](const Point &p) {
             const double st = std::sin(p.p.theta);
             return py::make_tuple(p.r * st * std::cos(p.p.phi), p.r * st * std::sin(p.p.phi),
                                   p.r * std::cos(p.p.theta));
           })
      .def_buffer([](Point &p) -> py::buffer_info {
        return py::buffer_info(&p.r, sizeof(double), py::format_descriptor<double>::format(), 1, {3},
                               {sizeof(double)});
      });

  py::class_<PointContainer>(m, "PointContainer", py::buffer_protocol(),
                             "Collection of Point; np.array(pc) gives an (n_points, 3) array of (r, theta, phi).")
      .def(py::init<>())
      .def(py::init(&container_from_array), py::arg("coordinates"), py::arg("coordinate_type") = "spherical")
      .def("__getitem__",
           [](const PointContainer &pc, size_t i) {
             if (i >= pc.size()) throw py::index_error("index out of bounds!");
             return pc[i];
           })
      .def("add_point", &PointContainer::add_point, py::arg("point"))
      .def("__len__", [](const PointContainer &pc) { return pc.size(); })
      .def_buffer([](PointContainer &pc) -> py::buffer_info {
        return py::buffer_info(pc.empty() ? nullptr : &pc.data()->r, sizeof(double),
                               py::format_descriptor<double>::format(), 2, {(py::ssize_t)pc.size(), (py::ssize_t)3},
                               {(py::ssize_t)(3 * sizeof(double)), (py::ssize_t)sizeof(double)});
      });

  py::class_<PixelizedPointContainer>(m, "PixelizedPointContainer",
                                      "HEALPix (RING) pixelised collection of radii; nside must be a power of 2.")
      .def(py::init<unsigned int>(), py::arg("nside"))
      .def(py::init<unsigned int, const PointContainer &>(), py::arg("nside"), py::arg("point_container"),
           py::call_guard<py::gil_scoped_release>())
      .def("get_countmap",
           [](const PixelizedPointContainer &p) {
             const std::vector<int> v = p.get_countmap();
             py::array_t<int> a((py::ssize_t)v.size());
             if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(int));
             return a;
           })
      .def("add_point", &PixelizedPointContainer::add_point, py::arg("point"))
      .def("get_nside", &PixelizedPointContainer::get_nside)
      .def("get_npoints", &PixelizedPointContainer::get_npoints)
      .def("__len__", &PixelizedPointContainer::size);

  py::class_<KClkk>(m, "KClkk", "Return type of the SFB decomposition methods")
      .def(py::init<unsigned short, unsigned short, double>(), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"))
      .def_property_readonly("k_ln", [](const KClkk &k) { return to_numpy(k.k_ln); },
                             "k values at which C_l(k,k) is evaluated, shape (lmax, nmax)")
      .def_property_readonly("c_ln", [](const KClkk &k) { return to_numpy(k.c_ln); },
                             "C_l(k_ln, k_ln), shape (lmax, nmax)")
      .def_property_readonly("f_lmn", &flmn_to_list, "list over l of (l+1, nmax) complex arrays (m >= 0)")
      .def("savetxt", &KClkk::savetxt, py::arg("filename_base"));

  m.def(
      "sfb_decomposition",
      [](const PointContainer &pc, unsigned short lmax, unsigned short nmax, double rmax, bool store, bool verbose) {
        return sfb_decomposition(pc, lmax, nmax, rmax, store, verbose);
      },
      py::arg("point_container"), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"), py::arg("store_flmn") = false,
      py::arg("verbose") = true, py::call_guard<py::gil_scoped_release>(),
      "Direct SFB transform of a PointContainer (FP64 on the GPU; Bessel table when N > 1000).");
  m.def(
      "sfb_decomposition",
      [](const PixelizedPointContainer &pc, unsigned short lmax, unsigned short nmax, double rmax, bool store,
         bool verbose) { return sfb_decomposition(pc, lmax, nmax, rmax, store, verbose); },
      py::arg("pixelized_point_container"), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"),
      py::arg("store_flmn") = false, py::arg("verbose") = true, py::call_guard<py::gil_scoped_release>(),
      "SFB transform of a PixelizedPointContainer (HEALPix path).");

  m.def(
      "sfb_decomposition_reverse",
      [](const PointContainer &pc, unsigned short lmax, double k_max, double rmax, unsigned short nmax,
         double window_volume, bool store, bool verbose, const std::string &precision) {
        const int pr = precision_code(precision);
        py::gil_scoped_release rel;
        return sfb_decomposition_reverse(pc, lmax, k_max, rmax, nmax, window_volume, store, verbose, pr);
      },
      py::arg("point_container"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"), py::arg("nmax"),
      py::arg("window_volume") = 0.0, py::arg("store_flmn") = false, py::arg("verbose") = true,
      py::arg("precision") = "mixed",
      "Direct SFB with per-l truncation n_l = #{z_ln < k_max rmax} (nmax columns) and an optional window "
      "volume (0: 4/3 pi rmax^3).  precision 'mixed' (matrix cores, FP64 reduction) or 'fp64'.");
  m.def(
      "sfb_decomposition_healpix",
      [](const PixelizedPointContainer &pc, unsigned short lmax, double k_max, double rmax, unsigned short nmax,
         double window_volume, bool store, bool verbose, const std::string &precision) {
        const int pr = precision_code(precision);
        py::gil_scoped_release rel;
        return sfb_decomposition_healpix(pc, lmax, k_max, rmax, nmax, window_volume, store, verbose, pr);
      },
      py::arg("pixelized_point_container"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"), py::arg("nmax"),
      py::arg("window_volume") = 0.0, py::arg("store_flmn") = false, py::arg("verbose") = true,
      py::arg("precision") = "fp64",
      "HEALPix SFB with per-l truncation and an optional window volume.  precision 'fp64' (the pixel path) or "
      "'mixed' (the ring path: FP32 maps, split-FP16 ring contraction).");
  m.def(
      "sfb_decomposition_healpix",
      [](const PointContainer &pc, unsigned int nside, unsigned short lmax, double k_max, double rmax,
         unsigned short nmax, double window_volume, bool store, bool verbose, const std::string &precision) {
        const int pr = precision_code(precision);
        py::gil_scoped_release rel;
        return sfb_decomposition_healpix(pc, nside, lmax, k_max, rmax, nmax, window_volume, store, verbose, pr);
      },
      py::arg("point_container"), py::arg("nside"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"),
      py::arg("nmax"), py::arg("window_volume") = 0.0, py::arg("store_flmn") = false, py::arg("verbose") = true,
      py::arg("precision") = "fp64");

  m.def(
      "sfb_decomposition_reverse_batched",
      [](py::list catalogs, bool store, bool verbose, const std::string &precision) {
        const int pr = precision_code(precision);
        std::vector<CatalogSpec> specs;
        for (py::handle item : catalogs) {
          py::tuple t = py::reinterpret_borrow<py::tuple>(item);
          if (t.size() < 5 || t.size() > 6)
            throw std::invalid_argument(
                "each catalogue is (point_container, lmax, k_max, rmax, nmax[, window_volume])");
          CatalogSpec c;
          c.points = t[0].cast<const PointContainer *>();
          c.lmax = t[1].cast<unsigned short>();
          c.k_max = t[2].cast<double>();
          c.rmax = t[3].cast<double>();
          c.nmax = t[4].cast<unsigned short>();
          c.window_volume = t.size() > 5 ? t[5].cast<double>() : 0.0;
          specs.push_back(c);
        }
        py::gil_scoped_release rel;
        return sfb_decomposition_reverse_batched(specs, store, verbose, pr);
      },
      py::arg("catalogs"), py::arg("store_flmn") = false, py::arg("verbose") = true, py::arg("precision") = "mixed",
      "sfb_decomposition_reverse of several catalogues, given as (point_container, lmax, k_max, rmax, nmax"
      "[, window_volume]) tuples, in one batched device call; returns a list of KClkk.");
  m.def(
      "_sfb_raw",
      [](const PointContainer &pc, unsigned short lmax, unsigned short nmax, double rmax, bool store, bool verbose,
         bool parallel, bool interpolated) {
        return _sfb_raw(pc, lmax, nmax, rmax, store, verbose, parallel, interpolated);
      },
      py::arg("point_container"), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"), py::arg("store_flmn"),
      py::arg("verbose"), py::arg("parallel"), py::arg("interpolated"), py::call_guard<py::gil_scoped_release>());
  m.def(
      "_sfb_reverse_fft",
      [](const PixelizedPointContainer &pc, unsigned short lmax, unsigned short nmax, double rmax, bool store,
         bool verbose, bool parallel, bool interpolated) {
        return _sfb_reverse_fft(pc, lmax, nmax, rmax, store, verbose, parallel, interpolated);
      },
      py::arg("pixelized_point_container"), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"), py::arg("store_flmn"),
      py::arg("verbose"), py::arg("parallel"), py::arg("interpolated"), py::call_guard<py::gil_scoped_release>());
  m.def(
      "_sfb_reverse",
      [](const PixelizedPointContainer &pc, unsigned short lmax, unsigned short nmax, double rmax, bool store,
         bool verbose, bool parallel, bool interpolated) {
        return _sfb_reverse(pc, lmax, nmax, rmax, store, verbose, parallel, interpolated);
      },
      py::arg("pixelized_point_container"), py::arg("lmax"), py::arg("nmax"), py::arg("rmax"), py::arg("store_flmn"),
      py::arg("verbose"), py::arg("parallel"), py::arg("interpolated"), py::call_guard<py::gil_scoped_release>());

  py::class_<FunctionInterpolator>(m, "FunctionInterpolator",
                                   "A function sampled at equidistant points on [x_min, x_max] and interpolated "
                                   "linearly; x outside [x_min, x_max) raises ValueError")
      .def(py::init([](const py::function &fct, unsigned int interpolation_points, double x_min, double x_max) {
             return FunctionInterpolator([&fct](double x) { return fct(x).cast<double>(); }, interpolation_points,
                                         x_min, x_max, false);
           }),
           py::arg("function"), py::arg("interpolation_points"), py::arg("x_min"), py::arg("x_max"))
      .def("__call__", [](const FunctionInterpolator &fi, double x) { return fi(x); }, py::arg("x"))
      .def("__call__", [](const FunctionInterpolator &fi, py::array_t<double> x) {
             return py::vectorize([&fi](double v) { return fi(v); })(x);
           },
           py::arg("x"));

  py::module_ bt = m.def_submodule("besseltools", "Spherical Bessel zeros and double Bessel integrals "
                                                  "(python/py_besseltools.cpp:12-80)");
  py::class_<besseltools::SphericalBesselZeros>(bt, "SphericalBesselZeros",
                                                "Zeros of the spherical Bessel function j_l, computed on demand")
      .def(py::init<unsigned int>(), py::arg("l"))
      .def("__getitem__", &besseltools::SphericalBesselZeros::operator[], py::arg("n"))
      .def("compute_up_to", &besseltools::SphericalBesselZeros::compute_up_to, py::arg("n"));
  bt.def(
      "double_sbessel_integrator",
      [](const FunctionInterpolator &fi, unsigned int l, double r_max, py::array_t<double> k1, py::array_t<double> k2) {
        return vectorized_k(k1, k2, l, [&](const std::vector<int32_t> &ls, const std::vector<double> &a,
                                           const std::vector<double> &b) {
          py::gil_scoped_release nogil;
          return besseltools::double_sbessel_integrals(fi, ls, a, b, r_max);
        });
      },
      "int_0^r_max w(r) j_l(k1 r) j_l(k2 r) dr with w a FunctionInterpolator (interpolated on the device); "
      "vectorised over k1, k2",
      py::arg("w"), py::arg("l"), py::arg("r_max"), py::arg("k1"), py::arg("k2"));
  bt.def(
      "double_sbessel_integrator",
      [](const py::function &w, unsigned int l, double r_max, py::array_t<double> k1, py::array_t<double> k2) {
        return vectorized_k(k1, k2, l, [&](const std::vector<int32_t> &ls, const std::vector<double> &a,
                                           const std::vector<double> &b) { return integrals_py(w, ls, a, b, r_max); });
      },
      "int_0^r_max w(r) j_l(k1 r) j_l(k2 r) dr, w evaluated at the quadrature nodes (one vectorised call when it "
      "accepts an array); vectorised over k1, k2",
      py::arg("w"), py::arg("l"), py::arg("r_max"), py::arg("k1"), py::arg("k2"));
  bt.def(
      "double_sbessel_integrals",
      [](const py::function &w, std::vector<int32_t> l, std::vector<double> k1, std::vector<double> k2, double r_max) {
        return integrals_py(w, l, k1, k2, r_max);
      },
      "The same for arrays of triples (l[i], k1[i], k2[i]) in one device call", py::arg("w"), py::arg("l"),
      py::arg("k1"), py::arg("k2"), py::arg("r_max"));

  m.def(
      "shot_noise",
      [](const FunctionInterpolator &w, unsigned short lmax, double k_max, double rmax, unsigned short nmax,
         double n_points, double window_volume) {
        Array2D a;
        {
          py::gil_scoped_release nogil;
          a = shot_noise(w, lmax, k_max, rmax, nmax, n_points, window_volume);
        }
        return to_numpy(a);
      },
      "Shot-noise term (V/N)(2/pi) k_ln^2 int_0^rmax r^2 w(r) j_l(k_ln r)^2 dr of c_ln, (lmax, nmax)",
      py::arg("w"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"), py::arg("nmax"), py::arg("n_points"),
      py::arg("window_volume") = 0.0);
  m.def(
      "shot_noise",
      [](const py::function &w, unsigned short lmax, double k_max, double rmax, unsigned short nmax, double n_points,
         double window_volume) {
        if (!(n_points > 0)) throw std::invalid_argument("shot_noise: n_points must be positive");
        std::vector<int32_t> ls, ns;
        std::vector<double> ks;
        kept_modes(lmax, k_max, rmax, nmax, &ls, &ns, &ks);
        if (ls.empty()) return to_numpy(Array2D(lmax, nmax));
        // the integrand r^2 w(r): w is evaluated at the nodes, r^2 applied here
        py::function w2 = py::cpp_function([w](py::array_t<double> r) {
          py::array_t<double> wr = py::array_t<double, py::array::forcecast>::ensure(w(r));
          auto rr = r.unchecked<1>();
          py::array_t<double> out(rr.shape(0));
          auto o = out.mutable_unchecked<1>();
          if (wr.size() == rr.shape(0)) {
            const double *wp = wr.data();
            for (py::ssize_t i = 0; i < rr.shape(0); ++i) o(i) = rr(i) * rr(i) * wp[i];
          } else {
            const double c = *wr.data();
            for (py::ssize_t i = 0; i < rr.shape(0); ++i) o(i) = rr(i) * rr(i) * c;
          }
          return out;
        });
        std::vector<double> I;
        try {
          I = integrals_py(w2, ls, ks, ks, rmax);
        } catch (py::error_already_set &) {
          // w does not take arrays: per-node calls
          py::function w1 = py::cpp_function([w](double r) { return r * r * w(r).cast<double>(); });
          I = integrals_py(w1, ls, ks, ks, rmax);
        }
        const double V = window_volume > 0 ? window_volume : 4.0 / 3.0 * M_PI * rmax * rmax * rmax;
        Array2D out(lmax, nmax);
        for (size_t i = 0; i < I.size(); ++i) out(ls[i], ns[i]) = V / n_points * (2.0 / M_PI) * ks[i] * ks[i] * I[i];
        return to_numpy(out);
      },
      "Shot-noise term (V/N)(2/pi) k_ln^2 int_0^rmax r^2 w(r) j_l(k_ln r)^2 dr of c_ln, (lmax, nmax)",
      py::arg("w"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"), py::arg("nmax"), py::arg("n_points"),
      py::arg("window_volume") = 0.0);

  // ---- catana.io (python/py_io.cpp) ----
  py::module_ io_m = m.def_submodule("io", "Sources, sinks and point filters (python/py_io.cpp); filtering runs on "
                                           "the device");
  io_m.def("init_random", py::overload_cast<unsigned int>(&init_random), py::arg("seed"));
  io_m.def("init_random", py::overload_cast<>(&init_random));
  m.def("init_random", py::overload_cast<unsigned int>(&init_random), py::arg("seed"));
  m.def("init_random", py::overload_cast<>(&init_random));
  py::class_<io::Source>(io_m, "Source", "(Abstract) parent class for Source implementations.")
      .def("reset", &io::Source::reset)
      .def("get_point_container", &io::Source::get_point_container)
      .def("get_pixelized_point_container", &io::Source::get_pixelized_point_container, py::arg("nside"));
  py::class_<io::Sink>(io_m, "Sink", "(Abstract) parent class for Sink implementations")
      .def("write", [](io::Sink &s, const PointContainer &pc) { return s.write(pc); }, py::arg("point_container"));
  py::class_<io::Filter>(io_m, "Filter", "(Abstract) parent class for point-filtering implementations.")
      .def("__call__", [](io::Filter &f, PointContainer &pc) { return f(pc); }, py::arg("point_container"));
  py::class_<io::PointContainerSource, io::Source>(io_m, "PointContainerSource")
      .def(py::init<const PointContainer &>(), py::arg("point_container"), py::keep_alive<1, 2>());
  py::class_<io::GadgetSource, io::Source>(io_m, "GadgetSource")
      .def(py::init<std::string, bool>(), py::arg("filename"), py::arg("verbose") = true);
  py::class_<CartesianTextSource, io::Source>(io_m, "CartesianTextSource")
      .def(py::init<std::string, double>(), py::arg("filename"), py::arg("shift") = 0.0);
  py::class_<SphericalTextSource, io::Source>(io_m, "SphericalTextSource")
      .def(py::init<std::string>(), py::arg("filename"));
  py::class_<SphericalTextSource3dex, io::Source>(io_m, "SphericalTextSource_3dex")
      .def(py::init<std::string>(), py::arg("filename"));
  py::class_<io::RawBinarySource, io::Source>(io_m, "RawBinarySource")
      .def(py::init([](std::string f, bool v) { return new io::RawBinarySource(f, v, io::RecordType::SPHERICAL_F32); }),
           py::arg("filename"), py::arg("verbose") = true);
  py::class_<io::PointContainerSink, io::Sink>(io_m, "PointContainerSink")
      .def(py::init<PointContainer &>(), py::arg("point_container"), py::keep_alive<1, 2>());
  py::class_<io::PixelizedPointContainerSink, io::Sink>(io_m, "PixelizedPointContainerSink")
      .def(py::init<PixelizedPointContainer &>(), py::arg("pixelized_point_container"), py::keep_alive<1, 2>());
  py::class_<CartesianTextSink, io::Sink>(io_m, "CartesianTextSink")
      .def(py::init<std::string, bool>(), py::arg("filename"), py::arg("verbose") = true)
      .def("close", &io::TextSink::close);
  py::class_<SphericalTextSink, io::Sink>(io_m, "SphericalTextSink")
      .def(py::init<std::string, bool>(), py::arg("filename"), py::arg("verbose") = true)
      .def("close", &io::TextSink::close);
  py::class_<SphericalTextSink3dex, io::Sink>(io_m, "SphericalTextSink_3dex")
      .def(py::init<std::string, bool>(), py::arg("filename"), py::arg("verbose") = true)
      .def("close", &io::TextSink::close);
  py::class_<io::RawBinarySink, io::Sink>(io_m, "RawBinarySink")
      .def(py::init([](std::string f, bool v, bool a) {
             return new io::RawBinarySink(f, v, a, io::RecordType::SPHERICAL_F32);
           }),
           py::arg("filename"), py::arg("verbose") = true, py::arg("append") = false)
      .def("close", &io::RawBinarySink::close);
  py::class_<io::TophatRadialWindowFunctionFilter, io::Filter>(io_m, "TophatRadialWindowFunctionFilter")
      .def(py::init<double>(), py::arg("R0"));
  py::class_<io::GenericRadialWindowFunctionFilter, io::Filter>(io_m, "GenericRadialWindowFunctionFilter")
      .def(py::init<std::function<double(double)>>(), py::arg("function"))
      .def(py::init([](std::function<double(double)> w, size_t n, double lo, double hi) {
             return new io::GenericRadialWindowFunctionFilter(std::move(w), n, lo, hi, false);
           }),
           py::arg("function"), py::arg("interpolation_points"), py::arg("min"), py::arg("max"));
  py::class_<io::GaussianRadialWindowFunctionFilter, io::GenericRadialWindowFunctionFilter>(
      io_m, "GaussianRadialWindowFunctionFilter")
      .def(py::init<double>(), py::arg("R0"));
  py::class_<io::AngularMaskFilter, io::Filter>(io_m, "AngularMaskFilter")
      .def(py::init<const std::string &>(), py::arg("mask_file"))
      .def(py::init([](py::array_t<float, py::array::c_style | py::array::forcecast> a) {
             return new io::AngularMaskFilter(std::vector<float>(a.data(), a.data() + a.size()));
           }),
           py::arg("map"));
  py::class_<io::FilterStream>(io_m, "FilterStream",
                               "Connects a Source to a Sink through Filters; the filtering runs on the device")
      .def(py::init<io::Source *, io::Sink *, size_t, bool>(), py::arg("source"), py::arg("sink"),
           py::arg("buffer_size") = 1000000, py::arg("verbose") = true, py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>())
      .def("add_filter", &io::FilterStream::add_filter, py::arg("filter"), py::keep_alive<1, 2>())
      .def("set_source", &io::FilterStream::set_source, py::arg("source"), py::keep_alive<1, 2>())
      .def("run", &io::FilterStream::run, py::arg("subsample_size") = 0,
           py::arg("temp_filename") = std::string("tmp.bin"))
      .def("run_totemp", &io::FilterStream::run_totemp, py::arg("temp_filename") = std::string("tmp.bin"),
           py::arg("append") = true)
      .def("run_fromtemp", &io::FilterStream::run_fromtemp, py::arg("temp_filename") = std::string("tmp.bin"),
           py::arg("subsample_size") = 0, py::arg("remove_temp") = true);

  py::class_<Analyzer>(m, "Analyzer", "Source + filters + SFB decomposition, the filtered catalogue kept on the device")
      .def(py::init<io::Source *>(), py::arg("source"), py::keep_alive<1, 2>())
      .def("set_source", &Analyzer::set_source, py::arg("source"), py::keep_alive<1, 2>())
      .def("set_subsample_size", &Analyzer::set_subsample_size, py::arg("n"))
      .def("add_filter", &Analyzer::add_filter, py::arg("Filter"), py::keep_alive<1, 2>())
      .def("compute_sfb", &Analyzer::compute_sfb, py::arg("lmax"), py::arg("nmax"), py::arg("rmax"),
           py::arg("store_flmn") = false, py::arg("verbose") = true)
      .def("compute_sfb_pixelized", &Analyzer::compute_sfb_pixelized, py::arg("lmax"), py::arg("nmax"),
           py::arg("rmax"), py::arg("nside"), py::arg("store_flmn") = false, py::arg("verbose") = true);

  m.def(
      "compute_kln",
      [](unsigned short lmax, unsigned short nmax, double rmax) { return to_numpy(compute_kln(lmax, nmax, rmax)); },
      py::arg("lmax"), py::arg("nmax"), py::arg("rmax"));
}
