// py_catana.cpp -- pybind11 module catana_amd._catana: the Python names of
// CatAna's basictypes and decomposition modules for the SFB path
// (python/py_basictypes.cpp:22-154, python/py_decomposition.cpp:27-153), plus
// the k_max / window-volume entry points sfb_decomposition_reverse and
// sfb_decomposition_healpix.  The GIL is released around device work.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

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
         double window_volume, bool store, bool verbose) {
        return sfb_decomposition_healpix(pc, lmax, k_max, rmax, nmax, window_volume, store, verbose);
      },
      py::arg("pixelized_point_container"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"), py::arg("nmax"),
      py::arg("window_volume") = 0.0, py::arg("store_flmn") = false, py::arg("verbose") = true,
      py::call_guard<py::gil_scoped_release>());
  m.def(
      "sfb_decomposition_healpix",
      [](const PointContainer &pc, unsigned int nside, unsigned short lmax, double k_max, double rmax,
         unsigned short nmax, double window_volume, bool store, bool verbose) {
        return sfb_decomposition_healpix(pc, nside, lmax, k_max, rmax, nmax, window_volume, store, verbose);
      },
      py::arg("point_container"), py::arg("nside"), py::arg("lmax"), py::arg("k_max"), py::arg("rmax"),
      py::arg("nmax"), py::arg("window_volume") = 0.0, py::arg("store_flmn") = false, py::arg("verbose") = true,
      py::call_guard<py::gil_scoped_release>());

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

  m.def(
      "compute_kln",
      [](unsigned short lmax, unsigned short nmax, double rmax) { return to_numpy(compute_kln(lmax, nmax, rmax)); },
      py::arg("lmax"), py::arg("nmax"), py::arg("rmax"));
}
