// sanitize_host.cpp -- drives the host-only parts of the C++ API
// (catana_amd/csrc/catana_api.cpp: points, the text / raw binary / Gadget
// sources and sinks, FunctionInterpolator, filter descriptions, KClkk::savetxt)
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (oracle/Makefile, target sanitize).  Nothing here touches a GPU.  Test
// infrastructure only.  Argument: the directory of the Gadget fixture.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "catana_api.hpp"

using namespace catana;

static int failures = 0;

static void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

static std::vector<Point> drain(io::Source &s, size_t chunk) {
  std::vector<Point> all;
  std::vector<Point> buf(chunk);
  for (;;) {
    const long long n = s.read(buf.data(), chunk);
    if (n < 0) break;
    all.insert(all.end(), buf.begin(), buf.begin() + n);
    if (n == 0) break;
  }
  return all;
}

int main(int argc, char **argv) {
  const std::string gold = argc > 1 ? argv[1] : "tests/golden";
  const std::string tmp = "/tmp/sanitize_host_";

  // points from cartesian and spherical coordinates
  PointContainer pc;
  for (int i = 0; i < 257; ++i) {
    const double x = std::sin(0.37 * i) * 50.0, y = std::cos(0.11 * i) * 40.0, z = 0.3 * i - 30.0;
    const Point p(x, y, z);
    check(std::fabs(p.r - std::sqrt(x * x + y * y + z * z)) < 1e-9, "cartesian radius");
    pc.add_point(p);
  }
  pc.add_point(point_from_spherical_position(10.0, 0.5, 1.5));
  pc.add_point(point_from_box_position(1.0, 2.0, 3.0, -1.0));

  // text round trips in the three formats, read back in odd chunk sizes
  const io::TextFormat fmts[3] = {io::TextFormat::CARTESIAN, io::TextFormat::SPHERICAL,
                                  io::TextFormat::SPHERICAL_3DEX};
  for (int k = 0; k < 3; ++k) {
    const std::string f = tmp + "text" + std::to_string(k) + ".txt";
    {
      io::TextSink sink(f, fmts[k]);
      check(sink.write(pc.data(), pc.size()) == (long long)pc.size(), "text write");
      sink.close();
    }
    io::TextSource src(f, fmts[k]);
    const std::vector<Point> back = drain(src, 37);
    check(back.size() == pc.size(), "text read count");
    for (size_t i = 0; i < back.size() && i < pc.size(); ++i)
      check(std::fabs(back[i].r - pc[i].r) <= 1e-6 * (1.0 + pc[i].r), "text radius");
    src.reset();
    check(drain(src, 1000).size() == pc.size(), "text reset");
    std::remove(f.c_str());
  }

  // raw binary round trips in FP64 and FP32 records
  const io::RecordType types[2] = {io::RecordType::SPHERICAL_F64, io::RecordType::SPHERICAL_F32};
  for (int k = 0; k < 2; ++k) {
    const std::string f = tmp + "raw" + std::to_string(k) + ".bin";
    {
      io::RawBinarySink sink(f, false, false, types[k]);
      check(sink.write(pc.data(), pc.size()) == (long long)pc.size(), "raw write");
      sink.close();
    }
    io::RawBinarySource src(f, false, types[k]);
    check(src.get_npoints() == pc.size(), "raw count");
    const std::vector<Point> back = drain(src, 64);
    check(back.size() == pc.size(), "raw read count");
    for (size_t i = 0; i < back.size() && i < pc.size(); ++i)
      check(std::fabs(back[i].r - pc[i].r) <= 1e-5 * (1.0 + pc[i].r), "raw radius");
    std::remove(f.c_str());
  }

  // the reference's Gadget fixture: two files of one snapshot
  {
    io::GadgetSource g(gold + "/mock_data_z0p000.0");
    const std::vector<Point> pts = drain(g, 100);
    check(!pts.empty() && pts.size() == g.get_npoints(), "gadget read");
    check(g.box_size() > 0.0, "gadget box size");
  }

  // a source over a container
  {
    io::PointContainerSource s(pc);
    check(drain(s, 50).size() == pc.size(), "container source");
  }

  // FunctionInterpolator: values, and the domain error outside [x_min, x_max)
  {
    FunctionInterpolator fi([](double x) { return x * x; }, 101, 0.0, 10.0);
    check(std::fabs(fi(5.0) - 25.0) < 1e-2, "interpolated value");
    bool thrown = false;
    try {
      (void)fi(10.5);
    } catch (const std::domain_error &) {
      thrown = true;
    }
    check(thrown, "interpolator domain error");
  }

  // filter descriptions evaluated on the host
  {
    io::TophatRadialWindowFunctionFilter top(20.0);
    check(top.spec().a == 20.0, "tophat spec");
    io::GaussianRadialWindowFunctionFilter gau(15.0);
    (void)gau.spec();
    io::GenericRadialWindowFunctionFilter gen([](double r) { return std::exp(-r / 30.0); });
    gen.prepare(pc.data(), pc.size());
    io::AngularMaskFilter mask(std::vector<float>(12 * 4 * 4, 1.0f));
    bool thrown = false;
    try {
      io::AngularMaskFilter fits(std::string("mask.fits"));
    } catch (const std::runtime_error &) {
      thrown = true;
    }
    check(thrown, "FITS mask needs cfitsio");
  }

  // KClkk::savetxt
  {
    KClkk kc;
    kc.k_ln = Array2D(4, 3);
    kc.c_ln = Array2D(4, 3);
    for (size_t i = 0; i < 12; ++i) {
      kc.k_ln.data[i] = 0.01 * (i + 1);
      kc.c_ln.data[i] = 1e3 / (i + 1);
    }
    kc.savetxt(tmp + "kclkk");
    std::remove((tmp + "kclkk.k_ln").c_str());
    std::remove((tmp + "kclkk.c_ln").c_str());
  }

  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("sanitize_host: all checks passed\n");
  return 0;
}
