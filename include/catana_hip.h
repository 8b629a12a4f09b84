/*
 * catana_hip.h -- C-ABI of the MI355X (gfx950) SFB decomposition library
 * (libcatana_hip.so).  Plain C types only: device pointers, sizes, a stream
 * passed as void* (a hipStream_t, NULL = the default stream).
 *
 * Conventions (those of the reference, catana/include/catana/decomposition/):
 *   - lmax is exclusive: l in [0, lmax)                       (sfb_raw.cpp:38)
 *   - k_ln, c_ln: lmax x nmax doubles, row-major, row = l     (return_types.hpp:19-38)
 *   - f_lmn: m >= 0 only, one (l+1) x nmax complex128 block per l, row-major
 *     (m, n), blocks consecutive in l; sfbhip_flmn_count() gives the length
 *   - points: AoS (r, theta, phi) float64, 24 bytes each -- the memory of a
 *     catana::PointContainer (Point.hpp:17-18)
 *   - f_lmn = sqrt(2/pi) (V/N) k_ln sum_i j_l(k_ln r_i) Y*_lm(theta_i, phi_i),
 *     c_ln = sum_{m=-l..l} |f_lmn|^2 / (2l+1), V = (4/3) pi rmax^3 unless a
 *     window volume is given.
 * Every entry point is synchronous with respect to the stream it is given:
 * it returns after its work on that stream has finished, so that domain
 * errors found on the device can be reported.
 */
#ifndef CATANA_HIP_H
#define CATANA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sfbhip_status {
  SFBHIP_OK = 0,
  SFBHIP_INVALID_ARGUMENT = 1, /* std::invalid_argument / std::runtime_error in the reference */
  SFBHIP_DOMAIN_ERROR = 2,     /* std::domain_error (SBesselInterpolator.cpp:20-22) */
  SFBHIP_HIP_ERROR = 3,
  SFBHIP_OUT_OF_MEMORY = 4,
  SFBHIP_NOT_SUPPORTED = 5
} sfbhip_status;

/* arithmetic of the point contraction */
enum {
  SFBHIP_PREC_FP64 = 0,   /* FP64 everywhere: the drop-in and the GPU oracle */
  SFBHIP_PREC_MIXED = 1   /* FP32 Bessel table (FP64 argument reduction), FP64 Legendre rows,
                             matrix-core contraction, FP64 tile sums; low l in FP64 */
};
/* matrix-core operand format of the mixed path */
enum {
  SFBHIP_OPERANDS_FP32 = 0,        /* FP32 operands, v_mfma_f32_16x16x4f32 */
  SFBHIP_OPERANDS_SPLIT_FP16 = 1,  /* operands split into two FP16 pieces (22 bits), three
                                      v_mfma_f32_16x16x32_f16 per product block */
  SFBHIP_OPERANDS_SPLIT3_FP16 = 2  /* three FP16 pieces (33 bits), six v_mfma_f32_16x16x32_f16:
                                      products as accurate as with FP32 operands */
};
/* how j_l(k r) is evaluated in the FP64 path */
enum {
  SFBHIP_BESSEL_EXACT = 0,     /* gsl_sf_bessel_jl semantics (interpolated = false) */
  SFBHIP_BESSEL_REFERENCE = 1, /* the reference's linear table, 1000*nmax nodes on [0, z_{l,nmax}] */
  SFBHIP_BESSEL_AUTO = 2       /* the reference's default: table when N > 1000 (sfb_decomposition.cpp:17) */
};
/* Bessel-zero finder */
enum {
  SFBHIP_ZEROS_REFERENCE = 0, /* SphericalBesselZeros as written (early-stopping Newton) */
  SFBHIP_ZEROS_CONVERGED = 1  /* Newton run to |delta|/z <= 1e-15 */
};

typedef struct sfbhip_params {
  int32_t lmax;          /* exclusive */
  int32_t nmax;          /* columns of k_ln / c_ln */
  double rmax;
  double k_max;          /* > 0: l keeps n_l = #{n < nmax : z_ln < k_max rmax} modes; <= 0: nmax */
  double window_volume;  /* > 0 replaces (4/3) pi rmax^3 in the normalisation */
  int32_t precision;     /* SFBHIP_PREC_* */
  int32_t bessel;        /* SFBHIP_BESSEL_* (FP64 path only) */
  int32_t zeros;         /* SFBHIP_ZEROS_* */
  int32_t alias_compat;  /* HEALPix path: 1 keeps the reference's unconjugated fold (sfb_reverse_fft.cpp:202-206) */
  int32_t verbose;       /* print a one-line summary per call to stdout */
  int32_t mixed_fp64_below; /* mixed path: multipoles l < this run in FP64; default 2 (at
                               2^25 points l = 1 reached 1.2e-6 in the mixed kernels) */
  int32_t mixed_operands;   /* mixed path: SFBHIP_OPERANDS_* */
  int32_t mixed_chunks;     /* mixed path: largest number of 16-wide m chunks one work item
                               accumulates (1, 2, 4 or 8; 0 = 4) -- more chunks reuse each Bessel
                               value more often but hold more accumulators in registers */
  int32_t mixed_kernel;     /* mixed path, split-FP16 operands: SFBHIP_MIXED_KERNEL_* */
} sfbhip_params;

/* which contraction kernel the mixed path runs (split-FP16 operands; the other operand formats
 * always run the 256-thread kernel) */
enum {
  SFBHIP_MIXED_KERNEL_AUTO = 0,  /* GROUPED (measured fastest on MI355X); batch plans: 256 */
  SFBHIP_MIXED_KERNEL_256 = 1,   /* 256-thread blocks, 64-row items, Bessel nodes in LDS */
  SFBHIP_MIXED_KERNEL_512 = 2,   /* 512-thread blocks, items of up to 128 rows, cubic Bessel
                                    coefficients in LDS, double-buffered Y* halves */
  SFBHIP_MIXED_KERNEL_GROUPED = 3 /* 256-thread blocks, 64-row items, one cubic per (row, 8-point
                                    group) for the Bessel values, two-operation Legendre steps,
                                    point-major Y* chunks read by transposed LDS loads */
};

typedef struct sfbhip_handle sfbhip_handle;

/* Fills defaults: precision FP64, bessel AUTO, zeros REFERENCE, alias_compat 1,
 * mixed_fp64_below 2, mixed_operands SPLIT_FP16, mixed_chunks 4, mixed_kernel AUTO. */
sfbhip_status sfbhip_default_params(sfbhip_params *p);

/* Library state bound to one device: workspace and cached k_ln tables.
 * No counterpart in the reference (it keeps no state between calls). */
sfbhip_status sfbhip_create(sfbhip_handle **out, int device);
sfbhip_status sfbhip_destroy(sfbhip_handle *h);

/* Number of complex entries in f_lmn for (lmax, nmax): sum_l (l+1)*nmax. */
size_t sfbhip_flmn_count(int32_t lmax, int32_t nmax);

/* k_ln and the per-l mode counts n_l on the host.
 * Replaces catana::compute_kln (return_types.cpp:10-19) / KClkk::KClkk (:22-23). */
sfbhip_status sfbhip_kln(sfbhip_handle *h, const sfbhip_params *p, double *k_ln_host, int32_t *nl_host);

/* Direct point-mode SFB on device-resident points.
 * Replaces catana::_sfb_raw (sfb_raw.cpp:13-108) and, with bessel = AUTO,
 * catana::sfb_decomposition(const PointContainer&, ...) (sfb_decomposition.cpp:15-19).
 * d_k_ln, d_c_ln: lmax*nmax doubles (device); d_f_lmn: NULL or 2*flmn_count doubles (device). */
sfbhip_status sfbhip_direct(sfbhip_handle *h, const double *d_points, int64_t n_points,
                            const sfbhip_params *p, double *d_k_ln, double *d_c_ln, double *d_f_lmn,
                            void *stream);

/* sfbhip_direct without any device allocation, for callers that cannot
 * allocate (serving loops that must not call hipMalloc, harnesses that forbid
 * it).  No counterpart in the reference.  Precision FP64 with exact j_l only
 * (bessel EXACT, or AUTO with at most 1000 points); results are bit-identical
 * to sfbhip_direct's for the same arguments.  Every device buffer beyond the
 * caller's points and outputs lives in one static 128 MiB array per GPU
 * (held until the process ends); a call that would need more -- about 48
 * bytes per point plus the partial sums -- fails with SFBHIP_NOT_SUPPORTED.
 * Runs on the current device.  Calls that overlap on one GPU (other host
 * threads, other streams) share that array and race: serialise them. */
sfbhip_status sfbhip_direct_noalloc(const double *d_points, int64_t n_points, const sfbhip_params *p,
                                    double *d_k_ln, double *d_c_ln, double *d_f_lmn, void *stream);

/* One catalogue of a batched call (device buffers, shapes as for sfbhip_direct). */
typedef struct sfbhip_catalog {
  const double *d_points;   /* AoS (r, theta, phi), n_points records */
  int64_t n_points;
  int32_t lmax, nmax;       /* lmax exclusive; columns of k_ln / c_ln */
  double rmax, k_max, window_volume;
  double *d_k_ln, *d_c_ln;  /* lmax*nmax doubles each */
  double *d_f_lmn;          /* NULL or 2*sfbhip_flmn_count(lmax, nmax) doubles */
} sfbhip_catalog;

/* Direct SFB of up to 128 catalogues with their own sizes, lmax, nmax, rmax,
 * k_max and window volume, in one grouped launch per kernel: the work items of
 * all catalogues are planned together, the Bessel and Legendre tables are
 * shared (they depend on l only), and one reduction writes every catalogue's
 * f_lmn.  precision, zeros, verbose and the mixed_* fields come from *p (its
 * lmax, nmax, rmax, k_max, window_volume are not used); j_l is exact in the
 * FP64 path.  No counterpart in the reference, where a caller loops over
 * catana::sfb_decomposition; each catalogue's result equals that of
 * sfbhip_direct on it alone, up to summation order. */
sfbhip_status sfbhip_direct_batched(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                    const sfbhip_params *p, void *stream);

/* ---- batch plans: sfbhip_direct_batched as a plan made once and replayed ----
 * A plan binds the catalogues' device buffers (points, k_ln, c_ln, f_lmn) and
 * shapes, and does all host work once: Bessel zeros, k_ln, tables, the work
 * items of every catalogue.  Enqueueing it launches only the library's own
 * kernels on one stream -- no host synchronisation, no allocation, no
 * library call -- so the enqueue can be captured in a hipGraph and the graph
 * replayed whenever the points in the bound buffers change.  The contraction
 * is one persistent grouped launch: every block pulls work items of any
 * catalogue, multipole and shape from a device counter, in order of
 * decreasing cost.  Each enqueue puts every catalogue's points in radius
 * order on the device (a 16-bit radius key, two 8-bit counting-sort passes,
 * one gather); each catalogue's Bessel tables cover r <= r_bound (default:
 * its rmax).
 * Precision MIXED with SPLIT_FP16 operands only (the FP64 and other operand
 * formats run through sfbhip_direct_batched).  No counterpart in the
 * reference, where a caller loops over catana::sfb_decomposition
 * (sfb_decomposition.cpp:15-19); each catalogue's result equals that of
 * sfbhip_direct on it alone, up to summation order. */
typedef struct sfbhip_batch_plan sfbhip_batch_plan;

/* Bands of l the tile table distinguishes: [0,16) [16,32) [32,64) [64,128) [128,256) [256,512). */
#define SFBHIP_TILE_BANDS 6
typedef struct sfbhip_tile_table {
  int32_t chunks[SFBHIP_TILE_BANDS];       /* 16-wide m chunks one work item accumulates: 1, 2 or 4 */
  int32_t point_split[SFBHIP_TILE_BANDS];  /* largest point split: 1, 2 or 4 (an item of <= 16 rows splits
                                              each point tile over 4 waves, of <= 32 rows over 2) */
  int32_t items_per_cu;                    /* work items planned per compute unit (1 .. 4096) */
} sfbhip_tile_table;

/* The tile table later plans of this handle use.  NULL restores the defaults
 * (4 chunks, point split 4 in every band, 64 items per compute unit). */
sfbhip_status sfbhip_set_tile_table(sfbhip_handle *h, const sfbhip_tile_table *t);
sfbhip_status sfbhip_get_tile_table(sfbhip_handle *h, sfbhip_tile_table *t);

/* Plans the batch (cats as for sfbhip_direct_batched; up to 128 catalogues).
 * r_bound: NULL, or a host array of ncat radii that bound each catalogue's
 * points (values below rmax count as rmax).  A point beyond its bound is
 * clamped to it and reported as a domain error by sfbhip_batch_plan_sync;
 * a point with r < 0 or a non-finite coordinate likewise.  Synchronous. */
sfbhip_status sfbhip_batch_plan_create(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                       const sfbhip_params *p, const double *r_bound, sfbhip_batch_plan **out);
sfbhip_status sfbhip_batch_plan_destroy(sfbhip_batch_plan *plan);

/* Enqueues the plan on stream (NULL: the plan's own stream) and returns
 * without waiting.  Capturable: called inside hipStreamBeginCapture /
 * hipStreamEndCapture it records the whole batched decomposition. */
sfbhip_status sfbhip_batch_plan_enqueue(sfbhip_batch_plan *plan, void *stream);

/* Captures one enqueue into a hipGraph owned by the plan (replacing an
 * earlier one) and instantiates it; sfbhip_batch_plan_launch replays it. */
sfbhip_status sfbhip_batch_plan_capture(sfbhip_batch_plan *plan, void *stream);
sfbhip_status sfbhip_batch_plan_launch(sfbhip_batch_plan *plan, void *stream);

/* Waits for the plan's work on stream and reports the device's error flags of
 * the last enqueue or launch: SFBHIP_DOMAIN_ERROR for a point beyond its
 * bound, r < 0 or a non-finite coordinate; SFBHIP_OK otherwise.  Also times
 * the last run: ms4[0] FP64 low-l kernels, [1] the grouped mixed contraction,
 * [2] reduction and outputs (these three after an eager enqueue only), [3] the
 * whole enqueue or graph launch (HIP events on the stream; ms4 may be NULL). */
sfbhip_status sfbhip_batch_plan_sync(sfbhip_batch_plan *plan, void *stream, float *ms4);

/* Autotuner: for the bands of l this batch reaches, times candidate
 * (chunks, point split) pairs band by band, then items per compute unit,
 * each on the whole batch (plans of this handle, `reps` timed runs of the
 * grouped contraction per candidate, the fastest run kept), installs the best
 * table in the handle and copies it to *out (may be NULL).  Writes the
 * catalogues' outputs as a side effect. */
sfbhip_status sfbhip_batch_autotune(sfbhip_handle *h, const sfbhip_catalog *cats, int32_t ncat,
                                    const sfbhip_params *p, const double *r_bound, int32_t reps,
                                    sfbhip_tile_table *out);

/* HEALPix RING binning of device points into CSR order by pixel: d_offsets
 * (12 nside^2 + 1 int64), d_radii (n_points doubles), radii of a pixel kept in
 * input order.  Replaces PixelizedPointContainer(nside, pc)
 * (PixelizedPointContainer.cpp:11-30); nside must be a power of 2. */
sfbhip_status sfbhip_pixelize(sfbhip_handle *h, const double *d_points, int64_t n_points, int32_t nside,
                              int64_t *d_offsets, double *d_radii, void *stream);

/* HEALPix RING pixel of one direction, on the host (the same function the
 * device binning uses).  Replaces Healpix_Base::ang2pix as called by
 * PixelizedPointContainer::add_point (PixelizedPointContainer.cpp:32-35). */
sfbhip_status sfbhip_ang2pix_ring(int32_t nside, double theta, double phi, int64_t *pix);

/* HEALPix SFB on a pixelised catalogue (CSR as produced by sfbhip_pixelize).
 * Replaces catana::_sfb_reverse_fft (sfb_reverse_fft.cpp:43-209) and
 * catana::sfb_decomposition(const PixelizedPointContainer&, ...) (sfb_decomposition.cpp:22-26);
 * with alias_compat = 0 it equals catana::_sfb_reverse (sfb_reverse.cpp:13-113). */
sfbhip_status sfbhip_healpix(sfbhip_handle *h, const int64_t *d_offsets, const double *d_radii,
                             int32_t nside, const sfbhip_params *p, double *d_k_ln, double *d_c_ln,
                             double *d_f_lmn, void *stream);

/* First n positive zeros of j_l on the host (SFBHIP_ZEROS_* mode).
 * Replaces besseltools::SphericalBesselZeros(l)[0..n-1] (SphericalBesselZeros.cpp:11-101). */
sfbhip_status sfbhip_bessel_zeros(int32_t l, int32_t n, int32_t mode, double *z_host);

/* Weight function w(r) of the double spherical-Bessel integrals. */
enum {
  SFBHIP_WEIGHT_NODES = 0, /* values[i]: w at node i of sfbhip_sbessel_nodes for the same triples */
  SFBHIP_WEIGHT_TABLE = 1  /* values[0..n): samples on [x_min, x_max], linearly interpolated;
                              an argument outside [x_min, x_max) is a domain error
                              (FunctionInterpolator.cpp:7-30) */
};
typedef struct sfbhip_weight {
  int32_t kind;          /* SFBHIP_WEIGHT_* */
  int32_t r_power;       /* the integrand carries w(r) r^r_power (0..8; 2 for the shot-noise term) */
  int64_t n;             /* NODES: node count of the call; TABLE: sample count (>= 2) */
  const double *values;  /* host array of n doubles */
  double x_min, x_max;   /* TABLE: the sampled range */
} sfbhip_weight;

/* Quadrature nodes of the batched integrals below, on the host: with
 * r_nodes == NULL only *n_nodes is set; otherwise r_nodes (host, *n_nodes
 * doubles) receives every node, triple by triple.  A caller with an arbitrary
 * w evaluates it there and passes the values as SFBHIP_WEIGHT_NODES.
 * l, k1, k2: host arrays of n triples (l >= 0, k >= 0). */
sfbhip_status sfbhip_sbessel_nodes(sfbhip_handle *h, int64_t n, const int32_t *l, const double *k1,
                                   const double *k2, double rmax, int64_t *n_nodes, double *r_nodes);

/* out[t] = int_0^rmax w(r) j_l(k1 r) j_l(k2 r) dr for n triples (host arrays),
 * computed on the device: ranges between the zeros of both Bessel functions,
 * 16-point Gauss-Legendre pieces, and the reference's stop after an outer
 * range that changes the sum by less than 1e-10 of itself.
 * Replaces besseltools::double_sbessel_integrator / _bz (sbessel_integrator.cpp:45-110),
 * which the reference's Python binding vectorises over k1, k2 (py_besseltools.cpp:45-78). */
sfbhip_status sfbhip_sbessel_integrals(sfbhip_handle *h, int64_t n, const int32_t *l, const double *k1,
                                       const double *k2, double rmax, const sfbhip_weight *w, double *out);

/* ---- ingestion before the path: record conversion, point filters, subsets ---- */
enum {
  SFBHIP_RECORD_POINT = 0,          /* (r, theta, phi) f64: catana::Point */
  SFBHIP_RECORD_CARTESIAN_F32 = 1,  /* (x, y, z) f32, + shift: io::CartesianRecord<float> */
  SFBHIP_RECORD_CARTESIAN_F64 = 2,  /* (x, y, z) f64, + shift: io::CartesianRecord<double> */
  SFBHIP_RECORD_SPHERICAL_F32 = 3,  /* (r, theta, phi) f32: io::SphericalRecord<float> */
  SFBHIP_RECORD_SPHERICAL_F64 = 4   /* (r, theta, phi) f64: io::SphericalRecord<double> */
};
enum {
  SFBHIP_FILTER_TOPHAT = 0,         /* keep r < a (TophatRadialWindowFunctionFilter) */
  SFBHIP_FILTER_GAUSSIAN = 1,       /* keep with probability exp(-(r/a)^2) (GaussianRadialWindowFunctionFilter) */
  SFBHIP_FILTER_WINDOW_TABLE = 2,   /* keep with probability w(r), w linearly interpolated from n samples on
                                       [x_min, x_max]; r outside [x_min, x_max) is a domain error
                                       (GenericRadialWindowFunctionFilter with interpolation) */
  SFBHIP_FILTER_WINDOW_VALUES = 3,  /* keep record i with probability d_values[i] (n = record count; the
                                       caller evaluated w at the records: GenericRadialWindowFunctionFilter) */
  SFBHIP_FILTER_ANGULAR_MASK = 4    /* keep if d_values[ang2pix_ring(nside, theta, phi)] == 1.0f; float map of
                                       n = 12 nside^2 pixels (AngularMaskFilter) */
};
typedef struct sfbhip_filter {
  int32_t kind;          /* SFBHIP_FILTER_* */
  int32_t nside;         /* ANGULAR_MASK */
  double a;              /* TOPHAT / GAUSSIAN radius */
  double x_min, x_max;   /* WINDOW_TABLE range */
  int64_t n;             /* WINDOW_TABLE samples, WINDOW_VALUES records, ANGULAR_MASK pixels */
  const void *d_values;  /* device: f64 samples / f64 probabilities / f32 mask */
} sfbhip_filter;

/* Converts n device records to points, applies up to 8 filters in order and
 * writes the points that pass all of them, in input order, to d_points_out
 * (AoS (r, theta, phi) f64, capacity n); *n_out (host) receives their count.
 * A random filter keeps a point when its acceptance probability exceeds a
 * uniform draw that depends only on (seed, filter position, index_base + i).
 * Replaces FilterStream::run_step's filter loop (FilterStream.cpp:103-126),
 * Filter::operator() (Filter.cpp:7-34) and the filters' filter(Point&)
 * (radial_filters.cpp:10-33, angular_filters.cpp:9-12); record conversion as
 * CartesianRecord / SphericalRecord::point (record_types.hpp, Point.cpp:5-27). */
sfbhip_status sfbhip_filter_points(sfbhip_handle *h, const void *d_records, int64_t n, int32_t record_kind,
                                   double shift, const sfbhip_filter *filters, int32_t nfilters, uint64_t seed,
                                   uint64_t index_base, double *d_points_out, int64_t *n_out, void *stream);

/* A uniform random subset of k of the n device points (AoS), in input order,
 * into d_out (capacity k).  Replaces RandomSubsetFilter (FilterStream.cpp:14-41)
 * as FilterStream::run_fromtemp applies it; k >= n copies all points. */
sfbhip_status sfbhip_random_subset(sfbhip_handle *h, const double *d_points, int64_t n, int64_t k, uint64_t seed,
                                   double *d_out, void *stream);

/* Duration in ms of the dominant kernel of the last call on this handle
 * (the mixed contraction kernels when the call ran any, else the FP64
 * contraction kernels; for sfbhip_sbessel_integrals the integral kernel),
 * measured with HIP events on the call's stream. */
sfbhip_status sfbhip_last_kernel_ms(sfbhip_handle *h, float *ms);

/* Durations in ms of the three phases of the last call, from HIP events on
 * its stream: [0] FP64 contraction kernels, [1] mixed contraction kernels,
 * [2] reduction of partial blocks and C_l assembly. */
sfbhip_status sfbhip_last_phase_ms(sfbhip_handle *h, float *ms3);

/* Message of the last error on the calling thread ("" if none). */
const char *sfbhip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
