"""ctypes binding of the C-ABI (include/catana_hip.h) for device-resident data.

torch provides device memory and streams; every computation runs in
libcatana_hip.so.  Loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcatana_hip.so")

OK, INVALID_ARGUMENT, DOMAIN_ERROR, HIP_ERROR, OUT_OF_MEMORY, NOT_SUPPORTED = range(6)
PREC_FP64, PREC_MIXED = 0, 1
BESSEL_EXACT, BESSEL_REFERENCE, BESSEL_AUTO = 0, 1, 2
ZEROS_REFERENCE, ZEROS_CONVERGED = 0, 1
OPERANDS_FP32, OPERANDS_SPLIT_FP16, OPERANDS_SPLIT3_FP16 = 0, 1, 2
MIXED_KERNEL_AUTO, MIXED_KERNEL_256, MIXED_KERNEL_512, MIXED_KERNEL_GROUPED = 0, 1, 2, 3

EXPORTED = [
    "sfbhip_default_params", "sfbhip_create", "sfbhip_destroy", "sfbhip_flmn_count", "sfbhip_kln",
    "sfbhip_direct", "sfbhip_direct_batched", "sfbhip_pixelize", "sfbhip_ang2pix_ring", "sfbhip_healpix", "sfbhip_last_kernel_ms",
    "sfbhip_last_phase_ms", "sfbhip_last_error", "sfbhip_bessel_zeros", "sfbhip_sbessel_nodes",
    "sfbhip_sbessel_integrals", "sfbhip_filter_points", "sfbhip_random_subset",
    "sfbhip_set_tile_table", "sfbhip_get_tile_table", "sfbhip_batch_plan_create", "sfbhip_batch_plan_destroy",
    "sfbhip_batch_plan_enqueue", "sfbhip_batch_plan_capture", "sfbhip_batch_plan_launch", "sfbhip_batch_plan_sync",
    "sfbhip_batch_autotune", "sfbhip_direct_noalloc",
]
TILE_BANDS = 6  # bands of l: [0,16) [16,32) [32,64) [64,128) [128,256) [256,512)
RECORD_POINT, RECORD_CARTESIAN_F32, RECORD_CARTESIAN_F64, RECORD_SPHERICAL_F32, RECORD_SPHERICAL_F64 = range(5)
FILTER_TOPHAT, FILTER_GAUSSIAN, FILTER_WINDOW_TABLE, FILTER_WINDOW_VALUES, FILTER_ANGULAR_MASK = range(5)
WEIGHT_NODES, WEIGHT_TABLE = 0, 1


class Params(ctypes.Structure):
    _fields_ = [
        ("lmax", ctypes.c_int32), ("nmax", ctypes.c_int32), ("rmax", ctypes.c_double),
        ("k_max", ctypes.c_double), ("window_volume", ctypes.c_double), ("precision", ctypes.c_int32),
        ("bessel", ctypes.c_int32), ("zeros", ctypes.c_int32), ("alias_compat", ctypes.c_int32),
        ("verbose", ctypes.c_int32), ("mixed_fp64_below", ctypes.c_int32), ("mixed_operands", ctypes.c_int32),
        ("mixed_chunks", ctypes.c_int32), ("mixed_kernel", ctypes.c_int32),
    ]


class Catalog(ctypes.Structure):
    _fields_ = [
        ("d_points", ctypes.c_void_p), ("n_points", ctypes.c_int64), ("lmax", ctypes.c_int32),
        ("nmax", ctypes.c_int32), ("rmax", ctypes.c_double), ("k_max", ctypes.c_double),
        ("window_volume", ctypes.c_double), ("d_k_ln", ctypes.c_void_p), ("d_c_ln", ctypes.c_void_p),
        ("d_f_lmn", ctypes.c_void_p),
    ]


class TileTable(ctypes.Structure):
    _fields_ = [
        ("chunks", ctypes.c_int32 * TILE_BANDS), ("point_split", ctypes.c_int32 * TILE_BANDS),
        ("items_per_cu", ctypes.c_int32),
    ]

    def as_dict(self):
        return {"chunks": list(self.chunks), "point_split": list(self.point_split),
                "items_per_cu": int(self.items_per_cu)}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for b in range(TILE_BANDS):
            t.chunks[b] = int(d["chunks"][b])
            t.point_split[b] = int(d["point_split"][b])
        t.items_per_cu = int(d["items_per_cu"])
        return t


class Weight(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("r_power", ctypes.c_int32), ("n", ctypes.c_int64), ("values", ctypes.c_void_p),
        ("x_min", ctypes.c_double), ("x_max", ctypes.c_double),
    ]


class Filter(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("nside", ctypes.c_int32), ("a", ctypes.c_double), ("x_min", ctypes.c_double),
        ("x_max", ctypes.c_double), ("n", ctypes.c_int64), ("d_values", ctypes.c_void_p),
    ]


class SfbError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libcatana_hip.so (no fallback: a missing library is an error)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or catana_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.sfbhip_default_params.argtypes = [ctypes.POINTER(Params)]
        L.sfbhip_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
        L.sfbhip_destroy.argtypes = [P]
        L.sfbhip_flmn_count.restype = ctypes.c_size_t
        L.sfbhip_flmn_count.argtypes = [i32, i32]
        L.sfbhip_kln.argtypes = [P, ctypes.POINTER(Params), P, P]
        L.sfbhip_direct.argtypes = [P, P, i64, ctypes.POINTER(Params), P, P, P, P]
        L.sfbhip_direct_batched.argtypes = [P, ctypes.POINTER(Catalog), i32, ctypes.POINTER(Params), P]
        L.sfbhip_direct_noalloc.argtypes = [P, i64, ctypes.POINTER(Params), P, P, P, P]
        L.sfbhip_pixelize.argtypes = [P, P, i64, i32, P, P, P]
        L.sfbhip_ang2pix_ring.argtypes = [i32, ctypes.c_double, ctypes.c_double, ctypes.POINTER(i64)]
        L.sfbhip_healpix.argtypes = [P, P, P, i32, ctypes.POINTER(Params), P, P, P, P]
        L.sfbhip_last_kernel_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
        L.sfbhip_last_phase_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
        L.sfbhip_bessel_zeros.argtypes = [i32, i32, i32, P]
        L.sfbhip_sbessel_nodes.argtypes = [P, i64, P, P, P, ctypes.c_double, ctypes.POINTER(i64), P]
        L.sfbhip_sbessel_integrals.argtypes = [P, i64, P, P, P, ctypes.c_double, ctypes.POINTER(Weight), P]
        L.sfbhip_filter_points.argtypes = [P, P, i64, i32, ctypes.c_double, ctypes.POINTER(Filter), i32,
                                           ctypes.c_uint64, ctypes.c_uint64, P, ctypes.POINTER(i64), P]
        L.sfbhip_random_subset.argtypes = [P, P, i64, i64, ctypes.c_uint64, P, P]
        L.sfbhip_set_tile_table.argtypes = [P, ctypes.POINTER(TileTable)]
        L.sfbhip_get_tile_table.argtypes = [P, ctypes.POINTER(TileTable)]
        L.sfbhip_batch_plan_create.argtypes = [P, ctypes.POINTER(Catalog), i32, ctypes.POINTER(Params), P,
                                               ctypes.POINTER(P)]
        L.sfbhip_batch_plan_destroy.argtypes = [P]
        L.sfbhip_batch_plan_enqueue.argtypes = [P, P]
        L.sfbhip_batch_plan_capture.argtypes = [P, P]
        L.sfbhip_batch_plan_launch.argtypes = [P, P]
        L.sfbhip_batch_plan_sync.argtypes = [P, P, P]
        L.sfbhip_batch_autotune.argtypes = [P, ctypes.POINTER(Catalog), i32, ctypes.POINTER(Params), P, i32,
                                            ctypes.POINTER(TileTable)]
        L.sfbhip_last_error.restype = ctypes.c_char_p
        L.sfbhip_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc, what):
    if rc == OK:
        return
    msg = lib().sfbhip_last_error().decode()
    if rc in (INVALID_ARGUMENT, DOMAIN_ERROR):
        raise ValueError(f"{what}: {msg}")
    raise SfbError(f"{what}: status {rc}: {msg}")


def make_params(lmax, nmax, rmax, k_max=0.0, window_volume=0.0, precision=PREC_FP64, bessel=BESSEL_AUTO,
                zeros=ZEROS_REFERENCE, alias_compat=True, verbose=False, mixed_fp64_below=None,
                mixed_operands=None, mixed_chunks=None, mixed_kernel=None):
    p = Params()
    lib().sfbhip_default_params(ctypes.byref(p))
    p.lmax, p.nmax, p.rmax = int(lmax), int(nmax), float(rmax)
    p.k_max, p.window_volume = float(k_max), float(window_volume)
    p.precision, p.bessel, p.zeros = int(precision), int(bessel), int(zeros)
    p.alias_compat, p.verbose = int(bool(alias_compat)), int(bool(verbose))
    if mixed_fp64_below is not None:
        p.mixed_fp64_below = int(mixed_fp64_below)
    if mixed_operands is not None:
        p.mixed_operands = int(mixed_operands)
    if mixed_chunks is not None:
        p.mixed_chunks = int(mixed_chunks)
    if mixed_kernel is not None:
        p.mixed_kernel = int(mixed_kernel)
    return p


def flmn_count(lmax, nmax):
    return int(lib().sfbhip_flmn_count(int(lmax), int(nmax)))


def split_flmn(flat, lmax, nmax):
    """Flat complex f_lmn -> list of (l+1, nmax) arrays (the reference's KClkk.f_lmn)."""
    c = np.asarray(flat).view(np.complex128).reshape(-1)
    out, off = [], 0
    for l in range(lmax):
        out.append(c[off:off + (l + 1) * nmax].reshape(l + 1, nmax).copy())
        off += (l + 1) * nmax
    return out


class Handle:
    """Owns an sfbhip_handle bound to one device."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(lib().sfbhip_create(ctypes.byref(self._h), int(device)), "sfbhip_create")
        self.device = device

    def close(self):
        if self._h:
            lib().sfbhip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ptr(self):
        return self._h

    def kln(self, params):
        k = np.zeros((params.lmax, params.nmax))
        nl = np.zeros(params.lmax, dtype=np.int32)
        check(lib().sfbhip_kln(self._h, ctypes.byref(params), k.ctypes.data_as(ctypes.c_void_p),
                               nl.ctypes.data_as(ctypes.c_void_p)), "sfbhip_kln")
        return k, nl

    def last_kernel_ms(self):
        ms = ctypes.c_float()
        check(lib().sfbhip_last_kernel_ms(self._h, ctypes.byref(ms)), "sfbhip_last_kernel_ms")
        return ms.value

    def last_phase_ms(self):
        """(FP64 contraction, mixed contraction, reduction + C_l) of the last call, ms."""
        ms = (ctypes.c_float * 3)()
        check(lib().sfbhip_last_phase_ms(self._h, ms), "sfbhip_last_phase_ms")
        return tuple(ms)


def _torch():
    import torch
    return torch


def sfb_direct_torch(handle, points, lmax, nmax, rmax, store_flmn=False, stream=None, **kw):
    """points: torch float64 tensor (N, 3) on a GPU, rows (r, theta, phi).
    Returns (k_ln, c_ln, f_lmn|None) as torch tensors on the same device."""
    torch = _torch()
    assert points.dtype == torch.float64 and points.dim() == 2 and points.shape[1] == 3
    pts = points.contiguous()
    p = make_params(lmax, nmax, rmax, **kw)
    dev = pts.device
    k = torch.empty((lmax, nmax), dtype=torch.float64, device=dev)
    c = torch.empty((lmax, nmax), dtype=torch.float64, device=dev)
    f = torch.empty(2 * flmn_count(lmax, nmax), dtype=torch.float64, device=dev) if store_flmn else None
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib().sfbhip_direct(handle.ptr, ctypes.c_void_p(pts.data_ptr()), ctypes.c_int64(pts.shape[0]),
                             ctypes.byref(p), ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(c.data_ptr()),
                             ctypes.c_void_p(f.data_ptr()) if f is not None else None,
                             ctypes.c_void_p(stream) if stream else None)
    check(rc, "sfbhip_direct")
    return k, c, f


def sfb_direct_batched_torch(handle, catalogs, store_flmn=False, stream=None, **kw):
    """catalogs: list of dicts with points (torch float64 (N, 3) on the GPU), lmax,
    nmax, rmax and optional k_max, window_volume.  Common options (precision,
    zeros, mixed_*) as for sfb_direct_torch.  Returns a list of (k_ln, c_ln,
    f_lmn|None) torch tensors, one per catalogue, from one batched call."""
    torch = _torch()
    p = make_params(1, 1, 1.0, **kw)
    arr = (Catalog * len(catalogs))()
    outs, keep = [], []
    for i, c in enumerate(catalogs):
        pts = c["points"].contiguous()
        assert pts.dtype == torch.float64 and pts.dim() == 2 and pts.shape[1] == 3
        lmax, nmax = int(c["lmax"]), int(c["nmax"])
        k = torch.empty((lmax, nmax), dtype=torch.float64, device=pts.device)
        cl = torch.empty((lmax, nmax), dtype=torch.float64, device=pts.device)
        f = torch.empty(2 * flmn_count(lmax, nmax), dtype=torch.float64, device=pts.device) if store_flmn else None
        arr[i] = Catalog(pts.data_ptr(), pts.shape[0], lmax, nmax, float(c["rmax"]), float(c.get("k_max", 0.0)),
                         float(c.get("window_volume", 0.0)), k.data_ptr(), cl.data_ptr(),
                         f.data_ptr() if f is not None else None)
        keep.append(pts)
        outs.append((k, cl, f))
    if stream is None:
        stream = torch.cuda.current_stream(keep[0].device if keep else None).cuda_stream
    rc = lib().sfbhip_direct_batched(handle.ptr, arr, len(catalogs), ctypes.byref(p),
                                     ctypes.c_void_p(stream) if stream else None)
    check(rc, "sfbhip_direct_batched")
    return outs


def _catalog_array(catalogs, store_flmn):
    """Catalog structs for the batched entry points, with freshly allocated
    outputs: (array, [(k_ln, c_ln, f_lmn|None)], points kept alive)."""
    torch = _torch()
    arr = (Catalog * len(catalogs))()
    outs, keep = [], []
    for i, c in enumerate(catalogs):
        pts = c["points"]
        assert pts.dtype == torch.float64 and pts.dim() == 2 and pts.shape[1] == 3 and pts.is_contiguous()
        lmax, nmax = int(c["lmax"]), int(c["nmax"])
        k = torch.empty((lmax, nmax), dtype=torch.float64, device=pts.device)
        cl = torch.empty((lmax, nmax), dtype=torch.float64, device=pts.device)
        f = torch.empty(2 * flmn_count(lmax, nmax), dtype=torch.float64, device=pts.device) if store_flmn else None
        arr[i] = Catalog(pts.data_ptr(), pts.shape[0], lmax, nmax, float(c["rmax"]), float(c.get("k_max", 0.0)),
                         float(c.get("window_volume", 0.0)), k.data_ptr(), cl.data_ptr(),
                         f.data_ptr() if f is not None else None)
        keep.append(pts)
        outs.append((k, cl, f))
    return arr, outs, keep


def _bounds(r_bound, n):
    if r_bound is None:
        return None, None
    rb = np.ascontiguousarray(np.broadcast_to(np.asarray(r_bound, dtype=np.float64), (n,)))
    return rb, ctypes.c_void_p(rb.ctypes.data)


class BatchPlan:
    """A plan of the batched mixed decomposition (sfbhip_batch_plan_*): binds
    the catalogues' point tensors (contiguous float64 (N, 3) on the GPU, kept
    alive here) and fresh output tensors.  Writing new points into the same
    tensors and enqueueing (or replaying the captured graph) recomputes
    `outputs`, a list of (k_ln, c_ln, f_lmn|None) per catalogue."""

    def __init__(self, handle, catalogs, store_flmn=False, r_bound=None, **kw):
        self._h = handle
        p = make_params(1, 1, 1.0, precision=PREC_MIXED, **kw)
        arr, self.outputs, self._keep = _catalog_array(catalogs, store_flmn)
        self._rb, rbp = _bounds(r_bound, len(catalogs))
        out = ctypes.c_void_p()
        check(lib().sfbhip_batch_plan_create(handle.ptr, arr, len(catalogs), ctypes.byref(p), rbp,
                                             ctypes.byref(out)), "sfbhip_batch_plan_create")
        self._p = out

    @staticmethod
    def _s(stream):
        return ctypes.c_void_p(stream) if stream else None

    def enqueue(self, stream=None):
        check(lib().sfbhip_batch_plan_enqueue(self._p, self._s(stream)), "sfbhip_batch_plan_enqueue")

    def capture(self, stream=None):
        check(lib().sfbhip_batch_plan_capture(self._p, self._s(stream)), "sfbhip_batch_plan_capture")

    def launch(self, stream=None):
        check(lib().sfbhip_batch_plan_launch(self._p, self._s(stream)), "sfbhip_batch_plan_launch")

    def sync(self, stream=None):
        """Waits, raises ValueError on the device's error flags, and returns the
        timings (FP64 low l, grouped contraction, reduction, whole run) in ms."""
        ms = (ctypes.c_float * 4)()
        check(lib().sfbhip_batch_plan_sync(self._p, self._s(stream), ms), "sfbhip_batch_plan_sync")
        return tuple(ms)

    def close(self):
        if getattr(self, "_p", None):
            lib().sfbhip_batch_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_tile_table(handle):
    t = TileTable()
    check(lib().sfbhip_get_tile_table(handle.ptr, ctypes.byref(t)), "sfbhip_get_tile_table")
    return t.as_dict()


def set_tile_table(handle, table=None):
    t = TileTable.from_dict(table) if table is not None else None
    check(lib().sfbhip_set_tile_table(handle.ptr, ctypes.byref(t) if t is not None else None), "sfbhip_set_tile_table")


def batch_autotune(handle, catalogs, reps=2, r_bound=None, **kw):
    """Runs the autotuner on these catalogues; installs and returns the tile table (a dict)."""
    p = make_params(1, 1, 1.0, precision=PREC_MIXED, **kw)
    arr, _outs, _keep = _catalog_array(catalogs, False)
    _rb, rbp = _bounds(r_bound, len(catalogs))
    t = TileTable()
    check(lib().sfbhip_batch_autotune(handle.ptr, arr, len(catalogs), ctypes.byref(p), rbp, int(reps),
                                      ctypes.byref(t)), "sfbhip_batch_autotune")
    return t.as_dict()


def pixelize_torch(handle, points, nside, stream=None):
    """HEALPix RING binning of device points: returns (offsets int64 [12 nside^2 + 1],
    radii float64 [N]) on the same device (PixelizedPointContainer(nside, pc))."""
    torch = _torch()
    pts = points.contiguous()
    npix = 12 * nside * nside
    off = torch.empty(npix + 1, dtype=torch.int64, device=pts.device)
    rad = torch.empty(max(pts.shape[0], 1), dtype=torch.float64, device=pts.device)
    if stream is None:
        stream = torch.cuda.current_stream(pts.device).cuda_stream
    rc = lib().sfbhip_pixelize(handle.ptr, ctypes.c_void_p(pts.data_ptr()), ctypes.c_int64(pts.shape[0]),
                               int(nside), ctypes.c_void_p(off.data_ptr()), ctypes.c_void_p(rad.data_ptr()),
                               ctypes.c_void_p(stream) if stream else None)
    check(rc, "sfbhip_pixelize")
    return off, rad[:pts.shape[0]]


def sfb_healpix_torch(handle, offsets, radii, nside, lmax, nmax, rmax, store_flmn=False, stream=None, **kw):
    """Pixelised SFB (catana::_sfb_reverse_fft semantics) on a device CSR."""
    torch = _torch()
    p = make_params(lmax, nmax, rmax, **kw)
    dev = offsets.device
    k = torch.empty((lmax, nmax), dtype=torch.float64, device=dev)
    c = torch.empty((lmax, nmax), dtype=torch.float64, device=dev)
    f = torch.empty(2 * flmn_count(lmax, nmax), dtype=torch.float64, device=dev) if store_flmn else None
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib().sfbhip_healpix(handle.ptr, ctypes.c_void_p(offsets.data_ptr()), ctypes.c_void_p(radii.data_ptr()),
                              int(nside), ctypes.byref(p), ctypes.c_void_p(k.data_ptr()),
                              ctypes.c_void_p(c.data_ptr()),
                              ctypes.c_void_p(f.data_ptr()) if f is not None else None,
                              ctypes.c_void_p(stream) if stream else None)
    check(rc, "sfbhip_healpix")
    return k, c, f


def bessel_zeros(l, n, mode=ZEROS_REFERENCE):
    """First n zeros of j_l (host), as SphericalBesselZeros(l)[0..n-1]."""
    z = np.zeros(int(n))
    check(lib().sfbhip_bessel_zeros(int(l), int(n), int(mode), z.ctypes.data), "sfbhip_bessel_zeros")
    return z


def _triples(l, k1, k2):
    l, k1, k2 = np.broadcast_arrays(np.asarray(l, dtype=np.int32), np.asarray(k1, dtype=np.float64),
                                    np.asarray(k2, dtype=np.float64))
    shape = l.shape
    return (shape, np.ascontiguousarray(l.ravel()), np.ascontiguousarray(k1.ravel()),
            np.ascontiguousarray(k2.ravel()))


def sbessel_nodes(handle, l, k1, k2, rmax):
    """Quadrature nodes of sbessel_integrals for the same triples (host array)."""
    _, l, k1, k2 = _triples(l, k1, k2)
    cnt = ctypes.c_int64()
    L = lib()
    check(L.sfbhip_sbessel_nodes(handle.ptr, l.size, l.ctypes.data, k1.ctypes.data, k2.ctypes.data, float(rmax),
                                 ctypes.byref(cnt), None), "sfbhip_sbessel_nodes")
    r = np.zeros(cnt.value)
    check(L.sfbhip_sbessel_nodes(handle.ptr, l.size, l.ctypes.data, k1.ctypes.data, k2.ctypes.data, float(rmax),
                                 ctypes.byref(cnt), r.ctypes.data), "sfbhip_sbessel_nodes")
    return r


def sbessel_integrals(handle, l, k1, k2, rmax, node_values=None, table=None, r_power=0):
    """int_0^rmax w(r) j_l(k1 r) j_l(k2 r) dr over broadcast (l, k1, k2).

    w is given either as its values at sbessel_nodes(...) (node_values) or as
    a table (samples, x_min, x_max) interpolated linearly on the device; the
    integrand carries w(r) r^r_power."""
    shape, l, k1, k2 = _triples(l, k1, k2)
    if (node_values is None) == (table is None):
        raise ValueError("give exactly one of node_values and table")
    if table is not None:
        v, x0, x1 = table
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = Weight(WEIGHT_TABLE, int(r_power), v.size, v.ctypes.data, float(x0), float(x1))
    else:
        v = np.ascontiguousarray(node_values, dtype=np.float64)
        w = Weight(WEIGHT_NODES, int(r_power), v.size, v.ctypes.data, 0.0, 0.0)
    out = np.zeros(l.size)
    check(lib().sfbhip_sbessel_integrals(handle.ptr, l.size, l.ctypes.data, k1.ctypes.data, k2.ctypes.data,
                                         float(rmax), ctypes.byref(w), out.ctypes.data), "sfbhip_sbessel_integrals")
    return out.reshape(shape)


def _stream_ptr(stream, device=None):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(device).cuda_stream
    return ctypes.c_void_p(stream) if stream else None


def filter_points_torch(handle, records, record_kind=RECORD_POINT, filters=(), seed=0, index_base=0, shift=0.0,
                        stream=None):
    """sfbhip_filter_points on a (N, 3) device tensor of records.  filters: sequence of
    dicts {"kind": FILTER_*, "a", "x_min", "x_max", "values" (device tensor), "nside"}.
    Returns the kept points, (M, 3) float64 on the device, in input order."""
    import torch
    n = records.shape[0]
    out = torch.empty((max(n, 1), 3), dtype=torch.float64, device=records.device)
    arr = (Filter * max(len(filters), 1))()
    keep = []
    for i, f in enumerate(filters):
        v = f.get("values")
        keep.append(v)
        arr[i] = Filter(int(f["kind"]), int(f.get("nside", 0)), float(f.get("a", 0.0)), float(f.get("x_min", 0.0)),
                        float(f.get("x_max", 0.0)), int(v.numel()) if v is not None else 0,
                        v.data_ptr() if v is not None else None)
    cnt = ctypes.c_int64()
    check(lib().sfbhip_filter_points(handle.ptr, records.data_ptr() if n else None, n, int(record_kind),
                                     float(shift), arr, len(filters), int(seed) & ((1 << 64) - 1),
                                     int(index_base), out.data_ptr(), ctypes.byref(cnt), _stream_ptr(stream, records.device)),
          "sfbhip_filter_points")
    return out[:cnt.value]


def random_subset_torch(handle, points, k, seed=0, stream=None):
    """sfbhip_random_subset: k of the (N, 3) device points, in input order."""
    import torch
    n = points.shape[0]
    k = min(int(k), n)
    out = torch.empty((max(k, 1), 3), dtype=torch.float64, device=points.device)
    check(lib().sfbhip_random_subset(handle.ptr, points.data_ptr(), n, k, int(seed) & ((1 << 64) - 1),
                                     out.data_ptr(), _stream_ptr(stream, points.device)), "sfbhip_random_subset")
    return out[:k]
