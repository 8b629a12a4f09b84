"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the shipped package.  Each wrapper names
the reference function it restates; the C file carries the file:line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ZEROS_REFERENCE = 0
ZEROS_CONVERGED = 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i, ll = ctypes.c_double, ctypes.c_int, ctypes.c_longlong
        P = ctypes.c_void_p
        L.orc_sbessel_jl.restype = d
        L.orc_sbessel_jl.argtypes = [i, d]
        L.orc_sph_plm.restype = d
        L.orc_sph_plm.argtypes = [i, i, d]
        L.orc_bessel_zeros.argtypes = [i, i, i, P]
        L.orc_kln.argtypes = [i, i, d, i, P]
        L.orc_count_nl.argtypes = [i, i, d, d, i, P]
        L.orc_interp_table.argtypes = [i, i, i, P, P, P, P]
        L.orc_sfb_raw.argtypes = [P, ll, i, i, d, P, d, i, i, i, P, P, P]
        L.orc_ang2pix_ring.restype = ll
        L.orc_ang2pix_ring.argtypes = [i, d, d]
        L.orc_pix2ang_ring.argtypes = [i, ll, P, P]
        L.orc_ring_info.argtypes = [i, i, P, P, P, P, P]
        L.orc_pixelize.argtypes = [P, ll, i, P, P]
        L.orc_sfb_reverse_fft.argtypes = [P, P, i, i, i, d, P, d, i, i, i, i, P, P, P]
        L.orc_sfb_reverse.argtypes = [P, P, i, i, i, d, P, d, i, i, i, P, P, P]
        L.orc_savetxt.argtypes = [ctypes.c_char_p, P, i, i]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc == 1:
        raise ValueError(f"{what}: invalid argument")
    if rc == 2:
        raise ValueError(f"{what}: argument (z) outside of defined range")
    if rc != 0:
        raise OracleError(f"{what}: status {rc}")


def sbessel_jl(l, x):
    f = lib().orc_sbessel_jl
    return np.vectorize(lambda xx: f(int(l), float(xx)), otypes=[float])(x)


def sph_plm(l, m, x):
    f = lib().orc_sph_plm
    return np.vectorize(lambda xx: f(int(l), int(m), float(xx)), otypes=[float])(x)


def bessel_zeros(l, n, mode=ZEROS_REFERENCE):
    z = np.zeros(n)
    _check(lib().orc_bessel_zeros(l, n, mode, _p(z)), "bessel_zeros")
    return z


def kln(lmax, nmax, rmax, mode=ZEROS_REFERENCE):
    k = np.zeros((lmax, nmax))
    _check(lib().orc_kln(lmax, nmax, rmax, mode, _p(k)), "kln")
    return k


def count_nl(lmax, nmax, rmax, k_max, mode=ZEROS_REFERENCE):
    nl = np.zeros(lmax, dtype=np.int32)
    _check(lib().orc_count_nl(lmax, nmax, rmax, k_max, mode, _p(nl)), "count_nl")
    return nl


def interp_table(l, nmax, mode=ZEROS_REFERENCE):
    y = np.zeros(1000 * nmax)
    P = ctypes.c_int()
    zmax = ctypes.c_double()
    dxi = ctypes.c_double()
    _check(lib().orc_interp_table(l, nmax, mode, _p(y), ctypes.byref(P), ctypes.byref(zmax),
                                  ctypes.byref(dxi)), "interp_table")
    return y, zmax.value, dxi.value


W_POWER_GAUSS = 0
W_TABLE = 1


def sbessel_integral(l, rmax, k1, k2, power=0.0, gauss=0.0, table=None):
    """int_0^rmax w(r) j_l(k1 r) j_l(k2 r) dr (double_sbessel_integrator,
    sbessel_integrator.cpp:45-110).  w = r^power exp(-(r/gauss)^2) (gauss <= 0:
    no Gaussian), or table = (samples, xmin, xmax) linearly interpolated."""
    L = lib()
    if not getattr(L, "_dsb_init", False):
        L.orc_sbessel_integral.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L._dsb_init = True
    out = ctypes.c_double()
    if table is None:
        rc = L.orc_sbessel_integral(int(l), float(rmax), float(k1), float(k2), W_POWER_GAUSS, float(power),
                                    float(gauss), None, 0, 0.0, 0.0, ctypes.byref(out))
    else:
        y, x0, x1 = table
        y = np.ascontiguousarray(y, dtype=np.float64)
        rc = L.orc_sbessel_integral(int(l), float(rmax), float(k1), float(k2), W_TABLE, 0.0, 0.0, _p(y), y.size,
                                    float(x0), float(x1), ctypes.byref(out))
    _check(rc, "sbessel_integral")
    return out.value


def _split_flmn(flat, lmax, nmax):
    out, off = [], 0
    c = flat.view(np.complex128)
    for l in range(lmax):
        out.append(c[off:off + (l + 1) * nmax].reshape(l + 1, nmax).copy())
        off += (l + 1) * nmax
    return out


def _nl_arg(nl, lmax):
    if nl is None:
        return None
    a = np.ascontiguousarray(nl, dtype=np.int32)
    assert a.shape == (lmax,)
    return a


def sfb_raw(points, lmax, nmax, rmax, nl=None, volume=0.0, interpolated=False,
            zeros_mode=ZEROS_REFERENCE, nthreads=0, store_flmn=False):
    """Restates catana::_sfb_raw (sfb_raw.cpp:13-108); returns (k_ln, c_ln, f_lmn|None)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    assert pts.ndim == 2 and pts.shape[1] == 3
    k = np.zeros((lmax, nmax))
    c = np.zeros((lmax, nmax))
    nflat = sum((l + 1) * nmax for l in range(lmax))
    f = np.zeros(2 * nflat) if store_flmn else None
    nla = _nl_arg(nl, lmax)
    rc = lib().orc_sfb_raw(_p(pts), pts.shape[0], lmax, nmax, float(rmax), _p(nla), float(volume),
                           int(interpolated), int(zeros_mode), int(nthreads), _p(k), _p(c), _p(f))
    _check(rc, "sfb_raw")
    return k, c, (_split_flmn(f, lmax, nmax) if store_flmn else None)


def ang2pix_ring(nside, theta, phi):
    f = lib().orc_ang2pix_ring
    return np.vectorize(lambda t, p: f(int(nside), float(t), float(p)), otypes=[np.int64])(theta, phi)


def pix2ang_ring(nside, pix):
    t = ctypes.c_double()
    p = ctypes.c_double()
    lib().orc_pix2ang_ring(int(nside), int(pix), ctypes.byref(t), ctypes.byref(p))
    return t.value, p.value


def ring_info(nside, ring):
    sp = ctypes.c_longlong()
    rp = ctypes.c_int()
    sh = ctypes.c_int()
    z = ctypes.c_double()
    ph = ctypes.c_double()
    lib().orc_ring_info(int(nside), int(ring), ctypes.byref(sp), ctypes.byref(rp), ctypes.byref(sh),
                        ctypes.byref(z), ctypes.byref(ph))
    return sp.value, rp.value, bool(sh.value), z.value, ph.value


def pixelize(points, nside):
    """PixelizedPointContainer(nside, pc) as CSR: (offsets[npix+1], radii[N])."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    npix = 12 * nside * nside
    off = np.zeros(npix + 1, dtype=np.int64)
    radii = np.zeros(max(pts.shape[0], 1))
    _check(lib().orc_pixelize(_p(pts), pts.shape[0], int(nside), _p(off), _p(radii)), "pixelize")
    return off, radii[:pts.shape[0]]


def sfb_reverse_fft(offsets, radii, nside, lmax, nmax, rmax, nl=None, volume=0.0, interpolated=False,
                    zeros_mode=ZEROS_REFERENCE, alias_compat=True, nthreads=0, store_flmn=False):
    """Restates catana::_sfb_reverse_fft (sfb_reverse_fft.cpp:43-209)."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    k = np.zeros((lmax, nmax))
    c = np.zeros((lmax, nmax))
    nflat = sum((l + 1) * nmax for l in range(lmax))
    f = np.zeros(2 * nflat) if store_flmn else None
    nla = _nl_arg(nl, lmax)
    rc = lib().orc_sfb_reverse_fft(_p(off), _p(rad), int(nside), lmax, nmax, float(rmax), _p(nla),
                                   float(volume), int(interpolated), int(zeros_mode), int(alias_compat),
                                   int(nthreads), _p(k), _p(c), _p(f))
    _check(rc, "sfb_reverse_fft")
    return k, c, (_split_flmn(f, lmax, nmax) if store_flmn else None)


def sfb_reverse(offsets, radii, nside, lmax, nmax, rmax, nl=None, volume=0.0, interpolated=False,
                zeros_mode=ZEROS_REFERENCE, nthreads=0, store_flmn=False):
    """Restates catana::_sfb_reverse (sfb_reverse.cpp:13-113)."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    k = np.zeros((lmax, nmax))
    c = np.zeros((lmax, nmax))
    nflat = sum((l + 1) * nmax for l in range(lmax))
    f = np.zeros(2 * nflat) if store_flmn else None
    nla = _nl_arg(nl, lmax)
    rc = lib().orc_sfb_reverse(_p(off), _p(rad), int(nside), lmax, nmax, float(rmax), _p(nla),
                               float(volume), int(interpolated), int(zeros_mode), int(nthreads),
                               _p(k), _p(c), _p(f))
    _check(rc, "sfb_reverse")
    return k, c, (_split_flmn(f, lmax, nmax) if store_flmn else None)


def savetxt(path, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    _check(lib().orc_savetxt(path.encode(), _p(a), a.shape[0], a.shape[1]), "savetxt")
