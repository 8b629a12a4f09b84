"""Generate the golden vectors in tests/golden/ from independent implementations
(mpmath at 40 digits, scipy.special, numpy.fft), not from the oracle.

    python tests/golden/make_golden.py

Outputs (all small):
  special_functions.npz  j_l(x) (mpmath), orthonormal P_l^m with Condon-Shortley
                         phase (explicit polynomial, 450 digits), converged zeros of j_l
                         (mpmath findroot), and zeros produced by the
                         reference's Newton loop (SphericalBesselZeros.cpp:17-68)
                         run on mpmath Bessel values rounded to double
  brute_force.npz        f_lmn / c_ln of a 7-point catalogue from scipy
                         sph_harm_y and spherical_jn
  alias_quirk.npz        c_ln of the HEALPix ring-FFT path with numpy.fft.rfft,
                         reading the half spectrum as sfb_reverse_fft.cpp:202-206
                         does (as written and conjugate-corrected), nside 4
The reference's own fixtures (gaussian_catalog*) are copied from
test/test_data/ unchanged.
"""
import os

import mpmath as mp
import numpy as np
from scipy.special import lpmv, spherical_jn, sph_harm_y

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40


def mp_jl(l, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) if l == 0 else mp.mpf(0)
    return mp.sqrt(mp.pi / (2 * x)) * mp.besselj(l + mp.mpf(1) / 2, x)


def mp_plm(l, m, x):
    """Orthonormal P_l^m(x) with the (-1)^m phase from the explicit polynomial
    P_l(x) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) x^(l-2k), differentiated m
    times, at 450 digits (the alternating sum cancels ~150 digits at l = 255)."""
    with mp.workdps(450):
        X = mp.mpf(x)
        d = mp.mpf(0)
        for k in range(l // 2 + 1):
            p = l - 2 * k
            if p < m:
                continue
            d += (-1) ** k * mp.binomial(l, k) * mp.binomial(2 * l - 2 * k, l) * \
                mp.factorial(p) / mp.factorial(p - m) * X ** (p - m)
        d /= mp.mpf(2) ** l
        val = (-1) ** m * (1 - X * X) ** (mp.mpf(m) / 2) * d
        norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - m) / mp.factorial(l + m))
        return +(norm * val)


def true_zeros(l, n):
    out = []
    x = mp.mpf(l) + 1 if l > 0 else mp.mpf(1)
    step = mp.mpf("0.1")
    f = lambda t: mp_jl(l, t)
    prev = f(x)
    while len(out) < n:
        x2 = x + step
        cur = f(x2)
        if prev * cur < 0:
            out.append(float(mp.findroot(f, (x, x2), solver="anderson")))
        x, prev = x2, cur
    return np.array(out)


def ref_newton_zeros(l, n):
    """SphericalBesselZeros as written: guesses (:17-46), Newton with the signed
    stopping test delta/z > 1e-15 (:49-61), extrapolation 2 z_k - z_{k-1} (:63-68);
    every operation rounded to double."""
    def jl(ll, z):
        return float(mp_jl(ll, z))

    def newton(z):
        while True:
            j = jl(l, z)
            d = j / (float(l) / z * j - jl(l + 1, z))
            z = z - d
            if not (d / z > 1e-15):
                return z

    def guess(which):
        if l == 0:
            return 3.141592653589793238463 if which == 0 else 6.283185307179586476925
        a = l + 0.5
        if which == 0:
            return (a + 1.8557571 * a ** (1 / 3) + 1.033150 * a ** (-1 / 3) - 0.00397 * a ** (-1.0)
                    - 0.0908 * a ** (-5 / 3) + 0.043 * a ** (-7 / 3))
        return (a + 3.2446076 * a ** (1 / 3) + 3.158244 * a ** (-1 / 3) - 0.08331 * a ** (-1.0)
                - 0.8437 * a ** (-5 / 3) + 0.864 * a ** (-7 / 3))

    z = [newton(guess(0)), newton(guess(1))]
    while len(z) < n:
        z.append(newton(2 * z[-1] - z[-2]))
    return np.array(z[:n])


def special_functions():
    rng = np.random.default_rng(2024)
    jl_l, jl_x, jl_v = [], [], []
    for l in [0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 128, 200, 255]:
        xs = np.concatenate([[1e-6, 1e-3, 0.3, 1.0, 2.5], rng.uniform(0, 600, 12),
                             l + np.array([-15.0, -3.0, -0.5, 0.0, 0.5, 3.0, 15.0])])
        for x in xs:
            if x <= 0:
                continue
            jl_l.append(l)
            jl_x.append(float(x))
            jl_v.append(float(mp_jl(l, x)))
    pl_l, pl_m, pl_x, pl_v = [], [], [], []
    for l in [0, 1, 2, 4, 7, 16, 40, 100, 255]:
        for m in sorted({m_ for m_ in (0, 1, 2, l // 3, l // 2, l - 1, l) if 0 <= m_ <= l}):
            for x in np.concatenate([[-1.0, -0.999999, 0.0, 0.5, 0.999], rng.uniform(-1, 1, 4)]):
                pl_l.append(l)
                pl_m.append(m)
                pl_x.append(float(x))
                pl_v.append(float(mp_plm(l, m, x)))
    zl = [0, 1, 2, 5, 10, 50, 100, 255]
    zt = np.array([true_zeros(l, 20) for l in zl])
    zr = np.array([ref_newton_zeros(l, 20) for l in zl])
    np.savez_compressed(os.path.join(HERE, "special_functions.npz"),
                        jl_l=np.array(jl_l), jl_x=np.array(jl_x), jl_v=np.array(jl_v),
                        pl_l=np.array(pl_l), pl_m=np.array(pl_m), pl_x=np.array(pl_x), pl_v=np.array(pl_v),
                        zeros_l=np.array(zl), zeros_true=zt, zeros_reference=zr)


def brute_force():
    """7 points, lmax 6, nmax 4, rmax 50; converged zeros; scipy special functions."""
    rng = np.random.default_rng(7)
    n, lmax, nmax, rmax = 7, 6, 4, 50.0
    pts = np.stack([rmax * rng.random(n) ** (1 / 3), np.arccos(rng.uniform(-1, 1, n)),
                    rng.uniform(0, 2 * np.pi, n)], axis=1)
    k = np.array([true_zeros(l, nmax) for l in range(lmax)]) / rmax
    norm = np.sqrt(2 / np.pi) * (4 / 3) * np.pi * rmax ** 3 / n
    c = np.zeros((lmax, nmax))
    flat = []
    for l in range(lmax):
        f = np.zeros((l + 1, nmax), dtype=np.complex128)
        for m in range(l + 1):
            y = np.conj(sph_harm_y(l, m, pts[:, 1], pts[:, 2]))
            for q in range(nmax):
                f[m, q] = norm * k[l, q] * np.sum(spherical_jn(l, k[l, q] * pts[:, 0]) * y)
        c[l] = (np.abs(f[0]) ** 2 + 2 * np.sum(np.abs(f[1:]) ** 2, axis=0)) / (2 * l + 1)
        flat.append(f.ravel())
    np.savez_compressed(os.path.join(HERE, "brute_force.npz"), points=pts, k_ln=k, c_ln=c,
                        f_lmn=np.concatenate(flat), lmax=lmax, nmax=nmax, rmax=rmax)


# ---- HEALPix RING in numpy (Gorski et al. 2005), vectorised ----
def ang2pix_ring(ns, th, ph):
    z = np.cos(th)
    za = np.abs(z)
    tt = np.mod(ph * 2 / np.pi, 4.0)
    npix, ncap = 12 * ns * ns, 2 * ns * (ns - 1)
    pix = np.empty(len(th), dtype=np.int64)
    eq = za <= 2 / 3
    t1 = ns * (0.5 + tt[eq])
    t2 = ns * z[eq] * 0.75
    jp = (t1 - t2).astype(np.int64)
    jm = (t1 + t2).astype(np.int64)
    ir = ns + 1 + jp - jm
    ks = 1 - (ir & 1)
    ip = ((jp + jm - ns + ks + 1) // 2) % (4 * ns)
    pix[eq] = ncap + (ir - 1) * 4 * ns + ip
    po = ~eq
    tp = tt[po] - np.floor(tt[po])
    tmp = ns * np.sqrt(3 * (1 - za[po]))
    jp = (tp * tmp).astype(np.int64)
    jm = ((1 - tp) * tmp).astype(np.int64)
    ir = jp + jm + 1
    ip = (tt[po] * ir).astype(np.int64) % (4 * ir)
    pix[po] = np.where(z[po] > 0, 2 * ir * (ir - 1) + ip, npix - 2 * ir * (ir + 1) + ip)
    return pix


def rings(ns):
    out = []
    npix, ncap = 12 * ns * ns, 2 * ns * (ns - 1)
    for i in range(1, 4 * ns):
        if i < ns:
            rp, sp, z, sh = 4 * i, 2 * i * (i - 1), 1 - i * i / (3 * ns * ns), True
        elif i < 3 * ns:
            rp, sp, z, sh = 4 * ns, ncap + (i - ns) * 4 * ns, (2 * ns - i) * 2 / (3 * ns), (i - ns) % 2 == 0
        else:
            nr = 4 * ns - i
            rp, sp, z, sh = 4 * nr, npix - 2 * nr * (nr + 1), -1 + nr * nr / (3 * ns * ns), True
        out.append((sp, rp, z, (0.5 if sh else 0.0) * 2 * np.pi / rp))
    return out


def sph_plm_scipy(l, m, x):
    from math import lgamma
    return np.sqrt((2 * l + 1) / (4 * np.pi) * np.exp(lgamma(l - m + 1) - lgamma(l + m + 1))) * lpmv(m, l, x)


def alias_quirk():
    rng = np.random.default_rng(11)
    n, ns, lmax, nmax, rmax = 2000, 4, 12, 3, 100.0
    pts = np.stack([rmax * rng.random(n) ** (1 / 3), np.arccos(rng.uniform(-1, 1, n)),
                    rng.uniform(0, 2 * np.pi, n)], axis=1)
    pix = ang2pix_ring(ns, pts[:, 1], pts[:, 2])
    k = np.array([true_zeros(l, nmax) for l in range(lmax)]) / rmax
    norm = np.sqrt(2 / np.pi) * (4 / 3) * np.pi * rmax ** 3 / n
    out = {}
    for mode in ("as_written", "conjugated"):
        c = np.zeros((lmax, nmax))
        for l in range(lmax):
            for q in range(nmax):
                vals = spherical_jn(l, k[l, q] * pts[:, 0])
                mp_ = np.bincount(pix, weights=vals, minlength=12 * ns * ns)
                f = np.zeros(l + 1, dtype=np.complex128)
                for (sp, rp, z, ph0) in rings(ns):
                    X = np.fft.rfft(mp_[sp:sp + rp])
                    for m in range(l + 1):
                        s = m % rp
                        mf = rp // 2 - abs(rp // 2 - s)
                        v = X[mf]
                        if m == 0:
                            v = complex(v.real, 0.0)
                        elif mode == "conjugated" and s > rp // 2:
                            v = np.conj(v)
                        f[m] += v * sph_plm_scipy(l, m, z) * np.exp(-1j * m * ph0)
                cc = 2 * np.sum(np.abs(f) ** 2) - np.abs(f[0]) ** 2
                c[l, q] = cc * (norm * k[l, q]) ** 2 / (2 * l + 1)
        out[mode] = c
    np.savez_compressed(os.path.join(HERE, "alias_quirk.npz"), points=pts, nside=ns, lmax=lmax, nmax=nmax,
                        rmax=rmax, k_ln=k, c_as_written=out["as_written"], c_conjugated=out["conjugated"])


if __name__ == "__main__":
    special_functions()
    brute_force()
    alias_quirk()
    print("golden vectors written to", HERE)
