"""Batched double spherical-Bessel integrals on the device (sfbhip_sbessel_nodes /
sfbhip_sbessel_integrals) against the oracle's restatement of
double_sbessel_integrator (sbessel_integrator.cpp:45-110), the reference's
known values, closed forms, and the shot-noise term against direct SFB of
Poisson catalogues."""
import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: the first import can take minutes)

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

# (l, k1, k2, R, power, gauss): w = r^power exp(-(r/gauss)^2)
CASES = [(0, 1, 1, 100, 0, 0), (0, 1, 2, 100, 0, 0), (4, 1, 4.2, 100, 0, 0), (6, 2.3, 4.5, 100, 2, 40),
         (6, 4.5, 2.3, 100, 2, 40), (5, 0.013, 0.037, 1000, 2, 0), (5, 0.2, 0.21, 1000, 2, 0),
         (255, 0.3, 0.3, 1000, 2, 0), (255, 0.27, 0.5, 1000, 2, 0), (0, 0.0, 0.3, 1000, 2, 0),
         (0, 0.0, 0.0, 1000, 2, 0), (3, 0.1, 0.1, 1000, 2, 50), (40, 0.11, 0.11, 1000, 2, 0),
         (12, 0.002, 0.4, 1000, 2, 300)]


def _w(power, gauss):
    return lambda r: r ** power * (np.exp(-(r / gauss) ** 2) if gauss > 0 else 1.0)


def _scale(l, k1, k2, R, power, gauss):
    # size of the integrand's envelope integral: r^power / (k1 r k2 r) over [0, R]
    a, b = max(k1, 1.0 / R), max(k2, 1.0 / R)
    return R ** max(power - 1, 0) / (a * b) if power >= 2 else R / (a * b * R * R) + 1e-300


@pytest.mark.parametrize("case", CASES)
def test_integral_vs_oracle(handle, case):
    l, k1, k2, R, pw, g = case
    print(f"case: sbessel integral l={l} k=({k1},{k2}) R={R} w=r^{pw} gauss={g}", flush=True)
    r = capi.sbessel_nodes(handle, l, k1, k2, R)
    v = float(capi.sbessel_integrals(handle, l, k1, k2, R, node_values=_w(pw, g)(r)))
    o = O.sbessel_integral(l, R, k1, k2, power=pw, gauss=g)
    tol = 1e-10 * max(abs(o), _scale(l, k1, k2, R, pw, g))
    print("gpu", v, "oracle", o)
    assert abs(v - o) <= tol


def test_batched_triples_and_r_power(handle):
    # many triples in one call; r_power = 2 on the device equals r^2 in the node values
    rng = np.random.default_rng(7)
    n = 300
    l = rng.integers(0, 60, n).astype(np.int32)
    k1 = rng.uniform(0.0, 0.3, n)
    k2 = rng.uniform(0.0, 0.3, n)
    R = 800.0
    print(f"case: {n} random triples, R={R}", flush=True)
    r = capi.sbessel_nodes(handle, l, k1, k2, R)
    w = np.exp(-(r / 300.0) ** 2)
    a = capi.sbessel_integrals(handle, l, k1, k2, R, node_values=w * r * r)
    b = capi.sbessel_integrals(handle, l, k1, k2, R, node_values=w, r_power=2)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(a).max())
    for i in range(0, n, 37):
        o = O.sbessel_integral(int(l[i]), R, k1[i], k2[i], power=2, gauss=300.0)
        assert abs(a[i] - o) <= 1e-10 * max(abs(o), R / max(k1[i], 1 / R) / max(k2[i], 1 / R)), i


def test_table_weight_vs_oracle(handle):
    # the device interpolates the table at each node; the oracle integrates the
    # piecewise-linear w exactly (split at the samples)
    R, l, k1, k2 = 300.0, 7, 0.07, 0.11
    x = np.linspace(0.0, 320.0, 3201)
    y = np.exp(-(x / 120.0) ** 2) * (1 + 0.3 * np.sin(x / 17.0))
    print("case: table weight l=7", flush=True)
    v = float(capi.sbessel_integrals(handle, l, k1, k2, R, table=(y, 0.0, 320.0)))
    o = O.sbessel_integral(l, R, k1, k2, table=(y, 0.0, 320.0))
    # the integral cancels to ~1e-2 of the integral of |integrand|, which sets the scale
    from scipy.special import spherical_jn as J
    rr = np.linspace(0.0, R, 200001)
    env = np.trapezoid(np.abs(np.interp(rr, x, y) * J(l, k1 * rr) * J(l, k2 * rr)), rr)
    print("gpu", v, "oracle", o, "envelope", env)
    assert abs(v - o) <= 1e-8 * env, (v, o)
    with pytest.raises(ValueError, match="FunctionInterpolator"):
        capi.sbessel_integrals(handle, l, k1, k2, R, table=(y[:1001], 0.0, 100.0))


def test_invalid_triples(handle):
    with pytest.raises(ValueError):
        capi.sbessel_integrals(handle, -1, 0.1, 0.1, 100.0, table=(np.ones(3), 0.0, 200.0))
    with pytest.raises(ValueError):
        capi.sbessel_integrals(handle, 2, -0.1, 0.1, 100.0, table=(np.ones(3), 0.0, 200.0))
    with pytest.raises(ValueError, match="node values"):
        capi.sbessel_integrals(handle, 2, 0.1, 0.1, 100.0, node_values=np.ones(5))


# ---- the Python API (catana.besseltools, test/python/besseltools_test.py:26-52) ----

def test_known_values_python_api():
    import catana_amd as catana
    print("case: besseltools known values", flush=True)
    w = lambda r: 1  # noqa: E731
    f = catana.besseltools.double_sbessel_integrator
    assert f(w, 0, 100, 1, 1) == pytest.approx(1.56582, abs=1e-5)
    assert f(w, 0, 100, 1, 2) == pytest.approx(0.785393, abs=1e-5)
    assert f(w, 4, 100, 1, 4.2) == pytest.approx(0.000133933, abs=1e-9)


def test_k_symmetric_python_api():
    import catana_amd as catana
    w = lambda r: np.exp(-(r / 40) ** 2) * r ** 2  # noqa: E731
    res1 = catana.besseltools.double_sbessel_integrator(w, 6, 100, 2.3, 4.5)
    res2 = catana.besseltools.double_sbessel_integrator(w, 6, 100, 4.5, 2.3)
    assert res1 == pytest.approx(res2, rel=1e-13)


def test_function_interpolator_python_api():
    import catana_amd as catana
    import math
    w = lambda r: np.exp(-(r / 40) ** 2) * r ** 2  # noqa: E731
    fi = catana.FunctionInterpolator(w, 100000, 0, 100)
    res1 = catana.besseltools.double_sbessel_integrator(w, 6, 100, 2.3, 4.5)
    res2 = catana.besseltools.double_sbessel_integrator(fi, 6, 100, 2.3, 4.5)
    assert res1 == pytest.approx(res2, rel=1e-5)
    # a scalar-only callable is evaluated node by node
    res3 = catana.besseltools.double_sbessel_integrator(lambda r: math.exp(-(r / 40) ** 2) * r * r, 6, 100, 2.3,
                                                        4.5)
    assert res3 == pytest.approx(res1, rel=1e-14)


def test_vectorised_over_k():
    import catana_amd as catana
    w = lambda r: r ** 2  # noqa: E731
    k1 = np.array([[0.05], [0.1], [0.2]])
    k2 = np.array([[0.05, 0.07, 0.1, 0.3]])
    out = catana.besseltools.double_sbessel_integrator(w, 3, 500.0, k1, k2)
    assert out.shape == (3, 4)
    for i in range(3):
        for j in range(4):
            s = catana.besseltools.double_sbessel_integrator(w, 3, 500.0, float(k1[i, 0]), float(k2[0, j]))
            assert isinstance(s, float)
            assert out[i, j] == pytest.approx(s, rel=1e-14, abs=1e-14 * abs(out).max())


# ---- shot noise (docs/catana_usage.rst:126-134) ----

def test_shot_noise_closed_form():
    import catana_amd as catana
    from scipy.special import spherical_jn as J
    lmax, nmax, R, kmax, N = 8, 10, 500.0, 0.1, 12345.0
    print("case: shot noise, tophat window", flush=True)
    sn = catana.shot_noise(lambda r: 1.0, lmax, kmax, R, nmax, N)
    k = catana.compute_kln(lmax, nmax, R)
    nl = O.count_nl(lmax, nmax, R, kmax)
    V = 4 / 3 * np.pi * R ** 3
    for l in range(lmax):
        for n in range(nmax):
            if n >= nl[l]:
                assert sn[l, n] == 0.0
                continue
            x = k[l, n] * R
            jm = np.cos(x) / x if l == 0 else J(l - 1, x)
            I = R ** 3 / 2 * (J(l, x) ** 2 - jm * J(l + 1, x))
            ref = V / N * 2 / np.pi * k[l, n] ** 2 * I
            assert sn[l, n] == pytest.approx(ref, rel=1e-9)
    # the interpolator overload runs the same integrals with w on the device
    fi = catana.FunctionInterpolator(lambda r: 1.0, 101, 0.0, 600.0)
    np.testing.assert_allclose(catana.shot_noise(fi, lmax, kmax, R, nmax, N), sn, rtol=1e-12, atol=0)


def test_shot_noise_matches_poisson_catalogues(handle):
    # for a uniform (Poisson) catalogue the expected c_ln at l >= 1 is the
    # shot-noise term alone (the ball's window only has l = 0); averaged over
    # catalogues and modes the direct SFB must reproduce it
    import catana_amd as catana
    lmax, nmax, R, kmax, N = 12, 12, 500.0, 0.08, 20000
    sn = catana.shot_noise(lambda r: 1.0, lmax, kmax, R, nmax, float(N))
    tot_c, tot_sn = 0.0, 0.0
    for seed in range(4):
        print(f"case: Poisson catalogue seed {seed}", flush=True)
        pts = torch.from_numpy(workloads.uniform_ball(N, R, seed=100 + seed)).cuda()
        _, c, _ = capi.sfb_direct_torch(handle, pts, lmax, nmax, R, k_max=kmax, precision=capi.PREC_FP64,
                                        bessel=capi.BESSEL_EXACT)
        c = c.cpu().numpy()
        tot_c += c[1:].sum()
        tot_sn += sn[1:].sum()
    ratio = tot_c / tot_sn
    print("sum c_ln / sum shot noise over l >= 1:", ratio)
    assert abs(ratio - 1.0) < 0.05
