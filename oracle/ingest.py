"""CPU restatement of the ingestion step before the SFB path (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the shipped package.

* record -> point: catana::Point(x, y, z) (Point.cpp:5-10) after
  point_from_box_position(x, y, z, shift) (Point.cpp:15-21); spherical records
  pass through (record_types.hpp SphericalRecordBase::point).
* filters (radial_filters.cpp:10-33, angular_filters.cpp:9-12): tophat r < R0;
  Gaussian / generic window: keep when w(r) exceeds a uniform draw; angular
  mask: keep when map[ang2pix_ring(nside, theta, phi)] == 1.
* the uniform draws: the reference takes them from a global std::mt19937 in
  processing order; the device path (and this restatement of it) hashes
  (seed, filter position, record index) with SplitMix64's finaliser (Steele,
  Lea, Flood 2014), so parity with the device is exact and parity with the
  reference is statistical (its own filter tests are statistical).
* GadgetSource (GadgetSource.cpp:38-177): format-1 snapshots, halo positions
  shifted by -BoxSize/2.
"""
import os

import numpy as np

from . import oracle as O

M64 = (1 << 64) - 1


def splitmix64(z):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def filter_uniform(seed, slot, index):
    s = splitmix64(np.uint64(seed) ^ np.uint64((0x9E3779B97F4A7C15 * (slot + 1)) & M64))
    with np.errstate(over="ignore"):
        h = splitmix64(s + np.asarray(index, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def cartesian_to_points(xyz, shift=0.0):
    """(N, 3) cartesian -> (N, 3) (r, theta, phi), phi in [0, 2 pi)."""
    x, y, z = (np.asarray(xyz[:, k], dtype=np.float64) + shift for k in range(3))
    r = np.sqrt(x * x + y * y + z * z)
    th = np.arctan2(np.sqrt(x * x + y * y), z)
    ph = np.where((x == 0) & (y == 0), 0.0, np.arctan2(y, x))
    ph = np.where(ph < 0, ph + 2 * np.pi, ph)
    ph = np.where(ph >= 2 * np.pi, ph - 2 * np.pi, ph)
    return np.stack([r, th, ph], axis=1)


def apply_filters(points, filters, seed=0, index_base=0):
    """filters: list of ("tophat", R0) | ("gaussian", R0) | ("table", samples, x_min, x_max)
    | ("values", probabilities) | ("mask", map).  Returns (kept points in input order, mask)."""
    p = np.asarray(points, dtype=np.float64)
    n = p.shape[0]
    keep = np.ones(n, dtype=bool)
    idx = np.arange(n, dtype=np.uint64) + np.uint64(index_base)
    r, th, ph = p[:, 0], p[:, 1], p[:, 2]
    for slot, f in enumerate(filters):
        kind = f[0]
        if kind == "tophat":
            ok = r < f[1]
        elif kind == "mask":
            m = np.asarray(f[1], dtype=np.float32)
            nside = int(round(np.sqrt(m.size / 12)))
            pix = np.array([O.ang2pix_ring(nside, a, b) for a, b in zip(th, ph)], dtype=np.int64)
            ok = m[pix] == np.float32(1.0)
        else:
            if kind == "gaussian":
                w = np.exp(-(r / f[1]) ** 2)
            elif kind == "values":
                w = np.asarray(f[1], dtype=np.float64)
            else:
                y, x0, x1 = np.asarray(f[1], dtype=np.float64), f[2], f[3]
                if np.any((r < x0) | (r >= x1)):
                    raise ValueError("argument outside range for which this FunctionInterpolator is defined")
                u = (r - x0) * ((y.size - 1) / (x1 - x0))
                j = np.minimum(u.astype(np.int64), y.size - 2)
                w = y[j] + (y[j + 1] - y[j]) * (u - j)
            ok = w > filter_uniform(seed, slot, idx)
        keep &= ok
    return p[keep], keep


def read_gadget(path):
    """All halo positions of a format-1 snapshot (one file or path.0 .. path.(N-1)), shifted
    by -BoxSize/2, as (r, theta, phi)."""
    def header(fn):
        with open(fn, "rb") as f:
            f.seek(4)
            h = f.read(256)
        npart = np.frombuffer(h[0:24], dtype=np.uint32)
        num_files = int(np.frombuffer(h[24 + 48 + 16 + 8 + 24 + 4:24 + 48 + 16 + 8 + 24 + 8], dtype=np.int32)[0])
        box = float(np.frombuffer(h[24 + 48 + 16 + 8 + 24 + 8:24 + 48 + 16 + 8 + 24 + 16], dtype=np.float64)[0])
        return npart, num_files, box

    def halos(fn):
        npart, _, _ = header(fn)
        with open(fn, "rb") as f:
            f.seek(4 + 256 + 4 + 4 + int(npart[0]) * 12)
            return np.frombuffer(f.read(int(npart[1]) * 12), dtype=np.float32).reshape(-1, 3)

    _, nfiles, box = header(path)
    if nfiles > 1:
        base = path[:path.rfind(".")]
        xyz = np.concatenate([halos(f"{base}.{i}") for i in range(nfiles)])
    else:
        xyz = halos(path)
    return cartesian_to_points(xyz, -box / 2.0)
