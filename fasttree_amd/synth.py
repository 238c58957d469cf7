"""Seeded synthetic particle snapshots (counter-based, reproducible anywhere).

Random numbers come from SplitMix64 applied to (seed, stream, index) counters,
so a value depends only on its counter and every machine produces the same
snapshot.

* ``uniform``: i.i.d. uniform positions in [0, L)^3.
* ``nfw_clustered``: 70% of the particles in ``n_halos`` NFW halos, 30% uniform
  background, periodic box of side L.  Halo centres uniform in the box;
  concentration c ~ U[4, 12]; halo masses from dN/dM ~ M^-1.9 on
  [M_MIN, M_MAX] = [1, 1000]; particle counts proportional to mass (largest
  remainders rounded up); virial radius R_VIR_MAX * (M / M_MAX)^(1/3) with
  R_VIR_MAX = 0.02 L; radii from the NFW enclosed-mass profile truncated at c
  (inverse CDF by bisection); isotropic directions; positions wrapped mod L.
  These mass-range and radius constants are choices of this generator.
"""
from __future__ import annotations

import numpy as np

M_MIN, M_MAX, MASS_SLOPE = 1.0, 1000.0, -1.9
R_VIR_MAX_FRAC = 0.02
HALO_FRACTION = 0.70

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from counters (seed, stream, start..start+n-1)."""
    with np.errstate(over="ignore"):
        base = _splitmix64(np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)
                           + np.uint64(stream) * np.uint64(0x8CB92BA72F3D8DD7))
        ctr = np.arange(start, start + n, dtype=np.uint64) + base
    r = _splitmix64(ctr)
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(n: int, seed: int, box: float = 1.0) -> tuple:
    return tuple(uniform01(seed, s, n) * box for s in (0, 1, 2))


def _nfw_mass(x):
    return np.log1p(x) - x / (1.0 + x)


def _nfw_radius(u, c):
    """Solves m(x) = u * m(c) for x in [0, c] (m monotone) by bisection."""
    target = u * _nfw_mass(c)
    lo = np.zeros_like(u)
    hi = c.copy()
    for _ in range(48):
        mid = 0.5 * (lo + hi)
        below = _nfw_mass(mid) < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return 0.5 * (lo + hi)


def nfw_clustered(n: int, seed: int, box: float = 100.0, n_halos: int = 2000) -> tuple:
    """Clustered snapshot: returns (x, y, z) float64 arrays in [0, box)."""
    n_halo_pts = int(round(HALO_FRACTION * n))
    n_bg = n - n_halo_pts
    # Halo properties (stream 10..16)
    u = uniform01(seed, 10, n_halos)
    a = 1.0 + MASS_SLOPE
    mass = (M_MIN ** a + u * (M_MAX ** a - M_MIN ** a)) ** (1.0 / a)
    conc = 4.0 + 8.0 * uniform01(seed, 11, n_halos)
    centres = [uniform01(seed, 12 + d, n_halos) * box for d in range(3)]
    rvir = R_VIR_MAX_FRAC * box * (mass / M_MAX) ** (1.0 / 3.0)
    # Counts proportional to mass, largest remainders rounded up.
    share = mass / mass.sum() * n_halo_pts
    counts = np.floor(share).astype(np.int64)
    short = n_halo_pts - int(counts.sum())
    if short > 0:
        order = np.argsort(-(share - counts), kind="stable")
        counts[order[:short]] += 1
    hid = np.repeat(np.arange(n_halos), counts)
    # Particle draws (streams 20..23)
    ur = uniform01(seed, 20, n_halo_pts)
    r = _nfw_radius(ur, conc[hid]) * (rvir[hid] / conc[hid])
    cos_t = 2.0 * uniform01(seed, 21, n_halo_pts) - 1.0
    phi = 2.0 * np.pi * uniform01(seed, 22, n_halo_pts)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    dirs = (sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t)
    out = []
    for d in range(3):
        h = np.mod(centres[d][hid] + r * dirs[d], box)
        bg = uniform01(seed, 30 + d, n_bg) * box
        v = np.concatenate([h, bg])
        v = np.where(v >= box, 0.0, v)  # mod can round up to box
        out.append(v)
    return tuple(out)


def to_fixed(v: np.ndarray, box: float, bits: int, dtype=np.uint64) -> np.ndarray:
    """Fixed-point coordinates floor(v / box * 2^bits), clamped to [0, 2^bits - 1]."""
    q = np.floor(v / box * float(1 << bits))
    q = np.clip(q, 0, (1 << bits) - 1)
    return q.astype(dtype)


def clustered_fixed(n: int, seed: int, bits: int = 21, box: float = 100.0, dtype=np.uint64,
                    n_halos: int = 2000) -> tuple:
    x, y, z = nfw_clustered(n, seed, box=box, n_halos=n_halos)
    return tuple(to_fixed(a, box, bits, dtype) for a in (x, y, z))


def batched_workload(seed: int = 0, n_trees: int = 64, n_min: int = 10_000,
                     n_max: int = 2_000_000, n_queries: int = 4_000_000, ks=(1, 8, 32, 128),
                     requests_per_tree: int = 4, bits: int = 21, box: float = 100.0,
                     dtype=np.uint32):
    """Uneven multi-tree workload.

    Tree sizes are log-uniform in [n_min, n_max].  Each tree is a clustered
    snapshot whose halo count keeps the 10M snapshot's 3,500 particles per
    halo, in fixed-point coordinates of `bits` bits per axis.  n_queries
    external queries are split over n_trees * requests_per_tree requests by a
    flat Dirichlet draw (largest remainders rounded up), each request with k
    drawn from `ks`; a query is a random particle of its tree displaced by a
    Gaussian of sigma = box / 1000, so queries follow the tree's density.

    Returns (sizes, points per tree, requests (tree, k, query_first,
    num_queries, result_first), (qx, qy, qz)).
    """
    u = uniform01(seed, 40, n_trees)
    sizes = np.floor(np.exp(np.log(n_min) + u * (np.log(n_max) - np.log(n_min)))).astype(np.int64)
    pts = []
    for t in range(n_trees):
        n = int(sizes[t])
        halos = max(4, int(round(n / 3500.0)))
        pts.append(clustered_fixed(n, seed * 1000 + t + 1, bits=bits, box=box, dtype=dtype,
                                   n_halos=halos))
    nr = n_trees * requests_per_tree
    g = -np.log(np.maximum(uniform01(seed, 41, nr), 1e-300))
    share = g / g.sum() * n_queries
    counts = np.floor(share).astype(np.int64)
    short = n_queries - int(counts.sum())
    if short > 0:
        order = np.argsort(-(share - counts), kind="stable")
        counts[order[:short]] += 1
    kidx = np.floor(uniform01(seed, 42, nr) * len(ks)).astype(np.int64)
    reqs, qs = [], [[], [], []]
    qf = rf = 0
    top = float((1 << bits) - 1)
    scale = float(1 << bits) / box
    for r in range(nr):
        t = r // requests_per_tree
        m = int(counts[r])
        k = int(ks[kidx[r]])
        reqs.append((t, k, qf, m, rf))
        src = pts[t]
        pick = np.floor(uniform01(seed, 1000 + r, m) * len(src[0])).astype(np.int64)
        for d in range(3):
            z1 = uniform01(seed, 100000 + 3 * r + d, m)
            z2 = uniform01(seed, 200000 + 3 * r + d, m)
            gauss = np.sqrt(-2.0 * np.log(np.maximum(z1, 1e-300))) * np.cos(2.0 * np.pi * z2)
            v = src[d][pick].astype(np.float64) + gauss * (box / 1000.0) * scale
            qs[d].append(np.clip(np.floor(v), 0, top).astype(dtype))
        qf += m
        rf += m * k
    q = tuple(np.concatenate(a) if a else np.zeros(0, dtype) for a in qs)
    return sizes, pts, reqs, q
