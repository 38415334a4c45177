"""Point sets for the benchmark and test workloads (host-side input generation).

``fibonacci_sphere`` and ``cube_sphere`` keep the reference's point order and
formulas (src/fluidlets/geometry/spheres.py:7-13 and 33-52) because the
wall-sphere resistance workload and the legacy drag fixtures are defined on
them.  The icosphere, ellipsoid, rod-carpet and biflagellate generators define
the other synthetic workloads.
"""
from __future__ import annotations

import numpy as np


def fibonacci_sphere(n: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n golden-angle spiral points on a sphere, shape (n, 3)
    (same point set and order as src/fluidlets/geometry/spheres.py:7-13)."""
    k = np.arange(n) + 0.5
    polar = np.arccos(1 - 2 * k / n)
    azim = np.pi * (1 + np.sqrt(5)) * k
    unit = np.c_[np.cos(azim) * np.sin(polar), np.sin(azim) * np.sin(polar), np.cos(polar)]
    return radius * unit + np.asarray(center, float)


def cube_sphere(n: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Six n x n cell-centred face grids of [-1,1]^3 projected onto the sphere,
    6 n^2 points (same set and order as src/fluidlets/geometry/spheres.py:33-52)."""
    s = np.linspace(-1.0 + 1.0 / n, 1.0 - 1.0 / n, n)
    A, B = np.meshgrid(s, s, indexing="ij")
    A, B, one = A.ravel(), B.ravel(), np.ones(n * n)
    faces = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            cols = [A, B]
            cols.insert(axis, sign * one)
            faces.append(np.stack(cols, axis=1))
    cube = np.vstack(faces)
    unit = cube / np.linalg.norm(cube, axis=1)[:, None]
    return radius * unit + np.asarray(center, float)


def icosphere(level: int, radius: float = 1.0) -> np.ndarray:
    """Icosahedron refined ``level`` times by edge midpoints projected to the
    sphere: 10 * 4^level + 2 points (level 3 -> 642)."""
    t = (1 + 5**0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        new = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new
    return radius * np.array(verts)


def mean_nearest_spacing(X: np.ndarray) -> float:
    """Mean distance from each point to its nearest other point."""
    from scipy.spatial import cKDTree
    d, _ = cKDTree(X).query(X, k=2)
    return float(d[:, 1].mean())


def ellipsoid(n: int, axes, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Fibonacci points of the unit sphere scaled by the semi-axes."""
    return fibonacci_sphere(n) * np.asarray(axes, float) + np.asarray(center, float)


def ellipsoid_area(axes) -> float:
    """Surface area of an ellipsoid (Knud Thomsen's approximation, p = 1.6075)."""
    a, b, c = (float(x) for x in axes)
    p = 1.6075
    return 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def rod_carpet(n_side: int = 32, spacing: float = 0.2, jitter: float = 0.02, radius: float = 0.05,
               height: float = 1.0, per_rod: int = 128, seed: int = 4) -> np.ndarray:
    """Rigid cylinders standing on the wall z = 0 on a jittered square grid.

    Rod (i, j) has its axis at (spacing i + U(+-jitter), spacing j + U(+-jitter));
    its k-th surface point sits at height (k + 1/2) height / per_rod and azimuth
    k times the golden angle (a helical lattice), so every point has z > 0.
    Returns (n_side^2 * per_rod, 3) points, rod-major."""
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    base = np.c_[spacing * ii.ravel(), spacing * jj.ravel()] + rng.uniform(-jitter, jitter, (n_side * n_side, 2))
    k = np.arange(per_rod)
    z = (k + 0.5) * height / per_rod
    theta = k * np.pi * (3 - np.sqrt(5))
    ring = np.c_[radius * np.cos(theta), radius * np.sin(theta), z]
    pts = base[:, None, :2] + ring[None, :, :2]
    zz = np.broadcast_to(ring[None, :, 2], (len(base), per_rod))
    return np.concatenate([pts, zz[..., None]], axis=2).reshape(-1, 3)


def spheroid(n: int, a: float, c: float) -> np.ndarray:
    """Prolate spheroid with equatorial semi-axis a and polar semi-axis c
    along z (Fibonacci points scaled)."""
    return ellipsoid(n, (a, a, c))


def helix_flagellum(n: int, t: float, amplitude: float = 0.2, wavelength: float = 1.0,
                    length: float = 2.0, base=(0.0, 0.0, 0.0), direction: float = 1.0, phase: float = 0.0,
                    omega: float = 2 * np.pi):
    """A flagellum along z with a planar travelling sinusoidal beat in x:
    x = A sin(2 pi s / lambda - omega t + phase), s in (0, length].
    Returns (points, velocities) in the body frame, shape (n, 3) each."""
    s = (np.arange(n) + 1) * length / n
    k = 2 * np.pi / wavelength
    arg = k * s - omega * t + phase
    pts = np.c_[amplitude * np.sin(arg), np.zeros(n), direction * s] + np.asarray(base, float)
    vel = np.c_[-amplitude * omega * np.cos(arg), np.zeros(n), np.zeros(n)]
    return pts, vel
