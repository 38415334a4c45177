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


def gs_biflagellate(variant: str = "paper", head_axes=(0.5, 0.6, 0.6), insertion_angle: float = 30.0,
                    n_head: int = 4, n_head_quad: int = 10, n_flag: int = 40, n_flag_quad: int = 400,
                    dt: float = 1e-5):
    """shape_at(t) of the model biflagellate alga of Gallagher & Smith, Phys.
    Rev. Fluids 3, 053101 (2018), Sec. V A, as the reference's example runs it
    (examples/biflagellate_swimmer/biflagellate.py, input3d): lengths in
    flagellar lengths, one beat every 2 pi.

    Body: a rigid ellipsoid with semi-axes ``head_axes``, its force and
    quadrature points the cube-sphere sets of 6 n_head^2 and 6 n_head_quad^2
    points.  Flagella: unit length, mirror images leaving the body at
    ``insertion_angle`` degrees from the forward (+x2) axis, their tangent
    angle (Eqs. 29-30)
        psi(s, t) = -2.5 s - (0.7 + 0.15 sin 2 pi s) cos(t - 2 pi s),
    integrated along the arclength by the trapezoidal rule on 4001 points;
    the chord from the base to s = 4/29 points forward.  ``variant``
    "authors_code" takes the published code's shorter flagellum instead (the
    arc 1/30 <= sigma <= 1, the chord to 5/30).  Force points are n_flag and
    quadrature points n_flag_quad equispaced in s per flagellum; velocities
    are central differences in t with step ``dt``."""
    from .swimming import SwimmerShape

    if variant not in ("paper", "authors_code"):
        raise ValueError(f"unknown variant {variant!r}")
    a1, a2, a3 = (float(v) for v in head_axes)
    ang = np.deg2rad(insertion_angle)
    sf = np.linspace(0.0, 1.0, 4001)
    ds = np.diff(sf)

    def plane(s, t):
        if variant == "authors_code":
            s = 1.0 / 30 + np.asarray(s) * 29.0 / 30
            s_ref, base = 5.0 / 30, 1.0 / 30
        else:
            s_ref, base = 4.0 / 29, 0.0
        psi = -2.5 * sf - (0.7 + 0.15 * np.sin(2 * np.pi * sf)) * np.cos(t - 2 * np.pi * sf)
        cx = np.concatenate([[0.0], np.cumsum(0.5 * (np.cos(psi[1:]) + np.cos(psi[:-1])) * ds)])
        cy = np.concatenate([[0.0], np.cumsum(0.5 * (np.sin(psi[1:]) + np.sin(psi[:-1])) * ds)])
        cx = cx - np.interp(base, sf, cx)
        cy = cy - np.interp(base, sf, cy)
        rot = np.pi / 2 - np.arctan2(np.interp(s_ref, sf, cy), np.interp(s_ref, sf, cx))
        X = np.cos(rot) * cx - np.sin(rot) * cy
        Y = np.sin(rot) * cx + np.cos(rot) * cy
        return np.interp(s, sf, X), np.interp(s, sf, Y)

    def rot_z(P, th):
        c, s_ = np.cos(th), np.sin(th)
        return P @ np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]]).T

    def flagella(s, t):
        X, Y = plane(s, t)
        z = np.zeros_like(X)
        cis = rot_z(np.stack([X, Y, z], axis=1), -ang) + [a1 * np.sin(ang), a2 * np.cos(ang), 0.0]
        trans = rot_z(np.stack([-X, Y, z], axis=1), ang) + [-a1 * np.sin(ang), a2 * np.cos(ang), 0.0]
        return np.vstack([cis, trans])

    axes = np.array([a1, a2, a3])
    head = cube_sphere(n_head) * axes
    head_quad = cube_sphere(n_head_quad) * axes
    s_force = np.linspace(0.0, 1.0, n_flag)
    s_quad = np.linspace(0.0, 1.0, n_flag_quad)

    def shape_at(t):
        flag = flagella(s_force, t)
        vel = (flagella(s_force, t + dt) - flagella(s_force, t - dt)) / (2 * dt)
        return SwimmerShape(x=np.vstack([head, flag]), v=np.vstack([np.zeros_like(head), vel]),
                            X=np.vstack([head_quad, flagella(s_quad, t)]))

    return shape_at
