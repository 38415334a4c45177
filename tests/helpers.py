"""Closed forms and small utilities shared by the tests (no product code)."""
import numpy as np


def singular_stokeslet(x, f):
    r = np.linalg.norm(x)
    return (f / r + (f @ x) * x / r**3) / (8 * np.pi)


def singular_brinkmanlet(x, f, a):
    """tests/test_brinkmanlet.py:62-67 restated: the eps -> 0 Brinkmanlet."""
    r = np.linalg.norm(x)
    z = a * r
    h1 = (np.exp(-z) * (1 + 1 / z + 1 / z**2) - 1 / z**2) / (4 * np.pi * r)
    h2 = (3 / z**2 - np.exp(-z) * (1 + 3 / z + 3 / z**2)) / (4 * np.pi * r**3)
    return h1 * f + h2 * (f @ x) * x


def blake(x, x0, f, mu=1.0):
    """Blake (1971) singular Stokeslet above a no-slip wall z = 0, from the
    closed form G = S(r) - S(R) + 2h (P grad_R) phi, with S(d) = I/|d| + d d^T/|d|^3,
    r = x - x0, R = x - x0*, phi_i = h R_i/|R|^3 - (delta_i3/|R| + R_i R_3/|R|^3),
    P = diag(1, 1, -1); the R-gradient is written out analytically here."""
    x = np.asarray(x, float); x0 = np.asarray(x0, float); f = np.asarray(f, float)
    h = x0[2]
    r = x - x0
    R = x - np.array([x0[0], x0[1], -h])

    def S(d):
        n = np.linalg.norm(d)
        return np.eye(3) / n + np.outer(d, d) / n**3

    n = np.linalg.norm(R)
    e3 = np.array([0.0, 0.0, 1.0])
    I = np.eye(3)
    # d/dR_k of h R_i / n^3
    g1 = h * (I / n**3 - 3 * np.outer(R, R) / n**5)
    # d/dR_k of delta_i3 / n
    g2 = -np.outer(e3, R) / n**3
    # d/dR_k of R_i R_3 / n^3
    g3 = (np.outer(R, e3) + R[2] * I) / n**3 - 3 * R[2] * np.outer(R, R) / n**5
    dphi = g1 - (g2 + g3)
    P = np.diag([1.0, 1.0, -1.0])
    G = S(r) - S(R) + 2 * h * dphi @ P
    return G @ f / (8 * np.pi * mu)


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def eps_rates(errors):
    e = np.asarray(errors, float)
    return np.log2(e[:-1] / e[1:])
