"""Deterministic synthetic point clouds for the registration workloads.

All generators use numpy's PCG64 with explicit seeds, so the same seed gives
the same clouds on every machine with this numpy. Every generator returns
``(X, Y, R_true, t_true)`` with X the target and Y the source, related by
X ~= R_true Y + t_true (the problem the reference solves, README.md:38).
"""
from __future__ import annotations

import numpy as np

SURFACE_SCALE = 1.0 / 2.4  # r(theta, phi) <= 1.2: the surface fits the unit box


def rotation(axis, angle):
    """Rodrigues rotation about a unit axis."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def random_rotation(rng, max_deg):
    axis = rng.normal(size=3)
    return rotation(axis, np.deg2rad(rng.uniform(0.0, max_deg)))


def geodesic(Ra, Rb):
    """Rotation angle of Ra^T Rb in radians, stable near 0."""
    d = np.linalg.norm(np.asarray(Ra, np.float64) - np.asarray(Rb, np.float64))
    return 2.0 * np.arcsin(min(1.0, d / (2.0 * np.sqrt(2.0))))


def octahedral_rotations():
    """The 24 proper signed permutation matrices (octahedral rotation group)."""
    import itertools
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3))
            for i in range(3):
                R[i, perm[i]] = signs[i]
            if np.linalg.det(R) > 0:
                out.append(R)
    return out


def _directions(rng, n, axis=None, half_angle_deg=None):
    """n unit vectors uniform on the sphere, optionally restricted to a cone."""
    out = np.empty((0, 3))
    while len(out) < n:
        v = rng.normal(size=(max(2 * (n - len(out)), 64), 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        if axis is not None:
            a = np.asarray(axis) / np.linalg.norm(axis)
            v = v[v @ a >= np.cos(np.deg2rad(half_angle_deg))]
        out = np.vstack([out, v])
    return out[:n]


def surface(rng, n, axis=None, half_angle_deg=None):
    """Points on r(theta, phi) = 1 + 0.2 sin(3 theta) cos(2 phi), scaled to
    the unit bounding box."""
    d = _directions(rng, n, axis, half_angle_deg)
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    r = 1.0 + 0.2 * np.sin(3 * theta) * np.cos(2 * phi)
    return d * r[:, None] * SURFACE_SCALE


def _move(rng, Ysrc, max_deg, noise):
    R = random_rotation(rng, max_deg)
    t = rng.uniform(-0.1, 0.1, size=3)
    Y = (Ysrc - t) @ R  # rows: R^T (y - t)
    if noise > 0:
        Y = Y + rng.normal(0.0, noise, size=Y.shape)
    return Y, R, t


def uniform_cube(n=1007, seed=0, max_deg=30.0, noise=0.001):
    """X uniform in [-0.5, 0.5]^3, Y = R^T (X - t) + noise, rows permuted
    (EM-ICP N = M = 1007 workload; at n = 10000 the 10k workload)."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.5, 0.5, size=(n, 3))
    Y, R, t = _move(rng, X, max_deg, noise)
    Y = Y[rng.permutation(n)]
    return X.astype(np.float32), Y.astype(np.float32), R, t


def surface_with_outliers(n=5032, seed=1, max_deg=45.0, outlier_frac=0.1):
    """Softassign N = M = 5032 workload: two independent samplings of the
    surface plus 10% uniform outliers each; Y moved by up to 45 degrees."""
    rng = np.random.default_rng(seed)
    n_out = int(round(n * outlier_frac))
    n_in = n - n_out
    X = np.vstack([surface(rng, n_in), rng.uniform(-0.5, 0.5, size=(n_out, 3))])
    Ysrc = np.vstack([surface(rng, n_in), rng.uniform(-0.5, 0.5, size=(n_out, 3))])
    Y, R, t = _move(rng, Ysrc, max_deg, 0.0)
    return X.astype(np.float32), Y.astype(np.float32), R, t


def range_scans(n=40000, m=None, seed=20, max_deg=30.0, noise=0.002, half_angle=70.0,
                separation=45.0):
    """Partial-overlap range-scan pair: X on the surface within a 70 degree
    cone around a random direction d, Y within a cone around d' at 45 degrees
    from d (about 60% overlap), N(0, 0.002^2) noise, moved as the cube case."""
    m = n if m is None else m
    rng = np.random.default_rng(seed)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    w = np.cross(d, rng.normal(size=3))
    w /= np.linalg.norm(w)
    d2 = rotation(w, np.deg2rad(separation)) @ d
    X = surface(rng, n, d, half_angle)
    Ysrc = surface(rng, m, d2, half_angle)
    X = X + rng.normal(0.0, noise, size=X.shape)
    Y, R, t = _move(rng, Ysrc, max_deg, noise)
    return X.astype(np.float32), Y.astype(np.float32), R, t


def batch_sizes(count=256, lo=500, hi=20000, seed=1000):
    """Log-uniform (N_k, M_k) for the batched workload."""
    rng = np.random.default_rng(seed)
    N = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    M = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    return N, M


def load_xyz(path):
    """Whitespace-separated x y z rows (CRLF tolerated), as data/P*.txt."""
    rows = []
    with open(path, "r", newline=None) as f:
        for line in f:
            parts = line.replace("\r", " ").split()
            if len(parts) >= 3:
                rows.append([float(parts[0]), float(parts[1]), float(parts[2])])
    return np.asarray(rows, dtype=np.float64)
