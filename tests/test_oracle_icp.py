"""Pins of the FP64 ICP oracle (oracle.icp_nearest, oracle.icp; src/icp.cpp:122-289).

- the exact nearest-neighbour search equals scipy's exact kd-tree query
  (an independent implementation of the search FLANN approximates);
- with correspondences that are exactly right, one iteration equals the
  weighted-Kabsch (SVD) solution of those pairs, which fixes the layout of S
  (a transposed S gives the inverse rotation) and the centring;
- on a noiseless pair ICP converges to the generating transform, and starting
  at it stays there.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from cuda_emicp_softassign_amd.clouds import geodesic, random_rotation
from oracle import oracle as O


def _pair(n, seed, deg, tmax=0.05):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.5, 0.5, size=(n, 3))
    R = random_rotation(rng, deg)
    t = rng.uniform(-tmax, tmax, 3)
    Y = (X - t) @ R  # X = R y + t
    return X, Y, R, t


def test_nearest_matches_exact_kdtree():
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, size=(1500, 3))
    P = rng.uniform(-1.2, 1.2, size=(1300, 3))
    idx, d2 = O.icp_nearest(X, P)
    dk, ik = cKDTree(X).query(P)
    assert np.array_equal(idx, ik)
    np.testing.assert_allclose(d2, dk ** 2, rtol=1e-12, atol=1e-15)


def test_one_iteration_with_true_correspondences_is_kabsch():
    X, Y, R, t = _pair(400, seed=2, deg=3.0)
    rng = np.random.default_rng(3)
    Xn = X + rng.normal(0, 1e-3, X.shape)   # noisy targets, correspondences exact
    # from the true pose every source point's nearest target is its own
    # partner (noise 1e-3 against ~0.07 mean spacing)
    idx, _ = O.icp_nearest(Xn, Y @ R.T + t)
    assert np.array_equal(idx, np.arange(len(X)))
    R1, t1, _ = O.icp(Xn, Y, R, t, maxIteration=1)
    # Kabsch: minimise sum |x_n - (R y + t)|^2 over R, t
    xc, yc = Xn.mean(0), Y.mean(0)
    H = (Y - yc).T @ (Xn - xc)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1, 1, np.sign(np.linalg.det(Vt.T @ U.T))])
    Rk = Vt.T @ D @ U.T
    tk = xc - Rk @ yc
    assert geodesic(R1, Rk) < 1e-10
    np.testing.assert_allclose(t1, tk, atol=1e-12)


def test_converges_to_generating_transform():
    X, Y, R, t = _pair(800, seed=4, deg=8.0, tmax=0.02)
    Rr, tr, trace = O.icp(X, Y, maxIteration=60)
    assert geodesic(Rr, R) < 1e-9
    np.testing.assert_allclose(tr, t, atol=1e-9)
    assert trace[-1, 1] < 1e-20  # mean squared correspondence distance
    R2, t2, _ = O.icp(X, Y, R, t, maxIteration=3)
    assert geodesic(R2, R) < 1e-12
    np.testing.assert_allclose(t2, t, atol=1e-12)


@pytest.mark.parametrize("n_it", [0, 1])
def test_iteration_count(n_it):
    X, Y, R, t = _pair(50, seed=5, deg=5.0)
    Rr, tr, trace = O.icp(X, Y, maxIteration=n_it)
    assert len(trace) == n_it
    if n_it == 0:
        assert np.array_equal(Rr, np.eye(3)) and not tr.any()
