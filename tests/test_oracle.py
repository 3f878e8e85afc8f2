"""Pins the CPU oracle (test infrastructure) without running the reference.

The reference ships no golden outputs and its CPU path cannot be built here
(it needs PCL and LAPACK), so each oracle function is fixed by checks chosen
so that a dropped term, a wrong sign or index, or a transposed operand fails:
  * Jacobi eigensolver   -> numpy.linalg.eigh (LAPACK);
  * Horn findRTfromS     -> SVD-Kabsch, the exact transform between the
                            reference's data/P*.txt and data/Qnew*.txt;
  * EM-ICP step          -> a literal materialised-matrix restatement in plain
                            Python loops (tests/_literal.py) with LAPACK's
                            eigensolver, and closed forms (uniform affinities
                            at huge sigma^2, floored rows);
  * EM-ICP loop          -> exact recovery of the P/Qnew transform, FP32 vs FP64;
  * Softassign           -> the literal loops (every sweep's row and column
                            sums included), Sinkhorn invariants, FP32 vs FP64,
                            exact recovery of the P/Qnew transform (no slack,
                            converged Sinkhorn, annealed below the spacing);
  * batched fixture      -> tests/golden/config3_oracle.npz re-derived from
                            the oracle on its smallest instance.
"""
import os

import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import geodesic, rotation, uniform_cube
from tests import _literal
from tests.conftest import GOLDEN


def eig_lapack(N):
    w, V = np.linalg.eigh(np.asarray(N, dtype=np.float64))
    return list(w), [list(row) for row in V]


def kabsch(P, Q):
    """R, t minimising |P - (R Q + t)| via SVD with the det correction."""
    Pc, Qc = P.mean(0), Q.mean(0)
    H = (Q - Qc).T @ (P - Pc)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1, 1, np.sign(np.linalg.det(Vt.T @ U.T))])
    R = Vt.T @ D @ U.T
    return R, Pc - R @ Qc


def s_matrix(P, Q, w=None):
    """h_S of findRTfromS for target P, source Q: S[i+3j] = sum w (P-Pc)_i (Q-Qc)_j."""
    w = np.ones(len(P)) if w is None else w
    Pc = (w[:, None] * P).sum(0) / w.sum()
    Qc = (w[:, None] * Q).sum(0) / w.sum()
    S = np.zeros(9)
    for i in range(3):
        for j in range(3):
            S[i + 3 * j] = (w * (P[:, i] - Pc[i]) * (Q[:, j] - Qc[j])).sum()
    return Pc, Qc, S


def test_sigma_schedule(oracle):
    s = oracle.sigma_schedule(0.01, 1e-5, 0.9)
    assert len(s) == 66
    assert s[0] == np.float32(0.01)
    assert abs(s[-1] - 1.0611e-5) < 1e-9
    assert oracle.sigma_schedule(1e-5, 1e-5, 0.9).size == 0  # loop never runs


@pytest.mark.parametrize("seed", range(6))
def test_jacobi_matches_lapack(oracle, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(4, 4))
    A = A + A.T
    if seed == 5:  # repeated eigenvalues
        A = np.diag([2.0, 2.0, -1.0, 0.5])
    w, V, rc = oracle.jacobi4(A)
    assert rc >= 0
    np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(A), atol=1e-12)
    np.testing.assert_allclose(A @ V, V @ np.diag(w), atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(4), atol=1e-12)


def test_jacobi_rejects_nan(oracle):
    A = np.eye(4)
    A[1, 2] = A[2, 1] = np.nan
    assert oracle.jacobi4(A)[2] < 0


@pytest.mark.parametrize("n", [101, 1007])
def test_horn_recovers_reference_fixture_transform(oracle, pq_pairs, n):
    P, Q = pq_pairs[n]  # Qnew = A P + b exactly (up to 6-decimal rounding)
    Pc, Qc, S = s_matrix(P, Q)
    R, t, rc = oracle.find_rt_from_s(Pc, Qc, S)
    assert rc == 0
    assert np.abs(P - (Q @ R.T + t)).max() < 2e-6
    assert abs(np.linalg.det(R) - 1) < 1e-12
    angle = np.degrees(geodesic(R, np.eye(3)))
    assert abs(angle - 91.824) < 5e-3
    Rk, tk = kabsch(P, Q)
    assert geodesic(R, Rk) < 1e-10
    np.testing.assert_allclose(t, tk, atol=1e-12)


@pytest.mark.parametrize("seed", range(5))
def test_horn_matches_kabsch_weighted(oracle, seed):
    rng = np.random.default_rng(seed)
    Q = rng.normal(size=(40, 3))
    Rt = rotation(rng.normal(size=3), rng.uniform(0, np.pi))
    P = Q @ Rt.T + rng.normal(size=3) + 0.05 * rng.normal(size=Q.shape)
    w = rng.uniform(0.1, 2.0, size=40)
    Pc, Qc, S = s_matrix(P, Q, w)
    R, t, _ = oracle.find_rt_from_s(Pc, Qc, S)
    # weighted Kabsch
    H = (w[:, None] * (Q - Qc)).T @ (P - Pc)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1, 1, np.sign(np.linalg.det(Vt.T @ U.T))])
    Rk = Vt.T @ D @ U.T
    assert geodesic(R, Rk) < 1e-9
    np.testing.assert_allclose(t, Pc - Rk @ Qc, atol=1e-9)
    R32, t32, _ = oracle.find_rt_from_s(Pc, Qc, S, precision="f32")
    assert geodesic(R32.astype(np.float64), Rk) < 2e-6


def test_horn_special_cases(oracle):
    rng = np.random.default_rng(3)
    X = rng.normal(size=(30, 3))
    Pc, Qc, S = s_matrix(X, X)  # Y = X -> identity
    R, t, _ = oracle.find_rt_from_s(Pc, Qc, S)
    assert geodesic(R, np.eye(3)) < 1e-12 and np.abs(t).max() < 1e-12
    Pc, Qc, S = s_matrix(X + [0.3, -0.2, 0.1], X)  # pure translation
    R, t, _ = oracle.find_rt_from_s(Pc, Qc, S)
    assert geodesic(R, np.eye(3)) < 1e-12
    np.testing.assert_allclose(t, [0.3, -0.2, 0.1], atol=1e-12)
    _, _, rc = oracle.find_rt_from_s(Pc, Qc, np.full(9, np.nan))
    assert rc != 0


def _tiny(seed, n=7, m=6):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.5, 0.5, size=(n, 3))
    Y = X[:m] @ rotation([1, 2, 3], 0.2).T + 0.05 + 0.01 * rng.normal(size=(m, 3))
    return X, Y


@pytest.mark.parametrize("sigma_p2", [0.3, 0.01, 2e-3])
def test_emicp_step_matches_literal_loops(oracle, sigma_p2):
    X, Y = _tiny(1)
    R0 = rotation([0, 0, 1], 0.1)
    t0 = np.array([0.01, -0.02, 0.03])
    f32 = lambda v: float(np.float32(v))  # the oracle takes sigma_p2, d_02 as float
    lit = _literal.emicp_step(X.tolist(), Y.tolist(), R0.reshape(9).tolist(), t0.tolist(),
                              f32(sigma_p2), f32(0.01), eig_lapack)
    o = oracle.emicp_step(X, Y, R0, t0, sigma_p2, 0.01)
    np.testing.assert_allclose(o["C"], lit["C"], rtol=1e-13)
    np.testing.assert_allclose(o["lam"], lit["lam"], rtol=1e-13)
    assert abs(o["Lambda"] - lit["Lambda"]) < 1e-12
    np.testing.assert_allclose(o["Xc"], lit["Xc"], atol=1e-13)
    np.testing.assert_allclose(o["Yc"], lit["Yc"], atol=1e-13)
    np.testing.assert_allclose(o["S"], lit["S"], atol=1e-13)
    assert geodesic(o["R"], np.reshape(lit["R"], (3, 3))) < 1e-10
    np.testing.assert_allclose(o["t"], lit["t"], atol=1e-10)


def test_emicp_uniform_affinity_closed_form(oracle):
    """sigma^2 -> huge: A -> 1, C = N + 1, lambda = N / sqrt(N+1), X' = mean(X)."""
    X, Y = _tiny(2, n=9, m=5)
    o = oracle.emicp_step(X, Y, np.eye(3), np.zeros(3), 1e12, 0.01)
    N = len(X)
    np.testing.assert_allclose(o["C"], N + 1, rtol=1e-9)
    np.testing.assert_allclose(o["lam"], N / np.sqrt(N + 1), rtol=1e-9)
    np.testing.assert_allclose(o["Xp"], np.tile(X.mean(0), (len(Y), 1)), atol=1e-9)


def test_emicp_floored_row(oracle):
    """A source point far from every target at small sigma^2 (outlier term
    below 1e-6): its row is set to 1/N (src/emicp_cpu.cpp:121-125)."""
    X, Y = _tiny(3, n=8, m=4)
    Y = Y.copy()
    Y[2] = [5.0, 5.0, 5.0]
    o = oracle.emicp_step(X, Y, np.eye(3), np.zeros(3), 1e-4, 0.01)
    N = len(X)
    fl = float(np.float32(1.0) / np.float32(N))
    assert o["C"][2] < 1e-6
    assert abs(o["lam"][2] - N * fl) < 1e-12
    np.testing.assert_allclose(o["Xp"][2], X.mean(0) * N * fl / (N * fl), atol=1e-12)


def test_emicp_recovers_fixture_transform(oracle, pq_pairs):
    """data/P1007 (target) vs data/Qnew1007 (source) from 15 degrees off the
    exact answer. With the default schedule (66 steps) the soft fixed point
    sits within 1e-3 rad of P = R* Qnew + t*; annealed on to sigma^2 = 1e-7
    (110 steps) the assignment becomes hard and the transform is exact."""
    P, Q = pq_pairs[1007]
    Pc, Qc, S = s_matrix(P, Q)
    Rs, ts, _ = oracle.find_rt_from_s(Pc, Qc, S)
    R0 = rotation([1, -1, 2], np.deg2rad(15)) @ Rs
    t0 = Pc - R0 @ Qc
    R, t, tr = oracle.emicp(P, Q, R0, t0)
    assert len(tr) == 66
    assert geodesic(R, Rs) < 1e-3
    R, t, tr = oracle.emicp(P, Q, R0, t0, sigma_inf=1e-7)
    assert len(tr) == 110
    assert geodesic(R, Rs) < 1e-7
    assert np.abs(t - ts).max() < 1e-8


def test_emicp_f32_tracks_f64(oracle):
    X, Y, Rt, tt = uniform_cube(1007, seed=0)
    R64, t64, tr64 = oracle.emicp(X, Y)
    R32, t32, tr32 = oracle.emicp(X, Y, precision="f32")
    assert geodesic(R32.astype(np.float64), R64) < 1e-5
    assert np.abs(t32 - t64).max() < 1e-5
    # the registration itself succeeded (error against the generating transform)
    assert geodesic(R64, Rt) < 2e-3


def _tiny_sa(seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.5, 0.5, size=(6, 3))
    Y = np.vstack([X[:4] @ rotation([0, 1, 1], 0.3).T + 0.02, rng.uniform(-0.5, 0.5, (1, 3))])
    return X, Y


@pytest.mark.parametrize("T", [100.0, 1.5, 0.3])
def test_softassign_inner_matches_literal_loops(oracle, T):
    X, Y = _tiny_sa(4)
    R0 = rotation([1, 0, 0], 0.05)
    t0 = np.array([0.0, 0.01, -0.01])
    mo = 0.0367879
    orow_l, ocol_l = [mo] * len(X), [mo] * len(Y)
    orow, ocol = np.full(len(X), mo), np.full(len(Y), mo)
    for _ in range(2):  # two inner iterations: the slack state carries over
        lit = _literal.softassign_inner(X.tolist(), Y.tolist(), R0.reshape(9).tolist(),
                                        t0.tolist(), float(np.float32(T)), 3.0, 3, orow_l,
                                        ocol_l, eig_lapack)
        o = oracle.softassign_inner(X, Y, R0, t0, T, 3.0, 3, orow, ocol)
        np.testing.assert_allclose(o["rowsum"], lit["rowsum"], rtol=1e-12)
        np.testing.assert_allclose(o["colsum"], lit["colsum"], rtol=1e-12)
        # every sweep's row and column sums before their division
        np.testing.assert_allclose(o["a"], lit["a"], rtol=1e-12)
        np.testing.assert_allclose(o["b"], lit["b"], rtol=1e-12)
        np.testing.assert_allclose(orow, orow_l, rtol=1e-12)
        np.testing.assert_allclose(ocol, ocol_l, rtol=1e-12)
        np.testing.assert_allclose(o["S"], lit["S"], atol=1e-12)
        np.testing.assert_allclose(o["Xc"], lit["Xc"], atol=1e-12)
        np.testing.assert_allclose(o["Yc"], lit["Yc"], atol=1e-12)
        assert geodesic(o["R"], np.reshape(lit["R"], (3, 3))) < 1e-9
        R0, t0 = o["R"], o["t"]


def test_softassign_sinkhorn_invariants(oracle):
    """After the last column pass every column sum plus its slack is 1."""
    rng = np.random.default_rng(7)
    X = rng.uniform(-0.5, 0.5, size=(30, 3))
    Y = rng.uniform(-0.5, 0.5, size=(25, 3))
    orow, ocol = np.full(30, 0.0367879), np.full(25, 0.0367879)
    o = oracle.softassign_inner(X, Y, np.eye(3), np.zeros(3), 2.0, 3.0, 3, orow, ocol)
    Mat = o["Mat"]
    np.testing.assert_allclose(Mat.sum(0) + ocol, 1.0, rtol=1e-12)
    np.testing.assert_allclose(Mat.sum(1), o["rowsum"], rtol=1e-12)
    assert abs(o["sumM"] - Mat.sum()) < 1e-10


def test_softassign_f32_tracks_f64(oracle):
    """The FP32 restatement follows the FP64 oracle inner iteration by inner
    iteration. (At temperatures far below the point spacing the iteration
    becomes ill-conditioned and any two roundings part ways; this schedule
    stays above that.)"""
    rng = np.random.default_rng(11)
    X = rng.uniform(-0.5, 0.5, size=(120, 3))
    Y = (X[:100] - 0.03) @ rotation([1, 1, 0], 0.2)
    kw = dict(JMAX=30, I0=5, I1=3, T_0=0.5, TFACTOR=0.9, alpha=0.05)
    R64, t64, tr = oracle.softassign(X, Y, **kw)
    R32, t32, tr32 = oracle.softassign(X, Y, precision="f32", **kw)
    assert len(tr) == 150
    d = [geodesic(tr[i, 17:26].reshape(3, 3), tr32[i, 17:26].reshape(3, 3))
         for i in range(len(tr))]
    assert max(d) < 1e-5


def test_softassign_recovers_fixture_transform(oracle, pq_pairs):
    """Known answer: the reference's data/P1007.txt and data/Qnew1007.txt
    (row correspondence, one exact rigid transform) from a start 10 degrees
    off. With no slack (moutlier 0) and Sinkhorn run to convergence (I1 = 30),
    annealed until even the closest non-partners no longer count (T < 1e-7
    against squared spacings >= 5e-7), the soft correspondences become the
    exact permutation and Softassign must return the least-squares (Kabsch)
    transform of the true correspondences. A wrong sign, index or transpose
    in the affinities, the Sinkhorn sweeps, the weighted centroids, S or
    Horn's step would not converge there."""
    P, Q = pq_pairs[1007]
    Rs, ts = kabsch(P, Q)  # P ~ Rs Q + ts
    R0 = Rs @ rotation([1, 2, 3], np.deg2rad(10))
    t0 = P.mean(0) - R0 @ Q.mean(0)
    R, t, tr = oracle.softassign(P, Q, R0, t0, JMAX=56, I0=2, I1=30, alpha=0.0, T_0=0.01,
                                 TFACTOR=0.8, moutlier=0.0)
    assert tr[-1, 0] < 1e-7
    assert geodesic(R, Rs) < 1e-9
    np.testing.assert_allclose(t, ts, atol=1e-10)


def test_config3_fixture_reproduces(oracle):
    """tests/golden/config3_oracle.npz holds what the oracle computes: the
    smallest instance's clouds regenerate bit for bit (SHA-256) and two of
    its starts re-run to the stored FP64 and FP32 results."""
    from tests.golden import make_config3_oracle as G
    d = np.load(os.path.join(GOLDEN, "config3_oracle.npz"), allow_pickle=False)
    i = int(np.argmin(d["N"] * d["M"]))
    X, Y = G.instance(int(d["instances"][i]))
    assert G.clouds_sha(X, Y) == str(d["sha"][i])
    assert (len(X), len(Y)) == (d["N"][i], d["M"][i])
    for s in (0, 13):
        rt0 = d["rt0"][i, s]
        R, t, _ = oracle.emicp(X.astype(np.float64), Y.astype(np.float64),
                               rt0[:9].astype(np.float64).reshape(3, 3), rt0[9:].astype(np.float64))
        np.testing.assert_allclose(np.concatenate([R.reshape(9), t]), d["oracle_f64"][i, s],
                                   rtol=0, atol=1e-12)
        R, t, _ = oracle.emicp(X, Y, rt0[:9].reshape(3, 3), rt0[9:], precision="f32")
        np.testing.assert_array_equal(np.concatenate([R.reshape(9), t]).astype(np.float64),
                                      d["restatement_f32"][i, s])
