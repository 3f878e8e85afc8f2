"""emsa_emicp_cpu, the library's FP32 host EM-ICP (the counterpart of the
reference's emicp_cpu, src/emicp_cpu.cpp:186-547), on the CPU only.

Bars: on the N = M = 1007 uniform-cube workload at the reference defaults
(66 sigma^2 steps) R within 1e-4 rad and t within 1e-5 of the extent of the
FP64 oracle of the identical iteration; within 1e-5 rad, step by step, of the
oracle's FP32 restatement (the same FP32 operations, sums in another order);
bit-identical results for any thread count; errors reported before any work.
"""
import os

import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import geodesic, rotation, uniform_cube


@pytest.fixture(scope="module")
def E():
    import cuda_emicp_softassign_amd as E
    E._lib.load()
    return E


def test_config0_matches_fp64_oracle(E, oracle):
    X, Y, Rt, tt = uniform_cube(1007, seed=0)
    Ro, to, tro = oracle.emicp(X.astype(np.float64), Y.astype(np.float64))
    R32, t32, tr32 = oracle.emicp(X, Y, precision="f32")
    R, t, tr = E.emicp_cpu(X, Y, np.eye(3), np.zeros(3), traced=True)
    assert len(tr) == len(tro) == 66
    extent = np.ptp(X, axis=0).max()
    print("emicp_cpu vs FP64 oracle: %.2e rad, |dt|/extent %.2e; vs FP32 restatement %.2e rad"
          % (geodesic(R.astype(float), Ro), np.abs(t - to).max() / extent,
             geodesic(R.astype(float), R32.astype(float))))
    assert geodesic(R.astype(np.float64), Ro) < 1e-4
    assert np.abs(t - to).max() < 1e-5 * extent
    for k in range(len(tr)):
        assert geodesic(tr[k, 17:26].reshape(3, 3), tr32[k, 17:26].reshape(3, 3)) < 1e-5, k
        assert abs(tr[k, 1] - tr32[k, 1]) <= 1e-5 * tr32[k, 1], k
    assert geodesic(R.astype(np.float64), Rt) < 2e-3


def test_thread_count_does_not_change_bits(E, monkeypatch):
    X, Y, _, _ = uniform_cube(600, seed=5)
    p = E.registrationParameters(sigma_inf=1e-4)
    out = []
    for n in ("1", "3", "8"):
        monkeypatch.setenv("EMSA_CPU_THREADS", n)
        out.append(E.emicp_cpu(X, Y, np.eye(3), np.zeros(3), p))
    for R, t in out[1:]:
        assert np.array_equal(R, out[0][0]) and np.array_equal(t, out[0][1])


def test_floored_rows_and_ragged_sizes(E, oracle):
    """Two source points far from every target point (C_r below the 1e-6
    floor, A' = 1/N) and uneven sizes."""
    rng = np.random.default_rng(2)
    X = rng.uniform(-0.5, 0.5, size=(301, 3)).astype(np.float32)
    Y = ((X[:150] - 0.01) @ rotation([0, 1, 0], 0.1)).astype(np.float32)
    Y[7] = [9.0, 9.0, 9.0]
    Y[100] = [-9.0, 9.0, 9.0]
    p = E.registrationParameters(sigma_inf=1e-4)
    Ro, to, _ = oracle.emicp(X.astype(np.float64), Y.astype(np.float64), sigma_inf=1e-4)
    R, t = E.emicp_cpu(X, Y, np.eye(3), np.zeros(3), p)
    assert geodesic(R.astype(np.float64), Ro) < 1e-4
    assert np.abs(t - to).max() < 1e-5


def test_errors_leave_rt_unchanged(E):
    X = np.zeros((4, 3), np.float32)
    hR = np.eye(3, dtype=np.float32)
    ht = np.ones(3, np.float32)
    with pytest.raises(E.EmsaError) as ei:
        E.emicp_cpu(np.zeros((0, 3)), X, hR, ht)
    assert ei.value.code == E._lib.EMSA_EINVAL
    bad = X.copy()
    bad[0, 0] = np.nan
    with pytest.raises(E.EmsaError) as ei:
        E.emicp_cpu(bad, X, hR, ht)
    assert ei.value.code == E._lib.EMSA_ENONFINITE
    with pytest.raises(E.EmsaError) as ei:
        E.emicp_cpu(X, X, hR, ht, E.registrationParameters(sigma_factor=1.0))
    assert ei.value.code == E._lib.EMSA_EINVAL
    assert np.array_equal(hR, np.eye(3)) and np.array_equal(ht, np.ones(3))
    R, t = E.emicp_cpu(X, X, hR, ht, E.registrationParameters(sigma_p2=1e-6, sigma_inf=1e-5))
    assert np.array_equal(R, np.eye(3)) and np.array_equal(t, np.ones(3))
