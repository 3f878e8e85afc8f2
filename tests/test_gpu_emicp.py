"""EM-ICP and Procrustes on the GPU (through the C ABI) against the FP64 oracle.

Tolerances (from the registration's own numerics):
  * R within 1e-4 rad geodesic, t within 1e-5 of the cloud extent of the FP64
    oracle of the identical iteration (asserted tighter where it holds);
  * per-step normalisers C_r and lambda_r within 1e-5 relative;
  * Horn's step from identical (Xc, Yc, S): R within 1e-6 rad.
"""
import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import geodesic, rotation, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import cuda_emicp_softassign_amd as E
    E._lib.load()
    return E


def _f32(a):
    return np.asarray(a, dtype=np.float32)


def test_find_rt_from_s_matches_oracle(E, oracle, pq_pairs):
    from tests.test_oracle import s_matrix
    rng = np.random.default_rng(0)
    cases = []
    for n in (101, 1007):
        P, Q = pq_pairs[n]
        cases.append(s_matrix(P, Q))
    for _ in range(6):
        Q = rng.normal(size=(50, 3))
        P = Q @ rotation(rng.normal(size=3), rng.uniform(0, 3)).T + rng.normal(size=3)
        cases.append(s_matrix(P + 0.01 * rng.normal(size=P.shape), Q))
    for Pc, Qc, S in cases:
        Pc, Qc, S = _f32(Pc), _f32(Qc), _f32(S)
        Ro, to, _ = oracle.find_rt_from_s(Pc.astype(np.float64), Qc.astype(np.float64),
                                          S.astype(np.float64))
        R, t = E.findRTfromS(Pc, Qc, S)
        assert geodesic(R.astype(np.float64), Ro) < 1e-6
        np.testing.assert_allclose(t, to, atol=2e-6)


def test_find_rt_from_s_rejects_nan(E):
    with pytest.raises(E.EmsaError):
        E.findRTfromS(np.zeros(3), np.zeros(3), np.full(9, np.nan))


@pytest.mark.parametrize("sigma_p2", [0.01, 1e-3, 1e-4, 1.0611e-5])
def test_step_normalisers_match_oracle(E, oracle, sigma_p2):
    X, Y, Rt, tt = uniform_cube(1007, seed=0)
    # a realistic state: a few oracle steps in
    R0 = rotation([1, 2, 3], 0.05) @ Rt
    t0 = tt + 0.002
    o = oracle.emicp_step(X.astype(np.float64), Y.astype(np.float64), _f32(R0).astype(float),
                          _f32(t0).astype(float), sigma_p2, 0.01)
    g = E.emicp_step(X, Y, R0, t0, sigma_p2, 0.01)
    keep = o["C"] > 2e-6  # away from the floor discontinuity at 1e-6
    np.testing.assert_allclose(g["C"][keep], o["C"][keep], rtol=1e-5)
    np.testing.assert_allclose(g["lam"], o["lam"], rtol=1e-5)
    rec = g["record"]
    assert abs(rec[1] - o["Lambda"]) <= 1e-5 * o["Lambda"]
    np.testing.assert_allclose(rec[2:5], o["Xc"], atol=1e-6)
    np.testing.assert_allclose(rec[5:8], o["Yc"], atol=1e-6)
    np.testing.assert_allclose(rec[8:17], o["S"], atol=1e-5 * np.abs(o["S"]).max())
    assert geodesic(g["R"].astype(np.float64), o["R"]) < 2e-5
    np.testing.assert_allclose(g["t"], o["t"], atol=2e-6)


@pytest.mark.parametrize("sigma_p2", [0.01, 4e-3, 2e-3])
def test_step_normalisers_expanded_form(E, oracle, sigma_p2):
    """The affinity's exponent expanded about each wave's row centre on the
    matrix cores (taken by a wave when ka^2 a^2 <= 32, a its rows' distance
    from the centre, ka^2 = log2 e / (2 sigma^2)) and the VALU difference
    (bound 0) each give per-row C_r and lambda_r within 1e-5 relative of the
    FP64 oracle, on a 4000-point cube where every wave qualifies at 1e-2 and
    some at 2e-3; the two forms differ (the expansion was taken)."""
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    X, Y, Rt, tt = uniform_cube(4000, seed=7)
    R0 = rotation([1, 2, 3], 0.02) @ Rt
    t0 = tt + 0.001
    o = oracle.emicp_step(X.astype(np.float64), Y.astype(np.float64), _f32(R0).astype(float),
                          _f32(t0).astype(float), sigma_p2, 0.01)
    keep = o["C"] > 2e-6
    got = {}
    try:
        for bound in (0.0, 32.0):
            _lib.check("expanded", L.emsa_probe_set_affinity_expanded(bound))
            g = E.emicp_step(X, Y, R0, t0, sigma_p2, 0.01)
            relC = np.max(np.abs(g["C"][keep] - o["C"][keep]) / o["C"][keep])
            relL = np.max(np.abs(g["lam"] - o["lam"]) / np.maximum(o["lam"], 1e-300))
            print("sigma^2 %g bound %g: rel err C %.1e lambda %.1e" % (sigma_p2, bound, relC, relL),
                  flush=True)
            assert relC <= 1e-5 and relL <= 1e-5
            assert abs(g["record"][1] - o["Lambda"]) <= 1e-5 * o["Lambda"]
            assert geodesic(g["R"].astype(np.float64), o["R"]) < 2e-5
            got[bound] = g["C"]
    finally:
        _lib.check("expanded", L.emsa_probe_set_affinity_expanded(32.0))
    assert not np.array_equal(got[0.0], got[32.0])


@pytest.mark.parametrize("k", [0, 13, 24, 30, 37])
def test_step_normalisers_partial_overlap(E, oracle, k):
    """A partial-overlap range-scan pair with few targets (60 x 1500): most
    rows' C_r are made of a handful of distant terms and, from step 25 on,
    hundreds of rows are floored. Step k of the sigma^2 = 0.01 * 0.9^k
    schedule from the oracle's own state after k steps (the same float32
    (R, t) to both sides): C_r (away from the floor), lambda_r and Lambda
    within 1e-5 relative."""
    from cuda_emicp_softassign_amd.clouds import range_scans
    X, Y, _, _ = range_scans(60, 1500, seed=100)
    Rt0 = E.multistart_rt0(X, Y, [np.eye(3)])[0].astype(np.float64)
    sig = oracle.sigma_schedule(0.01, 2e-4, 0.9)
    R, t = Rt0[:9].reshape(3, 3), Rt0[9:]
    if k > 0:
        R, t, _ = oracle.emicp(X.astype(np.float64), Y.astype(np.float64), R, t, sigma_inf=2e-4,
                               max_iter=k)
    R32, t32 = _f32(R), _f32(t)
    o = oracle.emicp_step(X.astype(np.float64), Y.astype(np.float64), R32.astype(float),
                          t32.astype(float), float(sig[k]), 0.01)
    g = E.emicp_step(X, Y, R32.copy(), t32.copy(), float(sig[k]), 0.01)
    keep = o["C"] > 2e-6
    relC = np.max(np.abs(g["C"][keep] - o["C"][keep]) / o["C"][keep])
    relL = np.max(np.abs(g["lam"] - o["lam"]) / np.maximum(o["lam"], 1e-300))
    print("step %d sigma^2 %.3g: rel err C %.1e lambda %.1e, floored %d" %
          (k, sig[k], relC, relL, (o["C"] <= 1e-6).sum()), flush=True)
    assert relC <= 1e-5 and relL <= 1e-5
    assert abs(g["record"][1] - o["Lambda"]) <= 1e-5 * o["Lambda"]
    assert geodesic(g["R"].astype(np.float64), o["R"]) < 2e-5


def test_floored_rows(E, oracle):
    X, Y, _, _ = uniform_cube(500, seed=3)
    Y = Y.copy()
    Y[7] = [9.0, 9.0, 9.0]
    Y[300] = [-9.0, 9.0, 9.0]
    o = oracle.emicp_step(X.astype(float), Y.astype(float), np.eye(3), np.zeros(3), 1e-4, 0.01)
    g = E.emicp_step(X, Y, np.eye(3), np.zeros(3), 1e-4, 0.01)
    assert o["C"][7] < 1e-6 and o["C"][300] < 1e-6
    np.testing.assert_allclose(g["lam"][[7, 300]], o["lam"][[7, 300]], rtol=1e-12)
    assert g["record"][30] >= 2  # floored-row count


def test_config0_parity(E, oracle):
    """N = M = 1007 uniform cube, defaults (66 sigma steps): GPU vs FP64 oracle
    of the identical iteration and vs the FP32 restatement of emicp_cpu."""
    X, Y, Rt, tt = uniform_cube(1007, seed=0)
    Ro, to, tro = oracle.emicp(X.astype(np.float64), Y.astype(np.float64))
    R32, t32, _ = oracle.emicp(X, Y, precision="f32")
    R, t, tr = E.emicp_traced(X, Y, np.eye(3, dtype=np.float32), np.zeros(3, np.float32))
    assert len(tr) == len(tro) == 66
    extent = np.ptp(X, axis=0).max()
    assert geodesic(R.astype(np.float64), Ro) < 1e-5
    assert np.abs(t - to).max() < 1e-5 * extent
    assert geodesic(R.astype(np.float64), R32.astype(np.float64)) < 1e-5
    for k in range(len(tr)):
        assert geodesic(tr[k, 17:26].reshape(3, 3), tro[k, 17:26].reshape(3, 3)) < 1e-5, k
        assert abs(tr[k, 1] - tro[k, 1]) <= 1e-5 * tro[k, 1], k
    assert geodesic(R.astype(np.float64), Rt) < 2e-3  # and the registration worked


def test_fixture_exact_recovery(E, oracle, pq_pairs):
    """The reference's data/P1007 vs data/Qnew1007 (one exact rigid transform)
    from 15 degrees off, annealed to sigma^2 = 1e-7: the exact transform."""
    from tests.test_oracle import s_matrix
    P, Q = pq_pairs[1007]
    Pc, Qc, S = s_matrix(P, Q)
    Rs, ts, _ = oracle.find_rt_from_s(Pc, Qc, S)
    R0 = rotation([1, -1, 2], np.deg2rad(15)) @ Rs
    t0 = Pc - R0 @ Qc
    p = E.registrationParameters(sigma_inf=1e-7)
    R, t = E.emicp(P, Q, R0, t0, p)
    assert geodesic(R.astype(np.float64), Rs) < 2e-6
    assert np.abs(t - ts).max() < 2e-6


@pytest.mark.parametrize("N,M", [(1, 1), (9, 1), (17, 3), (63, 65), (65, 63),
                                 (200, 1000), (1000, 200), (4097, 31)])
def test_ragged_sizes(E, oracle, N, M):
    rng = np.random.default_rng(N * 7919 + M)
    X = rng.uniform(-0.5, 0.5, size=(N, 3)).astype(np.float32)
    Y = ((X[rng.integers(0, N, size=M)] - 0.02) @ rotation([0, 1, 1], 0.1)).astype(np.float32)
    Y += rng.normal(0, 0.003, size=Y.shape).astype(np.float32)
    kw = dict(sigma_p2=0.01, sigma_inf=1e-4)
    Ro, to, tro = oracle.emicp(X.astype(float), Y.astype(float), **kw)
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), E.registrationParameters(**kw))
    assert len(tr) == len(tro)
    assert geodesic(R.astype(np.float64), Ro) < 1e-4
    assert np.abs(t - to).max() < 1e-5


def test_zero_iterations_leave_rt_unchanged(E):
    X, Y, _, _ = uniform_cube(100, seed=1)
    R0 = rotation([1, 0, 0], 0.3).astype(np.float32)
    t0 = np.array([0.1, 0.2, 0.3], np.float32)
    hR, ht = R0.copy(), t0.copy()
    R, t = E.emicp(X, Y, hR, ht, E.registrationParameters(sigma_p2=1e-5, sigma_inf=1e-5))
    assert np.array_equal(R, R0) and np.array_equal(t, t0)


def test_errors_leave_rt_unchanged(E):
    X, Y, _, _ = uniform_cube(100, seed=1)
    hR = np.eye(3, dtype=np.float32)
    ht = np.zeros(3, np.float32)
    with pytest.raises(E.EmsaError) as ei:
        E.emicp(np.zeros((0, 3)), Y, hR, ht)
    assert ei.value.code == E._lib.EMSA_EINVAL
    Xn = X.copy()
    Xn[3, 1] = np.nan
    with pytest.raises(E.EmsaError) as ei:
        E.emicp(Xn, Y, hR, ht)
    assert ei.value.code == E._lib.EMSA_ENONFINITE
    assert np.array_equal(hR, np.eye(3)) and not ht.any()


def test_reproducible(E):
    X, Y, _, _ = uniform_cube(3000, seed=5)
    R1, t1 = E.emicp(X, Y, np.eye(3), np.zeros(3))
    R2, t2 = E.emicp(X, Y, np.eye(3), np.zeros(3))
    assert np.array_equal(R1, R2) and np.array_equal(t1, t2)


def test_device_entry_point_with_torch(E, oracle):
    import torch
    X, Y, Rt, tt = uniform_cube(2000, seed=7)
    dX = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    dRt = torch.zeros(12, dtype=torch.float32, device="cuda")
    dRt[0] = dRt[4] = dRt[8] = 1.0
    E.emicp_dev(dX.data_ptr(), 2000, dY.data_ptr(), 2000, dRt.data_ptr(),
                stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out = dRt.cpu().numpy()
    R, t = E.emicp(X, Y, np.eye(3), np.zeros(3))
    assert geodesic(out[:9].reshape(3, 3).astype(float), R.astype(float)) < 1e-7
    np.testing.assert_allclose(out[9:], t, atol=1e-7)


def test_single_target_point():
    """N = 1: every source row averages to the one target, S vanishes and any
    rotation is a minimiser; the translation must still put R Yc + t on it."""
    import cuda_emicp_softassign_amd as E
    rng = np.random.default_rng(2)
    X = rng.uniform(-0.5, 0.5, size=(1, 3)).astype(np.float32)
    Y = (X + rng.normal(0, 0.01, size=(9, 3))).astype(np.float32)
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3),
                              E.registrationParameters(sigma_inf=1e-4))
    R = R.astype(np.float64)
    assert abs(np.linalg.det(R) - 1) < 1e-5
    Yc = tr[-1, 5:8]
    np.testing.assert_allclose(R @ Yc + t, X[0], atol=1e-6)


def test_large_matches_oracle_first_steps(E, oracle):
    """N = M = 10000: the first 3 sigma^2 steps against the FP64 oracle (the
    full 66-step oracle run is the benchmark's parity check), and the full
    registration returns a proper rotation with no failed Procrustes step."""
    X, Y, Rt, tt = uniform_cube(10000, seed=10)
    kw = dict(sigma_inf=float(np.float32(0.01) * np.float32(0.9) * np.float32(0.9) * 0.95))
    Ro, to, tro = oracle.emicp(X.astype(np.float64), Y.astype(np.float64), **kw)
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), E.registrationParameters(**kw))
    assert len(tr) == len(tro) == 3
    assert geodesic(R.astype(np.float64), Ro) < 1e-5
    assert np.abs(t - to).max() < 1e-5
    for k in range(3):
        assert abs(tr[k, 1] - tro[k, 1]) <= 1e-5 * tro[k, 1]
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3))
    R = R.astype(np.float64)
    assert len(tr) == 66
    assert abs(np.linalg.det(R) - 1) < 1e-5
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-5)
    assert geodesic(R, Rt) < 5e-3
    assert not tr[:, 31].any()  # no failed Procrustes step


def test_large_matches_oracle_full_schedule(E, oracle):
    """N = M = 10000 (the benchmark's pair 0), the full 66-step schedule
    against the FP64 oracle run from the same start: R within 1e-4 rad and t
    within 1e-5 of the extent after every step, the final registration within
    1e-6 rad, and Lambda of every step within 1e-5 relative."""
    X, Y, Rt, tt = uniform_cube(10000, seed=10)
    Ro, to, tro = oracle.emicp(X.astype(np.float64), Y.astype(np.float64))
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3))
    assert len(tr) == len(tro) == 66
    extent = float(np.ptp(X, axis=0).max())
    rot = [geodesic(tr[k, 17:26].reshape(3, 3), tro[k, 17:26].reshape(3, 3)) for k in range(66)]
    dt = [np.abs(tr[k, 26:29] - tro[k, 26:29]).max() / extent for k in range(66)]
    dl = [abs(tr[k, 1] - tro[k, 1]) / tro[k, 1] for k in range(66)]
    print("10k full schedule: max per-step rot err %.2e rad, |dt|/extent %.2e, Lambda %.2e; "
          "final %.2e rad" % (max(rot), max(dt), max(dl), geodesic(R.astype(np.float64), Ro)),
          flush=True)
    assert max(rot) < 1e-4 and max(dt) < 1e-5 and max(dl) < 1e-5
    assert geodesic(R.astype(np.float64), Ro) < 1e-6
    assert np.abs(t - to).max() < 1e-5 * extent


def test_dropin_cxx_caller():
    """tests/cpp/dropin_main.cpp, a caller of the reference API compiled
    against include/3dregistration.h and linked to libemsa.so."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "_build", "dropin_demo")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "dropin OK" in r.stdout


def test_zero_lambda_row_contributes_nothing(E, oracle):
    """The documented deviation at lambda_r = 0: a source row so far from the
    target that all its affinities underflow while the outlier term keeps C_r
    above the 1e-6 floor. The reference (src/emicp.cu:540-549) divides 0 by 0
    there: Xc and S turn NaN and Horn's step fails; the FP64 oracle and
    emicp_cpu (the reference's FP32 operations) keep that, so those steps
    leave (R, t) as they were (emicp_cpu flags them in its trace) until the
    row is floored at smaller sigma^2. The GPU step gives lambda_r = 0
    exactly, C_r = the outlier term, and otherwise the oracle's step without
    that row: the row contributes nothing, and no step fails."""
    X, Y, Rt, tt = uniform_cube(1007, seed=0)
    far = np.array([[3.0, 3.0, 3.0]], np.float32)
    Yf = np.vstack([Y, far]).astype(np.float32)
    R0 = rotation([1, 2, 3], 0.05) @ Rt
    t0 = tt + 0.002
    s2 = 0.01
    o_far = oracle.emicp_step(X.astype(np.float64), Yf.astype(np.float64), _f32(R0).astype(float),
                              _f32(t0).astype(float), s2, 0.01)
    assert np.isnan(o_far["Xc"]).all() and np.isnan(o_far["S"]).any()  # the reference's 0/0
    np.testing.assert_array_equal(o_far["R"], _f32(R0).astype(float))  # failed solve: R kept
    o = oracle.emicp_step(X.astype(np.float64), Y.astype(np.float64), _f32(R0).astype(float),
                          _f32(t0).astype(float), s2, 0.01)
    g = E.emicp_step(X, Yf, R0, t0, s2, 0.01)
    assert g["lam"][-1] == 0.0
    assert abs(g["C"][-1] - np.exp(-0.01 / s2)) < 1e-6
    np.testing.assert_allclose(g["lam"][:-1], o["lam"], rtol=1e-5)
    rec = g["record"]
    assert abs(rec[1] - o["Lambda"]) <= 1e-5 * o["Lambda"]
    np.testing.assert_allclose(rec[2:5], o["Xc"], atol=1e-6)
    np.testing.assert_allclose(rec[5:8], o["Yc"], atol=1e-6)
    assert geodesic(g["R"].astype(np.float64), o["R"]) < 2e-5
    np.testing.assert_allclose(g["t"], o["t"], atol=2e-6)
    # the whole registration: no failed step on the GPU; emicp_cpu (FP32, the
    # reference's order) fails every step before the row is floored
    R, t, tr = E.emicp_traced(X, Yf, np.eye(3), np.zeros(3))
    assert np.isfinite(R).all() and np.isfinite(t).all() and (tr[:, 31] == 0).all()
    _, _, trc = E.emicp_cpu(X, Yf, np.eye(3), np.zeros(3), traced=True)
    failed = trc[:, 31] != 0
    assert failed[0] and not failed[-1]
    assert (np.exp(-0.01 / trc[failed, 0]) > 1e-6).all()  # failed only while C_r is kept up
