"""Softassign on the GPU (both variants) against the FP64 oracle.

The reference has no CPU Softassign, so parity is against the FP64 oracle of
the identical iteration (oracle/emsa_oracle_impl.h, literal materialised
matrix with in-place divisions). Tolerances: the per-iteration normaliser
sum M within 1e-5 relative (measured <= 2.3e-7 over these cases), R within
2e-5 rad per inner iteration (measured <= 3.1e-6) and 1e-4 rad at the end,
t within 1e-5 of the cloud extent.
"""
import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import geodesic, rotation, surface_with_outliers

pytestmark = pytest.mark.gpu

VARIANTS = ["materialised", "logdomain"]


@pytest.fixture(scope="module")
def E():
    import cuda_emicp_softassign_amd as E
    E._lib.load()
    return E


def _case(seed=11, n=300, m=280):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.5, 0.5, size=(n, 3))
    Y = (X[:m] - 0.03) @ rotation([1, 1, 0], 0.2) + rng.normal(0, 0.002, size=(m, 3))
    return X.astype(np.float32), Y.astype(np.float32)


def _compare(E, oracle, X, Y, variant, tol=1e-4, step_tol=2e-5, norm_tol=1e-5, **kw):
    p = E.registrationParameters(**kw)
    Ro, to, tro = oracle.softassign(X.astype(np.float64), Y.astype(np.float64),
                                    JMAX=p.JMAX, I0=p.I0, I1=p.I1, alpha=p.alpha, T_0=p.T_0,
                                    TFACTOR=p.TFACTOR, moutlier=p.moutlier)
    R, t, tr = E.softassign_traced(X, Y, np.eye(3), np.zeros(3), p, variant=variant)
    assert len(tr) == len(tro) == p.JMAX * p.I0
    for k in range(len(tr)):
        assert geodesic(tr[k, 17:26].reshape(3, 3), tro[k, 17:26].reshape(3, 3)) < step_tol, k
        assert abs(tr[k, 1] - tro[k, 1]) <= norm_tol * abs(tro[k, 1]), k
    extent = np.ptp(X, axis=0).max()
    assert geodesic(R.astype(np.float64), Ro) < tol
    assert np.abs(t - to).max() < 1e-5 * extent
    return R, t, tr


@pytest.mark.parametrize("variant", VARIANTS)
def test_short_schedule_matches_oracle(E, oracle, variant):
    X, Y = _case()
    _compare(E, oracle, X, Y, variant, JMAX=30, I0=5, I1=3, T_0=0.5, TFACTOR=0.9, alpha=0.05)


@pytest.mark.parametrize("variant", VARIANTS)
def test_default_parameters_match_oracle(E, oracle, variant):
    """Reference defaults (JMAX 100, I0 5, I1 3, alpha 3, T_0 100, 0.95)."""
    X, Y = _case(seed=4, n=257, m=190)
    _compare(E, oracle, X, Y, variant)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("N,M", [(7, 4), (9, 12), (20, 300), (300, 20), (65, 64), (257, 513),
                                 (100, 1000), (1000, 100), (300, 2049), (600, 6144),
                                 (600, 6145)])
def test_ragged_sizes(E, oracle, variant, N, M):
    """Uneven and tiny sizes; Y is a moved, noisy resampling of X so that
    Horn's problem is well posed (for unrelated clouds, or a handful of points
    matched by hundreds, the top eigenvalue of Horn's matrix is nearly
    degenerate and the FP32 and FP64 oracles themselves part by radians).
    The materialised sweep holds up to 3 slabs of 2048 columns in registers:
    M = 2049 leaves a partly empty slab, 6144 fills all three, 6145 takes the
    separate row and column passes."""
    rng = np.random.default_rng(N + 31 * M)
    X = rng.uniform(-0.5, 0.5, size=(N, 3))
    Y = (X[rng.integers(0, N, size=M)] - 0.02) @ rotation([1, 0, 1], 0.15)
    Y = Y + rng.normal(0, 0.005, size=Y.shape)
    _compare(E, oracle, X.astype(np.float32), Y.astype(np.float32), variant, JMAX=8, I0=2,
             I1=2, T_0=1.0, TFACTOR=0.9)


@pytest.mark.parametrize("variant", VARIANTS)
def test_cold_schedule(E, oracle, variant):
    """T from 0.05 down to 0.0024: k1 max |x|^2 passes 64 after 12 of the
    30 temperatures, so the log domain evaluates the exponent expanded first
    and as the direct difference after; both variants match the oracle
    throughout."""
    X, Y = _case(seed=8, n=200, m=180)
    _compare(E, oracle, X, Y, variant, JMAX=30, I0=5, I1=3, T_0=0.05, TFACTOR=0.9,
             alpha=0.05)


def test_variants_agree_full_size(E):
    """N = M = 5032 surface + 10% outliers, reference defaults: the
    materialised and log-domain variants give the same registration."""
    X, Y, Rt, tt = surface_with_outliers(5032, seed=1)
    Rm, tm = E.softassign(X, Y, np.eye(3), np.zeros(3), variant="materialised")
    Rl, tl = E.softassign(X, Y, np.eye(3), np.zeros(3), variant="logdomain")
    assert geodesic(Rm.astype(np.float64), Rl.astype(np.float64)) < 1e-4
    assert np.abs(tm - tl).max() < 1e-5
    assert abs(np.linalg.det(Rm.astype(np.float64)) - 1) < 1e-5


def test_invalid_and_empty_schedules(E):
    X, Y = _case(n=50, m=40)
    hR = np.eye(3, dtype=np.float32)
    ht = np.zeros(3, np.float32)
    with pytest.raises(E.EmsaError) as ei:
        E.softassign(X, Y, hR, ht, E.registrationParameters(I1=0))
    assert ei.value.code == E._lib.EMSA_EINVAL
    R0 = rotation([0, 0, 1], 0.4).astype(np.float32)
    R, t = E.softassign(X, Y, R0.copy(), ht.copy(), E.registrationParameters(JMAX=0))
    assert np.array_equal(R, R0) and not t.any()


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def _steps_vs_oracle(E, oracle, X, Y, variant, temps, alpha=3.0, I1=3, R0=None, t0=None,
                     moutlier=None):
    """Inner iterations of emsa_softassign_step, each compared with the FP64
    oracle's inner iteration started from the same state (R, t, o_row,
    o_col): every sweep's row sums a_r and column sums b_c, the post-Sinkhorn
    row sums s_r and Sigma within 1e-5 relative; Xc, Yc and t within 1e-5 of
    the cloud extent; S within 1e-5 of its largest entry; R within 2e-5 rad.
    Returns the largest relative normaliser error seen."""
    mo = float(E.registrationParameters().moutlier) if moutlier is None else moutlier
    orow = np.full(len(X), mo, np.float32)
    ocol = np.full(len(Y), mo, np.float32)
    R = np.eye(3) if R0 is None else np.asarray(R0, np.float64)
    t = np.zeros(3) if t0 is None else np.asarray(t0, np.float64)
    ext = np.ptp(X, axis=0).max()
    worst = 0.0
    for k, T in enumerate(temps):
        o64r, o64c = orow.astype(np.float64), ocol.astype(np.float64)
        o = oracle.softassign_inner(X.astype(np.float64), Y.astype(np.float64),
                                    R.astype(np.float32).astype(np.float64),
                                    t.astype(np.float32).astype(np.float64), float(T), alpha,
                                    I1, o64r, o64c)
        g = E.softassign_step(X, Y, R.astype(np.float32), t.astype(np.float32), T, alpha, I1,
                              orow, ocol, variant=variant)
        errs = [_rel(g["a"], o["a"]), _rel(g["b"], o["b"]), _rel(g["s"], o["rowsum"]),
                abs(g["record"][1] - o["sumM"]) / abs(o["sumM"]),
                _rel(orow, o64r), _rel(ocol, o64c)]
        print("  %s step %d T=%g: rel err a %.1e b %.1e s %.1e Sigma %.1e o_row %.1e o_col %.1e"
              % (variant, k, T, *errs), flush=True)
        worst = max(worst, max(errs))
        assert max(errs) <= 1e-5, (k, errs)
        rec = g["record"]
        np.testing.assert_allclose(rec[2:5], o["Xc"], atol=1e-5 * ext)
        np.testing.assert_allclose(rec[5:8], o["Yc"], atol=1e-5 * ext)
        np.testing.assert_allclose(rec[8:17], o["S"], atol=1e-5 * np.abs(o["S"]).max())
        assert geodesic(g["R"].astype(np.float64), o["R"]) < 2e-5, k
        np.testing.assert_allclose(g["t"], o["t"], atol=1e-5 * ext)
        R, t = g["R"].astype(np.float64), g["t"].astype(np.float64)
    return worst


@pytest.mark.parametrize("variant", VARIANTS)
def test_step_normalisers_small(E, oracle, variant):
    """Per sweep and inner iteration on a 300 x 280 case: a hot step, a
    middle one and cold ones (slack state carried from step to step)."""
    X, Y = _case()
    _steps_vs_oracle(E, oracle, X, Y, variant, [100.0, 5.0, 0.5, 0.1, 0.05], alpha=0.05)


@pytest.mark.parametrize("variant", VARIANTS)
def test_step_normalisers_config1_size(E, oracle, variant):
    """N = M = 5032 (surface + 10% outliers, reference defaults): the first 5
    inner iterations of the schedule (T = 100) and 3 at the schedule's cold
    end (T = 0.7, 0.66, 0.6232)."""
    X, Y, _, _ = surface_with_outliers(5032, seed=1)
    print("config-1 size, %s" % variant, flush=True)
    _steps_vs_oracle(E, oracle, X, Y, variant, [100.0] * 5 + [0.7, 0.66, 0.6232])


def test_step_rejects_bad_arguments(E):
    X, Y = _case(n=50, m=40)
    orow = np.zeros(50, np.float32)
    ocol = np.zeros(40, np.float32)
    with pytest.raises(E.EmsaError) as ei:
        E.softassign_step(X, Y, np.eye(3), np.zeros(3), 1.0, 0.0, 0, orow, ocol)
    assert ei.value.code == E._lib.EMSA_EINVAL
    with pytest.raises(E.EmsaError) as ei:
        E.softassign_step(X, Y, np.eye(3), np.zeros(3), 0.0, 0.0, 3, orow, ocol)
    assert ei.value.code == E._lib.EMSA_ENONFINITE
    with pytest.raises(ValueError):
        E.softassign_step(X, Y, np.eye(3), np.zeros(3), 1.0, 0.0, 3, orow[:10], ocol)


@pytest.mark.parametrize("variant", VARIANTS)
def test_known_answer_fixture_transform(E, pq_pairs, variant):
    """The reference's data/P1007.txt / data/Qnew1007.txt from 10 degrees off,
    no slack, Sinkhorn run to convergence (I1 = 30), annealed to T < 1e-7:
    the soft correspondences become the exact permutation and the result is
    the least-squares transform of the true correspondences (the FP64 oracle
    reaches it to 1e-11 rad, tests/test_oracle.py); in FP32 to 1e-5 rad."""
    from tests.test_oracle import kabsch
    P, Q = pq_pairs[1007]
    Rs, ts = kabsch(P, Q)
    R0 = Rs @ rotation([1, 2, 3], np.deg2rad(10))
    t0 = P.mean(0) - R0 @ Q.mean(0)
    p = E.registrationParameters(JMAX=56, I0=2, I1=30, alpha=0.0, T_0=0.01, TFACTOR=0.8,
                                 moutlier=0.0)
    R, t = E.softassign(P.astype(np.float32), Q.astype(np.float32), R0.astype(np.float32),
                        t0.astype(np.float32), p, variant=variant)
    err = geodesic(R.astype(np.float64), Rs)
    print("  %s: rot err %.2e rad, |dt| %.2e" % (variant, err, np.abs(t - ts).max()), flush=True)
    assert err < 1e-5
    assert np.abs(t - ts).max() < 1e-5 * np.ptp(P, axis=0).max()
