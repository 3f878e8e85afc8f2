"""Batched multi-start EM-ICP, its hipGraph replay and the autotuner.

Bar: every registration of a batch and the plain run of the same pair and
start are within 1e-4 rad / 1e-5 of the extent of the FP64 oracle, and
equal each other to 1e-6 on the well-posed pairs; graph replays are
bit-identical to the eager batched run; the multi-start selection finds the
generating rotation.
"""
import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import (geodesic, octahedral_rotations, range_scans,
                                              rotation, uniform_cube)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import cuda_emicp_softassign_amd as E
    E._lib.load()
    return E


def _pairs():
    out = []
    for k, (n, m) in enumerate([(60, 1500), (2500, 700), (1007, 1007), (33, 17), (4100, 3900)]):
        X, Y, R, t = range_scans(n, m, seed=100 + k)
        out.append((X, Y))
    return out


def test_batched_matches_plain(E, oracle):
    """Every registration of the batch and the plain single-registration run
    of the same pair and start are each within 1e-4 rad and 1e-5 of the
    target's extent of the FP64 oracle of the identical iteration. The batch
    lays out pairs with 2048-column tiles, a plain call spreads one
    registration over the device (the partial sums are added in another FP32
    order); on the well-posed pairs 2-4 the two agree to 1e-6."""
    pairs = _pairs()
    starts = [np.eye(3), rotation([0, 0, 1], 0.3), rotation([1, 0, 0], -0.2)]
    Rt0 = np.stack([E.multistart_rt0(X, Y, starts) for X, Y in pairs])
    p = E.registrationParameters(sigma_inf=2e-4)
    Rt, res = E.emicp_batched(pairs, Rt0, p)
    assert Rt.shape == (len(pairs), 3, 12) and res.shape == (len(pairs), 3)
    assert np.isfinite(res).all() and (res >= 0).all()
    for k, (X, Y) in enumerate(pairs):
        ext = np.ptp(X, axis=0).max()
        for s in range(3):
            R0 = Rt0[k, s, :9].reshape(3, 3).copy()
            t0 = Rt0[k, s, 9:].copy()
            R, t = E.emicp(X, Y, R0.copy(), t0.copy(), p)  # in/out: keep R0, t0
            Ro, to, _ = oracle.emicp(X.astype(np.float64), Y.astype(np.float64),
                                     R0.astype(np.float64), t0.astype(np.float64),
                                     sigma_inf=p.sigma_inf)
            Rb = Rt[k, s, :9].reshape(3, 3).astype(float)
            eb, ep = geodesic(Rb, Ro), geodesic(R.astype(float), Ro)
            print("pair %d start %d: batched %.1e rad, plain %.1e rad from FP64" % (k, s, eb, ep),
                  flush=True)
            assert eb < 1e-4 and ep < 1e-4, (k, s, eb, ep)
            np.testing.assert_allclose(Rt[k, s, 9:], to, atol=1e-5 * ext)
            np.testing.assert_allclose(t, to, atol=1e-5 * ext)
            if k >= 2:
                assert geodesic(Rb, R.astype(float)) < 1e-6, (k, s)
                np.testing.assert_allclose(Rt[k, s, 9:], t, atol=1e-6)


def test_multistart_finds_large_rotation(E, pq_pairs):
    """The reference's bunny points (data/P1007.txt, asymmetric, unlike a
    uniform cube whose octahedral symmetry makes wrong starts look right)
    rotated by 150 degrees: outside EM-ICP's basin from R = I, recovered by
    the best of the 24 octahedral starts."""
    P, _ = pq_pairs[1007]
    X = P.astype(np.float32)
    Rg = rotation([1, 2, 2], np.deg2rad(150))
    Y = ((X - [0.01, -0.02, 0.005]) @ Rg).astype(np.float32)  # X = Rg Y + t
    p = E.registrationParameters(sigma_p2=0.001, sigma_inf=1e-6)
    Rt0 = E.multistart_rt0(X, Y, octahedral_rotations())
    Rt, res = E.emicp_batched([(X, Y)], Rt0, p)
    best, k = E.select_best(Rt, res)
    assert geodesic(best[0, :9].reshape(3, 3).astype(float), Rg) < 1e-3
    R1, _ = E.emicp(X, Y, np.eye(3), np.zeros(3), p)
    assert geodesic(R1.astype(float), Rg) > 0.1  # the single start does not get there


def _small_pairs():
    """Targets of at most 2048 points: with 2048-column tiles every
    registration is one chunk, so no row-finalize kernel is launched."""
    out = []
    for k, (n, m) in enumerate([(60, 1500), (2000, 900), (1007, 1007), (33, 17)]):
        X, Y, R, t = range_scans(n, m, seed=200 + k)
        out.append((X, Y))
    return out


@pytest.mark.parametrize("which", ["mixed", "single_chunk"])
def test_graph_replays_match_eager(E, which):
    """Replays of the recorded sequence equal the eager run bit for bit, with
    eager work between replays; "mixed": some registrations finalize their
    rows in the affinity kernel, the others in the row kernel; "single_chunk":
    all in the affinity kernel, no row kernel in the graph."""
    import torch
    pairs = _pairs() if which == "mixed" else _small_pairs()
    starts = [np.eye(3), rotation([0, 1, 0], 0.25)]
    Rt0 = np.stack([E.multistart_rt0(X, Y, starts) for X, Y in pairs]).astype(np.float32)
    p = E.registrationParameters(sigma_inf=2e-4)
    ref, ref_res = E.emicp_batched(pairs, Rt0, p)
    keep = []
    dpairs = []
    for X, Y in pairs:
        dx = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        dy = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
        keep += [dx, dy]
        dpairs.append((dx.data_ptr(), len(X), dy.data_ptr(), len(Y)))
    n = len(pairs) * 2
    d0 = torch.from_numpy(Rt0.reshape(-1)).cuda()
    dout = torch.zeros(12 * n, device="cuda")
    dres = torch.zeros(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        dout.fill_(float("nan"))
        E.emicp_batched_dev(dpairs, d0.data_ptr(), 2, dout.data_ptr(), dres.data_ptr(), p,
                            stream)
        torch.cuda.synchronize()
        got = dout.cpu().numpy().reshape(len(pairs), 2, 12)
        assert np.array_equal(got, ref), rep
        assert np.array_equal(dres.cpu().numpy().reshape(len(pairs), 2), ref_res), rep
        # eager work between replays
        E.emicp(pairs[0][0], pairs[0][1], np.eye(3), np.zeros(3))
        torch.cuda.synchronize()


def test_single_registration_graph_replays(E):
    import torch
    X, Y, _, _ = uniform_cube(3000, seed=8)
    R, t = E.emicp(X, Y, np.eye(3), np.zeros(3))
    dX = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    rt0 = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float32, device="cuda")
    dRt = rt0.clone()
    s = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        dRt.copy_(rt0)
        E.emicp_dev(dX.data_ptr(), 3000, dY.data_ptr(), 3000, dRt.data_ptr(), stream=s)
        torch.cuda.synchronize()
        out = dRt.cpu().numpy()
        assert np.array_equal(out[:9], R.reshape(9)) and np.array_equal(out[9:], t), rep
        dX.mul_(1.0)  # eager torch work in between
        torch.cuda.synchronize()


def test_autotune(E):
    import torch
    pairs = _pairs()
    keep, dpairs = [], []
    for X, Y in pairs:
        dx = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        dy = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
        keep += [dx, dy]
        dpairs.append((dx.data_ptr(), len(X), dy.data_ptr(), len(Y)))
    c = E.autotune_emicp(dpairs, 4)
    assert c in (512, 1024, 2048, 4096, 8192)
    c1 = E.autotune_emicp(dpairs[:1], 1)
    assert c1 in (2, 4, 8, 16, 32)
    X, Y = pairs[2]
    R, t = E.emicp(X, Y, np.eye(3), np.zeros(3))
    Rt, res = E.emicp_batched(pairs[2:3], np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]]))
    assert geodesic(Rt[0, 0, :9].reshape(3, 3).astype(float), R.astype(float)) < 1e-6
    E.graph_cache_clear()


def test_affinity_paths_agree(E):
    """B and W of the affinity on the matrix cores (default) and on the VALU
    give the same registrations up to FP32 rounding. Only well-conditioned
    pairs: the 33 x 17 and 60 x 1500 registrations annealed to 2e-4 amplify an
    FP32 summation-order difference into a different local optimum."""
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    pairs = [_pairs()[k] for k in (1, 2, 4)]
    p = E.registrationParameters(sigma_inf=2e-4)
    Rt0 = np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]], np.float32)
    try:
        _lib.check("path", L.emsa_probe_set_affinity_path(0))
        Rv, rv = E.emicp_batched(pairs, Rt0, p)
        R1v, t1v = E.emicp(pairs[2][0], pairs[2][1], np.eye(3), np.zeros(3), p)
    finally:
        _lib.check("path", L.emsa_probe_set_affinity_path(1))
    Rm, rm = E.emicp_batched(pairs, Rt0, p)
    R1m, t1m = E.emicp(pairs[2][0], pairs[2][1], np.eye(3), np.zeros(3), p)
    for k in range(len(pairs)):
        err = geodesic(Rv[k, 0, :9].reshape(3, 3).astype(float),
                       Rm[k, 0, :9].reshape(3, 3).astype(float))
        assert err < 1e-5, (k, err)
        np.testing.assert_allclose(Rv[k, 0, 9:], Rm[k, 0, 9:], atol=1e-5)
    np.testing.assert_allclose(rv, rm, rtol=1e-4)
    assert geodesic(R1v.astype(float), R1m.astype(float)) < 1e-5


def _cull_runs(E, mode):
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    X, Y, _, _ = uniform_cube(3000, seed=21)
    pairs = _pairs() + [(X, Y)]
    starts = [np.eye(3), rotation([0, 0, 1], 0.3)]
    Rt0 = np.stack([E.multistart_rt0(a, b, starts) for a, b in pairs])
    p = E.registrationParameters()  # full schedule down to sigma^2 = 1e-5
    try:
        _lib.check("cull", L.emsa_probe_set_affinity_cull(mode))
        out = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), p) + E.emicp_batched(pairs, Rt0, p)
    finally:
        _lib.check("cull", L.emsa_probe_set_affinity_cull(1))
    return out


def test_affinity_cull_exact_bit_identical(E):
    """Skipping only the target column groups whose affinities all underflow
    to exactly 0 (mode 2) changes no bit of any result: the per-step traces
    of a plain registration, and the batched multi-start results and
    residuals, equal those of the loop over every column in the same
    (Hilbert) order."""
    Rf, tf, trf, Rbf, rbf = _cull_runs(E, 0)
    Rc, tc, trc, Rbc, rbc = _cull_runs(E, 2)
    assert np.array_equal(trf, trc)
    assert np.array_equal(Rf, Rc) and np.array_equal(tf, tc)
    assert np.array_equal(Rbf, Rbc) and np.array_equal(rbf, rbc)


def test_affinity_cull_default_matches_full_loop(E):
    """The default skipping (mode 1: groups whose every affinity A = e^2 is
    below 2^-128, under FLT_MIN) against the loop over every column: per-step
    Lambda within 1e-7 relative and (R, t) within 1e-7 after every step of a
    plain registration; batched multi-start results within 1e-6 rad and
    residuals within 1e-6 relative on the well-posed pairs (2.5k x 700,
    1007 x 1007, 4100 x 3900 and the 3000-point cube)."""
    Rf, tf, trf, Rbf, rbf = _cull_runs(E, 0)
    Rc, tc, trc, Rbc, rbc = _cull_runs(E, 1)
    dl = np.max(np.abs(trc[:, 1] - trf[:, 1]) / trf[:, 1])
    dR = np.max(np.abs(trc[:, 17:29] - trf[:, 17:29]))
    print("default skipping vs every column: Lambda %.1e, (R, t) %.1e" % (dl, dR), flush=True)
    assert dl < 1e-7 and dR < 1e-7
    for k in (1, 2, 4, 5):
        for st in range(2):
            err = geodesic(Rbc[k, st, :9].reshape(3, 3).astype(float),
                           Rbf[k, st, :9].reshape(3, 3).astype(float))
            assert err < 1e-6, (k, st, err)
            assert abs(rbc[k, st] - rbf[k, st]) <= 1e-6 * abs(rbf[k, st]), (k, st)


def _balance_runs(E, on, tile_cols):
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    pairs = []
    for k in range(4):
        X, Y, _, _ = uniform_cube(5000, seed=40 + k)
        pairs.append((X, Y))
    starts = [np.eye(3), rotation([1, 0, 0], 0.4)]
    Rt0 = np.stack([E.multistart_rt0(a, b, starts) for a, b in pairs])
    p = E.registrationParameters()
    try:
        _lib.check("balance", L.emsa_probe_set_affinity_balance(on))
        E.set_tiling(0, tile_cols)
        out = E.emicp_traced(pairs[0][0], pairs[0][1], np.eye(3), np.zeros(3), p) + \
            E.emicp_batched(pairs, Rt0, p)
    finally:
        E.set_tiling(0, 0)
        _lib.check("balance", L.emsa_probe_set_affinity_balance(1))
    return out


@pytest.mark.parametrize("tile_cols", [2048, 5056])
def test_affinity_balance_bit_identical(E, tile_cols):
    """The balanced workgroup order (each step's affinity workgroups sorted by
    the previous step's work per tile) changes which workgroup computes a
    tile, not what it computes: a plain traced registration and 8 batched
    registrations (4 pairs of 5000 points x 2 starts; 2048-column tiles, 3
    chunks per row block and the row kernel, or whole-row tiles finalized in
    the affinity kernel; at most 4096 tiles, so the order is in use) give the
    same bits with the order on and off."""
    R0, t0, tr0, Rb0, rb0 = _balance_runs(E, 0, tile_cols)
    R1, t1, tr1, Rb1, rb1 = _balance_runs(E, 1, tile_cols)
    assert np.array_equal(tr0, tr1)
    assert np.array_equal(R0, R1) and np.array_equal(t0, t1)
    assert np.array_equal(Rb0, Rb1) and np.array_equal(rb0, rb1)
    assert np.all(np.isfinite(Rb1))


def test_affinity_cull_non_finite_inputs(E):
    """The device entry point takes clouds as they are. A NaN in the target
    makes its column group's box the whole space and a NaN in the source
    turns the skipping off for its wave, so a non-finite input propagates
    exactly as in the loop over every column: same outputs bit for bit (NaN
    included), with the skipping on and off."""
    import torch
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    X, Y, _, _ = uniform_cube(1500, seed=31)
    p = E.registrationParameters(sigma_inf=1e-4)
    stream = torch.cuda.current_stream().cuda_stream
    for where in ("target", "source", "target-inf"):
        Xb, Yb = X.copy(), Y.copy()
        if where == "target":
            Xb[700, 2] = np.nan
        elif where == "source":
            Yb[40, 0] = np.nan
        else:
            Xb[10, 1] = np.inf
        dX = torch.from_numpy(np.ascontiguousarray(Xb.T)).cuda()
        dY = torch.from_numpy(np.ascontiguousarray(Yb.T)).cuda()
        res = []
        for cull in (0, 1):
            _lib.check("cull", L.emsa_probe_set_affinity_cull(cull))
            dRt = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float32,
                               device="cuda")
            E.emicp_dev(dX.data_ptr(), len(Xb), dY.data_ptr(), len(Yb), dRt.data_ptr(), p,
                        stream)
            torch.cuda.synchronize()
            res.append(dRt.cpu().numpy())
        _lib.check("cull", L.emsa_probe_set_affinity_cull(1))
        assert np.array_equal(res[0], res[1], equal_nan=True), (where, res)


def test_row_fuse_bit_identical(E):
    """With tiles that span each registration's whole target (one column
    chunk), the affinity kernel finalizes the rows in its epilogue and the
    row-finalize kernel is not launched (or skips those registrations in a
    mixed batch). Same bits as the row kernel working from the partial sums
    in HBM: batched multi-start runs of pairs of uneven sizes with 2048-column
    tiles (mixed) and 4160-column tiles (every target in one chunk), and a
    traced plain 20,000-point registration (one workgroup per CU, so one chunk
    per row block)."""
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    pairs = _pairs()
    starts = [np.eye(3), rotation([0, 0, 1], 0.3)]
    Rt0 = np.stack([E.multistart_rt0(a, b, starts) for a, b in pairs])
    p = E.registrationParameters()
    X, Y, _, _ = uniform_cube(20000, seed=41)
    res = []
    try:
        for fuse in (0, 1):
            _lib.check("fuse", L.emsa_probe_set_row_fuse(fuse))
            E.set_tiling(0, 0)  # 2048-column tiles: targets of 60, 1007, 33 points
            Rm, rm = E.emicp_batched(pairs, Rt0, p)  # one chunk, 2500 and 4100: more
            E.set_tiling(1, 4160)
            Rb, rb = E.emicp_batched(pairs, Rt0, p)
            R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), p)
            res.append((Rm, rm, Rb, rb, R, t, tr))
    finally:
        _lib.check("fuse", L.emsa_probe_set_row_fuse(1))
        E.set_tiling(0, 0)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    assert len(res[1][6]) == 66 and not res[1][6][:, 31].any()
