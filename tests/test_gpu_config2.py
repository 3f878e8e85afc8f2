"""EM-ICP at the large plain size: N = M = 40,000 partial-overlap range scans
(70 degree view cones 45 degrees apart, about 60% overlap, N(0, 0.002^2)
noise, reference EM-ICP defaults: 66 sigma^2 steps of 1.6e9 affinities).

The FP64 C oracle needs minutes per step at this size, so the full-size check
is a plain PyTorch FP64 restatement of one sigma^2 step (the oracle's
formulas, emicp_cpu.cpp:342-460: row sums with the outlier term, the float
floor, lambda, X', Lambda, Xc, Yc), run on the GPU for every step from the
(R, t) the library's previous step produced; the restatement itself is
pinned against the C oracle at a small size (CPU test below). Tolerances:
per-step normaliser Lambda within 1e-5 relative, Xc and Yc within 1e-5 of
the cloud extent. Also at full size: skipping the all-zero column groups
changes no bit of the per-step trace or the result.
"""
import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import range_scans, uniform_cube


def step_fp64(torch, Xd, Yd, R, t, sigma_p2, d_02=0.01, chunk=2048):
    """One EM-ICP sigma^2 step from (R, t) in FP64 (emicp_cpu.cpp:342-460 as
    restated in oracle/emsa_oracle_impl.h). Xd: N x 3 targets, Yd: M x 3
    sources, FP64 tensors on one device. Returns Lambda, Xc, Yc."""
    N = Xd.shape[0]
    Rd = torch.as_tensor(np.asarray(R, np.float64).reshape(3, 3), device=Xd.device)
    td = torch.as_tensor(np.asarray(t, np.float64).reshape(3), device=Xd.device)
    outlier = np.exp(-np.float64(np.float32(d_02)) / np.float64(sigma_p2))
    floor_a = np.float64(np.float32(1.0) / np.float32(N))
    x2 = (Xd * Xd).sum(1)
    lam_sum = torch.zeros((), dtype=torch.float64, device=Xd.device)
    sx = torch.zeros(3, dtype=torch.float64, device=Xd.device)
    sy = torch.zeros(3, dtype=torch.float64, device=Xd.device)
    for r0 in range(0, Yd.shape[0], chunk):
        P = Yd[r0:r0 + chunk] @ Rd.T + td
        d2 = ((P * P).sum(1)[:, None] + x2[None, :] - 2.0 * (P @ Xd.T)).clamp_min(0.0)
        A = torch.exp(-d2 / sigma_p2)
        C = A.sum(1) + outlier
        floored = C.float() <= np.float32(10e-7)
        Ap = torch.sqrt(A / C[:, None])
        lam = torch.where(floored, torch.full_like(C, N * floor_a), Ap.sum(1))
        lamX = torch.where(floored[:, None], floor_a * Xd.sum(0)[None, :], Ap @ Xd)
        lam_sum += lam.sum()
        sx += lamX.sum(0)
        sy += (lam[:, None] * Yd[r0:r0 + chunk]).sum(0)
    return float(lam_sum), (sx / lam_sum).cpu().numpy(), (sy / lam_sum).cpu().numpy()


def test_fp64_step_restatement_matches_oracle(oracle):
    """CPU: the restatement's first two steps equal the C oracle's FP64 trace
    (the second from the oracle's own R, t after the first)."""
    import torch
    X, Y, _, _ = uniform_cube(400, seed=3)
    Y = Y[:300]
    p = dict(sigma_inf=float(np.float32(0.01) * np.float32(0.9) * 0.95))
    Ro, to, tro = oracle.emicp(X.astype(np.float64), Y.astype(np.float64), **p)
    assert len(tro) == 2
    Xd = torch.from_numpy(X.astype(np.float64))
    Yd = torch.from_numpy(Y.astype(np.float64))
    R, t = np.eye(3), np.zeros(3)
    for k in range(2):
        lam, Xc, Yc = step_fp64(torch, Xd, Yd, R, t, tro[k, 0], chunk=128)
        assert abs(lam - tro[k, 1]) <= 1e-12 * tro[k, 1]
        np.testing.assert_allclose(Xc, tro[k, 2:5], rtol=0, atol=1e-12)
        np.testing.assert_allclose(Yc, tro[k, 5:8], rtol=0, atol=1e-12)
        R, t = tro[k, 17:26].reshape(3, 3), tro[k, 26:29]


@pytest.mark.gpu
def test_range_scan_40k_full_size():
    import torch
    import cuda_emicp_softassign_amd as E
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    X, Y, _, _ = range_scans(40000, seed=20)
    print("case range_scans n=40000 seed=20", flush=True)
    p = E.registrationParameters()
    R, t, tr = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), p)
    try:
        _lib.check("cull", L.emsa_probe_set_affinity_cull(0))
        Rf, tf, trf = E.emicp_traced(X, Y, np.eye(3), np.zeros(3), p)
    finally:
        _lib.check("cull", L.emsa_probe_set_affinity_cull(1))
    assert np.array_equal(tr, trf)
    assert np.array_equal(R, Rf) and np.array_equal(t, tf)
    assert len(tr) == 66
    assert not tr[:, 31].any()  # no failed Procrustes step
    Rd = R.astype(np.float64)
    assert abs(np.linalg.det(Rd) - 1) < 1e-5
    np.testing.assert_allclose(Rd @ Rd.T, np.eye(3), atol=1e-5)

    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X.astype(np.float64)).to(dev)
    Yd = torch.from_numpy(Y.astype(np.float64)).to(dev)
    extent = np.ptp(X, axis=0).max()
    Rk, tk = np.eye(3), np.zeros(3)
    worst = [0.0, 0.0]
    for k in range(len(tr)):
        lam, Xc, Yc = step_fp64(torch, Xd, Yd, Rk, tk, tr[k, 0])
        rel = abs(tr[k, 1] - lam) / lam
        dc = max(np.abs(tr[k, 2:5] - Xc).max(), np.abs(tr[k, 5:8] - Yc).max()) / extent
        worst = [max(worst[0], rel), max(worst[1], dc)]
        assert rel <= 1e-5, (k, tr[k, 0], rel)
        assert dc <= 1e-5, (k, tr[k, 0], dc)
        Rk, tk = tr[k, 17:26].reshape(3, 3), tr[k, 26:29]
    print("worst Lambda rel %.2e, centroid %.2e of extent" % tuple(worst), flush=True)
