"""Unrolled Softassign on the GPU (csrc/sa_unrolled.hip) and its hand-written
adjoint behind torch.autograd.Function, against the plain FP64 PyTorch loop
(tests/_unrolled_ref.py: pinned to the C oracle's inner iteration and to FP64
central finite differences in test_unrolled_ref.py) and against the plain
Softassign entry point (materialised variant, same schedule).

Bars: forward rotation within 1e-5 rad and translation within 1e-5 of the FP64
loop; gradients (X, Y, initial transform, T_0) within 1e-3 of the FP64
gradients in relative 2-norm; batched gradients equal to 1e-4 of per-pair runs.
"""
import numpy as np
import pytest
import torch

from cuda_emicp_softassign_amd.clouds import geodesic, random_rotation, surface_with_outliers
from tests import _unrolled_ref as U

pytestmark = pytest.mark.gpu

IDENT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float64)


@pytest.fixture(scope="module")
def UR():
    from cuda_emicp_softassign_amd import unrolled
    return unrolled


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _pair(n, m, seed, deg=20.0):
    X = surface_with_outliers(n, seed=seed)[0].astype(np.float64)
    Ysrc = surface_with_outliers(m, seed=seed + 100)[0].astype(np.float64) if m != n else X
    rng = np.random.default_rng(seed)
    Rg = random_rotation(rng, deg)
    tg = rng.uniform(-0.05, 0.05, 3)
    Y = (Ysrc - tg) @ Rg  # Ysrc = Rg Y + tg
    return X.astype(np.float32), Y.astype(np.float32), np.concatenate([Rg.reshape(9), tg])


def _check(UR, X, Y, Rt_true, Rt0=IDENT, T_0=0.5, TFACTOR=0.9, alpha=0.05, moutlier=None,
           JMAX=20, I0=1, I1=3, tol=1e-3):
    print("case N=%d M=%d JMAX=%d I0=%d I1=%d moutlier=%s" % (len(X), len(Y), JMAX, I0, I1,
                                                             moutlier), flush=True)
    kw = dict(JMAX=JMAX, I0=I0, I1=I1, TFACTOR=TFACTOR, alpha=float(np.float32(alpha)))
    # reference: FP64, temperatures from the float T_0 (the GPU iterates T in float)
    args = [torch.from_numpy(np.asarray(a, np.float64)).requires_grad_(True)
            for a in (X, Y, Rt0, float(np.float32(T_0)))]
    Rr = U.softassign_unrolled(*args, moutlier=moutlier, **kw)
    L = ((Rr - torch.from_numpy(Rt_true)) ** 2).sum()
    gr = [g.numpy() for g in torch.autograd.grad(L, args)]
    Rr = Rr.detach().numpy()

    Xd = torch.tensor(X, device="cuda").requires_grad_(True)
    Yd = torch.tensor(Y, device="cuda").requires_grad_(True)
    R0 = torch.tensor(np.asarray(Rt0, np.float32), device="cuda").requires_grad_(True)
    Td = torch.tensor(float(T_0), dtype=torch.float64).requires_grad_(True)
    Rg = UR.softassign_unrolled(Xd, Yd, R0, Td, TFACTOR=TFACTOR, alpha=alpha,
                                moutlier=moutlier, JMAX=JMAX, I0=I0, I1=I1)
    ((Rg.double() - torch.tensor(Rt_true, device="cuda")) ** 2).sum().backward()
    torch.cuda.synchronize()
    Rg = Rg.detach().cpu().numpy()
    gg = [Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy(), R0.grad.cpu().numpy(), Td.grad.numpy()]
    assert geodesic(Rg[:9].reshape(3, 3).astype(float), Rr[:9].reshape(3, 3)) < 1e-5
    np.testing.assert_allclose(Rg[9:], Rr[9:], atol=1e-5)
    errs = [_rel(a, b) for a, b in zip(gg, gr)]
    print("  rel err gX %.2e gY %.2e gRt0 %.2e gT0 %.2e" % tuple(errs), flush=True)
    for name, e in zip(("X", "Y", "Rt0", "T_0"), errs):
        assert e < tol, (name, e)


def test_sa_adjoint_matches_fp64_autograd_256(UR):
    X, Y, Rt = _pair(256, 256, seed=3)
    _check(UR, X, Y, Rt)


def test_sa_adjoint_ragged_two_sweeps_two_inner(UR):
    X, Y, Rt = _pair(300, 200, seed=5)
    Rt0 = IDENT + np.r_[0.01 * np.ones(9), [0.01, 0.0, -0.01]]
    _check(UR, X, Y, Rt, Rt0=Rt0, JMAX=6, I0=2, I1=2)


def test_sa_adjoint_one_sweep_explicit_slack(UR):
    X, Y, Rt = _pair(130, 270, seed=7)
    _check(UR, X, Y, Rt, JMAX=10, I1=1, moutlier=0.2)


def test_sa_adjoint_four_sweeps(UR):
    X, Y, Rt = _pair(257, 257, seed=9)
    _check(UR, X, Y, Rt, JMAX=8, I1=4)


def test_sa_adjoint_cold_schedule(UR):
    """T from 0.02 down to 0.0063: k1 max|x|^2 (k1 = log2(e) / T, max|x|^2 =
    0.49) crosses 64 at T = 0.011, so the first 6 temperatures run the
    expanded exponent a_r + b_c + x'_r . p_c and the last 6 the direct
    difference, in the forward and in the adjoint. A float32 run of the FP64
    loop stays within 2e-5 of its gradients on this case."""
    X, Y, Rt = _pair(256, 256, seed=3)
    Ts = 0.02 * 0.9 ** np.arange(12)
    k1x = 1.4426950408889634 / Ts * (X.astype(np.float64) ** 2).sum(1).max()
    assert (k1x <= 64).any() and (k1x > 64).any()
    _check(UR, X, Y, Rt, T_0=0.02, TFACTOR=0.9, JMAX=12)


def test_sa_forward_matches_plain_softassign(UR):
    """The unrolled forward runs the iteration of the plain Softassign entry
    point (materialised variant), same float temperature schedule."""
    import cuda_emicp_softassign_amd as E
    X, Y, _ = _pair(500, 400, seed=11)
    p = E.registrationParameters(JMAX=15, I0=2, I1=3, T_0=0.5, TFACTOR=0.9, alpha=0.05)
    R, t = E.softassign(X, Y, np.eye(3), np.zeros(3), p)
    Rt = UR.softassign_unrolled(torch.tensor(X, device="cuda"), torch.tensor(Y, device="cuda"),
                                T_0=0.5, TFACTOR=0.9, alpha=0.05, moutlier=p.moutlier,
                                JMAX=15, I0=2, I1=3).cpu().numpy()
    assert geodesic(Rt[:9].reshape(3, 3).astype(float), R.astype(float)) < 1e-5
    np.testing.assert_allclose(Rt[9:], t, atol=1e-5)


def test_sa_adjoint_batch_equals_single(UR):
    pairs = [_pair(260, 240, seed=20 + k) for k in range(3)]
    X = torch.tensor(np.stack([p[0] for p in pairs]), device="cuda").requires_grad_(True)
    Y = torch.tensor(np.stack([p[1] for p in pairs]), device="cuda").requires_grad_(True)
    tgt = torch.tensor(np.stack([p[2] for p in pairs]), dtype=torch.float32, device="cuda")
    Rt = UR.softassign_unrolled(X, Y, JMAX=8)
    ((Rt - tgt) ** 2).sum().backward()
    for k in range(3):
        Xk = X[k].detach().clone().requires_grad_(True)
        Yk = Y[k].detach().clone().requires_grad_(True)
        Rk = UR.softassign_unrolled(Xk, Yk, JMAX=8)
        ((Rk - tgt[k]) ** 2).sum().backward()
        assert torch.allclose(Rk, Rt[k], atol=1e-6)
        assert _rel(X.grad[k].cpu(), Xk.grad.cpu()) < 1e-4
        assert _rel(Y.grad[k].cpu(), Yk.grad.cpu()) < 1e-4


def test_sa_unrolled_errors(UR):
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    assert L.emsa_softassign_unrolled_ckpt_bytes(1, 10, 10, 2, 1, 0) == _lib.EMSA_EINVAL
    assert L.emsa_softassign_unrolled_ckpt_bytes(1, 10, 10, 2, 1, 5) == _lib.EMSA_EINVAL
    assert L.emsa_softassign_unrolled_ckpt_bytes(2, 10, 7, 3, 1, 2) > 0
    X = torch.zeros(1, 8, 3, device="cuda")
    with pytest.raises(ValueError):
        UR.softassign_unrolled(X.cpu(), X.cpu())
    with pytest.raises(_lib.EmsaError):
        UR.softassign_unrolled(X, X, I1=0)
    with pytest.raises(_lib.EmsaError):
        UR.softassign_unrolled(X, X, T_0=-1.0)
