"""Pins the plain PyTorch reference of the unrolled EM-ICP (tests/_unrolled_ref.py)
that the GPU adjoint is checked against:

* its forward step equals the FP64 C oracle's EM-ICP step (which is pinned to
  the reference's fixtures and literal loops in test_oracle.py), including a
  floored row;
* its autograd gradients equal FP64 central finite differences (h = 1e-6) of
  the same loop at N = M = 256, 20 sigma steps, for the loss
  ||R - R_true||_F^2 + ||t - t_true||^2 with respect to X, Y, the initial
  transform and sigma2_0.
"""
import numpy as np
import pytest
import torch

from cuda_emicp_softassign_amd.clouds import uniform_cube
from tests import _unrolled_ref as U


def test_reference_step_matches_oracle(oracle):
    X, Y, _, _ = uniform_cube(120, seed=3)
    # a far source row: its affinities stay above FP64 underflow in the first two
    # steps and its C_r drops below the 1e-6 floor in the third (d_02 = 1)
    Y = np.vstack([Y, [[1.0, 1.0, 1.0]]])
    R = np.eye(3)
    t = np.zeros(3)
    Xt, Yt = torch.from_numpy(X.astype(np.float64)), torch.from_numpy(Y.astype(np.float64))
    for s2, d02 in [(0.01, 0.01), (2e-3, 0.01), (5e-4, 1.0)]:
        s2 = float(np.float32(s2))  # the oracle takes sigma^2 and d_02 as floats
        d02 = float(np.float32(d02))
        o = oracle.emicp_step(X.astype(np.float64), Y.astype(np.float64), R, t, s2, d_02=d02)
        C, L, Xc, Yc, S = U.step_moments(Xt, Yt, torch.from_numpy(R), torch.from_numpy(t),
                                         torch.tensor(s2, dtype=torch.float64), d02)
        np.testing.assert_allclose(C.numpy(), o["C"], rtol=1e-12)
        np.testing.assert_allclose(float(L), o["Lambda"], rtol=1e-12)
        np.testing.assert_allclose(Xc.numpy(), o["Xc"], atol=1e-12)
        np.testing.assert_allclose(Yc.numpy(), o["Yc"], atol=1e-12)
        np.testing.assert_allclose(S.numpy().reshape(9), o["S"], atol=1e-12)
        Ro, to = oracle.find_rt_from_s(Xc.numpy(), Yc.numpy(), S.numpy().reshape(9))[:2]
        Rr, tr = U.horn(Xc, Yc, S)
        np.testing.assert_allclose(Rr.numpy(), np.asarray(Ro).reshape(3, 3), atol=1e-12)
        np.testing.assert_allclose(tr.numpy(), np.asarray(to).reshape(3), atol=1e-12)
        R, t = Rr.numpy(), tr.numpy()
    assert not float(np.float32(o["C"][-1])) > 1e-6  # the far row was floored in the last step


def _loss(X, Y, Rt0, s2, Rt_true, steps=20, factor=0.9):
    Rt = U.unrolled(X, Y, Rt0, s2, factor, steps)
    return ((Rt - Rt_true) ** 2).sum()


@pytest.mark.parametrize("seed", [0, 1])
def test_reference_autograd_matches_central_differences(seed):
    X, Y, Rg, tg = uniform_cube(256, seed=40 + seed)
    X = torch.from_numpy(X.astype(np.float64))
    Y = torch.from_numpy(Y.astype(np.float64))
    Rt0 = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float64)
    s2 = torch.tensor(0.01, dtype=torch.float64)
    Rt_true = torch.from_numpy(np.concatenate([Rg.reshape(9), tg]))
    args = [t.clone().requires_grad_(True) for t in (X, Y, Rt0, s2)]
    L = _loss(*args, Rt_true)
    grads = torch.autograd.grad(L, args)
    rng = np.random.default_rng(seed)
    h = 1e-6
    for i, base in enumerate((X, Y, Rt0, s2)):
        for _ in range(2):
            d = torch.from_numpy(rng.standard_normal(tuple(base.shape)))
            if i == 3:
                d = d * 0.01  # sigma2_0 is 1e-2: a relative step
            plus = [t.clone() for t in (X, Y, Rt0, s2)]
            minus = [t.clone() for t in (X, Y, Rt0, s2)]
            plus[i] = plus[i] + h * d
            minus[i] = minus[i] - h * d
            with torch.no_grad():
                fd = (_loss(*plus, Rt_true) - _loss(*minus, Rt_true)) / (2 * h)
            an = (grads[i] * d).sum()
            assert abs(float(fd - an)) <= 1e-6 * max(1.0, abs(float(an))), (i, float(fd), float(an))


# ---- Softassign -------------------------------------------------------------

def test_reference_softassign_inner_matches_oracle(oracle):
    """Two inner iterations (so the carried slack vectors are covered) against
    the FP64 C oracle, which follows the reference's literal loops."""
    from cuda_emicp_softassign_amd.clouds import surface_with_outliers
    X = surface_with_outliers(150, seed=1)[0].astype(np.float64)
    Y = surface_with_outliers(130, seed=2)[1].astype(np.float64)
    R = np.eye(3)
    t = np.zeros(3)
    mo = 0.3
    orow_o = np.full(len(X), mo)
    ocol_o = np.full(len(Y), mo)
    orow_t = torch.full((len(X),), mo, dtype=torch.float64)
    ocol_t = torch.full((len(Y),), mo, dtype=torch.float64)
    Xt, Yt = torch.from_numpy(X), torch.from_numpy(Y)
    for T in (0.5, 0.45):
        T = float(np.float32(T))
        o = oracle.softassign_inner(X, Y, R, t, T, 0.05, 3, orow_o, ocol_o)
        Rr, tr, orow_t, ocol_t, rec = U.softassign_inner(
            Xt, Yt, torch.from_numpy(R), torch.from_numpy(t), torch.tensor(T, dtype=torch.float64),
            float(np.float32(0.05)), 3, orow_t, ocol_t)
        np.testing.assert_allclose(float(rec["Sigma"]), o["sumM"], rtol=1e-12)
        np.testing.assert_allclose(rec["Xc"].numpy(), o["Xc"], atol=1e-12)
        np.testing.assert_allclose(rec["Yc"].numpy(), o["Yc"], atol=1e-12)
        np.testing.assert_allclose(rec["S"].numpy().reshape(9), o["S"], atol=1e-12)
        np.testing.assert_allclose(orow_t.numpy(), orow_o, rtol=1e-12)
        np.testing.assert_allclose(ocol_t.numpy(), ocol_o, rtol=1e-12)
        np.testing.assert_allclose(Rr.numpy(), np.asarray(o["R"]).reshape(3, 3), atol=1e-12)
        np.testing.assert_allclose(tr.numpy(), np.asarray(o["t"]).reshape(3), atol=1e-12)
        R, t = Rr.numpy(), tr.numpy()


def _sa_loss(X, Y, Rt0, T0, Rt_true, **kw):
    Rt = U.softassign_unrolled(X, Y, Rt0, T0, **kw)
    return ((Rt - Rt_true) ** 2).sum()


def test_reference_softassign_autograd_matches_central_differences():
    """20 temperatures x 1 inner iteration x 3 Sinkhorn sweeps at N = M = 256,
    moutlier from T_0: gradients w.r.t. X, Y, the initial transform and T_0."""
    from cuda_emicp_softassign_amd.clouds import surface_with_outliers, random_rotation
    X = surface_with_outliers(256, seed=3)[0].astype(np.float64)
    Rg = random_rotation(np.random.default_rng(4), 20.0)
    tg = np.array([0.02, -0.03, 0.01])
    Y = (X - tg) @ Rg  # X = Rg Y + tg
    kw = dict(JMAX=20, I0=1, I1=3, TFACTOR=0.9, alpha=0.05)
    base = [torch.from_numpy(X), torch.from_numpy(Y),
            torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float64),
            torch.tensor(0.5, dtype=torch.float64)]
    Rt_true = torch.from_numpy(np.concatenate([Rg.reshape(9), tg]))
    args = [a.clone().requires_grad_(True) for a in base]
    grads = torch.autograd.grad(_sa_loss(*args, Rt_true, **kw), args)
    rng = np.random.default_rng(7)
    h = 1e-6
    for i in range(4):
        d = torch.from_numpy(rng.standard_normal(tuple(base[i].shape)))
        if i == 3:
            d = d * 0.5  # T_0 = 0.5: a relative step
        plus = [a.clone() for a in base]
        minus = [a.clone() for a in base]
        plus[i] = plus[i] + h * d
        minus[i] = minus[i] - h * d
        with torch.no_grad():
            fd = (_sa_loss(*plus, Rt_true, **kw) - _sa_loss(*minus, Rt_true, **kw)) / (2 * h)
        an = (grads[i] * d).sum()
        assert abs(float(fd - an)) <= 1e-6 * max(1.0, abs(float(an))), (i, float(fd), float(an))
