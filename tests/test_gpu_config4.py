"""The differentiable workloads at the benchmark's size (N = M = 4096, the
first two of the 32 pairs it times): batched hand-written adjoints against
autograd of the plain FP64 PyTorch reference loops (tests/_unrolled_ref.py,
pinned to the C oracle and to FP64 central differences in
test_unrolled_ref.py), the reference run on the GPU in FP64.

EM-ICP: 20 unrolled sigma^2 steps (0.01 * 0.9^k), gradients for X, Y, the
initial transform and sigma2_0. Softassign: JMAX = 20, I0 = 1, I1 = 3,
T_0 = 0.5, TFACTOR = 0.9, alpha = 0.05, gradients for X, Y, the initial
transform and T_0. Loss ||R - R_true||_F^2 + ||t - t_true||^2. Bars as at
256 points: forward rotation within 1e-5 rad and translation within 1e-5 of
FP64; each gradient within 1e-3 of the FP64 one in relative 2-norm (measured
on MI355X: at most 9e-7 for X, Y and the initial transform, 2e-6 for sigma2_0
and 4e-7 for T_0)."""
import numpy as np
import pytest
import torch

from cuda_emicp_softassign_amd.clouds import (geodesic, random_rotation, surface_with_outliers,
                                              uniform_cube)
from tests import _unrolled_ref as U

pytestmark = pytest.mark.gpu

N, B = 4096, 2


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _compare(Rg, gg, Rr, gr, names):
    assert geodesic(Rg[:9].reshape(3, 3).astype(float), Rr[:9].reshape(3, 3)) < 1e-5
    np.testing.assert_allclose(Rg[9:], Rr[9:], atol=1e-5)
    errs = [_rel(a, b) for a, b in zip(gg, gr)]
    print("  rel err " + " ".join("g%s %.2e" % (n, e) for n, e in zip(names, errs)), flush=True)
    for n, e in zip(names, errs):
        assert e < 1e-3, (n, e)


def test_emicp_unrolled_4096_matches_fp64():
    from cuda_emicp_softassign_amd import unrolled as UR
    pairs = [uniform_cube(N, seed=4000 + b) for b in range(B)]
    tgt = np.stack([np.concatenate([p[2].reshape(9), p[3]]) for p in pairs])
    X = torch.tensor(np.stack([p[0] for p in pairs]), device="cuda").requires_grad_(True)
    Y = torch.tensor(np.stack([p[1] for p in pairs]), device="cuda").requires_grad_(True)
    R0 = torch.tensor(np.tile(np.r_[np.eye(3).ravel(), np.zeros(3)], (B, 1)),
                      dtype=torch.float32, device="cuda").requires_grad_(True)
    s2 = torch.tensor(0.01, dtype=torch.float64).requires_grad_(True)
    print("EM-ICP unrolled %d x %d^2, 20 steps" % (B, N), flush=True)
    Rt = UR.emicp_unrolled(X, Y, R0, s2, factor=0.9, steps=20)
    ((Rt.double() - torch.tensor(tgt, device="cuda")) ** 2).sum().backward()
    torch.cuda.synchronize()
    Rg = Rt.detach().cpu().numpy()
    gs_total = float(s2.grad)
    gs_ref = 0.0
    for b in range(B):
        args = [torch.tensor(np.asarray(a, np.float64), device="cuda").requires_grad_(True)
                for a in (pairs[b][0], pairs[b][1], np.r_[np.eye(3).ravel(), np.zeros(3)])]
        sr = torch.tensor(0.01, dtype=torch.float64, device="cuda").requires_grad_(True)
        Rr = U.unrolled(args[0], args[1], args[2], sr, 0.9, 20, 0.01)
        L = ((Rr - torch.tensor(tgt[b], device="cuda")) ** 2).sum()
        gr = torch.autograd.grad(L, args + [sr])
        gs_ref += float(gr[3])
        print(" pair %d" % b, flush=True)
        _compare(Rg[b], [X.grad[b].cpu().numpy(), Y.grad[b].cpu().numpy(),
                         R0.grad[b].cpu().numpy()],
                 Rr.detach().cpu().numpy(), [g.cpu().numpy() for g in gr[:3]],
                 ("X", "Y", "Rt0"))
    print("  gsigma2_0 %.6e vs FP64 %.6e" % (gs_total, gs_ref), flush=True)
    assert abs(gs_total - gs_ref) <= 1e-3 * abs(gs_ref)


def test_softassign_unrolled_4096_matches_fp64():
    from cuda_emicp_softassign_amd import unrolled as UR
    Xs, Ys, tg = [], [], []
    for b in range(B):  # the benchmark's pairs 0 and 1
        Xb = surface_with_outliers(N, seed=4000 + b)[0].astype(np.float64)
        rng = np.random.default_rng(4000 + b)
        Rb = random_rotation(rng, 20.0)
        t = rng.uniform(-0.05, 0.05, 3)
        Xs.append(Xb.astype(np.float32))
        Ys.append(((Xb - t) @ Rb).astype(np.float32))
        tg.append(np.concatenate([Rb.reshape(9), t]))
    kw = dict(JMAX=20, I0=1, I1=3, TFACTOR=0.9, alpha=0.05)
    X = torch.tensor(np.stack(Xs), device="cuda").requires_grad_(True)
    Y = torch.tensor(np.stack(Ys), device="cuda").requires_grad_(True)
    R0 = torch.tensor(np.tile(np.r_[np.eye(3).ravel(), np.zeros(3)], (B, 1)),
                      dtype=torch.float32, device="cuda").requires_grad_(True)
    T0 = torch.tensor(0.5, dtype=torch.float64).requires_grad_(True)
    print("Softassign unrolled %d x %d^2, JMAX 20, I0 1, I1 3" % (B, N), flush=True)
    Rt = UR.softassign_unrolled(X, Y, R0, T0, **kw)
    ((Rt.double() - torch.tensor(np.stack(tg), device="cuda")) ** 2).sum().backward()
    torch.cuda.synchronize()
    Rg = Rt.detach().cpu().numpy()
    gT_ref = 0.0
    for b in range(B):
        args = [torch.tensor(np.asarray(a, np.float64), device="cuda").requires_grad_(True)
                for a in (Xs[b], Ys[b], np.r_[np.eye(3).ravel(), np.zeros(3)])]
        # temperatures from the float T_0, as the GPU iterates T in float
        Tr = torch.tensor(float(np.float32(0.5)), dtype=torch.float64,
                          device="cuda").requires_grad_(True)
        Rr = U.softassign_unrolled(args[0], args[1], args[2], Tr, JMAX=20, I0=1, I1=3,
                                   TFACTOR=0.9, alpha=float(np.float32(0.05)))
        L = ((Rr - torch.tensor(tg[b], device="cuda")) ** 2).sum()
        gr = torch.autograd.grad(L, args + [Tr])
        gT_ref += float(gr[3])
        print(" pair %d" % b, flush=True)
        _compare(Rg[b], [X.grad[b].cpu().numpy(), Y.grad[b].cpu().numpy(),
                         R0.grad[b].cpu().numpy()],
                 Rr.detach().cpu().numpy(), [g.cpu().numpy() for g in gr[:3]],
                 ("X", "Y", "Rt0"))
    print("  gT_0 %.6e vs FP64 %.6e" % (float(T0.grad), gT_ref), flush=True)
    assert abs(float(T0.grad) - gT_ref) <= 1e-3 * abs(gT_ref)
