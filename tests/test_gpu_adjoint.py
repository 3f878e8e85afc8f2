"""Hand-written adjoint of the unrolled EM-ICP (csrc/adjoint.hip) behind
torch.autograd.Function, against autograd of the plain FP64 PyTorch reference
loop (tests/_unrolled_ref.py, itself pinned to the C oracle and to FP64
central finite differences in test_unrolled_ref.py).

Bars: forward rotation within 1e-5 rad and translation within 1e-5 of the FP64
loop; each gradient (X, Y, initial transform, sigma2_0) within 1e-3 of the
FP64 gradient in relative 2-norm (the adjoint accumulates pair terms in FP32);
batched gradients equal to 1e-4 of the per-pair runs (FP32 summation order).
"""
import numpy as np
import pytest
import torch

from cuda_emicp_softassign_amd.clouds import geodesic, range_scans, uniform_cube
from tests import _unrolled_ref as U

pytestmark = pytest.mark.gpu

IDENT = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]


@pytest.fixture(scope="module")
def UR():
    from cuda_emicp_softassign_amd import unrolled
    return unrolled


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _ref_grads(X, Y, Rt0, s2, Rt_true, steps, factor, d_02=0.01):
    args = [torch.from_numpy(np.asarray(a, np.float64)).requires_grad_(True)
            for a in (X, Y, Rt0, s2)]
    Rt = U.unrolled(args[0], args[1], args[2], args[3], factor, steps, d_02)
    L = ((Rt - torch.from_numpy(Rt_true)) ** 2).sum()
    g = torch.autograd.grad(L, args)
    return Rt.detach().numpy(), [x.numpy() for x in g]


def _gpu_grads(UR, X, Y, Rt0, s2, Rt_true, steps, factor, d_02=0.01):
    dev = "cuda"
    Xd = torch.tensor(np.asarray(X, np.float32), device=dev).requires_grad_(True)
    Yd = torch.tensor(np.asarray(Y, np.float32), device=dev).requires_grad_(True)
    R0 = torch.tensor(np.asarray(Rt0, np.float32), device=dev).requires_grad_(True)
    sd = torch.tensor(float(s2), dtype=torch.float64).requires_grad_(True)
    Rt = UR.emicp_unrolled(Xd, Yd, R0, sd, factor=factor, steps=steps, d_02=d_02)
    L = ((Rt.double() - torch.tensor(Rt_true, device=dev)) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return (Rt.detach().cpu().numpy(),
            [Xd.grad.cpu().numpy(), Yd.grad.cpu().numpy(), R0.grad.cpu().numpy(),
             sd.grad.numpy()])


def _check(UR, X, Y, Rt_true, steps=20, factor=0.9, s2=0.01, d_02=0.01, Rt0=IDENT, tol=1e-3):
    print("case N=%d M=%d steps=%d s2=%g" % (len(X), len(Y), steps, s2), flush=True)
    Rt0 = np.asarray(Rt0, np.float64)
    Rr, gr = _ref_grads(X, Y, Rt0, s2, Rt_true, steps, factor, d_02)
    Rg, gg = _gpu_grads(UR, X, Y, Rt0, s2, Rt_true, steps, factor, d_02)
    assert geodesic(Rg[:9].reshape(3, 3).astype(float), Rr[:9].reshape(3, 3)) < 1e-5
    np.testing.assert_allclose(Rg[9:], Rr[9:], atol=1e-5)
    errs = [_rel(a, b) for a, b in zip(gg, gr)]
    print("  rel err gX %.2e gY %.2e gRt0 %.2e gsigma %.2e" % tuple(errs), flush=True)
    for name, e in zip(("X", "Y", "Rt0", "sigma2_0"), errs):
        assert e < tol, (name, e)
    return gg


@pytest.mark.parametrize("seed", [0, 1])
def test_adjoint_matches_fp64_autograd_256(UR, seed):
    X, Y, Rg, tg = uniform_cube(256, seed=40 + seed)
    _check(UR, X, Y, np.concatenate([Rg.reshape(9), tg]))


def test_adjoint_ragged_sizes_and_start(UR):
    X, Y, Rg, tg = range_scans(300, 257, seed=7)
    Rt0 = np.array([0.99, -0.1, 0.0, 0.1, 0.99, 0.0, 0.0, 0.0, 1.0, 0.01, -0.02, 0.0])
    _check(UR, X, Y, np.concatenate([Rg.reshape(9), tg]), steps=12, Rt0=Rt0)


def test_adjoint_with_floored_rows(UR):
    X, Y, Rg, tg = uniform_cube(200, seed=5)
    # far source rows: floored (C_r <= 1e-6) once sigma^2 is small (d_02 = 1)
    Y = np.vstack([Y, [[1.5, 1.5, 1.5], [-1.5, 1.2, 1.4]]]).astype(np.float32)
    _check(UR, X, Y, np.concatenate([Rg.reshape(9), tg]), steps=15, s2=2e-3, d_02=1.0)


def test_adjoint_small_sigma_compensated(UR):
    """sigma^2 from 2e-4 down to 6e-5: the later steps recompute the affinities
    with the compensated difference (below sigma^2 = 1e-4), the earlier without."""
    X, Y, Rg, tg = uniform_cube(256, seed=11)
    Rt_true = np.concatenate([Rg.reshape(9), tg])
    Rt0 = Rt_true + np.r_[0.002 * np.ones(9), -0.001 * np.ones(3)]  # near the optimum
    _check(UR, X, Y, Rt_true, steps=12, s2=2e-4, Rt0=Rt0)


def test_adjoint_batch_equals_single(UR):
    pairs = [uniform_cube(320, seed=60 + k) for k in range(3)]
    X = torch.tensor(np.stack([p[0] for p in pairs]), device="cuda").requires_grad_(True)
    Y = torch.tensor(np.stack([p[1] for p in pairs]), device="cuda").requires_grad_(True)
    tgt = torch.tensor(np.stack([np.concatenate([p[2].reshape(9), p[3]]) for p in pairs]),
                       dtype=torch.float32, device="cuda")
    Rt = UR.emicp_unrolled(X, Y, None, 0.01, steps=10)
    ((Rt - tgt) ** 2).sum().backward()
    for k in range(3):
        Xk = X[k].detach().clone().requires_grad_(True)
        Yk = Y[k].detach().clone().requires_grad_(True)
        Rk = UR.emicp_unrolled(Xk, Yk, None, 0.01, steps=10)
        ((Rk - tgt[k]) ** 2).sum().backward()
        assert torch.allclose(Rk, Rt[k], atol=1e-6)
        # the forward's column tiling differs (batched vs single plan): FP32 order only
        assert _rel(X.grad[k].cpu(), Xk.grad.cpu()) < 1e-4
        assert _rel(Y.grad[k].cpu(), Yk.grad.cpu()) < 1e-4


def test_adjoint_directional_fp32_difference(UR):
    """Directional derivative of the GPU forward itself (FP32, h = 1e-3 of the
    points' scale), so the adjoint is also checked against the code it inverts."""
    X, Y, Rg, tg = uniform_cube(256, seed=9)
    tgt = torch.tensor(np.concatenate([Rg.reshape(9), tg]), dtype=torch.float64, device="cuda")
    Xd = torch.tensor(X, device="cuda").requires_grad_(True)
    Yd = torch.tensor(Y, device="cuda")

    def loss(Xv):
        return ((UR.emicp_unrolled(Xv, Yd, steps=8).double() - tgt) ** 2).sum()

    loss(Xd).backward()
    d = torch.randn(X.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    h = 1e-3
    with torch.no_grad():
        fd = (loss(Xd + h * d) - loss(Xd - h * d)) / (2 * h)
    an = (Xd.grad * d).sum().double()
    assert abs(float(fd - an)) < 2e-2 * abs(float(an)) + 1e-6, (float(fd), float(an))


def test_ckpt_size_and_errors(UR):
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    arr = (_lib.emsa_pair * 2)(_lib.emsa_pair(1, 10, 1, 7), _lib.emsa_pair(1, 5, 1, 3))
    assert L.emsa_emicp_unrolled_ckpt_size(arr, 2, 4) == 4 * (32 + 5 * 7) + 4 * (32 + 5 * 3)
    X = torch.zeros(1, 8, 3, device="cuda")
    with pytest.raises(ValueError):
        UR.emicp_unrolled(X.cpu(), X.cpu())
    with pytest.raises(ValueError):
        UR.emicp_unrolled(X, X, steps=0)
    with pytest.raises(_lib.EmsaError):
        UR.emicp_unrolled(X, X, sigma2_0=-1.0)


def test_adjoint_ragged_batch_through_c_abi(UR):
    """Pairs of different sizes in one call of the C ABI (the torch wrapper
    only batches equal sizes): each pair's gradients equal its own run."""
    import ctypes as C
    from cuda_emicp_softassign_amd import _lib
    L = _lib.load()
    sizes = [(300, 257), (517, 410), (64, 90)]
    data = [range_scans(n, m, seed=30 + k) for k, (n, m) in enumerate(sizes)]
    dev = "cuda"
    Xs = [torch.tensor(np.ascontiguousarray(d[0].T), device=dev) for d in data]
    Ys = [torch.tensor(np.ascontiguousarray(d[1].T), device=dev) for d in data]
    B = len(sizes)
    pairs = (_lib.emsa_pair * B)(*[_lib.emsa_pair(Xs[k].data_ptr(), sizes[k][0], Ys[k].data_ptr(),
                                                  sizes[k][1]) for k in range(B)])
    rt0 = torch.tensor(np.tile(IDENT, B), dtype=torch.float32, device=dev)
    steps = 10
    n = L.emsa_emicp_unrolled_ckpt_size(pairs, B, steps)
    ckpt = torch.empty(n, dtype=torch.float64, device=dev)
    out = torch.empty(12 * B, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check("fwd", L.emsa_emicp_unrolled_fwd(pairs, B, C.c_void_p(rt0.data_ptr()), 0.01, 0.9,
                                                0.01, steps, C.c_void_p(out.data_ptr()),
                                                C.c_void_p(ckpt.data_ptr()), C.c_void_p(s)))
    g = torch.linspace(-1, 1, 12 * B, device=dev)
    gX = [torch.empty_like(x) for x in Xs]
    gY = [torch.empty_like(y) for y in Ys]
    grads = (_lib.emsa_pair_grad * B)(*[_lib.emsa_pair_grad(gX[k].data_ptr(), gY[k].data_ptr())
                                        for k in range(B)])
    gR0 = torch.empty(12 * B, device=dev)
    gs = torch.empty(B, dtype=torch.float64, device=dev)
    _lib.check("bwd", L.emsa_emicp_unrolled_bwd(pairs, B, C.c_void_p(rt0.data_ptr()), 0.01, 0.9,
                                                0.01, steps, C.c_void_p(ckpt.data_ptr()),
                                                C.c_void_p(g.data_ptr()), grads,
                                                C.c_void_p(gR0.data_ptr()),
                                                C.c_void_p(gs.data_ptr()), C.c_void_p(s)))
    torch.cuda.synchronize()
    for k in range(B):
        print("pair", k, sizes[k], flush=True)
        Xk = torch.tensor(data[k][0], device=dev).requires_grad_(True)
        Yk = torch.tensor(data[k][1], device=dev).requires_grad_(True)
        sk = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
        Rk = UR.emicp_unrolled(Xk, Yk, None, sk, steps=steps)
        (Rk * g[12 * k:12 * (k + 1)]).sum().backward()
        assert torch.allclose(Rk, out[12 * k:12 * (k + 1)], atol=1e-5)
        assert _rel(gX[k].T.cpu(), Xk.grad.cpu()) < 1e-4
        assert _rel(gY[k].T.cpu(), Yk.grad.cpu()) < 1e-4
        assert abs(float(gs[k]) - float(sk.grad)) <= 1e-4 * abs(float(sk.grad)) + 1e-9
