"""Differentiable unrolled EM-ICP behind ``torch.autograd.Function``.

``steps`` iterations of the EM-ICP step (the loop body of src/emicp.cu:443-668)
with sigma_k^2 = sigma2_0 * factor^k. The forward runs the fused HIP kernels
and checkpoints, per step, the step record and (C_r, lambda_r, lambda_r X'_r)
per source row; the backward is the hand-written adjoint of libemsa.so, which
recomputes the affinities instead of storing them. Gradients flow to the
target and source points, the initial transform and sigma2_0.

There is no CPU path: the tensors must be on a HIP device and libemsa.so must
be built.
"""
from __future__ import annotations

import ctypes as C

from . import _lib


def _check_inputs(X, Y, Rt0):
    import torch
    if X.dim() != 3 or Y.dim() != 3 or X.shape[2] != 3 or Y.shape[2] != 3:
        raise ValueError("X, Y must be (B, N, 3) and (B, M, 3)")
    if X.shape[0] != Y.shape[0] or Rt0.shape != (X.shape[0], 12):
        raise ValueError("batch sizes of X, Y and Rt0 (B, 12) differ")
    for name, T in (("X", X), ("Y", Y), ("Rt0", Rt0)):
        if T.device.type != "cuda":
            raise ValueError("%s must be a device tensor (no CPU path)" % name)
        if T.dtype != torch.float32:
            raise ValueError("%s must be float32" % name)


def _pairs(Xs, Ys):
    B, _, N = Xs.shape
    M = Ys.shape[2]
    arr = (_lib.emsa_pair * B)()
    for b in range(B):
        arr[b] = _lib.emsa_pair(Xs[b].data_ptr(), N, Ys[b].data_ptr(), M)
    return arr


class EmicpUnrolled:
    """Namespace for the autograd.Function (created lazily so that importing
    this module does not import torch)."""
    fn = None


def _function():
    if EmicpUnrolled.fn is not None:
        return EmicpUnrolled.fn
    import torch

    class _F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, X, Y, Rt0, sigma2_0, factor, d_02, steps):
            _check_inputs(X, Y, Rt0)
            L = _lib.load()
            s2 = float(sigma2_0)
            Xs = X.detach().transpose(1, 2).contiguous()  # SoA (B, 3, N)
            Ys = Y.detach().transpose(1, 2).contiguous()
            R0 = Rt0.detach().contiguous()
            B = X.shape[0]
            arr = _pairs(Xs, Ys)
            n = _lib.check("emicp_unrolled_ckpt_size", L.emsa_emicp_unrolled_ckpt_size(
                arr, B, steps))
            ckpt = torch.empty(n, dtype=torch.float64, device=X.device)
            out = torch.empty(B, 12, dtype=torch.float32, device=X.device)
            stream = torch.cuda.current_stream(X.device).cuda_stream
            _lib.check("emicp_unrolled_fwd", L.emsa_emicp_unrolled_fwd(
                arr, B, C.c_void_p(R0.data_ptr()), s2, float(factor), float(d_02), steps,
                C.c_void_p(out.data_ptr()), C.c_void_p(ckpt.data_ptr()), C.c_void_p(stream)))
            ctx.save_for_backward(Xs, Ys, R0, ckpt)
            ctx.args = (s2, float(factor), float(d_02), int(steps))
            ctx.sigma_like = sigma2_0 if torch.is_tensor(sigma2_0) else None
            return out

        @staticmethod
        def backward(ctx, gout):
            Xs, Ys, R0, ckpt = ctx.saved_tensors
            s2, factor, d_02, steps = ctx.args
            B = Xs.shape[0]
            g = gout.detach().to(torch.float32).contiguous()
            gXs = torch.empty_like(Xs)
            gYs = torch.empty_like(Ys)
            gR0 = torch.empty_like(R0)
            gsig = torch.empty(B, dtype=torch.float64, device=Xs.device)
            arr = _pairs(Xs, Ys)
            grads = (_lib.emsa_pair_grad * B)()
            for b in range(B):
                grads[b] = _lib.emsa_pair_grad(gXs[b].data_ptr(), gYs[b].data_ptr())
            stream = torch.cuda.current_stream(Xs.device).cuda_stream
            _lib.check("emicp_unrolled_bwd", _lib.load().emsa_emicp_unrolled_bwd(
                arr, B, C.c_void_p(R0.data_ptr()), s2, factor, d_02, steps,
                C.c_void_p(ckpt.data_ptr()), C.c_void_p(g.data_ptr()), grads,
                C.c_void_p(gR0.data_ptr()), C.c_void_p(gsig.data_ptr()), C.c_void_p(stream)))
            gs = None
            if ctx.sigma_like is not None:
                gs = gsig.sum().to(dtype=ctx.sigma_like.dtype, device=ctx.sigma_like.device)
            return (gXs.transpose(1, 2), gYs.transpose(1, 2), gR0, gs, None, None, None)

    EmicpUnrolled.fn = _F
    return _F


def emicp_unrolled(X, Y, Rt0=None, sigma2_0=0.01, factor=0.9, steps=20, d_02=0.01):
    """Unrolled EM-ICP, differentiable with respect to X (B, N, 3), Y (B, M, 3),
    Rt0 (B, 12) and sigma2_0 (a float or a 0-dim tensor; its gradient is the
    sum over the batch, which shares one schedule). Returns Rt (B, 12):
    R row-major, then t, for X ~ R Y + t."""
    import torch
    squeeze = X.dim() == 2
    if squeeze:
        X, Y = X.unsqueeze(0), Y.unsqueeze(0)
        if Rt0 is not None and Rt0.dim() == 1:
            Rt0 = Rt0.unsqueeze(0)
    if Rt0 is None:
        eye = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float32,
                           device=X.device)
        Rt0 = eye.expand(X.shape[0], 12).contiguous()
    if steps <= 0:
        raise ValueError("steps must be positive")
    out = _function().apply(X, Y, Rt0, sigma2_0, factor, d_02, steps)
    return out[0] if squeeze else out


def split_rt(Rt):
    """(..., 12) -> R (..., 3, 3), t (..., 3)."""
    return Rt[..., :9].reshape(Rt.shape[:-1] + (3, 3)), Rt[..., 9:]


# ---------------------------------------------------------------------------
# Softassign
# ---------------------------------------------------------------------------

class SoftassignUnrolled:
    """Namespace for the autograd.Function of the unrolled Softassign."""
    fn = None


def _sa_function():
    if SoftassignUnrolled.fn is not None:
        return SoftassignUnrolled.fn
    import torch

    class _F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, X, Y, Rt0, T_0, TFACTOR, alpha, moutlier, JMAX, I0, I1):
            _check_inputs(X, Y, Rt0)
            L = _lib.load()
            B, N, M = X.shape[0], X.shape[1], Y.shape[1]
            Xs = X.detach().transpose(1, 2).contiguous()
            Ys = Y.detach().transpose(1, 2).contiguous()
            R0 = Rt0.detach().contiguous()
            T0 = float(T_0)
            mo = -1.0 if moutlier is None else float(moutlier)
            nb = _lib.check("softassign_unrolled_ckpt_bytes", L.emsa_softassign_unrolled_ckpt_bytes(
                B, N, M, JMAX, I0, I1))
            ckpt = torch.empty((nb + 7) // 8, dtype=torch.float64, device=X.device)
            out = torch.empty(B, 12, dtype=torch.float32, device=X.device)
            stream = torch.cuda.current_stream(X.device).cuda_stream
            _lib.check("softassign_unrolled_fwd", L.emsa_softassign_unrolled_fwd(
                C.c_void_p(Xs.data_ptr()), C.c_void_p(Ys.data_ptr()), B, N, M,
                C.c_void_p(R0.data_ptr()), T0, float(TFACTOR), float(alpha), mo, JMAX, I0, I1,
                C.c_void_p(out.data_ptr()), C.c_void_p(ckpt.data_ptr()), C.c_void_p(stream)))
            ctx.save_for_backward(Xs, Ys, R0, ckpt)
            ctx.args = (T0, float(TFACTOR), float(alpha), mo, int(JMAX), int(I0), int(I1))
            ctx.T_like = T_0 if torch.is_tensor(T_0) else None
            return out

        @staticmethod
        def backward(ctx, gout):
            Xs, Ys, R0, ckpt = ctx.saved_tensors
            T0, TF, alpha, mo, JMAX, I0, I1 = ctx.args
            B, N, M = Xs.shape[0], Xs.shape[2], Ys.shape[2]
            g = gout.detach().to(torch.float32).contiguous()
            gXs = torch.empty_like(Xs)
            gYs = torch.empty_like(Ys)
            gR0 = torch.empty_like(R0)
            gT = torch.empty(B, dtype=torch.float64, device=Xs.device)
            stream = torch.cuda.current_stream(Xs.device).cuda_stream
            _lib.check("softassign_unrolled_bwd", _lib.load().emsa_softassign_unrolled_bwd(
                C.c_void_p(Xs.data_ptr()), C.c_void_p(Ys.data_ptr()), B, N, M,
                C.c_void_p(R0.data_ptr()), T0, TF, alpha, mo, JMAX, I0, I1,
                C.c_void_p(ckpt.data_ptr()), C.c_void_p(g.data_ptr()),
                C.c_void_p(gXs.data_ptr()), C.c_void_p(gYs.data_ptr()),
                C.c_void_p(gR0.data_ptr()), C.c_void_p(gT.data_ptr()), C.c_void_p(stream)))
            gt0 = None
            if ctx.T_like is not None:
                gt0 = gT.sum().to(dtype=ctx.T_like.dtype, device=ctx.T_like.device)
            return (gXs.transpose(1, 2), gYs.transpose(1, 2), gR0, gt0,
                    None, None, None, None, None, None)

    SoftassignUnrolled.fn = _F
    return _F


def softassign_unrolled(X, Y, Rt0=None, T_0=0.5, TFACTOR=0.9, alpha=0.05, moutlier=None,
                        JMAX=20, I0=1, I1=3):
    """Unrolled Softassign (JMAX temperatures x I0 inner iterations x I1
    Sinkhorn sweeps), differentiable with respect to X (B, N, 3), Y (B, M, 3),
    Rt0 (B, 12) and T_0 (a float or a 0-dim tensor; its gradient is the sum
    over the batch). moutlier None: exp(-1) / sqrt(T_0). Returns Rt (B, 12)."""
    import torch
    squeeze = X.dim() == 2
    if squeeze:
        X, Y = X.unsqueeze(0), Y.unsqueeze(0)
        if Rt0 is not None and Rt0.dim() == 1:
            Rt0 = Rt0.unsqueeze(0)
    if Rt0 is None:
        eye = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float32,
                           device=X.device)
        Rt0 = eye.expand(X.shape[0], 12).contiguous()
    out = _sa_function().apply(X, Y, Rt0, T_0, TFACTOR, alpha, moutlier, JMAX, I0, I1)
    return out[0] if squeeze else out
