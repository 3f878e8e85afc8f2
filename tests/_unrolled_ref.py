"""Plain PyTorch reference of the unrolled EM-ICP loop (test infrastructure).

Dense FP64 tensors, torch autograd for the gradients, torch.linalg.eigh for
Horn's eigenproblem. One iteration is the loop body of src/emicp_cpu.cpp:305-527
(== src/emicp.cu:443-668):

    A_rc  = exp(-|x_c - (R y_r + t)|^2 / sigma^2)            (updateA)
    C_r   = sum_c A_rc + exp(-d_02 / sigma^2)                 (sgemv + saxpy)
    A'_rc = sqrt(A_rc / C_r), or 1/N when C_r <= 1e-6         (normalizeRowsOfA)
    lambda_r = sum_c A'_rc, X'_r = sum_c A'_rc x_c / lambda_r
    Xc, Yc = lambda-weighted centroids, S = sum lambda (y - Yc)(X' - Xc)^T
    (R, t) = Horn(Xc, Yc, S)                                   (findRTfromS)

sqrt(A / C) is written e / sqrt(C) with e = exp(-|d|^2 / (2 sigma^2)): the same
value, but differentiable where A underflows to 0 (d sqrt(A)/dA is infinite
there).
"""
import torch


def horn(Xc, Yc, S):
    """S[a][b] = sum (source)_a (target)_b. Returns R (3, 3), t (3,)."""
    Sxx, Sxy, Sxz = S[0, 0], S[0, 1], S[0, 2]
    Syx, Syy, Syz = S[1, 0], S[1, 1], S[1, 2]
    Szx, Szy, Szz = S[2, 0], S[2, 1], S[2, 2]
    N = torch.stack([
        torch.stack([Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx]),
        torch.stack([Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz]),
        torch.stack([Szx - Sxz, Sxy + Syx, Syy - Sxx - Szz, Syz + Szy]),
        torch.stack([Sxy - Syx, Szx + Sxz, Syz + Szy, Szz - Sxx - Syy])])
    w, V = torch.linalg.eigh(N)
    q0, qx, qy, qz = V[0, 3], V[1, 3], V[2, 3], V[3, 3]
    R = torch.stack([
        torch.stack([q0 * q0 + qx * qx - qy * qy - qz * qz, 2 * (qx * qy - q0 * qz),
                     2 * (qx * qz + q0 * qy)]),
        torch.stack([2 * (qy * qx + q0 * qz), q0 * q0 - qx * qx + qy * qy - qz * qz,
                     2 * (qy * qz - q0 * qx)]),
        torch.stack([2 * (qz * qx - q0 * qy), 2 * (qz * qy + q0 * qx),
                     q0 * q0 - qx * qx - qy * qy + qz * qz])])
    return R, Xc - R @ Yc


def step_moments(X, Y, R, t, s2, d_02):
    """E-step of one sigma step; X (N, 3) target, Y (M, 3) source, s2 = sigma^2.
    Returns (C, Lambda, Xc, Yc, S)."""
    N = X.shape[0]
    P = Y @ R.T + t
    D = ((X[None, :, :] - P[:, None, :]) ** 2).sum(-1)  # (M, N)
    e = torch.exp(-D / (2 * s2))
    C = (e * e).sum(1) + torch.exp(-d_02 / s2)
    floored = ~(C.to(torch.float32) > 1e-6)
    Cs = torch.where(floored, torch.ones_like(C), C)
    rs = 1.0 / torch.sqrt(Cs)
    fl = float(torch.tensor(1.0 / N, dtype=torch.float32))
    lam = torch.where(floored, torch.full_like(C, N * fl), e.sum(1) * rs)
    V = torch.where(floored[:, None], (fl * X.sum(0))[None, :].expand(len(C), 3),
                    (e @ X) * rs[:, None])  # lambda_r X'_r
    L = lam.sum()
    Xc = V.sum(0) / L
    Yc = (lam[:, None] * Y).sum(0) / L
    S = Y.T @ V - L * torch.outer(Yc, Xc)  # S[a][b] = sum lambda y_a X'_b - L Yc_a Xc_b
    return C, L, Xc, Yc, S


def emicp_step(X, Y, R, t, s2, d_02):
    """One sigma step: (R, t) -> (R', t')."""
    _, _, Xc, Yc, S = step_moments(X, Y, R, t, s2, d_02)
    return horn(Xc, Yc, S)


def unrolled(X, Y, Rt0, sigma2_0, factor, steps, d_02=0.01):
    """steps iterations with sigma_k^2 = sigma2_0 * factor^k; returns Rt (12,)."""
    R = Rt0[:9].reshape(3, 3)
    t = Rt0[9:]
    for k in range(steps):
        R, t = emicp_step(X, Y, R, t, sigma2_0 * factor ** k, d_02)
    return torch.cat([R.reshape(9), t])


# ---------------------------------------------------------------------------
# Softassign: one inner iteration of src/softassign.cu:494-818 (== the C
# oracle's or_softassign_inner), rows r over the target X (N), columns c over
# the source Y (M), with the slack vectors o_row, o_col carried across the
# whole run (src/softassign.cu:385-390).
# ---------------------------------------------------------------------------

def softassign_inner(X, Y, R, t, T, alpha, I1, o_row, o_col):
    """Returns (R', t', o_row', o_col', record) with record =
    dict(Sigma, Xc, Yc, S) (S[a][b] = sum (source)_a (target)_b)."""
    P = Y @ R.T + t
    D2 = ((X[:, None, :] - P[None, :, :]) ** 2).sum(-1)  # (N, M)
    Mt = torch.exp(-(D2 - alpha) / T) / torch.sqrt(T)    # updateM
    b = None
    for _ in range(I1):
        a = Mt.sum(1) + o_row
        Mt = Mt / a[:, None]
        o_row = o_row / a
        b = Mt.sum(0) + o_col
        Mt = Mt / b[None, :]
        o_col = o_col / b
    s = Mt.sum(1)
    Sg = s.abs().sum()
    Xt = s[:, None] * X
    Yt = b[:, None] * Y
    Xc = Xt.sum(0) / Sg
    Yc = Yt.sum(0) / Sg
    Xh = Xt - Xc
    Yh = Yt - Yc
    Dm = Mt @ Yh                  # (N, 3)
    S = Dm.T @ Xh                 # S[a][b]: a over the source (D), b over the target
    Rn, tn = horn(Xc, Yc, S)
    return Rn, tn, o_row, o_col, dict(Sigma=Sg, Xc=Xc, Yc=Yc, S=S)


def softassign_unrolled(X, Y, Rt0, T_0, JMAX, I0, I1, TFACTOR, alpha, moutlier=None):
    """JMAX temperatures T_j = T_0 * TFACTOR^j, I0 inner iterations each;
    moutlier None: exp(-1) / sqrt(T_0) (differentiable in T_0). Returns Rt (12,)."""
    R = Rt0[:9].reshape(3, 3)
    t = Rt0[9:]
    kw = dict(dtype=X.dtype, device=X.device)
    mo = torch.exp(torch.tensor(-1.0, **kw)) / torch.sqrt(T_0) if moutlier is None \
        else torch.tensor(moutlier, **kw)
    o_row = mo * torch.ones(X.shape[0], **kw)
    o_col = mo * torch.ones(Y.shape[0], **kw)
    for j in range(JMAX):
        T = T_0 * TFACTOR ** j
        for _ in range(I0):
            R, t, o_row, o_col, _ = softassign_inner(X, Y, R, t, T, alpha, I1, o_row, o_col)
    return torch.cat([R.reshape(9), t])
