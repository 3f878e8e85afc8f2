"""Literal restatement of the reference iterations as plain Python loops, for
tiny clouds only. It keeps the reference's own structure -- the materialised
column-major matrix A[c*pitch + r] and one loop per BLAS call or kernel -- so
that it shares no code shape with the row-streamed C oracle it checks.

EM-ICP: src/emicp_cpu.cpp:305-527. Softassign: src/softassign.cu:486-822.
"""
import math


def _horn(Xc, Yc, S, eig):
    """findRTfromS (src/findRTfromS.cpp:33-97) with a caller-supplied
    symmetric eigensolver eig(N4x4) -> (w ascending, V columns)."""
    Sxx, Sxy, Sxz, Syx, Syy, Syz, Szx, Szy, Szz = S
    N = [[Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx],
         [Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz],
         [Szx - Sxz, Sxy + Syx, Syy - Sxx - Szz, Syz + Szy],
         [Sxy - Syx, Szx + Sxz, Syz + Szy, Szz - Sxx - Syy]]
    w, V = eig(N)
    q0, qx, qy, qz = [V[i][3] for i in range(4)]  # last column: largest eigenvalue
    R = [q0 * q0 + qx * qx - qy * qy - qz * qz, 2 * (qx * qy - q0 * qz), 2 * (qx * qz + q0 * qy),
         2 * (qy * qx + q0 * qz), q0 * q0 - qx * qx + qy * qy - qz * qz, 2 * (qy * qz - q0 * qx),
         2 * (qz * qx - q0 * qy), 2 * (qz * qy + q0 * qx), q0 * q0 - qx * qx - qy * qy + qz * qz]
    t = [Xc[0] - (R[0] * Yc[0] + R[1] * Yc[1] + R[2] * Yc[2]),
         Xc[1] - (R[3] * Yc[0] + R[4] * Yc[1] + R[5] * Yc[2]),
         Xc[2] - (R[6] * Yc[0] + R[7] * Yc[1] + R[8] * Yc[2])]
    return R, t


def emicp_step(X, Y, R, t, sigma_p2, d_02, eig):
    """One EM-ICP sigma step. X (N points), Y (M points) as lists of [x,y,z].
    Returns dict(C, lam, Xc, Yc, S, R, t)."""
    colsA, rowsA = len(X), len(Y)
    pitch = (rowsA // 4 + 1) * 4
    A = [0.0] * (pitch * colsA)
    for c in range(colsA):  # updateA
        for r in range(rowsA):
            tx = X[c][0] - (R[0] * Y[r][0] + R[1] * Y[r][1] + R[2] * Y[r][2] + t[0])
            ty = X[c][1] - (R[3] * Y[r][0] + R[4] * Y[r][1] + R[5] * Y[r][2] + t[1])
            tz = X[c][2] - (R[6] * Y[r][0] + R[7] * Y[r][1] + R[8] * Y[r][2] + t[2])
            A[c * pitch + r] = math.exp(-(tx * tx + ty * ty + tz * tz) / sigma_p2)
    C = [sum(A[c * pitch + r] for c in range(colsA)) for r in range(rowsA)]  # sgemv
    alpha = math.exp(-d_02 / sigma_p2)  # saxpy
    C = [v + alpha for v in C]
    for c in range(colsA):  # normalizeRowsOfA
        for r in range(rowsA):
            if C[r] > 1e-6:
                A[c * pitch + r] = math.sqrt(A[c * pitch + r] / C[r])
            else:
                A[c * pitch + r] = 1.0 / colsA
    lam = [sum(A[c * pitch + r] for c in range(colsA)) for r in range(rowsA)]
    sumLambda = sum(abs(v) for v in lam)  # sasum
    Xp = [[sum(A[c * pitch + r] * X[c][k] for c in range(colsA)) for k in range(3)]
          for r in range(rowsA)]  # sgemm
    Xp = [[Xp[r][k] / lam[r] for k in range(3)] for r in range(rowsA)]  # elementwiseDivision
    Xc = [sum(Xp[r][k] * lam[r] for r in range(rowsA)) / sumLambda for k in range(3)]
    Yc = [sum(Y[r][k] * lam[r] for r in range(rowsA)) / sumLambda for k in range(3)]
    Xcen = [[(Xp[r][k] - Xc[k]) * lam[r] for k in range(3)] for r in range(rowsA)]
    Ycen = [[Y[r][k] - Yc[k] for k in range(3)] for r in range(rowsA)]
    S = [0.0] * 9
    for j in range(3):
        for i in range(3):
            S[i + 3 * j] = sum(Xcen[r][i] * Ycen[r][j] for r in range(rowsA))
    Rn, tn = _horn(Xc, Yc, S, eig)
    return dict(C=C, lam=lam, Lambda=sumLambda, Xc=Xc, Yc=Yc, S=S, R=Rn, t=tn)


def softassign_inner(X, Y, R, t, T, alpha, I1, o_row, o_col, eig):
    """One Softassign inner iteration; o_row/o_col (lists) updated in place."""
    rowsM, colsM = len(X), len(Y)
    pitch = (rowsM // 4 + 1) * 4
    Mm = [0.0] * (pitch * colsM)
    for c in range(colsM):  # updateM
        for r in range(rowsM):
            tx = X[r][0] - (R[0] * Y[c][0] + R[1] * Y[c][1] + R[2] * Y[c][2] + t[0])
            ty = X[r][1] - (R[3] * Y[c][0] + R[4] * Y[c][1] + R[5] * Y[c][2] + t[1])
            tz = X[r][2] - (R[6] * Y[c][0] + R[7] * Y[c][1] + R[8] * Y[c][2] + t[2])
            Mm[c * pitch + r] = math.exp(-((tx * tx + ty * ty + tz * tz) - alpha) / T) / math.sqrt(T)
    colsum = None
    a_sweeps, b_sweeps = [], []
    for _ in range(I1):
        rowsum = [sum(Mm[c * pitch + r] for c in range(colsM)) + o_row[r] for r in range(rowsM)]
        a_sweeps.append(list(rowsum))
        for c in range(colsM):
            for r in range(rowsM):
                Mm[c * pitch + r] /= rowsum[r]
        for r in range(rowsM):
            o_row[r] /= rowsum[r]
        colsum = [sum(Mm[c * pitch + r] for r in range(rowsM)) + o_col[c] for c in range(colsM)]
        b_sweeps.append(list(colsum))
        for c in range(colsM):
            for r in range(rowsM):
                Mm[c * pitch + r] /= colsum[c]
        for c in range(colsM):
            o_col[c] /= colsum[c]
    rowsum = [sum(Mm[c * pitch + r] for c in range(colsM)) for r in range(rowsM)]
    sumM = sum(abs(v) for v in rowsum)
    Xr = [[rowsum[r] * X[r][k] for k in range(3)] for r in range(rowsM)]
    Yr = [[colsum[c] * Y[c][k] for k in range(3)] for c in range(colsM)]
    Xc = [sum(Xr[r][k] for r in range(rowsM)) * (1.0 / sumM) for k in range(3)]
    Yc = [sum(Yr[c][k] for c in range(colsM)) * (1.0 / sumM) for k in range(3)]
    Xr = [[Xr[r][k] - Xc[k] for k in range(3)] for r in range(rowsM)]
    Yr = [[Yr[c][k] - Yc[k] for k in range(3)] for c in range(colsM)]
    D = [[sum(Mm[c * pitch + r] * Yr[c][k] for c in range(colsM)) for k in range(3)]
         for r in range(rowsM)]
    S = [0.0] * 9
    for j in range(3):
        for i in range(3):
            S[i + 3 * j] = sum(Xr[r][i] * D[r][j] for r in range(rowsM))
    Rn, tn = _horn(Xc, Yc, S, eig)
    return dict(rowsum=rowsum, colsum=colsum, sumM=sumM, Xc=Xc, Yc=Yc, S=S, R=Rn, t=tn,
                a=a_sweeps, b=b_sweeps)
