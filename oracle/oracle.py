"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the benchmark's ``cpu_baseline``
leg import this module.  It restates the reference iterations on the CPU
(``src/emicp_cpu.cpp``, ``src/emicp.cu``, ``src/softassign.cu``,
``src/findRTfromS.cpp``) in FP64 (the oracle) and FP32 (the restatement of
``emicp_cpu`` used as the timed CPU baseline).  The product never calls it.

Point clouds are (n, 3) arrays here and are handed to C as SoA
``[x0..xn-1, y0.., z0..]`` exactly as ``cloud2data`` lays them out
(``src/cloud2data.cpp:29-47``).  R is 3x3 row-major, t has 3 entries.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMSA_ORACLE_LIB: another build of the same sources (the sanitizer test's)
_SO = os.environ.get("EMSA_ORACLE_LIB") or os.path.join(_HERE, "_build", "libemsa_oracle.so")
TRACE_W = 32
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.or_jacobi4.argtypes = [_dp, _dp, _dp]
        L.or_jacobi4.restype = C.c_int
        L.or_find_rt_from_s_f64.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.or_find_rt_from_s_f64.restype = C.c_int
        L.or_find_rt_from_s_f32.argtypes = [_fp, _fp, _fp, _fp, _fp]
        L.or_find_rt_from_s_f32.restype = C.c_int
        L.or_sigma_schedule.argtypes = [C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_int]
        L.or_sigma_schedule.restype = C.c_int
        for sfx, p in (("f64", _dp), ("f32", _fp)):
            f = getattr(L, "or_emicp_" + sfx)
            f.argtypes = [p, C.c_int, p, C.c_int, p, p, C.c_float, C.c_float, C.c_float,
                          C.c_float, C.c_int, C.c_void_p]
            f.restype = C.c_int
            f = getattr(L, "or_emicp_step_" + sfx)
            f.argtypes = [p, C.c_int, p, C.c_int, p, p, C.c_float, C.c_float, p, p, p, p, p, p]
            f.restype = C.c_double
            f = getattr(L, "or_softassign_" + sfx)
            f.argtypes = [p, C.c_int, p, C.c_int, p, p, C.c_int, C.c_int, C.c_int, C.c_float,
                          C.c_float, C.c_float, C.c_float, C.c_void_p]
            f.restype = C.c_int
            f = getattr(L, "or_softassign_inner_" + sfx)
            f.argtypes = [p, C.c_int, p, C.c_int, p, p, C.c_float, C.c_float, C.c_int, p, p, p,
                          p, p, p, p, p, C.c_void_p, C.c_void_p]
            f.restype = C.c_double if sfx == "f64" else C.c_float
        _lib = L
    return _lib


def _dt(precision):
    return {"f64": np.float64, "f32": np.float32}[precision]


def soa(P, dtype):
    """(n,3) points -> contiguous SoA (3n,) as cloud2data builds it."""
    P = np.asarray(P)
    return np.ascontiguousarray(P.T.reshape(-1), dtype=dtype)


def jacobi4(A):
    A = np.array(A, dtype=np.float64).reshape(16).copy()  # the C routine destroys A
    V = np.zeros(16)
    w = np.zeros(4)
    rc = lib().or_jacobi4(A, V, w)
    return w, V.reshape(4, 4), rc


def find_rt_from_s(Xc, Yc, S, precision="f64"):
    """Horn's Procrustes. S is the 9-vector h_S of findRTfromS (S[i+3j])."""
    dt = _dt(precision)
    Xc = np.ascontiguousarray(Xc, dtype=dt)
    Yc = np.ascontiguousarray(Yc, dtype=dt)
    S = np.ascontiguousarray(np.asarray(S).reshape(9), dtype=dt)
    R = np.zeros(9, dt)
    t = np.zeros(3, dt)
    rc = getattr(lib(), "or_find_rt_from_s_" + precision)(Xc, Yc, S, R, t)
    return R.reshape(3, 3), t, rc


def sigma_schedule(sigma_p2=0.01, sigma_inf=1e-5, sigma_factor=0.9):
    L = lib()
    n = L.or_sigma_schedule(sigma_p2, sigma_inf, sigma_factor, None, 0)
    out = np.zeros(max(n, 1), np.float32)
    L.or_sigma_schedule(sigma_p2, sigma_inf, sigma_factor, out.ctypes.data, n)
    return out[:n]


def emicp(X, Y, R0=None, t0=None, sigma_p2=0.01, sigma_inf=1e-5, sigma_factor=0.9, d_02=0.01,
          precision="f64", max_iter=-1):
    """Full EM-ICP registration. Returns (R, t, trace) with trace rows
    [sigma_p2, Lambda, Xc(3), Yc(3), S(9), R(9), t(3), pad]."""
    dt = _dt(precision)
    X = np.asarray(X)
    Y = np.asarray(Y)
    R = np.ascontiguousarray(np.eye(3) if R0 is None else np.asarray(R0).reshape(3, 3),
                             dtype=dt).reshape(9).copy()
    t = np.ascontiguousarray(np.zeros(3) if t0 is None else t0, dtype=dt).copy()
    n_it = lib().or_sigma_schedule(sigma_p2, sigma_inf, sigma_factor, None, 0)
    if max_iter >= 0:
        n_it = min(n_it, max_iter)
    trace = np.zeros((max(n_it, 1), TRACE_W))
    rc = getattr(lib(), "or_emicp_" + precision)(
        soa(X, dt), len(X), soa(Y, dt), len(Y), R, t, sigma_p2, sigma_inf, sigma_factor, d_02,
        max_iter, trace.ctypes.data)
    if rc < 0:
        raise ValueError("oracle emicp: bad arguments")
    return R.reshape(3, 3), t, trace[:rc]


def emicp_step(X, Y, R, t, sigma_p2, d_02=0.01, precision="f64"):
    """One sigma step from (R, t). Returns a dict with C, lam (per Y row),
    Xp (M,3), Lambda, Xc, Yc, S (9-vector), R, t (updated)."""
    dt = _dt(precision)
    M = len(Y)
    R = np.ascontiguousarray(np.asarray(R).reshape(9), dtype=dt).copy()
    t = np.ascontiguousarray(t, dtype=dt).copy()
    Cr = np.zeros(M, dt)
    lam = np.zeros(M, dt)
    Xp = np.zeros(3 * M, dt)
    Xc = np.zeros(3, dt)
    Yc = np.zeros(3, dt)
    S = np.zeros(9, dt)
    L = getattr(lib(), "or_emicp_step_" + precision)(
        soa(X, dt), len(X), soa(Y, dt), M, R, t, sigma_p2, d_02, Cr, lam, Xp, Xc, Yc, S)
    return dict(C=Cr, lam=lam, Xp=Xp.reshape(3, M).T, Lambda=L, Xc=Xc, Yc=Yc, S=S,
                R=R.reshape(3, 3), t=t)


def softassign(X, Y, R0=None, t0=None, JMAX=100, I0=5, I1=3, alpha=3.0, T_0=100.0,
               TFACTOR=0.95, moutlier=None, precision="f64"):
    """Full Softassign registration (materialised, literal). Returns (R, t, trace)."""
    dt = _dt(precision)
    if moutlier is None:
        moutlier = float(np.float32(1.0 / np.sqrt(np.float32(T_0)) * np.exp(np.float32(-1.0))))
    R = np.ascontiguousarray(np.eye(3) if R0 is None else np.asarray(R0).reshape(3, 3),
                             dtype=dt).reshape(9).copy()
    t = np.ascontiguousarray(np.zeros(3) if t0 is None else t0, dtype=dt).copy()
    n = max(JMAX * I0, 1)
    trace = np.zeros((n, TRACE_W))
    rc = getattr(lib(), "or_softassign_" + precision)(
        soa(X, dt), len(X), soa(Y, dt), len(Y), R, t, JMAX, I0, I1, alpha, T_0, TFACTOR,
        moutlier, trace.ctypes.data)
    if rc < 0:
        raise ValueError("oracle softassign: bad arguments (%d)" % rc)
    return R.reshape(3, 3), t, trace[:rc]


def softassign_inner(X, Y, R, t, T, alpha, I1, o_row, o_col, precision="f64"):
    """One inner iteration; o_row/o_col are updated in place (pass float arrays
    of the chosen precision). Returns dict(rowsum (post-Sinkhorn s_r), colsum
    (last sweep's b_c), a (I1, N) and b (I1, M): every sweep's row and column
    sums before their division, Xc, Yc, S, R, t, sumM)."""
    dt = _dt(precision)
    N, M = len(X), len(Y)
    R = np.ascontiguousarray(np.asarray(R).reshape(9), dtype=dt).copy()
    t = np.ascontiguousarray(t, dtype=dt).copy()
    Mat = np.zeros(N * M, dt)
    rs = np.zeros(N, dt)
    cs = np.zeros(M, dt)
    Xc = np.zeros(3, dt)
    Yc = np.zeros(3, dt)
    S = np.zeros(9, dt)
    asw = np.zeros(I1 * N, dt)
    bsw = np.zeros(I1 * M, dt)
    assert o_row.dtype == dt and o_col.dtype == dt
    sumM = getattr(lib(), "or_softassign_inner_" + precision)(
        soa(X, dt), N, soa(Y, dt), M, R, t, T, alpha, I1, o_row, o_col, Mat, rs, cs, Xc, Yc, S,
        asw.ctypes.data, bsw.ctypes.data)
    return dict(rowsum=rs, colsum=cs, Xc=Xc, Yc=Yc, S=S, R=R.reshape(3, 3), t=t, sumM=sumM,
                Mat=Mat.reshape(N, M), a=asw.reshape(I1, N), b=bsw.reshape(I1, M))


def icp_nearest(X, P):
    """Exact nearest neighbour in FP64: for each row of P, the index of the
    nearest row of X (squared Euclidean distance; the lowest index among
    equal distances) and that distance. The reference searches with FLANN's
    kd-tree (src/icp.cpp:175, :218-222, k = 1); this is the exact search it
    approximates."""
    X = np.asarray(X, np.float64)
    P = np.asarray(P, np.float64)
    idx = np.zeros(len(P), np.int64)
    d2 = np.zeros(len(P))
    chunk = max(1, 2_000_000 // max(len(X), 1))
    for s in range(0, len(P), chunk):
        Q = P[s:s + chunk]
        # |x - p|^2 per pair, formed from the differences (no expansion)
        D = ((X[None, :, :] - Q[:, None, :]) ** 2).sum(2)
        k = np.argmin(D, axis=1)
        idx[s:s + chunk] = k
        d2[s:s + chunk] = ((X[k] - Q) ** 2).sum(1)
    return idx, d2


def icp(X, Y, R0=None, t0=None, maxIteration=30):
    """ICP in FP64 with exact nearest neighbours (src/icp.cpp:122-289):
    Yc = mean Y once (:162); per iteration p = R y + t (:208-210), the nearest
    target point of each p (:218-229), Xc = mean of those (:234),
    S[i + 3j] = sum (x_n - Xc)_i (y - Yc)_j (:254-258), then findRTfromS (:268)
    through the C oracle's Horn step. Returns (R, t, trace) with trace rows
    [iteration, mean squared distance, Xc(3), Yc(3), h_S(9), R(9), t(3), 0, 0, 0]."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    R = np.eye(3) if R0 is None else np.asarray(R0, np.float64).reshape(3, 3)
    t = np.zeros(3) if t0 is None else np.asarray(t0, np.float64).reshape(3)
    Yc = Y.mean(0)
    trace = []
    for it in range(maxIteration):
        idx, d2 = icp_nearest(X, Y @ R.T + t)
        Xn = X[idx]
        Xc = Xn.mean(0)
        C = (Xn - Xc).T @ (Y - Yc)          # C[i, j] = sum (x - Xc)_i (y - Yc)_j
        S = C.T.reshape(9)                  # S[i + 3j] = C[i, j]
        Rn, tn, rc = find_rt_from_s(Xc, Yc, S)
        if rc >= 0 and np.all(np.isfinite(Rn)):
            R, t = Rn, tn
        trace.append(np.concatenate([[it, d2.mean()], Xc, Yc, S, R.reshape(9), t, [0, 0, 0]]))
    return R, t, np.array(trace).reshape(-1, TRACE_W)
