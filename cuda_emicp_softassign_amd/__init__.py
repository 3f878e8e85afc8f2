"""emsa -- MI355X-native EM-ICP and Softassign rigid registration.

Python mirror of the reference's registration API (``src/3dregistration.h``):
``emicp``, ``softassign``, ``findRTfromS``, ``cloud2data``, ``printRT`` and
``registrationParameters``, with the same names, argument meaning and
in/out convention: ``h_R`` (3x3 row-major, 9 floats) and ``h_t`` (3 floats)
hold the initial estimate on entry and the result on exit; the solved problem
is min ||X - (R Y + t)|| with X = cloud_target and Y = cloud_source.

Point clouds are (n, 3) float arrays (the x, y, z of ``pcl::PointXYZ``).
Every call runs the hand-written HIP kernels of ``libemsa.so`` on the GPU; a
failure raises :class:`EmsaError` and leaves ``h_R``/``h_t`` untouched (the
reference returns void and ignores CUDA errors).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (EMSA_OK, SOFTASSIGN_LOGDOMAIN, SOFTASSIGN_MATERIALISED, TRACE_W,
                   EmsaError, emsa_params)

__all__ = [
    "registrationParameters", "emicp", "softassign", "findRTfromS", "cloud2data", "printRT",
    "emicp_traced", "emicp_step", "emicp_cpu", "softassign_traced", "softassign_step", "emicp_iterations",
    "emicp_dev",
    "EmsaError", "SOFTASSIGN_MATERIALISED", "SOFTASSIGN_LOGDOMAIN", "TRACE_W",
    "emicp_batched", "emicp_batched_dev", "autotune_emicp", "multistart_rt0", "select_best",
    "graph_cache_clear", "set_tiling",
]


def _default_moutlier(T_0=100.0):
    # src/main.cpp:286: (float)(1/sqrtf(T_0)*expf(-1.0f))
    f = np.float32
    return float(f(f(1.0) / f(math.sqrt(f(T_0)))) * f(math.exp(f(-1.0))))


@dataclass
class registrationParameters:
    """Fields of registrationParameters (src/3dregistration.h:39-76) with the
    defaults of src/main.cpp:280-315. Viewer members are not modelled."""
    JMAX: int = 100
    I0: int = 5
    I1: int = 3
    alpha: float = 3.0
    T_0: float = 100.0
    TFACTOR: float = 0.95
    moutlier: float = None  # None: 1/sqrtf(T_0)*expf(-1), from this T_0 (main.cpp:286)
    sigma_p2: float = 0.01
    sigma_inf: float = 0.00001
    sigma_factor: float = 0.9
    d_02: float = 0.01
    maxIteration: int = 30
    noviewer: int = 1
    nostop: int = 0
    notimer: int = 1

    def __post_init__(self):
        if self.moutlier is None:
            self.moutlier = _default_moutlier(self.T_0)

    def to_c(self) -> emsa_params:
        p = emsa_params()
        for name, _ in emsa_params._fields_:
            setattr(p, name, getattr(self, name))
        return p


def cloud2data(cloud):
    """(n,3) cloud -> (SoA float32 array [x.., y.., z..], n), as
    src/cloud2data.cpp:29-47 lays it out."""
    P = np.asarray(cloud, dtype=np.float32).reshape(-1, 3)
    return np.ascontiguousarray(P.T).reshape(-1), P.shape[0]


def printRT(R, t):
    """src/main.cpp:184-195."""
    R = np.asarray(R, dtype=np.float32).reshape(9)
    t = np.asarray(t, dtype=np.float32).reshape(3)
    print("R")
    for r in range(9):
        print("%f " % R[r], end="")
        if (r + 1) % 3 == 0:
            print()
    print("t")
    print("".join("%f " % v for v in t))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _inout(h_R, h_t):
    R = np.ascontiguousarray(np.asarray(h_R, dtype=np.float32).reshape(9)).copy()
    t = np.ascontiguousarray(np.asarray(h_t, dtype=np.float32).reshape(3)).copy()
    return R, t


def _writeback(h_R, h_t, R, t):
    if isinstance(h_R, np.ndarray) and h_R.size == 9:
        h_R.reshape(-1)[:] = R
    if isinstance(h_t, np.ndarray) and h_t.size == 3:
        h_t.reshape(-1)[:] = t
    return R.reshape(3, 3).copy(), t.copy()


def emicp(cloud_target, cloud_source, h_R, h_t, param=None):
    """EM-ICP on the GPU (src/emicp.cu:247-716). Returns (R, t) and also
    writes them into h_R, h_t when those are numpy arrays."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    _lib.check("emicp", _lib.load().emsa_emicp(_fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t),
                                               C.byref(p)))
    return _writeback(h_R, h_t, R, t)


def emicp_iterations(param=None):
    p = (param or registrationParameters()).to_c()
    return _lib.load().emsa_emicp_iterations(C.byref(p))


def emicp_traced(cloud_target, cloud_source, h_R, h_t, param=None):
    """As emicp, also returning the per-sigma-step trace (n, 32): sigma_p2,
    Lambda, Xc(3), Yc(3), S(9), R(9), t(3), residual, floored rows, status."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    n = max(emicp_iterations(param), 1)
    tr = np.zeros((n, TRACE_W))
    rc = _lib.check("emicp_traced", _lib.load().emsa_emicp_traced(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), C.byref(p), _dptr(tr), n))
    Rr, tt = _writeback(h_R, h_t, R, t)
    return Rr, tt, tr[:rc]


def emicp_cpu(cloud_target, cloud_source, h_R, h_t, param=None, traced=False):
    """EM-ICP on the host CPU in FP32 (src/emicp_cpu.cpp:186-547): the
    reference's iteration and operation order, no GPU involved. Returns (R, t)
    (and the per-step trace when traced) and writes R, t into h_R, h_t when
    those are numpy arrays."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    L = _lib.load()
    n = max(_lib.check("emicp_iterations", L.emsa_emicp_iterations(C.byref(p))), 1)
    tr = np.zeros((n, TRACE_W))
    rc = _lib.check("emicp_cpu", L.emsa_emicp_cpu(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), C.byref(p),
        _dptr(tr) if traced else None, n if traced else 0))
    Rr, tt = _writeback(h_R, h_t, R, t)
    return (Rr, tt, tr[:rc]) if traced else (Rr, tt)


def emicp_step(cloud_target, cloud_source, R, t, sigma_p2, d_02=0.01):
    """One sigma^2 step from (R, t): returns dict(C, lam, record, R, t)."""
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    Rf, tf = _inout(R, t)
    Cr = np.zeros(M)
    lam = np.zeros(M)
    rec = np.zeros(TRACE_W)
    _lib.check("emicp_step", _lib.load().emsa_emicp_step(
        _fptr(X), N, _fptr(Y), M, _fptr(Rf), _fptr(tf), sigma_p2, d_02, _dptr(Cr), _dptr(lam),
        _dptr(rec)))
    return dict(C=Cr, lam=lam, record=rec, R=Rf.reshape(3, 3), t=tf)


def icp(cloud_target, cloud_source, h_R, h_t, param=None):
    """ICP with exact nearest-neighbour correspondences on the GPU
    (src/icp.cpp:122-289, param.maxIteration iterations). Returns (R, t) and
    also writes them into h_R, h_t when those are numpy arrays."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    _lib.check("icp", _lib.load().emsa_icp(_fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t),
                                           C.byref(p)))
    return _writeback(h_R, h_t, R, t)


def icp_traced(cloud_target, cloud_source, h_R, h_t, param=None):
    """As icp, also returning the per-iteration trace (n, 32): iteration, mean
    squared correspondence distance, Xc(3), Yc(3), S(9), R(9), t(3), -, -,
    status."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    n = max(int(param.maxIteration), 1)
    tr = np.zeros((n, TRACE_W))
    rc = _lib.check("icp_traced", _lib.load().emsa_icp_traced(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), C.byref(p), _dptr(tr), n))
    Rr, tt = _writeback(h_R, h_t, R, t)
    return Rr, tt, tr[:rc]


def icp_nearest(cloud_target, cloud_source, R, t):
    """The correspondences one ICP iteration uses: for each source point y_r,
    the index of the target point nearest to R y_r + t and the squared
    distance (FP32)."""
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    Rf, tf = _inout(R, t)
    idx = np.zeros(M, np.int32)
    d2 = np.zeros(M, np.float32)
    _lib.check("icp_nearest", _lib.load().emsa_icp_nearest(
        _fptr(X), N, _fptr(Y), M, _fptr(Rf), _fptr(tf),
        idx.ctypes.data_as(C.POINTER(C.c_int)), _fptr(d2)))
    return idx, d2


def _variant(v):
    if isinstance(v, str):
        return {"materialised": SOFTASSIGN_MATERIALISED, "logdomain": SOFTASSIGN_LOGDOMAIN,
                "log": SOFTASSIGN_LOGDOMAIN}[v]
    return int(v)


def softassign(cloud_target, cloud_source, h_R, h_t, param=None, variant="materialised"):
    """Softassign on the GPU (src/softassign.cu:256-886). variant:
    "materialised" (FP32 matrix in HBM) or "logdomain" (recomputed)."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    _lib.check("softassign", _lib.load().emsa_softassign(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), C.byref(p), _variant(variant)))
    return _writeback(h_R, h_t, R, t)


def softassign_traced(cloud_target, cloud_source, h_R, h_t, param=None, variant="materialised"):
    """As softassign with the per-inner-iteration trace (n, 32): T, sumM,
    Xc(3), Yc(3), S(9), R(9), t(3)."""
    param = param or registrationParameters()
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    p = param.to_c()
    n = max(param.JMAX * param.I0, 1)
    tr = np.zeros((n, TRACE_W))
    rc = _lib.check("softassign_traced", _lib.load().emsa_softassign_traced(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), C.byref(p), _variant(variant), _dptr(tr),
        n))
    Rr, tt = _writeback(h_R, h_t, R, t)
    return Rr, tt, tr[:rc]


def softassign_step(cloud_target, cloud_source, h_R, h_t, T, alpha, I1, o_row, o_col,
                    variant="materialised"):
    """One Softassign inner iteration on the GPU (src/softassign.cu:494-818)
    from (h_R, h_t) at temperature T, with the persistent slack vectors o_row
    (N) and o_col (M) (float32 arrays, updated in place). Returns dict(R, t,
    a (I1, N) and b (I1, M): every Sinkhorn sweep's row and column sums before
    their division, slack included; s (N): the post-Sinkhorn row sums;
    record: T, Sigma, Xc(3), Yc(3), S(9), R(9), t(3))."""
    X, N = cloud2data(cloud_target)
    Y, M = cloud2data(cloud_source)
    R, t = _inout(h_R, h_t)
    for v, n in ((o_row, N), (o_col, M)):
        if not (isinstance(v, np.ndarray) and v.dtype == np.float32 and v.size == n
                and v.flags.c_contiguous):
            raise ValueError("o_row / o_col: contiguous float32 arrays of N / M entries")
    a = np.zeros((I1, N))
    b = np.zeros((I1, M))
    sr = np.zeros(N)
    rec = np.zeros(TRACE_W)
    _lib.check("softassign_step", _lib.load().emsa_softassign_step(
        _fptr(X), N, _fptr(Y), M, _fptr(R), _fptr(t), float(T), float(alpha), int(I1),
        _fptr(o_row), _fptr(o_col), _variant(variant), _dptr(a), _dptr(b), _dptr(sr),
        _dptr(rec)))
    Rr, tt = _writeback(h_R, h_t, R, t)
    return dict(R=Rr, t=tt, a=a, b=b, s=sr, record=rec)


def findRTfromS(h_Xc, h_Yc, h_S, h_R=None, h_t=None):
    """Horn's Procrustes step on the device in FP64 (src/findRTfromS.cpp:33-97).
    h_S[3a+b] = sum (source)_a (target)_b."""
    Xc = np.ascontiguousarray(h_Xc, dtype=np.float32).reshape(3)
    Yc = np.ascontiguousarray(h_Yc, dtype=np.float32).reshape(3)
    S = np.ascontiguousarray(h_S, dtype=np.float32).reshape(9)
    R = np.zeros(9, np.float32)
    t = np.zeros(3, np.float32)
    _lib.check("findRTfromS", _lib.load().emsa_find_rt_from_s(_fptr(Xc), _fptr(Yc), _fptr(S),
                                                              _fptr(R), _fptr(t)))
    return _writeback(h_R, h_t, R, t)


def emicp_dev(X_soa, N, Y_soa, M, Rt, param=None, stream=0):
    """EM-ICP on device buffers (raw device pointers as ints): X_soa/Y_soa
    are SoA float arrays, Rt 12 floats (R row-major then t), in/out.
    Asynchronous on `stream` (a hipStream_t as int, 0 = null stream)."""
    p = (param or registrationParameters()).to_c()
    return _lib.check("emicp_dev", _lib.load().emsa_emicp_dev(
        C.c_void_p(X_soa), N, C.c_void_p(Y_soa), M, C.c_void_p(Rt), C.byref(p),
        C.c_void_p(stream)))


# ---------------------------------------------------------------------------
# Batched EM-ICP (no counterpart in the reference)
# ---------------------------------------------------------------------------

def multistart_rt0(X, Y, rotations):
    """Initial transforms for a multi-start run: each rotation R0 with the
    translation that aligns the centroids, t0 = mean(X) - R0 mean(Y).
    Returns (n_starts, 12) float32."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    out = []
    for R0 in rotations:
        R0 = np.asarray(R0, np.float64).reshape(3, 3)
        out.append(np.concatenate([R0.reshape(9), X.mean(0) - R0 @ Y.mean(0)]))
    return np.asarray(out, np.float32)


def select_best(Rt, residual):
    """Per pair, the (12,) result of the start with the lowest residual."""
    Rt = np.asarray(Rt)
    k = np.argmin(np.asarray(residual), axis=1)
    return Rt[np.arange(Rt.shape[0]), k], k


def emicp_batched(pairs, Rt0, param=None):
    """Registers every (X, Y) in `pairs` from every initial transform in Rt0
    ((count, n_starts, 12), or (n_starts, 12) for the same starts everywhere)
    in one batched run. Returns (Rt (count, n_starts, 12), residual
    (count, n_starts))."""
    param = param or registrationParameters()
    count = len(pairs)
    Rt0 = np.asarray(Rt0, np.float32)
    if Rt0.ndim == 2:
        Rt0 = np.broadcast_to(Rt0, (count,) + Rt0.shape)
    Rt0 = np.ascontiguousarray(Rt0)
    n_starts = Rt0.shape[1]
    keep = []
    arr = (_lib.emsa_pair * count)()
    for k, (X, Y) in enumerate(pairs):
        xs, n = cloud2data(X)
        ys, m = cloud2data(Y)
        keep += [xs, ys]
        arr[k] = _lib.emsa_pair(xs.ctypes.data, n, ys.ctypes.data, m)
    out = np.zeros((count, n_starts, 12), np.float32)
    res = np.zeros((count, n_starts))
    p = param.to_c()
    _lib.check("emicp_batched", _lib.load().emsa_emicp_batched(
        arr, count, _fptr(Rt0), n_starts, C.byref(p), _fptr(out), _dptr(res)))
    return out, res


def _pairs_dev(dpairs):
    arr = (_lib.emsa_pair * len(dpairs))()
    for k, (x, n, y, m) in enumerate(dpairs):
        arr[k] = _lib.emsa_pair(x, n, y, m)
    return arr


def emicp_batched_dev(dpairs, dRt0, n_starts, dRt_out, dResid=0, param=None, stream=0):
    """Device-buffer batched EM-ICP. dpairs: list of (X_ptr, N, Y_ptr, M) with
    device SoA pointers; dRt0/dRt_out: device [count*n_starts*12] float;
    dResid: device [count*n_starts] double or 0. Replays a captured hipGraph
    when called again with the same buffers."""
    p = (param or registrationParameters()).to_c()
    arr = _pairs_dev(dpairs)
    return _lib.check("emicp_batched_dev", _lib.load().emsa_emicp_batched_dev(
        arr, len(dpairs), C.c_void_p(dRt0), n_starts, C.byref(p), C.c_void_p(dRt_out),
        C.c_void_p(dResid), C.c_void_p(stream)))


def autotune_emicp(dpairs, n_starts=1, param=None, stream=0):
    """Times the candidate decompositions on these device buffers and keeps
    the fastest for later calls on this device; returns the choice."""
    p = (param or registrationParameters()).to_c()
    arr = _pairs_dev(dpairs)
    choice = C.c_int(0)
    _lib.check("autotune_emicp", _lib.load().emsa_autotune_emicp(
        arr, len(dpairs), n_starts, C.byref(p), C.c_void_p(stream), C.byref(choice)))
    return choice.value


def set_tiling(blocks_per_cu=0, tile_cols=0):
    """Manual override of the autotuner's choice on the current device
    (0 restores the built-in heuristic)."""
    return _lib.check("set_tiling", _lib.load().emsa_set_tiling(blocks_per_cu, tile_cols))


def graph_cache_clear():
    return _lib.check("graph_cache_clear", _lib.load().emsa_graph_cache_clear())
