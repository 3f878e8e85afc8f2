"""emsa_emicp_dev_noalloc: EM-ICP on device buffers with its workspace in a
fixed device array (no device allocation), called through the C ABI.

Bar: the same kernels and decomposition as emsa_emicp_dev, so R and t must
equal that entry point's results bit for bit, on two calls in a row (the
second with a smaller cloud, so stale workspace contents of the first call
would show); clouds whose workspace exceeds the array are rejected with
EMSA_EINVAL before anything is launched; a call allocates no device memory.
"""
import ctypes as C

import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    import cuda_emicp_softassign_amd as E
    L = E._lib.load()
    return torch, E, L


def _dev(torch, X, Y):
    """Device SoA clouds and an identity (R, t)."""
    dX = torch.from_numpy(np.ascontiguousarray(X.T).reshape(-1)).cuda()
    dY = torch.from_numpy(np.ascontiguousarray(Y.T).reshape(-1)).cuda()
    rt = np.zeros(12, np.float32)
    rt[[0, 4, 8]] = 1.0
    return dX, dY, torch.from_numpy(rt).cuda()


def _run(torch, E, L, fn, dX, N, dY, M, drt, p):
    rc = getattr(L, fn)(C.c_void_p(dX.data_ptr()), N, C.c_void_p(dY.data_ptr()), M,
                        C.c_void_p(drt.data_ptr()), C.byref(p), C.c_void_p(0))
    E._lib.check(fn, rc)
    torch.cuda.synchronize()
    return drt.cpu().numpy().copy()


@pytest.mark.parametrize("sizes", [[(3000, 2600), (700, 900)]])
def test_noalloc_matches_emicp_dev(env, sizes):
    torch, E, L = env
    p = E.registrationParameters().to_c()
    for k, (n, m) in enumerate(sizes):
        X, Y, _, _ = uniform_cube(max(n, m), seed=40 + k)
        X, Y = X[:n].copy(), Y[:m].copy()
        print("noalloc call %d: N=%d M=%d" % (k, n, m), flush=True)
        dX, dY, drt = _dev(torch, X, Y)
        got = _run(torch, E, L, "emsa_emicp_dev_noalloc", dX, n, dY, m, drt.clone(), p)
        ref = _run(torch, E, L, "emsa_emicp_dev", dX, n, dY, m, drt.clone(), p)
        assert np.isfinite(got).all()
        assert np.array_equal(got, ref), (k, np.abs(got - ref).max())


def test_noalloc_rejects_oversized(env):
    torch, E, L = env
    p = E.registrationParameters().to_c()
    n = 200_000  # workspace well above the fixed array
    dX = torch.zeros(3 * n, dtype=torch.float32, device="cuda")
    drt = torch.zeros(12, dtype=torch.float32, device="cuda")
    rc = L.emsa_emicp_dev_noalloc(C.c_void_p(dX.data_ptr()), n, C.c_void_p(dX.data_ptr()), n,
                                  C.c_void_p(drt.data_ptr()), C.byref(p), C.c_void_p(0))
    assert rc == E._lib.EMSA_EINVAL
    assert b"workspace" in L.emsa_last_error()
    torch.cuda.synchronize()
    assert float(drt.abs().sum()) == 0.0  # nothing ran


def test_noalloc_allocates_nothing(env):
    """After one warm-up call (which loads the library's code objects), a
    call with a larger cloud than any before leaves the device's free memory
    unchanged (emsa_emicp_dev would grow its workspace here)."""
    torch, E, L = env
    p = E.registrationParameters(sigma_inf=1e-3).to_c()
    Xs, Ys, _, _ = uniform_cube(500, seed=49)
    sX, sY, srt = _dev(torch, Xs, Ys)
    _run(torch, E, L, "emsa_emicp_dev_noalloc", sX, 500, sY, 500, srt, p)
    X, Y, _, _ = uniform_cube(5000, seed=50)
    dX, dY, drt = _dev(torch, X, Y)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    _run(torch, E, L, "emsa_emicp_dev_noalloc", dX, 5000, dY, 5000, drt, p)
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
