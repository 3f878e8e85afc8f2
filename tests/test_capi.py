"""CPU-side checks of the product boundary: the library loads, exports every
entry point include/emsa.h declares, rejects bad arguments before touching a
GPU, and the Python/C++ parameter defaults match the reference's main.cpp."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    import glob
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "emsa*.h")):
        names |= set(re.findall(r"\b(emsa_[a-z_0-9]+)\s*\(", open(h).read()))
    return sorted(names)


@pytest.fixture(scope="module")
def L():
    from cuda_emicp_softassign_amd import _lib
    return _lib.load()


def test_exports_every_declared_symbol(L):
    names = header_functions()
    assert len(names) >= 10
    out = subprocess.run(["nm", "-D", "--defined-only",
                          os.path.join(ROOT, "cuda_emicp_softassign_amd", "libemsa.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    from cuda_emicp_softassign_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(names)


def test_dropin_cxx_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only",
                          os.path.join(ROOT, "cuda_emicp_softassign_amd", "libemsa.so")],
                         capture_output=True, text=True, check=True).stdout
    for sig in ["emicp(std::shared_ptr<pcl::PointCloud<pcl::PointXYZ> >",
                "emicp_cpu(std::shared_ptr<pcl::PointCloud<pcl::PointXYZ> >",
                "softassign(std::shared_ptr<pcl::PointCloud<pcl::PointXYZ> >",
                "findRTfromS(float const*, float const*, float const*, float*, float*)",
                "printRT(float const*, float const*)", "cloud2data("]:
        assert sig in out, sig


def test_default_params_match_reference_main(L):
    import ctypes
    import cuda_emicp_softassign_amd as E
    p = E._lib.emsa_params()
    L.emsa_default_params(ctypes.byref(p))
    py = E.registrationParameters()
    # src/main.cpp:280-315
    assert (p.JMAX, p.I0, p.I1) == (100, 5, 3) == (py.JMAX, py.I0, py.I1)
    assert p.alpha == 3.0 and p.T_0 == 100.0 and p.TFACTOR == np.float32(0.95)
    assert p.moutlier == np.float32(py.moutlier)
    assert abs(p.moutlier - np.exp(-1) / 10) < 1e-7
    assert p.sigma_p2 == np.float32(0.01) and p.sigma_inf == np.float32(1e-5)
    assert p.sigma_factor == np.float32(0.9) and p.d_02 == np.float32(0.01)
    assert L.emsa_emicp_iterations(ctypes.byref(p)) == 66


def test_invalid_arguments_rejected_on_host():
    """Argument checks run before any device call, so they hold on a CPU-only
    host; R and t stay untouched."""
    import cuda_emicp_softassign_amd as E
    X = np.zeros((4, 3), np.float32)
    hR = np.eye(3, dtype=np.float32)
    ht = np.ones(3, np.float32)
    with pytest.raises(E.EmsaError) as ei:
        E.emicp(np.zeros((0, 3)), X, hR, ht)
    assert ei.value.code == E._lib.EMSA_EINVAL
    with pytest.raises(E.EmsaError) as ei:
        E.softassign(X, X, hR, ht, E.registrationParameters(I1=0))
    assert ei.value.code == E._lib.EMSA_EINVAL
    bad = X.copy()
    bad[1, 2] = np.inf
    with pytest.raises(E.EmsaError) as ei:
        E.emicp(bad, X, hR, ht)
    assert ei.value.code == E._lib.EMSA_ENONFINITE
    assert np.array_equal(hR, np.eye(3)) and np.array_equal(ht, np.ones(3))


def test_zero_iteration_schedules_need_no_device():
    import cuda_emicp_softassign_amd as E
    X = np.random.default_rng(0).uniform(size=(10, 3)).astype(np.float32)
    R, t = E.emicp(X, X, np.eye(3), np.zeros(3),
                   E.registrationParameters(sigma_p2=1e-6, sigma_inf=1e-5))
    assert np.array_equal(R, np.eye(3))
    R, t = E.softassign(X, X, np.eye(3), np.zeros(3), E.registrationParameters(JMAX=0))
    assert np.array_equal(R, np.eye(3))


def test_cloud2data_layout():
    import cuda_emicp_softassign_amd as E
    P = np.arange(12, dtype=np.float32).reshape(4, 3)
    X, n = E.cloud2data(P)
    assert n == 4
    np.testing.assert_array_equal(X, [0, 3, 6, 9, 1, 4, 7, 10, 2, 5, 8, 11])


def test_dropin_header_compiles_without_pcl(tmp_path):
    """A caller written against the reference header compiles unchanged
    against include/3dregistration.h (no PCL in this image)."""
    src = tmp_path / "caller.cpp"
    src.write_text(open(os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")).read())
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_endless_sigma_schedule_rejected(L):
    """sigma_factor outside (0, 1) (or sigma_inf < 0) with sigma_p2 > sigma_inf:
    the reference's loop (src/emicp.cu:443, :668) never ends; here every
    EM-ICP entry point returns EMSA_EINVAL before touching a device."""
    import ctypes
    import cuda_emicp_softassign_amd as E
    X = np.random.default_rng(1).uniform(size=(10, 3)).astype(np.float32)
    for kw in (dict(sigma_factor=1.0), dict(sigma_factor=1.5), dict(sigma_factor=0.0),
               dict(sigma_factor=-0.5), dict(sigma_inf=-1e-5)):
        p = E.registrationParameters(**kw)
        assert L.emsa_emicp_iterations(ctypes.byref(p.to_c())) == E._lib.EMSA_EINVAL
        with pytest.raises(E.EmsaError) as ei:
            E.emicp(X, X, np.eye(3), np.zeros(3), p)
        assert ei.value.code == E._lib.EMSA_EINVAL
        assert "sigma_factor" in str(ei.value)
    # a schedule that runs no step is fine whatever the factor
    p = E.registrationParameters(sigma_p2=1e-6, sigma_inf=1e-5, sigma_factor=2.0)
    assert L.emsa_emicp_iterations(ctypes.byref(p.to_c())) == 0
