"""Host entry points called from several threads at once.

The reference's functions are not reentrant (a static q_pre in
src/findRTfromS.cpp:110, process-global cublasInit/cublasShutdown in
src/emicp.cu:438 and :690). libemsa serialises the calls of one device under
its context lock, keeps the last error per thread, and returns only after its
results are on the host. Here 4 threads run a mix of EM-ICP and Softassign
registrations of different sizes (so the shared workspaces grow between calls)
plus failing calls, and every result must equal the same call made alone, bit
for bit, with each failing call reporting its own error.
"""
import threading

import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import surface_with_outliers, uniform_cube

pytestmark = pytest.mark.gpu


def test_concurrent_host_calls_match_sequential():
    import cuda_emicp_softassign_amd as E

    jobs = []
    for k, n in enumerate((300, 1007, 2500, 600)):
        X, Y, _, _ = uniform_cube(n, seed=50 + k)
        jobs.append(("emicp", X, Y))
    X, Y, _, _ = surface_with_outliers(400, seed=7)
    jobs.append(("softassign", X, Y))
    X, Y, _, _ = surface_with_outliers(700, seed=8)
    jobs.append(("logdomain", X, Y))
    p_sa = E.registrationParameters(JMAX=10, I0=2, I1=3, T_0=0.5, TFACTOR=0.9, alpha=0.05)

    def run(job):
        kind, X, Y = job
        if kind == "emicp":
            return E.emicp(X, Y, np.eye(3), np.zeros(3))
        variant = "logdomain" if kind == "logdomain" else "materialised"
        return E.softassign(X, Y, np.eye(3), np.zeros(3), p_sa, variant=variant)

    print("sequential reference of %d jobs" % len(jobs), flush=True)
    ref = [run(j) for j in jobs]
    Xbad = jobs[0][1].copy()
    Xbad[5, 2] = np.nan

    results, failures = {}, []

    def worker(tid):
        try:
            for rep in range(3):
                for i in range(len(jobs)):
                    j = (i + tid) % len(jobs)
                    results[(tid, rep, j)] = run(jobs[j])
                    if (i + tid) % 3 == 0:
                        with pytest.raises(E.EmsaError) as ei:
                            E.emicp(Xbad, jobs[0][2], np.eye(3), np.zeros(3))
                        assert ei.value.code == E._lib.EMSA_ENONFINITE
                        assert "non-finite" in str(ei.value)
        except BaseException as e:  # reported by the main thread
            failures.append((tid, repr(e)))

    print("4 threads x 3 rounds", flush=True)
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
    assert not failures, failures
    assert len(results) == 4 * 3 * len(jobs)
    for (tid, rep, j), (R, t) in results.items():
        assert np.array_equal(R, ref[j][0]), (tid, rep, jobs[j][0])
        assert np.array_equal(t, ref[j][1]), (tid, rep, jobs[j][0])
