"""Batched multi-start EM-ICP (the path behind the headline number) against
the FP64 oracle of the identical iteration, on 12 instances of the batched
range-scan workload (256 partial-overlap pairs, N_k, M_k log-uniform in
[500, 20000], seeds 1000 + k) with all 24 octahedral starts, at the
reference's default parameters (66 sigma^2 steps, src/main.cpp:312-315).

The oracle results are golden fixtures (tests/golden/config3_oracle.npz,
written by tests/golden/make_config3_oracle.py): per registration the final
(R, t) of the FP64 oracle and of the FP32 restatement of emicp_cpu (the
reference's own arithmetic, src/emicp_cpu.cpp). The batch runs through
emsa_emicp_batched_dev with the autotuned tile, as the benchmark runs it.

Bar: R within 1e-4 rad (geodesic) and t within 1e-5 of the target's extent
of the FP64 oracle, for every registration. The FP32 restatement is reported
beside it: on these partial-overlap scans the reference's own FP32
arithmetic misses that bar on most registrations (the late sigma^2 steps
floor rows whose C_r is near 1e-6, a discontinuity that roundings resolve
differently, and plain FP32 sums over thousands of affinities). The fused
kernels (compensated differences, FP64 row and moment sums) stay within
1e-7 rad of FP64 on all of them (measured on MI355X: median 2.9e-8, max
9.0e-8 rad). Had any registration missed the bar, the rule would have been
to excuse only registrations whose FP32 restatement also misses it; none
needs excusing, so none is excused.
"""
import hashlib
import os

import numpy as np
import pytest

from cuda_emicp_softassign_amd.clouds import batch_sizes, geodesic, range_scans

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config3_oracle.npz")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(FIXTURE, allow_pickle=False))


def _instance(k):
    Ns, Ms = batch_sizes(256)
    return range_scans(int(Ns[k]), int(Ms[k]), seed=1000 + k)[:2]


def _sha(X, Y):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(X, np.float32).tobytes())
    h.update(np.ascontiguousarray(Y, np.float32).tobytes())
    return h.hexdigest()


def test_batched_config3_matches_fp64_oracle(golden):
    import torch

    import cuda_emicp_softassign_amd as E
    inst = golden["instances"]
    S = golden["rt0"].shape[1]
    keep, dpairs, ext = [], [], []
    for i, k in enumerate(inst):
        X, Y = _instance(int(k))
        assert _sha(X, Y) == str(golden["sha"][i]), "input generator drifted (instance %d)" % k
        ext.append(float((X.max(0) - X.min(0)).max()))
        dx = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        dy = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
        keep += [dx, dy]
        dpairs.append((dx.data_ptr(), len(X), dy.data_ptr(), len(Y)))
    p = E.registrationParameters()
    stream = torch.cuda.current_stream().cuda_stream
    tile = E.autotune_emicp(dpairs, S, p, stream)
    print("config-3 subset: %d instances x %d starts, tile %d" % (len(inst), S, tile), flush=True)
    d0 = torch.from_numpy(np.ascontiguousarray(golden["rt0"]).reshape(-1)).cuda()
    dout = torch.zeros_like(d0)
    dres = torch.zeros(len(inst) * S, dtype=torch.float64, device="cuda")
    E.emicp_batched_dev(dpairs, d0.data_ptr(), S, dout.data_ptr(), dres.data_ptr(), p, stream)
    torch.cuda.synchronize()
    gpu = dout.cpu().numpy().reshape(len(inst), S, 12).astype(np.float64)
    f64, f32 = golden["oracle_f64"], golden["restatement_f32"]

    def rot(a, b):
        return geodesic(a[:9].reshape(3, 3), b[:9].reshape(3, 3))

    e_gpu = np.array([[rot(gpu[i, s], f64[i, s]) for s in range(S)] for i in range(len(inst))])
    e_f32 = np.array([[rot(f32[i, s], f64[i, s]) for s in range(S)] for i in range(len(inst))])
    t_gpu = np.array([[np.abs(gpu[i, s, 9:] - f64[i, s, 9:]).max() / ext[i] for s in range(S)]
                      for i in range(len(inst))])
    well = e_f32 <= 1e-4
    print("GPU vs FP64: median %.2e rad, max %.2e; FP32 restatement vs FP64: median %.2e, "
          "max %.2e" % (np.median(e_gpu), e_gpu.max(), np.median(e_f32), e_f32.max()), flush=True)
    print("well conditioned (FP32 restatement within 1e-4 rad): %d of %d; GPU within 1e-4 rad on "
          "those: %d; on all: %d" % (well.sum(), well.size, (e_gpu[well] <= 1e-4).sum(),
                                     (e_gpu <= 1e-4).sum()), flush=True)
    for i in range(len(inst)):
        print("  instance %d (%d x %d): GPU max %.2e rad, t/extent max %.2e; FP32 restatement "
              "max %.2e rad, %d of %d starts beyond 1e-4" % (
                  inst[i], golden["N"][i], golden["M"][i], e_gpu[i].max(), t_gpu[i].max(),
                  e_f32[i].max(), (~well[i]).sum(), S), flush=True)
    assert (e_gpu <= 1e-4).all(), np.argwhere(e_gpu > 1e-4)
    assert (t_gpu <= 1e-5).all(), np.argwhere(t_gpu > 1e-5)
