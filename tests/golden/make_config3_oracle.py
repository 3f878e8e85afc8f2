"""Writes tests/golden/config3_oracle.npz: the FP64 oracle and the FP32
restatement of emicp_cpu on 12 instances of the batched range-scan workload
(256 pairs, N_k, M_k log-uniform in [500, 20000], seeds 1000 + k, the 24
octahedral starts), at the reference's default EM-ICP parameters
(src/main.cpp:312-315: sigma_p2 0.01, sigma_inf 1e-5, factor 0.9, d_02 0.01;
66 sigma^2 steps).

The instances span both sizes from about 500 to about 19,000 points with
N_k M_k <= 1.2e7 so that the oracle finishes in minutes on 8 cores. Stored per
instance and start: the float32 initial transform the GPU receives, the final
(R, t) of the FP64 oracle of the identical iteration and of the FP32
restatement (same operation order as src/emicp_cpu.cpp, libm expf), and the
SHA-256 of the float32 clouds, so a test can check it regenerated the same
input. Run: python tests/golden/make_config3_oracle.py  (about 10 minutes).
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cuda_emicp_softassign_amd.clouds import (batch_sizes, octahedral_rotations,  # noqa: E402
                                              range_scans)
from oracle import oracle as O  # noqa: E402

INSTANCES = [220, 219, 136, 26, 23, 210, 238, 152, 199, 34, 248, 65]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config3_oracle.npz")


def instance(k):
    """Clouds of config-3 instance k, exactly as the benchmark builds them."""
    Ns, Ms = batch_sizes(256)
    X, Y, _, _ = range_scans(int(Ns[k]), int(Ms[k]), seed=1000 + k)
    return X, Y


def clouds_sha(X, Y):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(X, np.float32).tobytes())
    h.update(np.ascontiguousarray(Y, np.float32).tobytes())
    return h.hexdigest()


def multistart_rt0(X, Y, rotations):
    """(R0, t0 = mean X - R0 mean Y) per start, float32, as
    cuda_emicp_softassign_amd.multistart_rt0 forms them."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    return np.asarray([np.concatenate([R.reshape(9), X.mean(0) - R @ Y.mean(0)])
                       for R in rotations], np.float32)


def main():
    O.build()
    starts = octahedral_rotations()
    S = len(starts)
    n = len(INSTANCES)
    rt0 = np.zeros((n, S, 12), np.float32)
    f64 = np.zeros((n, S, 12))
    f32 = np.zeros((n, S, 12))
    sha, NN, MM = [], [], []
    t0 = time.time()
    for i, k in enumerate(INSTANCES):
        X, Y = instance(k)
        NN.append(len(X))
        MM.append(len(Y))
        sha.append(clouds_sha(X, Y))
        rt0[i] = multistart_rt0(X, Y, starts)
        for s in range(S):
            R0 = rt0[i, s, :9].astype(np.float64).reshape(3, 3)
            t0_ = rt0[i, s, 9:].astype(np.float64)
            R, t, _ = O.emicp(X.astype(np.float64), Y.astype(np.float64), R0, t0_)
            f64[i, s] = np.concatenate([R.reshape(9), t])
            R, t, _ = O.emicp(X, Y, rt0[i, s, :9].reshape(3, 3), rt0[i, s, 9:],
                              precision="f32")
            f32[i, s] = np.concatenate([R.reshape(9), t]).astype(np.float64)
        print("instance %d (%d x %d) done, %.0f s" % (k, len(X), len(Y), time.time() - t0),
              flush=True)
    np.savez(OUT, instances=np.asarray(INSTANCES), N=np.asarray(NN), M=np.asarray(MM),
             sha=np.asarray(sha), rt0=rt0, oracle_f64=f64, restatement_f32=f32)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
