"""Benchmark of the soft-correspondence registration hot path on one GPU.

Default workload: EM-ICP, N = M = 10,000 FP32 points (uniform cube, seed 10),
reference defaults (sigma_p2 0.01 -> 1e-5, factor 0.9: 66 sigma^2 steps).
A "step" is one full registration with inputs already resident in HBM.

Prints ONE JSON line with the throughput (registrations/s), the roofline of
the dominant kernel (the fused EM-ICP affinity step), and the CPU baseline
(the FP32 restatement of emicp_cpu in oracle/, on the host's cores, over a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 vector peak
FLOPS_PER_PAIR = 18           # 3 sub, 1 mul + 2 fma (d^2), 1 scale, 2 (C fma), 1 (B), 6 (W fma)
VALU_PER_PAIR = 15            # issued per pair: 6 sub (compensated), 3 (d^2), 1 scale, 5 accumulate
METRIC = "registrations/s batched vs plain, N=M=10k FP32 @ rot err<=1e-4 rad; fwd+bwd ms"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 oracle parity check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how the registration would shard)", file=sys.stderr)
        sys.exit(2)

    import torch

    import cuda_emicp_softassign_amd as E
    from cuda_emicp_softassign_amd.clouds import geodesic, uniform_cube

    n = args.n
    X, Y, Rtrue, ttrue = uniform_cube(n, seed=10)
    p = E.registrationParameters()
    n_sigma = E.emicp_iterations(p)
    dev = torch.device("cuda:0")
    dX = torch.from_numpy(np.ascontiguousarray(X.T)).to(dev)
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).to(dev)
    rt0 = torch.zeros(12, dtype=torch.float32, device=dev)
    rt0[0] = rt0[4] = rt0[8] = 1.0
    dRt = rt0.clone()
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def register(params):
        dRt.copy_(rt0)
        E.emicp_dev(dX.data_ptr(), n, dY.data_ptr(), n, dRt.data_ptr(), params, sptr)

    # one sigma step only: sigma_inf just below sigma_p2
    p1 = E.registrationParameters(sigma_inf=float(np.float32(p.sigma_p2) * np.float32(0.95)))
    assert E.emicp_iterations(p1) == 1

    for _ in range(args.warmup):
        register(p)
        register(p1)
    torch.cuda.synchronize()

    # ---- timed region: K full registrations ----
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        register(p)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms_total = ev0.elapsed_time(ev1)
    ms_per_reg = ms_total / args.steps
    R = dRt[:9].cpu().numpy().reshape(3, 3).astype(np.float64)
    t = dRt[9:].cpu().numpy().astype(np.float64)

    # ---- dominant kernel: average duration of one sigma step (HIP events on
    # the launch stream; full minus one-step registrations, divided by 65) ----
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    evs[0].record(stream)
    for _ in range(args.steps):
        register(p1)
    evs[1].record(stream)
    torch.cuda.synchronize()
    ms_one = evs[0].elapsed_time(evs[1]) / args.steps
    step_ms = (ms_per_reg - ms_one) / (n_sigma - 1)
    pairs = float(n) * n
    achieved = pairs * FLOPS_PER_PAIR / (step_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_reg,
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_reg,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform cube, seed 10; Y = R^T (X - t) + N(0, 1e-3^2), rows permuted)",
        "config": {"workload": "EM-ICP plain registration, N=M=%d, %d sigma^2 steps" % (n, n_sigma),
                   "N": n, "M": n, "sigma_steps": n_sigma, "mode": "plain (one registration per call)"},
        "iterations_per_s": n_sigma * 1000.0 / ms_per_reg,
        "pairs_per_s": pairs * n_sigma * 1000.0 / ms_per_reg,
        "wall_s_timed": t_wall,
        "rot_err_vs_true_rad": geodesic(R, Rtrue),
        "roofline": {
            "bound": "mfma",
            "compute": "FP32 VALU + v_exp_f32 (no MFMA: contraction depth 3)",
            "kernel": "emicp_iter_kernel",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "avg_launch_us": step_ms * 1e3,
            "algorithmic": "%d FP32 flops + 1 exp per (x, y) pair, N*M = %.3g pairs per launch"
                           % (FLOPS_PER_PAIR, pairs),
        },
    }

    # binding compute ceiling of the fused affinity, from probes of this device
    import ctypes
    pe, pf = ctypes.c_double(), ctypes.c_double()
    E._lib.check("probe", E._lib.load().emsa_probe_rates(ctypes.byref(pe), ctypes.byref(pf)))
    pair_rate = pairs / (step_ms * 1e-3)
    ceiling = 1.0 / (VALU_PER_PAIR / pf.value + 1.0 / pe.value)  # shared issue slots
    out["roofline"].update({
        "exp_rate_per_s": pe.value, "fma_rate_per_s": pf.value,
        "valu_per_pair": VALU_PER_PAIR, "pair_rate_per_s": pair_rate,
        "pair_ceiling_per_s": ceiling, "frac_of_binding_ceiling": pair_rate / ceiling,
    })

    if not args.no_check:
        from oracle import oracle as O
        Ro, to, _ = O.emicp(X.astype(np.float64), Y.astype(np.float64))
        out["rot_err_vs_fp64_oracle_rad"] = geodesic(R, Ro)
        out["t_err_vs_fp64_oracle"] = float(np.abs(t - to).max())

    if not args.no_cpu:
        from oracle import oracle as O
        k = 3
        O.emicp(X, Y, precision="f32", max_iter=1)  # warm the library and pages
        c0 = time.perf_counter()
        O.emicp(X, Y, precision="f32", max_iter=k)
        c_s = (time.perf_counter() - c0) / k
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {
            "value": 1.0 / (c_s * n_sigma),
            "unit": "registrations/s",
            "cores": cores,
            "kind": "port",
            "sample": "first %d of %d sigma^2 steps of the same N=M=%d registration, FP32 "
                      "OpenMP restatement of emicp_cpu; %.3f s per step, scaled to one "
                      "registration" % (k, n_sigma, n, c_s),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
