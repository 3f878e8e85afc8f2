// Throughput probes for the compute ceilings the registration kernels are
// measured against: the v_exp_f32 (transcendental) issue rate and the FP32
// FMA issue rate of the whole device. Each lane runs 8 independent
// dependency chains so that latency is hidden and issue rate is measured.
#include "emsa_device.h"
#include "emsa_internal.h"

#include "../../include/emsa_probe.h"

#include <vector>

namespace emsa {
namespace {

constexpr int kChains = 8;

__global__ void __launch_bounds__(256) probe_exp_kernel(float* __restrict__ out, int iters) {
    float a[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = 0.25f + 0.01f * (threadIdx.x & 7) + 0.05f * i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = __builtin_amdgcn_exp2f(-a[i]);  // stays in (0.5, 1)
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < kChains; ++i) s += a[i];
    if (s == 12345.f) out[blockIdx.x] = s;  // keeps the chains live; never true
}

__global__ void __launch_bounds__(256) probe_fma_kernel(float* __restrict__ out, int iters, float b,
                                                        float c) {
    float a[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = 0.5f + 0.01f * (threadIdx.x & 7) + 0.05f * i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = fmaf(a[i], b, c);  // fixed point 1: bounded
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < kChains; ++i) s += a[i];
    if (s == 12345.f) out[blockIdx.x] = s;
}

// Issue cost of the instructions the fused kernels mix: per loop trip 16
// independent FMA chains, plus (SHAPE 1) 8 independent v_mfma_f32_4x4x1f32,
// (SHAPE 2) 2 v_mfma_f32_16x16x4f32 (the same 2048 multiply-adds), or
// (SHAPE 3) 16 v_exp_f32 chains instead of the FMAs. The matrix operands are
// loop-invariant, so nothing waits on a result.
template <int SHAPE>
__global__ void __launch_bounds__(256) probe_mix_kernel(float* __restrict__ out, int iters, float b,
                                                        float c) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    float a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = 0.5f + 0.01f * (threadIdx.x & 7) + 0.03f * i;
    f4 acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < iters; ++it) {
        if (SHAPE == 3) {
#pragma unroll
            for (int i = 0; i < 16; ++i) a[i] = __builtin_amdgcn_exp2f(-a[i]);  // stays in (0.5, 1)
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) a[i] = fmaf(a[i], b, c);  // fixed point 1: bounded
        }
        if (SHAPE == 1) {
#pragma unroll
            for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_4x4x1f32(b, c, acc[i], 0, 0, 0);
        }
        if (SHAPE == 2) {
#pragma unroll
            for (int i = 0; i < 2; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(b, c, acc[i], 0, 0, 0);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += a[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    if (s == 12345.f) out[blockIdx.x] = s;
}

}  // namespace
}  // namespace emsa

using namespace emsa;

extern "C" int emsa_probe_issue_cycles(double* exp_cycles, double* mfma4_cycles,
                                       double* mfma16_cycles) {
    clear_error();
    if (!exp_cycles || !mfma4_cycles || !mfma16_cycles)
        return set_error("probe: null pointer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc;
    if ((rc = ctx->misc.reserve(sizeof(float) * 65536)) != EMSA_OK) return rc;
    hipStream_t s = ctx->stream;
    const int blocks = std::max(1, ctx->numCU) * 16, threads = 256, iters = 1024;
    hipEvent_t e0, e1;
    EMSA_HIP_TRY(hipEventCreate(&e0));
    EMSA_HIP_TRY(hipEventCreate(&e1));
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    auto run = [&](auto kern, int k) -> int {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(), 32,
                           0.5f, 0.5f);
        EMSA_HIP_TRY(hipEventRecord(e0, s));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(), iters,
                           0.5f, 0.5f);
        EMSA_HIP_TRY(hipEventRecord(e1, s));
        EMSA_HIP_TRY(hipEventSynchronize(e1));
        EMSA_HIP_TRY(hipEventElapsedTime(&t[k], e0, e1));
        return EMSA_OK;
    };
    if ((rc = run(probe_mix_kernel<0>, 0)) == EMSA_OK && (rc = run(probe_mix_kernel<1>, 1)) == EMSA_OK &&
        (rc = run(probe_mix_kernel<2>, 2)) == EMSA_OK)
        rc = run(probe_mix_kernel<3>, 3);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != EMSA_OK) return rc;
    if (!(t[0] > 0.f)) return set_error("probe_issue_cycles: no time measured"), EMSA_EHIP;
    // 16 FMA per trip take 16 x 4 SIMD cycles
    const double cyc = 64.0 / (double)t[0];  // SIMD cycles per ms of a trip's share
    *mfma4_cycles = ((double)t[1] - t[0]) * cyc / 8.0;
    *mfma16_cycles = ((double)t[2] - t[0]) * cyc / 2.0;
    *exp_cycles = (double)t[3] * cyc / 16.0;
    return EMSA_OK;
}

extern "C" int emsa_probe_rates(double* exp_per_s, double* fma_per_s) {
    clear_error();
    if (!exp_per_s || !fma_per_s) return set_error("probe: null pointer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc;
    if ((rc = ctx->misc.reserve(sizeof(float) * 65536)) != EMSA_OK) return rc;
    hipStream_t s = ctx->stream;
    const int blocks = std::max(1, ctx->numCU) * 16, threads = 256, iters = 2048;
    hipEvent_t e0, e1;
    EMSA_HIP_TRY(hipEventCreate(&e0));
    EMSA_HIP_TRY(hipEventCreate(&e1));
    float ms = 0.f;
    const double work = (double)blocks * threads * iters * kChains;
    // warm-up, then timed
    hipLaunchKernelGGL(probe_exp_kernel, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(),
                       64);
    EMSA_HIP_TRY(hipEventRecord(e0, s));
    hipLaunchKernelGGL(probe_exp_kernel, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(),
                       iters);
    EMSA_HIP_TRY(hipEventRecord(e1, s));
    EMSA_HIP_TRY(hipEventSynchronize(e1));
    EMSA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *exp_per_s = work / (ms * 1e-3);
    hipLaunchKernelGGL(probe_fma_kernel, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(),
                       64, 0.5f, 0.5f);
    EMSA_HIP_TRY(hipEventRecord(e0, s));
    hipLaunchKernelGGL(probe_fma_kernel, dim3(blocks), dim3(threads), 0, s, ctx->misc.as<float>(),
                       iters, 0.5f, 0.5f);
    EMSA_HIP_TRY(hipEventRecord(e1, s));
    EMSA_HIP_TRY(hipEventSynchronize(e1));
    EMSA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *fma_per_s = work / (ms * 1e-3);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return EMSA_OK;
}

extern "C" int emsa_probe_emicp_affinity(const emsa_pair* pairs, int count, int n_starts,
                                         float sigma_p2, int reps, double* us_per_launch,
                                         double* pairs_per_launch) {
    clear_error();
    if (!pairs || count <= 0 || n_starts <= 0 || reps <= 0 || !us_per_launch ||
        !pairs_per_launch || !(sigma_p2 > 0.f))
        return set_error("probe_emicp_affinity: bad arguments"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    std::vector<EmPair> pv(count);
    double np = 0.0;
    for (int k = 0; k < count; ++k) {
        if (!pairs[k].X || !pairs[k].Y || pairs[k].N <= 0 || pairs[k].M <= 0)
            return set_error("probe_emicp_affinity: null or empty cloud"), EMSA_EINVAL;
        pv[k] = EmPair{pairs[k].X, pairs[k].N, pairs[k].Y, pairs[k].M};
        np += (double)pairs[k].N * pairs[k].M * n_starts;
    }
    const bool single = (count == 1 && n_starts == 1);
    int rc = em_prepare(*ctx, ctx->em, pv.data(), count, n_starts,
                        single ? 0 : default_tile_cols(*ctx), nullptr, nullptr, s);
    if (rc != EMSA_OK) return rc;
    const size_t nregs = (size_t)count * n_starts;
    if ((rc = ctx->rt.reserve(sizeof(float) * 12 * nregs)) != EMSA_OK) return rc;
    std::vector<float> rt0(12 * nregs, 0.f);
    for (size_t i = 0; i < nregs; ++i) rt0[12 * i] = rt0[12 * i + 4] = rt0[12 * i + 8] = 1.f;
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->rt.ptr, rt0.data(), sizeof(float) * rt0.size(),
                                hipMemcpyHostToDevice, s));
    if ((rc = em_launch_setup(ctx->em, ctx->rt.as<float>(), s)) != EMSA_OK) return rc;
    if ((rc = em_launch_affinity(*ctx, ctx->em, sigma_p2, s)) != EMSA_OK) return rc;  // warm-up
    if (ctx->emAffinityMfma) {  // pairs evaluated by one launch (skipped groups excluded)
        if ((rc = ctx->misc.reserve(2 * sizeof(unsigned long long))) != EMSA_OK) return rc;
        unsigned long long* dv = ctx->misc.as<unsigned long long>();
        EMSA_HIP_TRY(hipMemsetAsync(dv, 0, 2 * sizeof(unsigned long long), s));
        if ((rc = em_launch_affinity(*ctx, ctx->em, sigma_p2, s, dv)) != EMSA_OK) return rc;
        unsigned long long hv[2] = {0, 0};
        EMSA_HIP_TRY(hipMemcpyAsync(hv, dv, sizeof(hv), hipMemcpyDeviceToHost, s));
        EMSA_HIP_TRY(hipStreamSynchronize(s));
        np = (double)hv[0];
    }
    hipEvent_t e0, e1;
    EMSA_HIP_TRY(hipEventCreate(&e0));
    EMSA_HIP_TRY(hipEventCreate(&e1));
    EMSA_HIP_TRY(hipEventRecord(e0, s));
    for (int i = 0; i < reps && rc == EMSA_OK; ++i)
        rc = em_launch_affinity(*ctx, ctx->em, sigma_p2, s);
    EMSA_HIP_TRY(hipEventRecord(e1, s));
    EMSA_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    EMSA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != EMSA_OK) return rc;
    *us_per_launch = 1e3 * (double)ms / reps;
    *pairs_per_launch = np;
    return EMSA_OK;
}

extern "C" int emsa_probe_set_affinity_path(int use_matrix_cores) {
    clear_error();
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->emAffinityMfma = use_matrix_cores ? 1 : 0;
    }
    return emsa_graph_cache_clear();  // recorded sequences hold the previous kernel
}

extern "C" int emsa_probe_set_affinity_cull(int mode) {
    clear_error();
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->emCull = (mode == 2) ? 2 : (mode ? 1 : 0);
    }
    return emsa_graph_cache_clear();  // recorded sequences hold the previous setting
}

extern "C" int emsa_probe_set_affinity_expanded(float max_ka2) {
    clear_error();
    if (!(max_ka2 >= 0.f) || !(max_ka2 < 1e30f))
        return set_error("probe_set_affinity_expanded: bound must be finite and >= 0"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->emExpanded = max_ka2;
    }
    return emsa_graph_cache_clear();  // recorded sequences hold the previous setting
}

extern "C" int emsa_probe_set_row_fuse(int enable) {
    clear_error();
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->emRowFuse = enable ? 1 : 0;
    }
    return emsa_graph_cache_clear();  // recorded sequences hold the previous setting
}

extern "C" int emsa_probe_set_affinity_balance(int enable) {
    clear_error();
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->emBalance = enable ? 1 : 0;
    }
    return emsa_graph_cache_clear();  // recorded sequences hold the previous setting
}

extern "C" int emsa_probe_emicp_sequence(const emsa_pair* pairs, int count, int n_starts,
                                         const emsa_params* params, double* affinity_us,
                                         double* pairs_evaluated, int* launches,
                                         double* per_step, int per_step_cap) {
    clear_error();
    if (!pairs || count <= 0 || n_starts <= 0 || !params || !affinity_us || !pairs_evaluated ||
        !launches)
        return set_error("probe_emicp_sequence: bad arguments"), EMSA_EINVAL;
    if (!finite_params(*params)) return set_error("probe_emicp_sequence: bad parameters"), EMSA_EINVAL;
    const std::vector<float> sig = sigma_schedule(*params);
    if (sig.empty()) return set_error("probe_emicp_sequence: empty schedule"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    std::vector<EmPair> pv(count);
    for (int k = 0; k < count; ++k) {
        if (!pairs[k].X || !pairs[k].Y || pairs[k].N <= 0 || pairs[k].M <= 0)
            return set_error("probe_emicp_sequence: null or empty cloud"), EMSA_EINVAL;
        pv[k] = EmPair{pairs[k].X, pairs[k].N, pairs[k].Y, pairs[k].M};
    }
    const bool single = (count == 1 && n_starts == 1);
    int rc = em_prepare(*ctx, ctx->em, pv.data(), count, n_starts,
                        single ? 0 : default_tile_cols(*ctx), nullptr, nullptr, s);
    if (rc != EMSA_OK) return rc;
    const size_t nregs = (size_t)count * n_starts;
    if ((rc = ctx->rt.reserve(sizeof(float) * 24 * nregs)) != EMSA_OK) return rc;
    if ((rc = ctx->misc.reserve(2 * sizeof(unsigned long long) * sig.size())) != EMSA_OK) return rc;
    std::vector<float> rt0(12 * nregs, 0.f);
    for (size_t i = 0; i < nregs; ++i) rt0[12 * i] = rt0[12 * i + 4] = rt0[12 * i + 8] = 1.f;
    float* drt = ctx->rt.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(drt, rt0.data(), sizeof(float) * rt0.size(),
                                hipMemcpyHostToDevice, s));
    // pass 1: evaluated pairs of every affinity launch of the sequence
    unsigned long long* dv = ctx->misc.as<unsigned long long>();
    EMSA_HIP_TRY(hipMemsetAsync(dv, 0, 2 * sizeof(unsigned long long) * sig.size(), s));
    EmStepProbe count_probe{nullptr, ctx->emAffinityMfma ? dv : nullptr};
    if ((rc = em_enqueue(*ctx, ctx->em, drt, drt + 12 * nregs, nullptr, *params, sig, s,
                         &count_probe)) != EMSA_OK)
        return rc;
    std::vector<unsigned long long> hv(2 * sig.size(), 0ull);
    EMSA_HIP_TRY(hipMemcpyAsync(hv.data(), dv, 2 * sizeof(unsigned long long) * sig.size(),
                                hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    double dense = 0.0;  // pairs of one launch that visits every column
    for (int k = 0; k < count; ++k) dense += (double)pairs[k].N * pairs[k].M * n_starts;
    // pass 2: the same sequence with events around each affinity launch
    std::vector<hipEvent_t> ev(2 * sig.size(), nullptr);
    for (auto& e : ev) EMSA_HIP_TRY(hipEventCreate(&e));
    EmStepProbe time_probe{ev.data(), nullptr};
    rc = em_enqueue(*ctx, ctx->em, drt, drt + 12 * nregs, nullptr, *params, sig, s, &time_probe);
    double us = 0.0, np = 0.0;
    if (rc == EMSA_OK && hipStreamSynchronize(s) == hipSuccess) {
        for (size_t i = 0; i < sig.size(); ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) {
                set_error("probe_emicp_sequence: event timing failed");
                rc = EMSA_EHIP;
                break;
            }
            const double pi = ctx->emAffinityMfma ? (double)hv[2 * i] : dense;
            us += 1e3 * (double)ms;
            np += pi;
            if (per_step && (int)i < per_step_cap) {
                per_step[4 * i] = (double)sig[i];
                per_step[4 * i + 1] = 1e3 * (double)ms;
                per_step[4 * i + 2] = pi;
                per_step[4 * i + 3] = ctx->emAffinityMfma ? (double)hv[2 * i + 1] : 0.0;
            }
        }
    } else if (rc == EMSA_OK) {
        set_error("probe_emicp_sequence: stream synchronize failed");
        rc = EMSA_EHIP;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc != EMSA_OK) return rc;
    *affinity_us = us;
    *pairs_evaluated = np;
    *launches = (int)sig.size();
    return EMSA_OK;
}

extern "C" int emsa_probe_hilbert_order(const float* dP, int n, int* dPerm) {
    clear_error();
    if (!dP || !dPerm || n <= 0) return set_error("probe_hilbert_order: bad arguments"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    // one registration with the cloud as target and source: the ordering
    // kernel writes the target permutation
    EmPair pr{dP, n, dP, n};
    int rc = em_prepare(*ctx, ctx->em, &pr, 1, 1, 0, nullptr, nullptr, s);
    if (rc != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 12)) != EMSA_OK) return rc;
    const float rt0[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->rt.ptr, rt0, sizeof(rt0), hipMemcpyHostToDevice, s));
    if ((rc = em_launch_setup(ctx->em, ctx->rt.as<float>(), s)) != EMSA_OK) return rc;
    EMSA_HIP_TRY(hipMemcpyAsync(dPerm, ctx->em.regs[0].colperm, sizeof(int) * (size_t)n,
                                hipMemcpyDeviceToDevice, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    return EMSA_OK;
}
