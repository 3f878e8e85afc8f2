// ICP with hard correspondences on gfx950: exact nearest neighbours, FP64
// moments, Horn's Procrustes on the device.
//
// Reference semantics (src/icp.cpp:122-289), rows r over the source Y (M),
// columns c over the target X (N):
//   Yc = mean of Y, once                                     :162 (findCenter)
//   maxIteration times                                       :199
//     p_r = R y_r + t                                        :208-210
//     n(r) = the target point nearest to p_r (k = 1)         :218-222
//     Xc = mean of x_n(r)                                    :229-234
//     S[i + 3j] = sum_r (x_n(r) - Xc)_i (y_r - Yc)_j         :254-258 (sgemm "t","n")
//     (R, t) = findRTfromS(Xc, Yc, S)                        :268
// The reference finds n(r) with FLANN's randomized kd-tree (KDTreeIndexParams
// and SearchParams defaults, :175, :222), which may return a point that is
// not the nearest. Here the search is exact: every target point is compared,
// with the difference x - p_r compensated for the rounding of p_r to FP32
// (p_r = p_hi + p_lo, d = (x - p_hi) - p_lo) and |d|^2 in FP32; among equal
// distances the lowest target index wins. S is formed as
// sum_r x_n(r) y_r^T - M Xc Yc^T from FP64 row-block moments, which equals
// the reference's centred product in exact arithmetic.
//
// Per iteration three launches: nearest candidates per (64-row block, column
// chunk) tile, then per row block the best candidate and the block's FP64
// moments, then one workgroup sums the moments in a fixed order and runs Horn.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace emsa {

namespace {

constexpr int kIcpRows = 64;         // source rows per tile (one per lane)
constexpr int kIcpWaves = 4;         // waves per tile, splitting the chunk's columns
constexpr int kIcpMaxChunk = 2048;   // target columns per tile (staged in LDS)
constexpr int kIcpColStep = 4 * kIcpWaves;  // columns per round of the 4 waves
constexpr int kIcpMom = 16;          // moments per row block (see icp_rows_kernel)
constexpr int kIcpSolveWaves = 4;

struct IcpPlan {
    int rowBlocks = 0;  // ceil(M / 64)
    int chunks = 0;     // column chunks per row block
    int chunkCols = 0;  // columns per chunk (multiple of kIcpColStep, <= kIcpMaxChunk)
};

// About 4 tiles per CU, chunks of at most kIcpMaxChunk columns.
IcpPlan plan_icp(int N, int M, int numCU) {
    IcpPlan p;
    p.rowBlocks = (M + kIcpRows - 1) / kIcpRows;
    const int maxChunks = (N + kIcpColStep - 1) / kIcpColStep;
    int chunks = (4 * std::max(1, numCU) + p.rowBlocks - 1) / p.rowBlocks;
    chunks = std::max(chunks, (N + kIcpMaxChunk - 1) / kIcpMaxChunk);
    chunks = std::min(std::max(chunks, 1), maxChunks);
    int cols = (N + chunks - 1) / chunks;
    cols = (cols + kIcpColStep - 1) / kIcpColStep * kIcpColStep;
    p.chunkCols = std::min(cols, kIcpMaxChunk);
    p.chunks = (N + p.chunkCols - 1) / p.chunkCols;
    return p;
}

__global__ void icp_setup_kernel(const float* __restrict__ Rt, double* __restrict__ state) {
    if (threadIdx.x < 12) state[threadIdx.x] = (double)Rt[threadIdx.x];
    if (threadIdx.x == 12) state[kStStatus] = 0.0;
}

__global__ void icp_output_kernel(const double* __restrict__ state, float* __restrict__ Rt) {
    if (threadIdx.x < 12) Rt[threadIdx.x] = (float)state[threadIdx.x];
}

// The nearest of the chunk's target points for each of the tile's 64 source
// rows. The chunk's columns are staged in LDS (padded with +inf to a multiple
// of kIcpColStep); wave w visits columns 4w .. 4w+3 of every group of 16, in
// increasing order, keeping the first strict minimum; the 4 waves are then
// combined in order (smaller distance, or equal distance and lower index).
// Output: dist[tile][64], idx[tile][64] (idx of a row past M is unused).
__global__ void __launch_bounds__(kIcpRows * kIcpWaves)
icp_nn_kernel(const float* __restrict__ X, int N, const float* __restrict__ Y, int M,
              const double* __restrict__ state, int chunks, int chunkCols,
              float* __restrict__ dist, int* __restrict__ idx) {
    const int rb = blockIdx.x / chunks;
    const int g = blockIdx.x - rb * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int c0 = g * chunkCols;
    const int nc = min(chunkCols, N - c0);
    const int ncPad = (nc + kIcpColStep - 1) / kIcpColStep * kIcpColStep;

    __shared__ __attribute__((aligned(16))) float st[3][kIcpMaxChunk];
    for (int i = threadIdx.x; i < ncPad; i += blockDim.x) {
        const bool in = i < nc;
        st[0][i] = in ? X[c0 + i] : INFINITY;
        st[1][i] = in ? X[(size_t)N + c0 + i] : INFINITY;
        st[2][i] = in ? X[2 * (size_t)N + c0 + i] : INFINITY;
    }

    const int r = rb * kIcpRows + lane;
    float yx = 0.f, yy = 0.f, yz = 0.f;
    if (r < M) {
        yx = Y[r];
        yy = Y[(size_t)M + r];
        yz = Y[2 * (size_t)M + r];
    }
    double Rm[9], tv[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Rm[i] = state[kStR + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = state[kStT + i];
    const double pxd = Rm[0] * yx + Rm[1] * yy + Rm[2] * yz + tv[0];
    const double pyd = Rm[3] * yx + Rm[4] * yy + Rm[5] * yz + tv[1];
    const double pzd = Rm[6] * yx + Rm[7] * yy + Rm[8] * yz + tv[2];
    const float px = (float)pxd, py = (float)pyd, pz = (float)pzd;
    const float plx = (float)(pxd - (double)px), ply = (float)(pyd - (double)py),
                plz = (float)(pzd - (double)pz);
    __syncthreads();

    float best = INFINITY;
    int bi = c0;  // a valid index even if every distance is +inf
    for (int i = 4 * wave; i < ncPad; i += kIcpColStep) {
        const float4 xs = *reinterpret_cast<const float4*>(&st[0][i]);
        const float4 ys = *reinterpret_cast<const float4*>(&st[1][i]);
        const float4 zs = *reinterpret_cast<const float4*>(&st[2][i]);
        const float cx[4] = {xs.x, xs.y, xs.z, xs.w};
        const float cy[4] = {ys.x, ys.y, ys.z, ys.w};
        const float cz[4] = {zs.x, zs.y, zs.z, zs.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float dx = (cx[j] - px) - plx;
            const float dy = (cy[j] - py) - ply;
            const float dz = (cz[j] - pz) - plz;
            float s = dx * dx;
            s = fmaf(dy, dy, s);
            s = fmaf(dz, dz, s);
            if (s < best) {
                best = s;
                bi = c0 + i + j;
            }
        }
    }
    __shared__ float rs[kIcpWaves][kIcpRows];
    __shared__ int ri[kIcpWaves][kIcpRows];
    rs[wave][lane] = best;
    ri[wave][lane] = bi;
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int w = 1; w < kIcpWaves; ++w) {
        const float s = rs[w][lane];
        const int k = ri[w][lane];
        if (s < best || (s == best && k < bi)) {
            best = s;
            bi = k;
        }
    }
    dist[(size_t)blockIdx.x * kIcpRows + lane] = best;
    idx[(size_t)blockIdx.x * kIcpRows + lane] = bi;
}

// One wave per row block: the best of the chunks' candidates per row (chunks
// in column order; equal distances keep the lower index), then the block's
// FP64 moments, each a wave sum in a fixed order:
//   0..2 sum x_n(r), 3..5 sum y_r, 6 + 3a + b sum y_a x_b, 15 sum |d|^2,
// stored moment-major (mom[q * rowBlocks + rb]). rowIdx / rowDist
// (nullable): each row's neighbour and squared distance.
__global__ void __launch_bounds__(kWave)
icp_rows_kernel(const float* __restrict__ X, int N, const float* __restrict__ Y, int M,
                int chunks, const float* __restrict__ dist, const int* __restrict__ idx,
                double* __restrict__ mom, int rowBlocks, int* __restrict__ rowIdx,
                float* __restrict__ rowDist) {
    const int rb = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = rb * kIcpRows + lane;
    const size_t base = (size_t)rb * chunks * kIcpRows + lane;
    float best = dist[base];
    int bi = idx[base];
    for (int g = 1; g < chunks; ++g) {
        const float s = dist[base + (size_t)g * kIcpRows];
        const int k = idx[base + (size_t)g * kIcpRows];
        if (s < best) {
            best = s;
            bi = k;
        }
    }
    bi = min(max(bi, 0), N - 1);
    double m[kIcpMom];
#pragma unroll
    for (int q = 0; q < kIcpMom; ++q) m[q] = 0.0;
    if (r < M) {
        const double x[3] = {X[bi], X[(size_t)N + bi], X[2 * (size_t)N + bi]};
        const double y[3] = {Y[r], Y[(size_t)M + r], Y[2 * (size_t)M + r]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            m[a] = x[a];
            m[3 + a] = y[a];
#pragma unroll
            for (int b = 0; b < 3; ++b) m[6 + 3 * a + b] = y[a] * x[b];
        }
        m[15] = best;
        if (rowIdx) rowIdx[r] = bi;
        if (rowDist) rowDist[r] = best;
    }
#pragma unroll
    for (int q = 0; q < kIcpMom; ++q) {
        const double v = wave_sum(m[q]);
        if (lane == 0) mom[(size_t)q * rowBlocks + rb] = v;
    }
}

// One workgroup: the row-block moments summed (wave w: moments w, w + 4, ...;
// lanes stride over the row blocks, then a wave sum), Xc, Yc, S, Horn, the
// new (R, t) in the state (kept on failure, with the status counted) and the
// trace record.
__global__ void __launch_bounds__(kWave * kIcpSolveWaves)
icp_solve_kernel(const double* __restrict__ mom, int rowBlocks, int M, double* __restrict__ state,
                 double* __restrict__ trace, int iter) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    __shared__ double tot[kIcpMom];
    constexpr int kPer = kIcpMom / kIcpSolveWaves;
    double acc[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
    for (int b = lane; b < rowBlocks; b += kWave)
#pragma unroll
        for (int k = 0; k < kPer; ++k) acc[k] += mom[(size_t)(wave + kIcpSolveWaves * k) * rowBlocks + b];
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < kPer; ++k) tot[wave + kIcpSolveWaves * k] = acc[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double inv = 1.0 / (double)M;
    double Xc[3], Yc[3], hS[9], Rn[9], tn[3];
    for (int a = 0; a < 3; ++a) {
        Xc[a] = tot[a] * inv;
        Yc[a] = tot[3 + a] * inv;
    }
    // h_S[3a + b] = sum (y - Yc)_a (x - Xc)_b: S[i + 3j] of src/icp.cpp:254-258
    // with i over the target, j over the source
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) hS[3 * a + b] = tot[6 + 3 * a + b] - (double)M * Yc[a] * Xc[b];
    const bool ok = horn_rt(Xc, Yc, hS, Rn, tn);
    if (ok) {
        for (int i = 0; i < 9; ++i) state[kStR + i] = Rn[i];
        for (int i = 0; i < 3; ++i) state[kStT + i] = tn[i];
    } else {
        state[kStStatus] += 1.0;
    }
    if (trace) {
        double* o = trace + (size_t)iter * EMSA_TRACE_W;
        o[0] = (double)iter;
        o[1] = tot[15] * inv;
        for (int i = 0; i < 3; ++i) {
            o[2 + i] = Xc[i];
            o[5 + i] = Yc[i];
        }
        for (int i = 0; i < 9; ++i) o[8 + i] = hS[i];
        for (int i = 0; i < 9; ++i) o[17 + i] = state[kStR + i];
        for (int i = 0; i < 3; ++i) o[26 + i] = state[kStT + i];
        o[29] = 0.0;
        o[30] = 0.0;
        o[31] = ok ? 0.0 : 1.0;
    }
}

struct IcpBufs {
    float* dist;
    int* idx;
    double* mom;
    double* state;
};

int icp_reserve(DeviceCtx& ctx, const IcpPlan& pl, int M, bool rows, IcpBufs& b) {
    int rc;
    const size_t tiles = (size_t)pl.rowBlocks * pl.chunks;
    if ((rc = ctx.icpDist.reserve(sizeof(float) * tiles * kIcpRows)) != EMSA_OK) return rc;
    if ((rc = ctx.icpIdx.reserve(sizeof(int) * tiles * kIcpRows)) != EMSA_OK) return rc;
    if ((rc = ctx.icpMom.reserve(sizeof(double) * kIcpMom * (size_t)pl.rowBlocks)) != EMSA_OK)
        return rc;
    if ((rc = ctx.icpState.reserve(sizeof(double) * kStateW)) != EMSA_OK) return rc;
    if (rows) {
        if ((rc = ctx.icpRowIdx.reserve(sizeof(int) * (size_t)M)) != EMSA_OK) return rc;
        if ((rc = ctx.icpRowDist.reserve(sizeof(float) * (size_t)M)) != EMSA_OK) return rc;
    }
    b.dist = ctx.icpDist.as<float>();
    b.idx = ctx.icpIdx.as<int>();
    b.mom = ctx.icpMom.as<double>();
    b.state = ctx.icpState.as<double>();
    return EMSA_OK;
}

// Enqueues setup, `iters` iterations (iters == 0 with rowIdx: one
// correspondence query, no solve) and the output kernel.
int icp_enqueue(const IcpPlan& pl, const IcpBufs& b, const float* dX, int N, const float* dY, int M,
                const float* dRtIn, float* dRtOut, int iters, double* trace, int* rowIdx,
                float* rowDist, hipStream_t s) {
    const unsigned tiles = (unsigned)((size_t)pl.rowBlocks * pl.chunks);
    hipLaunchKernelGGL(icp_setup_kernel, dim3(1), dim3(64), 0, s, dRtIn, b.state);
    const int rounds = iters > 0 ? iters : 1;
    for (int it = 0; it < rounds; ++it) {
        hipLaunchKernelGGL(icp_nn_kernel, dim3(tiles), dim3(kIcpRows * kIcpWaves), 0, s, dX, N, dY,
                           M, b.state, pl.chunks, pl.chunkCols, b.dist, b.idx);
        hipLaunchKernelGGL(icp_rows_kernel, dim3((unsigned)pl.rowBlocks), dim3(kWave), 0, s, dX, N,
                           dY, M, pl.chunks, b.dist, b.idx, b.mom, pl.rowBlocks,
                           iters > 0 ? nullptr : rowIdx, iters > 0 ? nullptr : rowDist);
        if (iters > 0)
            hipLaunchKernelGGL(icp_solve_kernel, dim3(1), dim3(kWave * kIcpSolveWaves), 0, s,
                               b.mom, pl.rowBlocks, M, b.state, trace, it);
    }
    hipLaunchKernelGGL(icp_output_kernel, dim3(1), dim3(64), 0, s, b.state, dRtOut);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

int icp_check(const float* X, int N, const float* Y, int M, const float* R, const float* t,
              const char* who) {
    if (!X || !Y || !R || !t || N <= 0 || M <= 0)
        return set_error(std::string(who) + ": null pointer or empty cloud"), EMSA_EINVAL;
    if ((long long)N * 3 > 0x7fffffffLL || (long long)M * 3 > 0x7fffffffLL)
        return set_error(std::string(who) + ": cloud too large"), EMSA_EINVAL;
    if (!finite_cloud(X, N) || !finite_cloud(Y, M))
        return set_error(std::string(who) + ": non-finite input"), EMSA_ENONFINITE;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error(std::string(who) + ": non-finite R"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error(std::string(who) + ": non-finite t"), EMSA_ENONFINITE;
    return EMSA_OK;
}

// Host clouds in, (R, t) or the correspondences out; under the device lock,
// synchronous.
int icp_host(const float* X, int N, const float* Y, int M, float* R, float* t, int iters,
             double* htrace, int trace_cap, int* hIdx, float* hDist, const char* who) {
    int rc = icp_check(X, N, Y, M, R, t, who);
    if (rc != EMSA_OK) return rc;
    const bool query = hIdx || hDist;
    if (!query && iters <= 0) return EMSA_OK;  // loop body never runs: R, t unchanged
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    const IcpPlan pl = plan_icp(N, M, ctx->numCU);
    IcpBufs b;
    if ((rc = icp_reserve(*ctx, pl, M, query, b)) != EMSA_OK) return rc;
    if ((rc = ctx->X.reserve(sizeof(float) * 3 * (size_t)N)) != EMSA_OK) return rc;
    if ((rc = ctx->Y.reserve(sizeof(float) * 3 * (size_t)M)) != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 32)) != EMSA_OK) return rc;
    const int nrec = query ? 0 : iters;
    if (htrace && nrec > 0 &&
        (rc = ctx->trace.reserve(sizeof(double) * EMSA_TRACE_W * (size_t)nrec)) != EMSA_OK)
        return rc;
    float rt[12];
    std::memcpy(rt, R, 36);
    std::memcpy(rt + 9, t, 12);
    float* drt = ctx->rt.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->X.ptr, X, sizeof(float) * 3 * (size_t)N,
                                hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->Y.ptr, Y, sizeof(float) * 3 * (size_t)M,
                                hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(drt, rt, sizeof(rt), hipMemcpyHostToDevice, s));
    double* dtrace = (htrace && nrec > 0) ? ctx->trace.as<double>() : nullptr;
    if ((rc = icp_enqueue(pl, b, ctx->X.as<float>(), N, ctx->Y.as<float>(), M, drt, drt + 16,
                          query ? 0 : iters, dtrace, query ? ctx->icpRowIdx.as<int>() : nullptr,
                          query ? ctx->icpRowDist.as<float>() : nullptr, s)) != EMSA_OK)
        return rc;
    float out[12];
    EMSA_HIP_TRY(hipMemcpyAsync(out, drt + 16, sizeof(out), hipMemcpyDeviceToHost, s));
    std::vector<double> tr;
    if (dtrace) {
        tr.resize((size_t)EMSA_TRACE_W * nrec);
        EMSA_HIP_TRY(hipMemcpyAsync(tr.data(), dtrace, sizeof(double) * tr.size(),
                                    hipMemcpyDeviceToHost, s));
    }
    if (hIdx)
        EMSA_HIP_TRY(hipMemcpyAsync(hIdx, ctx->icpRowIdx.ptr, sizeof(int) * (size_t)M,
                                    hipMemcpyDeviceToHost, s));
    if (hDist)
        EMSA_HIP_TRY(hipMemcpyAsync(hDist, ctx->icpRowDist.ptr, sizeof(float) * (size_t)M,
                                    hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    if (query) return EMSA_OK;
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(out[i])) return set_error(std::string(who) + ": non-finite result"), EMSA_ENONFINITE;
    std::memcpy(R, out, 36);
    std::memcpy(t, out + 9, 12);
    if (htrace) std::memcpy(htrace, tr.data(), sizeof(double) * EMSA_TRACE_W * (size_t)std::min(nrec, trace_cap));
    return EMSA_OK;
}

}  // namespace
}  // namespace emsa

using namespace emsa;

extern "C" {

int emsa_icp(const float* X, int N, const float* Y, int M, float* R, float* t,
             const emsa_params* p) {
    clear_error();
    if (!p) return set_error("icp: null params"), EMSA_EINVAL;
    if (!iteration_hook_set())
        return icp_host(X, N, Y, M, R, t, p->maxIteration, nullptr, 0, nullptr, nullptr, "icp");
    const int nrec = std::max(0, p->maxIteration);
    std::vector<double> trace(std::max(nrec, 1) * (size_t)EMSA_TRACE_W);
    const int rc =
        icp_host(X, N, Y, M, R, t, p->maxIteration, trace.data(), nrec, nullptr, nullptr, "icp");
    if (rc == EMSA_OK) report_iterations(EMSA_ALG_ICP, trace.data(), nrec);
    return rc;
}

int emsa_icp_traced(const float* X, int N, const float* Y, int M, float* R, float* t,
                    const emsa_params* p, double* trace, int trace_cap) {
    clear_error();
    if (!p || !trace || trace_cap < 0) return set_error("icp: bad trace args"), EMSA_EINVAL;
    const int rc = icp_host(X, N, Y, M, R, t, p->maxIteration, trace, trace_cap, nullptr,
                            nullptr, "icp");
    return rc == EMSA_OK ? std::max(0, std::min(p->maxIteration, trace_cap)) : rc;
}

int emsa_icp_nearest(const float* X, int N, const float* Y, int M, const float* R,
                     const float* t, int* idx, float* dist2) {
    clear_error();
    if (!idx || !dist2) return set_error("icp_nearest: null output"), EMSA_EINVAL;
    float Rc[9], tc[3];
    if (R) std::memcpy(Rc, R, sizeof(Rc));
    if (t) std::memcpy(tc, t, sizeof(tc));
    return icp_host(X, N, Y, M, R ? Rc : nullptr, t ? tc : nullptr, 0, nullptr, 0, idx, dist2,
                    "icp_nearest");
}

}  // extern "C"
