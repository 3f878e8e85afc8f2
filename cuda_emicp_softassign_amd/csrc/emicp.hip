// EM-ICP on gfx950: one fused kernel launch per sigma^2 step.
//
// Reference semantics (src/emicp.cu:443-668, src/emicp_cpu.cpp:305-527), with
// rows r over the source Y (M) and columns c over the target X (N):
//   A_rc  = exp(-|x_c - (R y_r + t)|^2 / sigma^2)               emicp.cu:110-124
//   C_r   = sum_c A_rc + exp(-d_02 / sigma^2)                     emicp.cu:477-489
//   A'_rc = sqrt(A_rc / C_r)  (or 1/N when C_r <= 1e-6)           emicp.cu:155-159
//   lambda_r = sum_c A'_rc,  X'_r = sum_c A'_rc x_c / lambda_r     emicp.cu:509-549
//   Xc, Yc, S from lambda-weighted moments, then Horn.            emicp.cu:564-635
//
// The kernel never forms A. With e_rc = exp(-|d|^2 / (2 sigma^2)) (so A = e^2
// and A' = e / sqrt(C_r)) each lane (one source row) accumulates in registers
//   C_r = sum e^2,  B_r = sum e,  W_r = sum e x_c
// and then lambda_r = B_r / sqrt(C_r), lambda_r X'_r = W_r / sqrt(C_r).
// Target points are wave-uniform, so they are read with scalar loads and
// enter the VALU as SGPR operands; one exp2 and 12 VALU ops per pair.
//
// Work decomposition per launch: a workgroup = 64 source rows x one column
// chunk; its 4 waves split the chunk. Partial sums go to HBM; the last
// workgroup to finish a row block (atomic ticket) reduces the chunks in a
// fixed order in FP64, forms the per-row normalisers and the 18 FP64 moments
// of its rows; the last row block to finish sums the moments in a fixed order
// and solves Horn's problem in FP64 on the device. No workgroup ever waits on
// another, so the launch cannot deadlock, and the reduction order does not
// depend on scheduling, so results are reproducible run to run.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>

namespace emsa {

namespace {

constexpr float kSentinel = 1.0e15f;  // padding target: exp2 underflows to 0 for any sigma

// Moments (FP64), per row block and in total:
//  [0] Lambda            [1..3] sum lambda X'      [4..6] sum lambda y
//  [7 + 3i + j] sum lambda X'_i y_j (i target axis, j source axis)
//  [16] sum lambda |X' - (R y + t)|^2   [17] floored rows
__global__ void __launch_bounds__(1024)
emicp_setup_kernel(const float* __restrict__ X, int N, int ldx, float* __restrict__ Xpad,
                   const float* __restrict__ Rt, double* __restrict__ state) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c = threadIdx.x; c < ldx; c += blockDim.x) {
        float x = kSentinel, y = kSentinel, z = kSentinel;
        if (c < N) {
            x = X[c];
            y = X[(size_t)N + c];
            z = X[2 * (size_t)N + c];
            sx += x;
            sy += y;
            sz += z;
        }
        Xpad[c] = x;
        Xpad[(size_t)ldx + c] = y;
        Xpad[2 * (size_t)ldx + c] = z;
    }
    __shared__ double lds[16 * 3];
    double v[3] = {sx, sy, sz};
    double tot;
    block_sum<3>(v, lds, blockDim.x / kWave, tot);
    if (threadIdx.x < 3) state[kStXsum + threadIdx.x] = tot;
    if (threadIdx.x < 12) state[threadIdx.x] = (double)Rt[threadIdx.x];
    if (threadIdx.x == 12) state[kStStatus] = 0.0;
}

__global__ void emicp_output_kernel(const double* __restrict__ state, float* __restrict__ Rt) {
    if (threadIdx.x < 12) Rt[threadIdx.x] = (float)state[threadIdx.x];
}

// COMP: subtract the rounding residual of p_r in every difference,
// d = (x - p_hi) - p_lo, where x - p_hi is exact for the pairs that matter
// (Sterbenz), so d carries no error from rounding p_r to FP32.
template <bool COMP>
__global__ void __launch_bounds__(kEmRows * kEmWaves)
emicp_iter_kernel(const float* __restrict__ Xpad, const float* __restrict__ Y, int N, int M,
                  int ldx, int chunks, int waveCols, float negk, float sigma_p2, double outlier,
                  double floorA, double* __restrict__ state, float* __restrict__ partial,
                  double* __restrict__ moments, unsigned* __restrict__ counters,
                  double* __restrict__ trace, double* __restrict__ rows, int iter) {
    const int rowBlocks = gridDim.x / chunks;
    const int rb = blockIdx.x / chunks;
    const int g = blockIdx.x - rb * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    // wave index made provably uniform so that target loads become scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r = rb * kEmRows + lane;
    const bool valid = r < M;

    // p_r = R y_r + t in FP64, rounded once.
    double Rm[9], tv[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Rm[i] = state[kStR + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = state[kStT + i];
    float yx = 0.f, yy = 0.f, yz = 0.f;
    if (valid) {
        yx = Y[r];
        yy = Y[(size_t)M + r];
        yz = Y[2 * (size_t)M + r];
    }
    const double pxd = Rm[0] * yx + Rm[1] * yy + Rm[2] * yz + tv[0];
    const double pyd = Rm[3] * yx + Rm[4] * yy + Rm[5] * yz + tv[1];
    const double pzd = Rm[6] * yx + Rm[7] * yy + Rm[8] * yz + tv[2];
    const float px = (float)pxd, py = (float)pyd, pz = (float)pzd;
    const float plx = (float)(pxd - (double)px), ply = (float)(pyd - (double)py),
                plz = (float)(pzd - (double)pz);

    const float* __restrict__ Xx = Xpad;
    const float* __restrict__ Xy = Xpad + ldx;
    const float* __restrict__ Xz = Xpad + 2 * (size_t)ldx;
    const int cb = (g * kEmWaves + wave) * waveCols;
    const int ce = min(cb + waveCols, ldx);

    float C = 0.f, B = 0.f, Wx = 0.f, Wy = 0.f, Wz = 0.f;
    for (int c = cb; c < ce; c += kColGroup) {
#pragma unroll
        for (int u = 0; u < kColGroup; ++u) {
            const float xx = Xx[c + u], xy = Xy[c + u], xz = Xz[c + u];
            float dx = xx - px, dy = xy - py, dz = xz - pz;
            if (COMP) {
                dx -= plx;
                dy -= ply;
                dz -= plz;
            }
            float s = dx * dx;
            s = fmaf(dy, dy, s);
            s = fmaf(dz, dz, s);
            const float e = __builtin_amdgcn_exp2f(s * negk);
            C = fmaf(e, e, C);
            B += e;
            Wx = fmaf(e, xx, Wx);
            Wy = fmaf(e, xy, Wy);
            Wz = fmaf(e, xz, Wz);
        }
    }

    // Combine the 4 waves of the chunk in LDS, fixed order.
    __shared__ float red[kEmWaves][5][kEmRows];
    red[wave][0][lane] = C;
    red[wave][1][lane] = B;
    red[wave][2][lane] = Wx;
    red[wave][3][lane] = Wy;
    red[wave][4][lane] = Wz;
    __syncthreads();
    __shared__ unsigned s_ticket;
    if (wave == 0) {
        float* P = partial + (size_t)blockIdx.x * 5 * kEmRows;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            float acc = red[0][q][lane];
#pragma unroll
            for (int w = 1; w < kEmWaves; ++w) acc += red[w][q][lane];
            P[q * kEmRows + lane] = acc;
        }
        __threadfence();
        if (lane == 0) s_ticket = atomicAdd(&counters[rb], 1u);
    }
    __syncthreads();
    if (s_ticket != (unsigned)(chunks - 1)) return;  // not the last chunk of this row block

    // ---- last workgroup of row block rb: per-row normalisers and moments ----
    __shared__ double mom[kMomW];
    if (wave == 0) {
        __threadfence();
        if (lane == 0) counters[rb] = 0u;
        double Cs = 0.0, Bs = 0.0, Xs = 0.0, Ys = 0.0, Zs = 0.0;
        const float* Pr = partial + (size_t)rb * chunks * 5 * kEmRows;
        for (int k = 0; k < chunks; ++k) {
            const float* q = Pr + (size_t)k * 5 * kEmRows;
            Cs += q[lane];
            Bs += q[kEmRows + lane];
            Xs += q[2 * kEmRows + lane];
            Ys += q[3 * kEmRows + lane];
            Zs += q[4 * kEmRows + lane];
        }
        Cs += outlier;
        double lam, Vx, Vy, Vz, fl = 0.0;
        if ((float)Cs > 10e-7f) {  // emicp.cu:155, compared in float
            const double rs = 1.0 / sqrt(Cs);
            lam = Bs * rs;
            Vx = Xs * rs;
            Vy = Ys * rs;
            Vz = Zs * rs;
        } else {  // emicp.cu:159: A' = 1/N for the whole row
            lam = (double)N * floorA;
            Vx = floorA * state[kStXsum + 0];
            Vy = floorA * state[kStXsum + 1];
            Vz = floorA * state[kStXsum + 2];
            fl = 1.0;
        }
        if (!valid) lam = Vx = Vy = Vz = fl = 0.0;
        if (rows && valid) {
            rows[r] = Cs;
            rows[(size_t)M + r] = lam;
        }
        double m[kMomW];
        m[0] = lam;
        m[1] = Vx;
        m[2] = Vy;
        m[3] = Vz;
        m[4] = lam * yx;
        m[5] = lam * yy;
        m[6] = lam * yz;
        const double V[3] = {Vx, Vy, Vz};
        const double yv[3] = {(double)yx, (double)yy, (double)yz};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) m[7 + 3 * i + j] = V[i] * yv[j];
        double res = 0.0;
        if (lam > 0.0) {
            const double ex = Vx / lam - pxd, ey = Vy / lam - pyd, ez = Vz / lam - pzd;
            res = lam * (ex * ex + ey * ey + ez * ez);
        }
        m[16] = res;
        m[17] = fl;
#pragma unroll
        for (int k = 0; k < kMomW; ++k) m[k] = wave_sum(m[k]);
        if (lane < kMomW) {
            double v = m[0];
#pragma unroll
            for (int k = 1; k < kMomW; ++k)
                if (lane == k) v = m[k];
            moments[(size_t)rb * kMomW + lane] = v;
        }
        __threadfence();
        if (lane == 0) s_ticket = atomicAdd(&counters[rowBlocks], 1u);
    }
    __syncthreads();
    if (s_ticket != (unsigned)(rowBlocks - 1)) return;

    // ---- last row block: total moments, S, Horn, new (R, t) ----
    if (wave == 0) {
        __threadfence();
        if (lane < kMomW) {
            double s = 0.0;
            for (int b = 0; b < rowBlocks; ++b) s += moments[(size_t)b * kMomW + lane];
            mom[lane] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        counters[rowBlocks] = 0u;
        const double L = mom[0];
        double Xc[3], Yc[3], hS[9], Rn[9], tn[3];
        for (int i = 0; i < 3; ++i) {
            Xc[i] = mom[1 + i] / L;
            Yc[i] = mom[4 + i] / L;
        }
        // h_S[3a+b] = sum lambda (y - Yc)_a (X' - Xc)_b   (emicp.cu:621-624)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) hS[3 * a + b] = mom[7 + 3 * b + a] - L * Xc[b] * Yc[a];
        const bool ok = horn_rt(Xc, Yc, hS, Rn, tn);
        if (ok) {
            for (int i = 0; i < 9; ++i) state[kStR + i] = Rn[i];
            for (int i = 0; i < 3; ++i) state[kStT + i] = tn[i];
        } else {
            state[kStStatus] += 1.0;
        }
        state[kStResid] = mom[16] / L;
        state[kStLambda] = L;
        if (trace) {
            double* o = trace + (size_t)iter * EMSA_TRACE_W;
            o[0] = sigma_p2;
            o[1] = L;
            for (int i = 0; i < 3; ++i) {
                o[2 + i] = Xc[i];
                o[5 + i] = Yc[i];
            }
            for (int i = 0; i < 9; ++i) o[8 + i] = hS[i];
            for (int i = 0; i < 9; ++i) o[17 + i] = state[kStR + i];
            for (int i = 0; i < 3; ++i) o[26 + i] = state[kStT + i];
            o[29] = mom[16] / L;
            o[30] = mom[17];
            o[31] = ok ? 0.0 : 1.0;
        }
    }
}

}  // namespace

EmPlan plan_emicp(int N, int M, int numCU, int chunks_override) {
    EmPlan pl;
    pl.rowBlocks = (M + kEmRows - 1) / kEmRows;
    pl.ldx = (N + kColGroup - 1) / kColGroup * kColGroup;
    const int groups = pl.ldx / kColGroup;  // column groups of 16
    int chunks = chunks_override;
    if (chunks <= 0) {
        const int target = std::max(1, numCU) * 8;  // workgroups to keep every CU busy
        chunks = (target + pl.rowBlocks - 1) / pl.rowBlocks;
    }
    chunks = std::max(1, std::min(chunks, (groups + kEmWaves - 1) / kEmWaves));
    const int gpw = (groups + chunks * kEmWaves - 1) / (chunks * kEmWaves);  // groups per wave
    pl.waveCols = gpw * kColGroup;
    pl.chunks = (pl.ldx + pl.waveCols * kEmWaves - 1) / (pl.waveCols * kEmWaves);
    return pl;
}

std::vector<float> sigma_schedule(const emsa_params& p) {
    std::vector<float> out;
    float s = p.sigma_p2;
    while (s > p.sigma_inf) {  // emicp.cu:443
        out.push_back(s);
        s *= p.sigma_factor;  // emicp.cu:668
        if (out.size() > 1000000) break;
    }
    return out;
}

int emicp_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                 const float* dRtIn, float* dRtOut, const emsa_params& p,
                 const std::vector<float>& sigmas, hipStream_t s, double* trace, double* rows) {
    const EmPlan pl = plan_emicp(N, M, ctx.numCU);
    const size_t nblocks = (size_t)pl.rowBlocks * pl.chunks;
    int rc;
    if ((rc = ctx.Xpad.reserve(sizeof(float) * 3 * (size_t)pl.ldx)) != EMSA_OK) return rc;
    if ((rc = ctx.partial.reserve(sizeof(float) * 5 * kEmRows * nblocks)) != EMSA_OK) return rc;
    if ((rc = ctx.moments.reserve(sizeof(double) * kMomW * (size_t)pl.rowBlocks)) != EMSA_OK)
        return rc;
    const size_t oldc = ctx.counters.bytes;
    if ((rc = ctx.counters.reserve(sizeof(unsigned) * ((size_t)pl.rowBlocks + 1))) != EMSA_OK)
        return rc;
    if (ctx.counters.bytes != oldc)
        EMSA_HIP_TRY(hipMemsetAsync(ctx.counters.ptr, 0, ctx.counters.bytes, s));
    if ((rc = ctx.state.reserve(sizeof(double) * kStateW)) != EMSA_OK) return rc;

    double* state = ctx.state.as<double>();
    hipLaunchKernelGGL(emicp_setup_kernel, dim3(1), dim3(1024), 0, s, dX, N, pl.ldx,
                       ctx.Xpad.as<float>(), dRtIn, state);
    EMSA_HIP_TRY(hipGetLastError());
    const double floorA = (double)(1.0f / (float)N);
    for (size_t it = 0; it < sigmas.size(); ++it) {
        const float sig = sigmas[it];
        const float negk = (float)(-kLog2e / (2.0 * (double)sig));
        const double outlier = (double)expf(-p.d_02 / sig);  // emicp.cu:489, host float
        const bool last = (it + 1 == sigmas.size());
        hipLaunchKernelGGL(emicp_iter_kernel<true>, dim3((unsigned)nblocks),
                           dim3(kEmRows * kEmWaves), 0, s, ctx.Xpad.as<float>(), dY, N, M, pl.ldx, pl.chunks, pl.waveCols,
                           negk, sig, outlier, floorA, state, ctx.partial.as<float>(),
                           ctx.moments.as<double>(), ctx.counters.as<unsigned>(), trace,
                           last ? rows : nullptr, (int)it);
        EMSA_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(emicp_output_kernel, dim3(1), dim3(64), 0, s, state, dRtOut);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // namespace emsa
