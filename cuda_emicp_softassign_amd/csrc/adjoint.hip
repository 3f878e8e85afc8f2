// Differentiable unrolled EM-ICP: forward with per-step checkpoints and the
// hand-written adjoint (no counterpart in the reference; it differentiates the
// iteration of src/emicp.cu:443-668).
//
// Forward (steps k = 0..K-1, sigma_k^2 = sigma_0^2 f^k): the ordinary fused
// kernels, with every registration also storing per step its trace record
// (Lambda, Xc, Yc, S, R_{k+1}, t_{k+1}) and per source row (C_r, lambda_r,
// lambda_r X'_r). No affinity is stored.
//
// Backward, per step from K-1 down to 0, given dL/dR_{k+1}, dL/dt_{k+1}:
//  1. Horn's step backwards -- t = Xc - R Yc, R = R(q), q the top
//     eigenvector of N(S) (dq = sum_{j != top} v_j v_j^T dN q / (w_top - w_j),
//     evaluated as one 4 x 4 solve), S = sum lambda X' y^T - Lambda Xc Yc^T,
//     Xc = SX / Lambda, Yc = SY / Lambda -- giving the 16 moment coefficients
//     dL/d(Lambda, SX, SY, SXY). A kernel of its own (adj_solve) for the last
//     step; for the others the column-0 thread of the previous step's adj_cols
//     runs it, so a step is three launches.
//  2. adj_pairs: per source row, dL/dB, dL/dC, dL/dW from lambda = B / sqrt(C),
//     lambda X' = W / sqrt(C); then every pair recomputes
//     e = exp(-|x - p|^2 / (2 sigma^2)) (compensated difference, as the forward)
//     and accumulates g_e = dB + 2 e dC + dW . x, h = -g_e e / sigma^2:
//     dL/dp_r -= h d, dL/dx_c += h d + e dW, dL/dsigma^2 += g_e e |d|^2 / (2 sigma^4).
//     Lanes hold 4 source rows each, so the per-column cross-lane sums are
//     shared by 256 pairs.
//  3. adj_rows: per row, dL/dy_r = R^T dL/dp_r + the direct moment terms, and
//     the registration partials dL/dR_k = sum dp y^T, dL/dt_k = sum dp.
//  4. adj_cols: per target column, dL/dx_c summed over row tiles in a fixed
//     order (+ the floored rows' 1/N share); its column-0 thread of each
//     registration also writes the new (dL/dR_k, dL/dt_k) and
//     dL/dsigma_0^2 += f^k dL/dsigma_k^2, then step k-1's stage 1.
// All reductions have a fixed order; no workgroup waits on another.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace emsa {

namespace {

constexpr int kAdjRPL = 4;                    // source rows per lane
constexpr int kAdjRows = kAdjRPL * kWave;      // 256 rows per tile (one wave covers them)
constexpr int kAdjWaves = 4;                   // waves per workgroup split the column chunk
constexpr int kAdjMomW = 16;                   // per-tile partials: dR(9) dt(3) dsig(1) Gfl(3)
constexpr int kAdjStW = 24;                    // adjoint state: dR(9) dt(3) dsig0(1) Gfl(3) ...
constexpr int kAdjStageW = 3 * kColGroup;      // staged column values per wave (48)
constexpr int kAdjStageLd = kWave + 1;         // LDS row stride of the staging tile
static_assert(kAdjStageW <= kWave, "one summing lane per staged value");
static_assert(kAdjWaves * kAdjStageW * kAdjStageLd >= kAdjWaves * 4 * kAdjRows,
              "the row combine reuses the staging memory");
constexpr float kSentinelA = 1.0e15f;
// Below this sigma^2 the recomputed affinities subtract the FP32 rounding
// residual of R y + t (as the forward always does); above it that residual
// changes e by less than ~1e-5 relative and is left out.
constexpr float kCompSigma2 = 1e-4f;
typedef const __attribute__((address_space(4))) float* cfloat_p;

struct AdjDesc {
    const float* X;     // target, SoA [3][N]
    float* Xpad;        // target, SoA [3][ldx], padded
    const float* Y;     // source, SoA [3][M]
    const double* rec;  // forward records [steps][EMSA_TRACE_W]
    const double* ck;   // forward row checkpoints [steps][M][5]
    const float* Rt0;   // initial transform, 12 floats
    double* adj;        // adjoint state, kAdjStW doubles
    double* coef;       // 16 moment coefficients of the current step
    float* colpart;     // [rowTiles][ldx][3]
    float* rowpart;     // [rowTiles * chunks][4][kAdjRows]
    double* rtmom;      // [rowTiles][kAdjMomW]
    float* gX;          // [3][N] accumulated over steps
    float* gY;          // [3][M] accumulated over steps
    double floorA;
    int N, M, ldx, rowTiles, chunks, waveCols, tileBase, rowBase, colBase, pad_;
};

// Orders LDS writes of some lanes before reads by other lanes of the same
// wave (LDS executes a wave's instructions in order; this keeps the compiler
// from moving accesses across and waits for outstanding LDS operations).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class F>
__device__ inline int find_by(const AdjDesc* __restrict__ d, int n, int b, F key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(d[mid]) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// (R_k, t_k): the transform step k started from.
__device__ inline void step_input(const AdjDesc& d, int k, double* R, double* t) {
    if (k == 0) {
        for (int i = 0; i < 9; ++i) R[i] = d.Rt0[i];
        for (int i = 0; i < 3; ++i) t[i] = d.Rt0[9 + i];
    } else {
        const double* o = d.rec + (size_t)(k - 1) * EMSA_TRACE_W;
        for (int i = 0; i < 9; ++i) R[i] = o[17 + i];
        for (int i = 0; i < 3; ++i) t[i] = o[26 + i];
    }
}

// Horn's step backwards (see the file comment). gR (row-major), gt: the
// gradient of the step's outputs. coef = dL/d of the step's moments in their
// forward order: [Lambda, SX(3), SY(3), SXY(9)], SXY[i][j] = sum lambda X'_i y_j
// at 7 + 3i + j.
__device__ inline void horn_backward(const double* rec, const double* gR, const double* gt,
                                     double* coef) {
    const double L = rec[1];
    const double Xc[3] = {rec[2], rec[3], rec[4]};
    const double Yc[3] = {rec[5], rec[6], rec[7]};
    const double* S = rec + 8;  // h_S[3a+b] = sum (source)_a (target)_b
    for (int i = 0; i < 16; ++i) coef[i] = 0.0;
    if (rec[31] != 0.0 || !(L > 0.0)) return;  // the forward step failed: no gradient
    double A[4][4];
    A[0][0] = S[0] + S[4] + S[8];
    A[0][1] = S[5] - S[7];
    A[0][2] = S[6] - S[2];
    A[0][3] = S[1] - S[3];
    A[1][1] = S[0] - S[4] - S[8];
    A[1][2] = S[1] + S[3];
    A[1][3] = S[6] + S[2];
    A[2][2] = S[4] - S[0] - S[8];
    A[2][3] = S[5] + S[7];
    A[3][3] = S[8] - S[0] - S[4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < i; ++j) A[i][j] = A[j][i];
    // q: the top eigenvector of A, recovered from the step's R = R(q) (the
    // sign of q does not matter: R is quadratic in q and gA below is too)
    const double* Rs = rec + 17;
    double q[4];
    {
        // the largest of 4 q_i^2 = 1 + {tr, 2 R00 - tr, 2 R11 - tr, 2 R22 - tr}
        const double tr = Rs[0] + Rs[4] + Rs[8];
        const double m[4] = {tr, 2.0 * Rs[0] - tr, 2.0 * Rs[4] - tr, 2.0 * Rs[8] - tr};
        int big = 0;
        for (int i = 1; i < 4; ++i)
            if (m[i] > m[big]) big = i;
        if (big == 0) {
            const double h = 2.0 * sqrt(1.0 + tr);  // 4 q0
            q[0] = 0.25 * h;
            q[1] = (Rs[7] - Rs[5]) / h;
            q[2] = (Rs[2] - Rs[6]) / h;
            q[3] = (Rs[3] - Rs[1]) / h;
        } else if (big == 1) {
            const double h = 2.0 * sqrt(1.0 + Rs[0] - Rs[4] - Rs[8]);  // 4 qx
            q[0] = (Rs[7] - Rs[5]) / h;
            q[1] = 0.25 * h;
            q[2] = (Rs[1] + Rs[3]) / h;
            q[3] = (Rs[2] + Rs[6]) / h;
        } else if (big == 2) {
            const double h = 2.0 * sqrt(1.0 - Rs[0] + Rs[4] - Rs[8]);  // 4 qy
            q[0] = (Rs[2] - Rs[6]) / h;
            q[1] = (Rs[1] + Rs[3]) / h;
            q[2] = 0.25 * h;
            q[3] = (Rs[5] + Rs[7]) / h;
        } else {
            const double h = 2.0 * sqrt(1.0 - Rs[0] - Rs[4] + Rs[8]);  // 4 qz
            q[0] = (Rs[3] - Rs[1]) / h;
            q[1] = (Rs[2] + Rs[6]) / h;
            q[2] = (Rs[5] + Rs[7]) / h;
            q[3] = 0.25 * h;
        }
        const double nq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int i = 0; i < 4; ++i) q[i] /= nq;
    }
    const double q0 = q[0], qx = q[1], qy = q[2], qz = q[3];
    const double R[9] = {q0 * q0 + qx * qx - qy * qy - qz * qz, 2 * (qx * qy - q0 * qz),
                         2 * (qx * qz + q0 * qy),           2 * (qy * qx + q0 * qz),
                         q0 * q0 - qx * qx + qy * qy - qz * qz, 2 * (qy * qz - q0 * qx),
                         2 * (qz * qx - q0 * qy),           2 * (qz * qy + q0 * qx),
                         q0 * q0 - qx * qx - qy * qy + qz * qz};
    // t = Xc - R Yc
    double gXc[3], gYc[3], G[9];
    for (int i = 0; i < 3; ++i) gXc[i] = gt[i];
    for (int j = 0; j < 3; ++j) gYc[j] = -(R[j] * gt[0] + R[3 + j] * gt[1] + R[6 + j] * gt[2]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) G[3 * i + j] = gR[3 * i + j] - gt[i] * Yc[j];
    // R = R(q)
    double gq[4];
    gq[0] = 2 * (q0 * G[0] - qz * G[1] + qy * G[2] + qz * G[3] + q0 * G[4] - qx * G[5] -
                 qy * G[6] + qx * G[7] + q0 * G[8]);
    gq[1] = 2 * (qx * G[0] + qy * G[1] + qz * G[2] + qy * G[3] - qx * G[4] - q0 * G[5] +
                 qz * G[6] + q0 * G[7] - qx * G[8]);
    gq[2] = 2 * (-qy * G[0] + qx * G[1] + q0 * G[2] + qx * G[3] + qy * G[4] + qz * G[5] -
                 q0 * G[6] + qz * G[7] - qy * G[8]);
    gq[3] = 2 * (-qz * G[0] - q0 * G[1] + qx * G[2] + q0 * G[3] - qz * G[4] + qy * G[5] +
                 qx * G[6] + qy * G[7] + qz * G[8]);
    // q = top eigenvector of A: gA = sum_{j != top} (v_j . gq) / (w_top - w_j) v_j q^T
    // = z q^T with z = (w_top I - A)^+ gq, i.e. z solves
    // (w_top I - A + s q q^T) z = gq - (q . gq) q (s > 0 lifts the null
    // direction q; z comes out orthogonal to q), by Gaussian elimination
    // with partial pivoting. A degenerate top eigenvalue (zero pivot) gives
    // no gradient through q, as the eigenvector sum skips zero gaps.
    double gA[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) gA[a][b] = 0.0;
    {
        double Aq[4], wtop = 0.0, qg = 0.0, an = 0.0;
        for (int i = 0; i < 4; ++i) {
            Aq[i] = A[i][0] * q[0] + A[i][1] * q[1] + A[i][2] * q[2] + A[i][3] * q[3];
            wtop += q[i] * Aq[i];
            qg += q[i] * gq[i];
            for (int j = 0; j < 4; ++j) an = fmax(an, fabs(A[i][j]));
        }
        const double sl = fmax(fabs(wtop), an);  // the lift: of the order of A's eigenvalues
        double Bm[4][5];
        for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) Bm[i][j] = (i == j ? wtop : 0.0) - A[i][j] + sl * q[i] * q[j];
            Bm[i][4] = gq[i] - qg * q[i];
        }
        bool ok = sl > 0.0;
        for (int c = 0; c < 4 && ok; ++c) {
            int pv = c;
            for (int i = c + 1; i < 4; ++i)
                if (fabs(Bm[i][c]) > fabs(Bm[pv][c])) pv = i;
            if (!(fabs(Bm[pv][c]) > 1e-13 * sl)) {
                ok = false;
                break;
            }
            if (pv != c)
                for (int j = c; j < 5; ++j) {
                    const double t = Bm[c][j];
                    Bm[c][j] = Bm[pv][j];
                    Bm[pv][j] = t;
                }
            for (int i = c + 1; i < 4; ++i) {
                const double f = Bm[i][c] / Bm[c][c];
                for (int j = c; j < 5; ++j) Bm[i][j] -= f * Bm[c][j];
            }
        }
        if (ok) {
            double z[4];
            for (int i = 3; i >= 0; --i) {
                double t = Bm[i][4];
                for (int j = i + 1; j < 4; ++j) t -= Bm[i][j] * z[j];
                z[i] = t / Bm[i][i];
            }
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) gA[a][b] = z[a] * q[b];
        }
    }
    // A = N(S): gS[m] = sum_{ab} gA[a][b] dA_ab / dS_m
    double gS[9];
    const double s01 = gA[0][1] + gA[1][0], s02 = gA[0][2] + gA[2][0], s03 = gA[0][3] + gA[3][0];
    const double s12 = gA[1][2] + gA[2][1], s13 = gA[1][3] + gA[3][1], s23 = gA[2][3] + gA[3][2];
    const double d0 = gA[0][0], d1 = gA[1][1], d2 = gA[2][2], d3 = gA[3][3];
    gS[0] = d0 + d1 - d2 - d3;
    gS[4] = d0 - d1 + d2 - d3;
    gS[8] = d0 - d1 - d2 + d3;
    gS[5] = s01 + s23;
    gS[7] = -s01 + s23;
    gS[6] = s02 + s13;
    gS[2] = -s02 + s13;
    gS[1] = s03 + s12;
    gS[3] = -s03 + s12;
    // h_S[3j+i] = SXY[i][j] - L Xc_i Yc_j
    double gL = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double g = gS[3 * j + i];
            coef[7 + 3 * i + j] = g;  // dL/dSXY[i][j]
            gXc[i] -= L * g * Yc[j];
            gYc[j] -= L * g * Xc[i];
            gL -= g * Xc[i] * Yc[j];
        }
    // Xc = SX / L, Yc = SY / L
    for (int i = 0; i < 3; ++i) {
        coef[1 + i] = gXc[i] / L;
        coef[4 + i] = gYc[i] / L;
        gL -= (gXc[i] * Xc[i] + gYc[i] * Yc[i]) / L;
    }
    coef[0] = gL;
}

// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
adj_setup_kernel(const AdjDesc* __restrict__ regs, const float* __restrict__ gRt) {
    const AdjDesc& d = regs[blockIdx.x];
    const int N = d.N, M = d.M, ldx = d.ldx;
    for (int c = threadIdx.x; c < ldx; c += blockDim.x) {
        float x = kSentinelA, y = kSentinelA, z = kSentinelA;
        if (c < N) {
            x = d.X[c];
            y = d.X[(size_t)N + c];
            z = d.X[2 * (size_t)N + c];
            d.gX[c] = 0.f;
            d.gX[(size_t)N + c] = 0.f;
            d.gX[2 * (size_t)N + c] = 0.f;
        }
        d.Xpad[c] = x;
        d.Xpad[(size_t)ldx + c] = y;
        d.Xpad[2 * (size_t)ldx + c] = z;
    }
    for (int r = threadIdx.x; r < 3 * M; r += blockDim.x) d.gY[r] = 0.f;
    if (threadIdx.x < kAdjStW) {
        const int i = threadIdx.x;
        d.adj[i] = (i < 12) ? (double)gRt[(size_t)blockIdx.x * 12 + i] : 0.0;
    }
}

__global__ void adj_solve_kernel(const AdjDesc* __restrict__ regs, int k) {
    if (threadIdx.x != 0) return;
    const AdjDesc& d = regs[blockIdx.x];
    horn_backward(d.rec + (size_t)k * EMSA_TRACE_W, d.adj, d.adj + 9, d.coef);
}

// Per-row coefficients of one source row for step k.
struct RowCoef {
    double gB, gC, gW[3], gLam, gV[3], lam, V[3];
    bool floored;
};

__device__ inline RowCoef row_coef(const AdjDesc& d, const double* coef, int k, int r, float yx,
                                   float yy, float yz) {
    RowCoef o;
    const double* ck = d.ck + ((size_t)k * d.M + r) * 5;
    const double C = ck[0];
    o.lam = ck[1];
    o.V[0] = ck[2];
    o.V[1] = ck[3];
    o.V[2] = ck[4];
    const double y[3] = {yx, yy, yz};
    o.gLam = coef[0] + coef[4] * y[0] + coef[5] * y[1] + coef[6] * y[2];
    for (int i = 0; i < 3; ++i)
        o.gV[i] = coef[1 + i] + coef[7 + 3 * i] * y[0] + coef[8 + 3 * i] * y[1] +
                  coef[9 + 3 * i] * y[2];
    o.floored = !((float)C > 10e-7f);
    if (o.floored) {
        o.gB = o.gC = 0.0;
        o.gW[0] = o.gW[1] = o.gW[2] = 0.0;
    } else {
        const double rs = 1.0 / sqrt(C);
        o.gB = o.gLam * rs;
        for (int i = 0; i < 3; ++i) o.gW[i] = o.gV[i] * rs;
        o.gC = -(o.lam * o.gLam + o.V[0] * o.gV[0] + o.V[1] * o.gV[1] + o.V[2] * o.gV[2]) /
               (2.0 * C);
    }
    return o;
}

// Stage 2: a workgroup = 256 source rows x one column chunk; lanes hold 4
// rows each, the 4 waves split the chunk. The differences are taken in units
// of the exponent scale, as the forward takes them: d' = ka x - ka p with
// ka = sqrt(log2 e / (2 sigma^2)), so e = exp2(-|d'|^2) needs no scaling
// multiply; the sums of h d' and h |d'|^2 are ka and ka^2 times those of
// h d and h |d|^2 (the rows and columns kernels divide in FP64), and the
// columns' e gW terms take ka gW. Per pair: 21 VALU + 1 exp (24 with the
// compensated difference COMP, used at small sigma) and ~1 more for the
// column sums. The row coefficients are pre-scaled by -1/sigma^2 so that
// h = ge e needs one multiply.
template <bool COMP>
__global__ void __launch_bounds__(kWave * kAdjWaves)
adj_pairs_kernel(const AdjDesc* __restrict__ regs, int nregs, int k, float ka, float minv_s2) {
    const int ri = find_by(regs, nregs, blockIdx.x, [](const AdjDesc& a) { return a.tileBase; });
    const AdjDesc& d = regs[ri];
    const int M = d.M, ldx = d.ldx, chunks = d.chunks, waveCols = d.waveCols;
    const int local = blockIdx.x - d.tileBase;
    const int rt = local / chunks;
    const int g = local - rt * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double Rk[9], tk[3];
    step_input(d, k, Rk, tk);
    double coef[16];
    for (int i = 0; i < 16; ++i) coef[i] = d.coef[i];

    // per row: p (hi, lo), m * (dB, 2 dC, dW) with m = -1/sigma^2, and dW
    float px[kAdjRPL], py[kAdjRPL], pz[kAdjRPL], lx[kAdjRPL], ly[kAdjRPL], lz[kAdjRPL];
    float mB[kAdjRPL], mC2[kAdjRPL], mWx[kAdjRPL], mWy[kAdjRPL], mWz[kAdjRPL];
    float gWx[kAdjRPL], gWy[kAdjRPL], gWz[kAdjRPL];
#pragma unroll
    for (int j = 0; j < kAdjRPL; ++j) {
        const int r = rt * kAdjRows + j * kWave + lane;
        float yx = 0.f, yy = 0.f, yz = 0.f;
        if (r < M) {
            yx = d.Y[r];
            yy = d.Y[(size_t)M + r];
            yz = d.Y[2 * (size_t)M + r];
        }
        const double kad = (double)ka;  // the row in units of the exponent scale
        const double pxd = kad * (Rk[0] * yx + Rk[1] * yy + Rk[2] * yz + tk[0]);
        const double pyd = kad * (Rk[3] * yx + Rk[4] * yy + Rk[5] * yz + tk[1]);
        const double pzd = kad * (Rk[6] * yx + Rk[7] * yy + Rk[8] * yz + tk[2]);
        px[j] = (float)pxd;
        py[j] = (float)pyd;
        pz[j] = (float)pzd;
        lx[j] = (float)(pxd - (double)px[j]);
        ly[j] = (float)(pyd - (double)py[j]);
        lz[j] = (float)(pzd - (double)pz[j]);
        mB[j] = mC2[j] = mWx[j] = mWy[j] = mWz[j] = 0.f;
        gWx[j] = gWy[j] = gWz[j] = 0.f;
        if (r < M) {
            const RowCoef o = row_coef(d, coef, k, r, yx, yy, yz);
            const double m = (double)minv_s2;
            mB[j] = (float)(m * o.gB);
            mC2[j] = (float)(m * 2.0 * o.gC);
            mWx[j] = (float)(m * o.gW[0]);
            mWy[j] = (float)(m * o.gW[1]);
            mWz[j] = (float)(m * o.gW[2]);
            gWx[j] = (float)(o.gW[0] * kad);
            gWy[j] = (float)(o.gW[1] * kad);
            gWz[j] = (float)(o.gW[2] * kad);
        }
    }

    const cfloat_p Xx = (cfloat_p)d.Xpad;
    const cfloat_p Xy = (cfloat_p)(d.Xpad + ldx);
    const cfloat_p Xz = (cfloat_p)(d.Xpad + 2 * (size_t)ldx);
    const int cb = (g * kAdjWaves + wave) * waveCols;
    const int ce = min(cb + waveCols, ldx);
    float gpx[kAdjRPL], gpy[kAdjRPL], gpz[kAdjRPL], gs[kAdjRPL];
#pragma unroll
    for (int j = 0; j < kAdjRPL; ++j) gpx[j] = gpy[j] = gpz[j] = gs[j] = 0.f;
    float* __restrict__ colpart = d.colpart + ((size_t)rt * ldx) * 3;
    // Column sums over the wave's 256 rows: each lane adds its 4 rows in
    // registers, stages (cx, cy, cz) of 16 columns in LDS ([value][lane],
    // stride 65: conflict-free both ways), then lanes 0..47 each sum one
    // value over the 64 lanes and store the 48 results contiguously.
    __shared__ float lds[kAdjWaves * kAdjStageW * kAdjStageLd];
    float* __restrict__ stage = lds + wave * kAdjStageW * kAdjStageLd;
    for (int c = cb; c < ce; c += kColGroup) {
#pragma unroll 4
        for (int u = 0; u < kColGroup; ++u) {
            const float xx = Xx[c + u], xy = Xy[c + u], xz = Xz[c + u];
            float cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll
            for (int j = 0; j < kAdjRPL; ++j) {
                float dx = fmaf(xx, ka, -px[j]), dy = fmaf(xy, ka, -py[j]),
                      dz = fmaf(xz, ka, -pz[j]);
                if (COMP) {
                    dx -= lx[j];
                    dy -= ly[j];
                    dz -= lz[j];
                }
                float s = dx * dx;
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                const float e = __builtin_amdgcn_exp2f(-s);
                // h = -(dB + 2 e dC + dW . x) e / sigma^2
                const float mg = fmaf(mWz[j], xz, fmaf(mWy[j], xy, fmaf(mWx[j], xx, fmaf(e, mC2[j], mB[j]))));
                const float h = mg * e;
                gpx[j] = fmaf(h, dx, gpx[j]);
                gpy[j] = fmaf(h, dy, gpy[j]);
                gpz[j] = fmaf(h, dz, gpz[j]);
                gs[j] = fmaf(h, s, gs[j]);
                cx = fmaf(h, dx, fmaf(e, gWx[j], cx));
                cy = fmaf(h, dy, fmaf(e, gWy[j], cy));
                cz = fmaf(h, dz, fmaf(e, gWz[j], cz));
            }
            stage[(3 * u + 0) * kAdjStageLd + lane] = cx;
            stage[(3 * u + 1) * kAdjStageLd + lane] = cy;
            stage[(3 * u + 2) * kAdjStageLd + lane] = cz;
        }
        wave_lds_sync();
        if (lane < kAdjStageW) {
            const float* __restrict__ row = stage + lane * kAdjStageLd;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
            for (int i = 0; i < kWave; i += 4) {
                a0 += row[i];
                a1 += row[i + 1];
                a2 += row[i + 2];
                a3 += row[i + 3];
            }
            colpart[(size_t)c * 3 + lane] = (a0 + a1) + (a2 + a3);
        }
        wave_lds_sync();
    }
    // row partials: combine the 4 waves (fixed order), reusing the staging
    // memory. gs holds sum h s = -(1/sigma^2) sum ge e s; the rows kernel
    // rescales it.
    __syncthreads();
    float* red = lds;  // [kAdjWaves][4][kAdjRows]
#pragma unroll
    for (int j = 0; j < kAdjRPL; ++j) {
        red[(wave * 4 + 0) * kAdjRows + j * kWave + lane] = gpx[j];
        red[(wave * 4 + 1) * kAdjRows + j * kWave + lane] = gpy[j];
        red[(wave * 4 + 2) * kAdjRows + j * kWave + lane] = gpz[j];
        red[(wave * 4 + 3) * kAdjRows + j * kWave + lane] = gs[j];
    }
    __syncthreads();
    float* P = d.rowpart + (size_t)local * 4 * kAdjRows;
    for (int q = 0; q < 4; ++q) {
        float acc = red[q * kAdjRows + threadIdx.x];
#pragma unroll
        for (int w = 1; w < kAdjWaves; ++w) acc += red[(w * 4 + q) * kAdjRows + threadIdx.x];
        P[q * kAdjRows + threadIdx.x] = acc;
    }
}

// Stage 3: one workgroup per 256-row tile, one thread per row.
__global__ void __launch_bounds__(kAdjRows)
adj_rows_kernel(const AdjDesc* __restrict__ regs, int nregs, int k, double gsig_scale,
                double outlier_dsig, double inv_ka) {
    const int ri = find_by(regs, nregs, blockIdx.x, [](const AdjDesc& a) { return a.rowBase; });
    const AdjDesc& d = regs[ri];
    const int M = d.M, chunks = d.chunks;
    const int rt = blockIdx.x - d.rowBase;
    const int r = rt * kAdjRows + threadIdx.x;
    double m[kAdjMomW];
#pragma unroll
    for (int i = 0; i < kAdjMomW; ++i) m[i] = 0.0;
    if (r < M) {
        double gp[3] = {0.0, 0.0, 0.0}, gsum = 0.0;
        const float* P = d.rowpart + (size_t)rt * chunks * 4 * kAdjRows + threadIdx.x;
        for (int c = 0; c < chunks; ++c) {
            const float* q = P + (size_t)c * 4 * kAdjRows;
            gp[0] += q[0];
            gp[1] += q[kAdjRows];
            gp[2] += q[2 * kAdjRows];
            gsum += q[3 * kAdjRows];
        }
        // accumulated +h d' = ka h d; dL/dp = -sum h d
        gp[0] = -gp[0] * inv_ka;
        gp[1] = -gp[1] * inv_ka;
        gp[2] = -gp[2] * inv_ka;
        gsum *= inv_ka * inv_ka;  // sum h |d'|^2 = ka^2 sum h |d|^2
        const float yx = d.Y[r], yy = d.Y[(size_t)M + r], yz = d.Y[2 * (size_t)M + r];
        const RowCoef o = row_coef(d, d.coef, k, r, yx, yy, yz);
        double Rk[9], tk[3];
        step_input(d, k, Rk, tk);
        const double y[3] = {yx, yy, yz};
        double gy[3];
        for (int j = 0; j < 3; ++j) {
            // moments: lambda y_j and V_i y_j
            double g = o.lam * d.coef[4 + j];
            for (int i = 0; i < 3; ++i) g += d.coef[7 + 3 * i + j] * o.V[i];
            // p = R y + t
            g += Rk[j] * gp[0] + Rk[3 + j] * gp[1] + Rk[6 + j] * gp[2];
            gy[j] = g;
        }
        d.gY[r] += (float)gy[0];
        d.gY[(size_t)M + r] += (float)gy[1];
        d.gY[2 * (size_t)M + r] += (float)gy[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m[3 * i + j] = gp[i] * y[j];
        for (int i = 0; i < 3; ++i) m[9 + i] = gp[i];
        m[12] = gsum * gsig_scale + (o.floored ? 0.0 : o.gC * outlier_dsig);
        if (o.floored)
            for (int i = 0; i < 3; ++i) m[13 + i] = o.gV[i];
    }
    __shared__ double lds[4 * kAdjMomW];
    double tot;
    block_sum<kAdjMomW>(m, lds, kAdjRows / kWave, tot);
    if (threadIdx.x < kAdjMomW) d.rtmom[(size_t)rt * kAdjMomW + threadIdx.x] = tot;
}

constexpr int kAdjTileRound = 8;  // row tiles whose partials a thread loads at once

// Stage 4: one thread per target column (workgroups 0 .. colBlocks - 1), and
// one workgroup per registration after them that finishes the step: the row
// tiles' moments summed in tile order (threads 0..15, one moment each) give
// the next step's dL/dR, dL/dt, the dL/dsigma_0^2 term and the floored rows'
// dL/dV sum, then Horn's step backwards for the next (earlier) step (its
// coefficients for the pair pass that follows; the pair pass of this step has
// read the current ones). Every column thread sums the floored rows' dL/dV
// itself. Sums over the row tiles load kAdjTileRound tiles at once
// (predicated) and add them in tile order. fk = factor^k. A step whose Horn
// solve failed in the forward left (R, t) unchanged, so its gradient passes
// through.
__global__ void __launch_bounds__(256)
adj_cols_kernel(const AdjDesc* __restrict__ regs, int nregs, long long totalCols, int k,
                double fk, double inv_ka, int next, int colBlocks) {
    static_assert(kAdjMomW >= 16, "moments 0..12 and the floored dL/dV 13..15");
    if ((int)blockIdx.x >= colBlocks) {
        const AdjDesc& d = regs[blockIdx.x - colBlocks];
        __shared__ double msum[16];
        const int q = threadIdx.x, rowTiles = d.rowTiles;
        if (q < 16) {
            double a = 0.0;
            for (int t0 = 0; t0 < rowTiles; t0 += kAdjTileRound) {
                double v[kAdjTileRound];
#pragma unroll
                for (int u = 0; u < kAdjTileRound; ++u)
                    v[u] = (t0 + u < rowTiles) ? d.rtmom[(size_t)(t0 + u) * kAdjMomW + q] : 0.0;
#pragma unroll
                for (int u = 0; u < kAdjTileRound; ++u) a += v[u];
            }
            msum[q] = a;
        }
        __syncthreads();
        if (q == 0) {
            const bool failed = d.rec[(size_t)k * EMSA_TRACE_W + 31] != 0.0;
            if (!failed)
                for (int i = 0; i < 12; ++i) d.adj[i] = msum[i];  // dL/dR_k, dL/dt_k
            d.adj[12] += fk * msum[12];                           // dL/dsigma_0^2
            for (int i = 0; i < 3; ++i) d.adj[13 + i] = msum[13 + i];
            if (next >= 0)
                horn_backward(d.rec + (size_t)next * EMSA_TRACE_W, d.adj, d.adj + 9, d.coef);
        }
        return;
    }
    const long long gc = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gc >= totalCols) return;
    const int ri = find_by(regs, nregs, (int)gc, [](const AdjDesc& a) { return a.colBase; });
    const AdjDesc& d = regs[ri];
    const int c = (int)gc - d.colBase;
    if (c >= d.N) return;
    const int rowTiles = d.rowTiles;
    double gfl[3] = {0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < rowTiles; t0 += kAdjTileRound) {
        double v[kAdjTileRound][3];
#pragma unroll
        for (int u = 0; u < kAdjTileRound; ++u)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                v[u][i] = (t0 + u < rowTiles) ? d.rtmom[(size_t)(t0 + u) * kAdjMomW + 13 + i] : 0.0;
#pragma unroll
        for (int u = 0; u < kAdjTileRound; ++u)
#pragma unroll
            for (int i = 0; i < 3; ++i) gfl[i] += v[u][i];
    }
    double g[3] = {0.0, 0.0, 0.0};  // the pair terms, ka times dL/dx (see adj_pairs_kernel)
    for (int t0 = 0; t0 < rowTiles; t0 += kAdjTileRound) {
        float v[kAdjTileRound][3];
#pragma unroll
        for (int u = 0; u < kAdjTileRound; ++u) {
            const float* o = d.colpart + ((size_t)(t0 + u) * d.ldx + c) * 3;
#pragma unroll
            for (int i = 0; i < 3; ++i) v[u][i] = (t0 + u < rowTiles) ? o[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kAdjTileRound; ++u)
#pragma unroll
            for (int i = 0; i < 3; ++i) g[i] += v[u][i];
    }
    for (int i = 0; i < 3; ++i) g[i] = g[i] * inv_ka + d.floorA * gfl[i];
    d.gX[c] += (float)g[0];
    d.gX[(size_t)d.N + c] += (float)g[1];
    d.gX[2 * (size_t)d.N + c] += (float)g[2];
}

__global__ void adj_output_kernel(const AdjDesc* __restrict__ regs, float* __restrict__ gRt0,
                                  double* __restrict__ gsig) {
    const AdjDesc& d = regs[blockIdx.x];
    if (gRt0 && threadIdx.x < 12) gRt0[(size_t)blockIdx.x * 12 + threadIdx.x] = (float)d.adj[threadIdx.x];
    if (gsig && threadIdx.x == 12) gsig[blockIdx.x] = d.adj[12];
}

struct AdjWork {
    DevBuf xpad, adj, coef, colpart, rowpart, rtmom, desc;
    std::vector<AdjDesc> regs;
    long long tiles = 0, rowTiles = 0, cols = 0;
    WorkFence fence;  // the last backward that used this workspace, on any stream
};

constexpr int kMaxDevices = 64;
AdjWork* g_adj[kMaxDevices];  // per device, created under that device's lock, never freed

AdjWork* adj_work(const DeviceCtx& ctx) {
    if (ctx.device < 0 || ctx.device >= kMaxDevices) return nullptr;
    if (!g_adj[ctx.device]) g_adj[ctx.device] = new AdjWork();
    return g_adj[ctx.device];
}

std::vector<float> unrolled_sigmas(float sigma2_0, float factor, int steps) {
    std::vector<float> s(steps);
    for (int k = 0; k < steps; ++k) s[k] = (float)((double)sigma2_0 * std::pow((double)factor, k));
    return s;
}

size_t ckpt_doubles(const emsa_pair* pairs, int count, int steps, std::vector<size_t>* offs) {
    size_t tot = 0;
    for (int k = 0; k < count; ++k) {
        if (offs) offs->push_back(tot);
        tot += (size_t)steps * (EMSA_TRACE_W + 5 * (size_t)pairs[k].M);
    }
    return tot;
}

int check(const emsa_pair* pairs, int count, int steps, float sigma2_0, float factor) {
    if (!pairs || count <= 0 || steps <= 0) return set_error("unrolled: bad sizes"), EMSA_EINVAL;
    if (!(sigma2_0 > 0.f) || !(factor > 0.f) || !std::isfinite(sigma2_0) || !std::isfinite(factor))
        return set_error("unrolled: sigma2_0 and factor must be positive"), EMSA_EINVAL;
    for (int k = 0; k < count; ++k)
        if (!pairs[k].X || !pairs[k].Y || pairs[k].N <= 0 || pairs[k].M <= 0)
            return set_error("unrolled: null or empty cloud"), EMSA_EINVAL;
    return EMSA_OK;
}

}  // namespace

}  // namespace emsa

using namespace emsa;

extern "C" {

long long emsa_emicp_unrolled_ckpt_size(const emsa_pair* pairs, int count, int steps) {
    clear_error();
    if (!pairs || count <= 0 || steps <= 0) return set_error("unrolled: bad sizes"), EMSA_EINVAL;
    return (long long)ckpt_doubles(pairs, count, steps, nullptr);
}

int emsa_emicp_unrolled_fwd(const emsa_pair* pairs, int count, const float* dRt0, float sigma2_0,
                            float factor, float d_02, int steps, float* dRtOut, double* dckpt,
                            void* stream) {
    clear_error();
    int rc;
    if ((rc = check(pairs, count, steps, sigma2_0, factor)) != EMSA_OK) return rc;
    if (!dRt0 || !dRtOut || !dckpt) return set_error("unrolled_fwd: null buffer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    std::vector<size_t> offs;
    ckpt_doubles(pairs, count, steps, &offs);
    std::vector<EmAux> aux(count);
    std::vector<EmPair> pv(count);
    for (int k = 0; k < count; ++k) {
        pv[k] = EmPair{pairs[k].X, pairs[k].N, pairs[k].Y, pairs[k].M};
        aux[k].trace = dckpt + offs[k];
        aux[k].ckpt = dckpt + offs[k] + (size_t)steps * EMSA_TRACE_W;
    }
    emsa_params p;
    emsa_default_params(&p);
    p.d_02 = d_02;
    rc = em_prepare(*ctx, ctx->em, pv.data(), count, 1, count == 1 ? 0 : default_tile_cols(*ctx),
                    nullptr, nullptr, s, aux.data());
    if (rc != EMSA_OK) return rc;
    return em_enqueue(*ctx, ctx->em, dRt0, dRtOut, nullptr, p,
                      unrolled_sigmas(sigma2_0, factor, steps), s);
}

int emsa_emicp_unrolled_bwd(const emsa_pair* pairs, int count, const float* dRt0, float sigma2_0,
                            float factor, float d_02, int steps, const double* dckpt,
                            const float* dgRt, const emsa_pair_grad* grads, float* dgRt0,
                            double* dgsigma, void* stream) {
    clear_error();
    int rc;
    if ((rc = check(pairs, count, steps, sigma2_0, factor)) != EMSA_OK) return rc;
    if (!dRt0 || !dckpt || !dgRt || !grads)
        return set_error("unrolled_bwd: null buffer"), EMSA_EINVAL;
    for (int k = 0; k < count; ++k)
        if (!grads[k].gX || !grads[k].gY)
            return set_error("unrolled_bwd: null gradient buffer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    AdjWork* wp = adj_work(*ctx);
    if (!wp) return set_error("unrolled_bwd: device index out of range"), EMSA_ENODEV;
    AdjWork& w = *wp;
    // a backward on another stream may still read this workspace
    if ((rc = w.fence.wait()) != EMSA_OK) return rc;
    FenceMark mark(w.fence, s);
    std::vector<size_t> offs;
    ckpt_doubles(pairs, count, steps, &offs);
    const int tileCols = 1024;
    w.regs.assign(count, AdjDesc{});
    size_t xpadF = 0, colF = 0, rowF = 0, momD = 0;
    long long tiles = 0, rowTiles = 0, cols = 0;
    for (int k = 0; k < count; ++k) {
        AdjDesc& d = w.regs[k];
        d.X = pairs[k].X;
        d.Y = pairs[k].Y;
        d.N = pairs[k].N;
        d.M = pairs[k].M;
        d.rec = dckpt + offs[k];
        d.ck = dckpt + offs[k] + (size_t)steps * EMSA_TRACE_W;
        d.Rt0 = dRt0 + (size_t)k * 12;
        d.gX = grads[k].gX;
        d.gY = grads[k].gY;
        d.floorA = (double)(1.0f / (float)d.N);
        d.ldx = (d.N + kColGroup - 1) / kColGroup * kColGroup;
        d.rowTiles = (d.M + kAdjRows - 1) / kAdjRows;
        d.waveCols = tileCols / kAdjWaves;
        d.chunks = (d.ldx + tileCols - 1) / tileCols;
        d.tileBase = (int)tiles;
        d.rowBase = (int)rowTiles;
        d.colBase = (int)cols;
        tiles += (long long)d.rowTiles * d.chunks;
        rowTiles += d.rowTiles;
        cols += d.ldx;
        d.Xpad = reinterpret_cast<float*>(xpadF);
        xpadF += 3 * (size_t)d.ldx;
        d.colpart = reinterpret_cast<float*>(colF);
        colF += (size_t)d.rowTiles * d.ldx * 3;
        d.rowpart = reinterpret_cast<float*>(rowF);
        rowF += (size_t)d.rowTiles * d.chunks * 4 * kAdjRows;
        d.rtmom = reinterpret_cast<double*>(momD);
        momD += (size_t)d.rowTiles * kAdjMomW;
    }
    if (tiles > 0x7fffffffLL || cols > 0x7fffffffLL)
        return set_error("unrolled_bwd: batch too large"), EMSA_EINVAL;
    if ((rc = w.xpad.reserve(sizeof(float) * xpadF)) != EMSA_OK) return rc;
    if ((rc = w.colpart.reserve(sizeof(float) * colF)) != EMSA_OK) return rc;
    if ((rc = w.rowpart.reserve(sizeof(float) * rowF)) != EMSA_OK) return rc;
    if ((rc = w.rtmom.reserve(sizeof(double) * momD)) != EMSA_OK) return rc;
    if ((rc = w.adj.reserve(sizeof(double) * kAdjStW * count)) != EMSA_OK) return rc;
    if ((rc = w.coef.reserve(sizeof(double) * 16 * count)) != EMSA_OK) return rc;
    if ((rc = w.desc.reserve(sizeof(AdjDesc) * count)) != EMSA_OK) return rc;
    for (int k = 0; k < count; ++k) {
        AdjDesc& d = w.regs[k];
        d.Xpad = w.xpad.as<float>() + reinterpret_cast<size_t>(d.Xpad);
        d.colpart = w.colpart.as<float>() + reinterpret_cast<size_t>(d.colpart);
        d.rowpart = w.rowpart.as<float>() + reinterpret_cast<size_t>(d.rowpart);
        d.rtmom = w.rtmom.as<double>() + reinterpret_cast<size_t>(d.rtmom);
        d.adj = w.adj.as<double>() + (size_t)k * kAdjStW;
        d.coef = w.coef.as<double>() + (size_t)k * 16;
    }
    EMSA_HIP_TRY(hipMemcpyAsync(w.desc.ptr, w.regs.data(), sizeof(AdjDesc) * count,
                                hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    const AdjDesc* desc = w.desc.as<AdjDesc>();
    hipLaunchKernelGGL(adj_setup_kernel, dim3(count), dim3(256), 0, s, desc, dgRt);
    EMSA_HIP_TRY(hipGetLastError());
    const std::vector<float> sig = unrolled_sigmas(sigma2_0, factor, steps);
    for (int k = steps - 1; k >= 0; --k) {
        const double s2 = (double)sig[k];
        const float ka = (float)std::sqrt(kLog2e / (2.0 * s2));  // as the forward's
        const double inv_ka = 1.0 / (double)ka;
        const float minv = (float)(-1.0 / s2);
        const double out = (double)expf(-d_02 / sig[k]);
        const double outlier_dsig = out * (double)d_02 / (s2 * s2);  // d exp(-d02/s2) / d s2
        // the pairs kernel accumulates sum h s = -(1/sigma^2) sum ge e s, and
        // dL/dsigma^2 gets sum ge e s / (2 sigma^4) from the affinities
        const double gsig_scale = -1.0 / (2.0 * s2);
        const double fk = std::pow((double)factor, k);
        if (k == steps - 1)  // later steps: the previous columns kernel runs it
            hipLaunchKernelGGL(adj_solve_kernel, dim3(count), dim3(64), 0, s, desc, k);
        if (sig[k] < kCompSigma2)
            hipLaunchKernelGGL(adj_pairs_kernel<true>, dim3((unsigned)tiles),
                               dim3(kWave * kAdjWaves), 0, s, desc, count, k, ka, minv);
        else
            hipLaunchKernelGGL(adj_pairs_kernel<false>, dim3((unsigned)tiles),
                               dim3(kWave * kAdjWaves), 0, s, desc, count, k, ka, minv);
        hipLaunchKernelGGL(adj_rows_kernel, dim3((unsigned)rowTiles), dim3(kAdjRows), 0, s, desc,
                           count, k, gsig_scale, outlier_dsig, inv_ka);
        const unsigned colBlocks = (unsigned)((cols + 255) / 256);  // then one per registration
        hipLaunchKernelGGL(adj_cols_kernel, dim3(colBlocks + (unsigned)count), dim3(256), 0, s,
                           desc, count, cols, k, fk, inv_ka, k - 1, (int)colBlocks);
        EMSA_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(adj_output_kernel, dim3(count), dim3(64), 0, s, desc, dgRt0, dgsigma);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // extern "C"
