// Softassign on gfx950: annealing x inner iterations x Sinkhorn sweeps.
//
// Reference semantics (src/softassign.cu:486-822), rows r over the target X
// (N), columns c over the source Y (M):
//   M_rc = exp(-(|x_r - (R y_c + t)|^2 - alpha) / T) / sqrt(T)     :106-131
//   I1 times: a_r = sum_c M_rc + o_row[r]; M /= a_r; o_row /= a_r   :538-561
//             b_c = sum_r M_rc + o_col[c]; M /= b_c; o_col /= b_c   :579-597
//   s_r = sum_c M_rc, Sigma = sum s_r, X~ = s.*X, Y~ = b.*Y,        :626-666
//   Xc = sum X~ / Sigma, Yc = sum Y~ / Sigma (both over Sigma),     :687-713
//   X^ = X~ - Xc, Y^ = Y~ - Yc, D = M Y^, S = X^T D, Horn.          :725-787
// o_row and o_col persist across the whole run (:385-390); b_c is the last
// column sum including o_col.
//
// Both variants keep the current matrix as M_rc = K_rc * u_r * v_c and never
// rewrite it: a Sinkhorn half-sweep is one read pass that produces the row
// (or column) sums and updates the scale vector.
//  * materialised: K (FP32, N x ldk, row-major) is written once per inner
//    iteration by the first row pass; every later pass reads it once
//    (4 bytes per element per pass, 4 * (1 + 2*I1) bytes per element per
//    inner iteration). u, v are linear scales.
//  * log-domain: K is never stored; every pass recomputes
//    exp2(log2 K_rc + g_c) from the points, with f_r = log2 u_r and
//    g_c = log2 v_c kept as potentials (one exp2 per element per pass, no
//    HBM traffic for the matrix).
// S is assembled from FP64 moments of the final pass:
//   D_r = D'_r - s_r Yc with D'_r = sum_c M_rc b_c y_c, and
//   S_ij = sum s x_i D'_j - Yc_j sum s^2 x_i - Xc_i sum D'_j + Xc_i Yc_j Sigma.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>

namespace emsa {

namespace {

constexpr int kSaRW = 4;              // rows per wave in row passes
constexpr int kSaWaves = 4;           // waves per workgroup
constexpr int kSaRowsPB = kSaRW * kSaWaves;  // rows per row-pass workgroup
constexpr int kSaColsPB = 256;        // columns per column-pass workgroup (64 lanes x 4)
constexpr int kSaMomW = 22;           // moments (see sa_rows_kernel)
constexpr float kSaSentinel = 1.0e15f;

struct SaPlan {
    int ldk = 0;        // padded column count (multiple of 256)
    int rowBlocks = 0;  // ceil(N / kSaRowsPB)
    int G = 0;          // column chunks per row block
    int chunkCols = 0;  // columns per chunk (multiple of 256)
    int colBlocks = 0;  // ldk / 256
    int H = 0;          // row chunks per column block
    int chunkRows = 0;  // rows per chunk (multiple of kSaWaves)
};

SaPlan plan_sa(int N, int M, int numCU) {
    SaPlan p;
    p.ldk = (M + kSaColsPB - 1) / kSaColsPB * kSaColsPB;
    p.rowBlocks = (N + kSaRowsPB - 1) / kSaRowsPB;
    p.colBlocks = p.ldk / kSaColsPB;
    const int target = std::max(1, numCU) * 4;
    int G = std::max(1, std::min((target + p.rowBlocks - 1) / p.rowBlocks, p.colBlocks));
    const int cpc = (p.colBlocks + G - 1) / G;  // 256-column groups per chunk
    p.chunkCols = cpc * kSaColsPB;
    p.G = (p.ldk + p.chunkCols - 1) / p.chunkCols;
    int H = std::max(1, (target + p.colBlocks - 1) / p.colBlocks);
    H = std::min(H, (N + kSaWaves - 1) / kSaWaves);
    p.chunkRows = ((N + H - 1) / H + kSaWaves - 1) / kSaWaves * kSaWaves;
    p.H = (N + p.chunkRows - 1) / p.chunkRows;
    return p;
}

struct SaBufs {
    const float* X;   // [3][N]
    const float* Y;   // [3][M]
    float* P;         // [3][ldk] transformed source points R y + t
    float* K;         // [N][ldk] (materialised only)
    float* u;         // [N] row scale (log: f = log2 u)
    float* v;         // [ldk] column scale (log: g = log2 v)
    float* orow;      // [N]
    float* ocol;      // [M]
    float* bcol;      // [ldk] last column sums b_c
    float* rpart;     // [rowBlocks*G][4][kSaRowsPB]
    float* cpart;     // [H][ldk]
    double* mom;      // [rowBlocks][kSaMomW]
    double* cmom;     // [colBlocks][3]
    unsigned* rcnt;   // [rowBlocks + 1]
    unsigned* ccnt;   // [colBlocks + 1]
    double* state;    // kStateW
    double* trace;    // nullable
};

// Transform the source points with the state's (R, t) into P; columns >= M
// get the sentinel (their affinities underflow to exactly 0).
__device__ inline void sa_transform(const SaBufs& b, int M, int ldk, const double* R,
                                    const double* t, int tid, int nthreads) {
    for (int c = tid; c < ldk; c += nthreads) {
        float px = kSaSentinel, py = kSaSentinel, pz = kSaSentinel;
        if (c < M) {
            const double y0 = b.Y[c], y1 = b.Y[(size_t)M + c], y2 = b.Y[2 * (size_t)M + c];
            px = (float)(R[0] * y0 + R[1] * y1 + R[2] * y2 + t[0]);
            py = (float)(R[3] * y0 + R[4] * y1 + R[5] * y2 + t[1]);
            pz = (float)(R[6] * y0 + R[7] * y1 + R[8] * y2 + t[2]);
        }
        b.P[c] = px;
        b.P[(size_t)ldk + c] = py;
        b.P[2 * (size_t)ldk + c] = pz;
    }
}

__global__ void __launch_bounds__(1024)
sa_setup_kernel(SaBufs b, int N, int M, int ldk, const float* __restrict__ Rt, float moutlier) {
    double R[9], t[3];
    for (int i = 0; i < 9; ++i) R[i] = Rt[i];
    for (int i = 0; i < 3; ++i) t[i] = Rt[9 + i];
    if (threadIdx.x < 12) b.state[threadIdx.x] = (double)Rt[threadIdx.x];
    if (threadIdx.x == 12) b.state[kStStatus] = 0.0;
    for (int r = threadIdx.x; r < N; r += blockDim.x) b.orow[r] = moutlier;
    for (int c = threadIdx.x; c < M; c += blockDim.x) b.ocol[c] = moutlier;
    sa_transform(b, M, ldk, R, t, threadIdx.x, blockDim.x);
}

__global__ void sa_output_kernel(const double* __restrict__ state, float* __restrict__ Rt) {
    if (threadIdx.x < 12) Rt[threadIdx.x] = (float)state[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Row passes. MODE 0: first row pass of an inner iteration (materialised: also
// writes K); MODE 1: Sinkhorn row pass; MODE 2: final pass (s_r, D'_r,
// moments, Horn, next P). A workgroup = kSaRowsPB rows x one column chunk,
// lanes along columns (4 consecutive columns per lane).
// ---------------------------------------------------------------------------
template <bool LOG, int MODE>
__global__ void __launch_bounds__(kWave * kSaWaves)
sa_rows_kernel(SaBufs b, int N, int M, int ldk, int G, int chunkCols, int colBlocks, float k1,
               float k0, float T, int iter) {
    constexpr int Q = (MODE == 2) ? 4 : 1;
    const int rowBlocks = gridDim.x / G;
    const int rb = blockIdx.x / G;
    const int g = blockIdx.x - rb * G;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r0 = rb * kSaRowsPB + wave * kSaRW;
    const int cb = g * chunkCols;
    const int ce = min(cb + chunkCols, ldk);

    float xr[kSaRW][3];
#pragma unroll
    for (int j = 0; j < kSaRW; ++j) {
        const int r = min(r0 + j, N - 1);
        xr[j][0] = b.X[r];
        xr[j][1] = b.X[(size_t)N + r];
        xr[j][2] = b.X[2 * (size_t)N + r];
    }
    float acc[kSaRW][Q];
#pragma unroll
    for (int j = 0; j < kSaRW; ++j)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[j][q] = 0.f;

    for (int c = cb + 4 * lane; c < ce; c += 4 * kWave) {
        float w[4][Q];  // per-column weights
        float p[3][4];
        float kg[4];
        if (LOG || MODE == 0) {
            const float4 P0 = *reinterpret_cast<const float4*>(b.P + c);
            const float4 P1 = *reinterpret_cast<const float4*>(b.P + (size_t)ldk + c);
            const float4 P2 = *reinterpret_cast<const float4*>(b.P + 2 * (size_t)ldk + c);
            p[0][0] = P0.x; p[0][1] = P0.y; p[0][2] = P0.z; p[0][3] = P0.w;
            p[1][0] = P1.x; p[1][1] = P1.y; p[1][2] = P1.z; p[1][3] = P1.w;
            p[2][0] = P2.x; p[2][1] = P2.y; p[2][2] = P2.z; p[2][3] = P2.w;
        }
        // column scale of a fresh inner iteration: v = 1 (log: g = 0)
        float vv[4] = {LOG ? 0.f : 1.f, LOG ? 0.f : 1.f, LOG ? 0.f : 1.f, LOG ? 0.f : 1.f};
        if (MODE != 0) {
            const float4 V = *reinterpret_cast<const float4*>(b.v + c);
            vv[0] = V.x; vv[1] = V.y; vv[2] = V.z; vv[3] = V.w;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) kg[i] = LOG ? (k0 + vv[i]) : k0;  // log: v holds g_c
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i][0] = LOG ? 1.f : vv[i];
        if (MODE == 2) {
            const float4 Bc = *reinterpret_cast<const float4*>(b.bcol + c);
            const float bb[4] = {Bc.x, Bc.y, Bc.z, Bc.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int cc = min(c + i, M - 1);
                const float base = LOG ? bb[i] : vv[i] * bb[i];
                const bool in = (c + i) < M;
                w[i][1 % Q] = in ? base * b.Y[cc] : 0.f;
                w[i][2 % Q] = in ? base * b.Y[(size_t)M + cc] : 0.f;
                w[i][3 % Q] = in ? base * b.Y[2 * (size_t)M + cc] : 0.f;
            }
        }
#pragma unroll
        for (int j = 0; j < kSaRW; ++j) {
            const int r = r0 + j;
            if (r >= N) continue;  // wave-uniform
            float val[4];
            if (!LOG && MODE != 0) {
                const float4 Kv = *reinterpret_cast<const float4*>(b.K + (size_t)r * ldk + c);
                val[0] = Kv.x; val[1] = Kv.y; val[2] = Kv.z; val[3] = Kv.w;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float dx = xr[j][0] - p[0][i];
                    const float dy = xr[j][1] - p[1][i];
                    const float dz = xr[j][2] - p[2][i];
                    float s = dx * dx;
                    s = fmaf(dy, dy, s);
                    s = fmaf(dz, dz, s);
                    val[i] = __builtin_amdgcn_exp2f(fmaf(s, -k1, kg[i]));
                }
                if (!LOG)  // MODE 0 materialised: store K
                    *reinterpret_cast<float4*>(b.K + (size_t)r * ldk + c) =
                        make_float4(val[0], val[1], val[2], val[3]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[j][q] = fmaf(val[i], w[i][q], acc[j][q]);
        }
    }

    // wave sums per row -> partials
    __shared__ unsigned s_ticket;
    float* Pp = b.rpart + (size_t)blockIdx.x * 4 * kSaRowsPB;
#pragma unroll
    for (int j = 0; j < kSaRW; ++j)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float s = wave_sum(acc[j][q]);
            if (lane == 0) Pp[q * kSaRowsPB + wave * kSaRW + j] = s;
        }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_ticket = atomicAdd(&b.rcnt[rb], 1u);
    __syncthreads();
    if (s_ticket != (unsigned)(G - 1)) return;
    __threadfence();

    const int tid = threadIdx.x;
    const int r = rb * kSaRowsPB + tid;
    if (tid == 0) b.rcnt[rb] = 0u;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    if (tid < kSaRowsPB) {
        const float* Pr = b.rpart + (size_t)rb * G * 4 * kSaRowsPB;
        for (int k = 0; k < G; ++k)
#pragma unroll
            for (int q = 0; q < Q; ++q) tot[q] += Pr[(size_t)k * 4 * kSaRowsPB + q * kSaRowsPB + tid];
    }
    if (MODE != 2) {
        if (tid < kSaRowsPB && r < N) {
            // a_r = (row scale) * sum + o_row[r]; scale /= a_r; o_row /= a_r
            const float su = (MODE == 0) ? (LOG ? 0.f : 1.f) : b.u[r];
            const float scale = LOG ? exp2f(su) : su;
            const float a = (float)(scale * tot[0]) + b.orow[r];
            b.orow[r] = b.orow[r] / a;
            b.u[r] = LOG ? (su - log2f(a)) : (su / a);
        }
        return;
    }

    // ---- MODE 2: per-row moments ----
    double m[kSaMomW];
#pragma unroll
    for (int k = 0; k < kSaMomW; ++k) m[k] = 0.0;
    if (tid < kSaRowsPB && r < N) {
        const float su = b.u[r];
        const double scale = LOG ? (double)exp2f(su) : (double)su;
        const double s = scale * tot[0];
        const double D[3] = {scale * tot[1], scale * tot[2], scale * tot[3]};
        const double x[3] = {b.X[r], b.X[(size_t)N + r], b.X[2 * (size_t)N + r]};
        m[0] = s;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            m[1 + i] = s * x[i];
            m[13 + i] = s * s * x[i];
            m[16 + i] = D[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) m[4 + 3 * i + j] = s * x[i] * D[j];
        }
    }
    __shared__ double lds[kSaWaves * kSaMomW];
    double total;
    block_sum<kSaMomW>(m, lds, kSaWaves, total);
    if (tid < kSaMomW) b.mom[(size_t)rb * kSaMomW + tid] = total;
    __threadfence();
    __syncthreads();
    if (tid == 0) s_ticket = atomicAdd(&b.rcnt[rowBlocks], 1u);
    __syncthreads();
    if (s_ticket != (unsigned)(rowBlocks - 1)) return;
    __threadfence();

    // ---- last row block: totals, S, Horn, next P ----
    __shared__ double mt[kSaMomW + 3];
    __shared__ double newRt[12];
    if (tid < kSaMomW) {
        double s = 0.0;
        for (int k = 0; k < rowBlocks; ++k) s += b.mom[(size_t)k * kSaMomW + tid];
        mt[tid] = s;
    } else if (tid >= 64 && tid < 67) {
        double s = 0.0;
        for (int k = 0; k < colBlocks; ++k) s += b.cmom[(size_t)k * 3 + (tid - 64)];
        mt[kSaMomW + tid - 64] = s;
    }
    __syncthreads();
    if (tid == 0) {
        b.rcnt[rowBlocks] = 0u;
        const double Sg = mt[0];
        double Xc[3], Yc[3], S[9], Rn[9], tn[3];
        for (int i = 0; i < 3; ++i) {
            Xc[i] = mt[1 + i] / Sg;
            Yc[i] = mt[kSaMomW + i] / Sg;
        }
        // S[i + 3j], i over X (target), j over Y (source): softassign.cu:769-772
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                S[i + 3 * j] = mt[4 + 3 * i + j] - Yc[j] * mt[13 + i] - Xc[i] * mt[16 + j] +
                               Xc[i] * Yc[j] * Sg;
        const bool ok = horn_rt(Xc, Yc, S, Rn, tn);
        if (ok) {
            for (int i = 0; i < 9; ++i) b.state[kStR + i] = Rn[i];
            for (int i = 0; i < 3; ++i) b.state[kStT + i] = tn[i];
        } else {
            b.state[kStStatus] += 1.0;
            for (int i = 0; i < 9; ++i) Rn[i] = b.state[kStR + i];
            for (int i = 0; i < 3; ++i) tn[i] = b.state[kStT + i];
        }
        for (int i = 0; i < 9; ++i) newRt[i] = Rn[i];
        for (int i = 0; i < 3; ++i) newRt[9 + i] = tn[i];
        if (b.trace) {
            double* o = b.trace + (size_t)iter * EMSA_TRACE_W;
            o[0] = T;
            o[1] = Sg;
            for (int i = 0; i < 3; ++i) {
                o[2 + i] = Xc[i];
                o[5 + i] = Yc[i];
            }
            for (int i = 0; i < 9; ++i) o[8 + i] = S[i];
            for (int i = 0; i < 9; ++i) o[17 + i] = Rn[i];
            for (int i = 0; i < 3; ++i) o[26 + i] = tn[i];
            o[29] = 0.0;
            o[30] = 0.0;
            o[31] = ok ? 0.0 : 1.0;
        }
    }
    __syncthreads();
    sa_transform(b, M, ldk, newRt, newRt + 9, tid, blockDim.x);
}

// ---------------------------------------------------------------------------
// Column pass: a workgroup = 256 columns x one row chunk; lanes own 4
// consecutive columns, the 4 waves split the chunk's rows. FIRST: scales of
// a fresh inner iteration (u from the first row pass, v = 1).
// ---------------------------------------------------------------------------
template <bool LOG, bool FIRST>
__global__ void __launch_bounds__(kWave * kSaWaves)
sa_cols_kernel(SaBufs b, int N, int M, int ldk, int H, int chunkRows, float k1, float k0) {
    const int colBlocks = gridDim.x / H;
    const int cbk = blockIdx.x / H;
    const int h = blockIdx.x - cbk * H;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int c = cbk * kSaColsPB + 4 * lane;
    const int q = chunkRows / kSaWaves;
    const int rs = h * chunkRows + wave * q;
    const int re = min(rs + q, N);

    float p[3][4];
    if (LOG) {
        const float4 P0 = *reinterpret_cast<const float4*>(b.P + c);
        const float4 P1 = *reinterpret_cast<const float4*>(b.P + (size_t)ldk + c);
        const float4 P2 = *reinterpret_cast<const float4*>(b.P + 2 * (size_t)ldk + c);
        p[0][0] = P0.x; p[0][1] = P0.y; p[0][2] = P0.z; p[0][3] = P0.w;
        p[1][0] = P1.x; p[1][1] = P1.y; p[1][2] = P1.z; p[1][3] = P1.w;
        p[2][0] = P2.x; p[2][1] = P2.y; p[2][2] = P2.z; p[2][3] = P2.w;
    }
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int r = rs; r < re; ++r) {
        const float ur = b.u[r];  // log: f_r
        if (LOG) {
            const float x0 = b.X[r], x1 = b.X[(size_t)N + r], x2 = b.X[2 * (size_t)N + r];
            const float kf = k0 + ur;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float dx = x0 - p[0][i], dy = x1 - p[1][i], dz = x2 - p[2][i];
                float s = dx * dx;
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                acc[i] += __builtin_amdgcn_exp2f(fmaf(s, -k1, kf));
            }
        } else {
            const float4 Kv = *reinterpret_cast<const float4*>(b.K + (size_t)r * ldk + c);
            acc[0] = fmaf(Kv.x, ur, acc[0]);
            acc[1] = fmaf(Kv.y, ur, acc[1]);
            acc[2] = fmaf(Kv.z, ur, acc[2]);
            acc[3] = fmaf(Kv.w, ur, acc[3]);
        }
    }
    __shared__ float red[kSaWaves][kSaColsPB];
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][4 * lane + i] = acc[i];
    __syncthreads();
    __shared__ unsigned s_ticket;
    {
        const int tid = threadIdx.x;
        float s = red[0][tid];
#pragma unroll
        for (int w = 1; w < kSaWaves; ++w) s += red[w][tid];
        b.cpart[(size_t)h * ldk + cbk * kSaColsPB + tid] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_ticket = atomicAdd(&b.ccnt[cbk], 1u);
    __syncthreads();
    if (s_ticket != (unsigned)(H - 1)) return;
    __threadfence();
    const int tid = threadIdx.x;
    if (tid == 0) b.ccnt[cbk] = 0u;
    const int cc = cbk * kSaColsPB + tid;
    double tot = 0.0;
    for (int k = 0; k < H; ++k) tot += b.cpart[(size_t)k * ldk + cc];
    double by[3] = {0.0, 0.0, 0.0};
    if (cc < M) {
        const float sv = FIRST ? (LOG ? 0.f : 1.f) : b.v[cc];
        const float scale = LOG ? exp2f(sv) : sv;
        const float bc = (float)(scale * tot) + b.ocol[cc];
        b.ocol[cc] = b.ocol[cc] / bc;
        b.v[cc] = LOG ? (sv - log2f(bc)) : (sv / bc);
        b.bcol[cc] = bc;
        by[0] = (double)bc * b.Y[cc];
        by[1] = (double)bc * b.Y[(size_t)M + cc];
        by[2] = (double)bc * b.Y[2 * (size_t)M + cc];
    } else {
        b.v[cc] = LOG ? 0.f : 0.f;
        b.bcol[cc] = 0.f;
    }
    __shared__ double lds[kSaWaves * 3];
    double total;
    block_sum<3>(by, lds, kSaWaves, total);
    if (tid < 3) b.cmom[(size_t)cbk * 3 + tid] = total;
    (void)colBlocks;
}

}  // namespace

int softassign_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                      const float* dRtIn, float* dRtOut, const emsa_params& p, int variant,
                      hipStream_t s, double* trace) {
    const bool LOGV = (variant == EMSA_SOFTASSIGN_LOGDOMAIN);
    const SaPlan pl = plan_sa(N, M, ctx.numCU);
    int rc;
    SaBufs b{};
    b.X = dX;
    b.Y = dY;
    if ((rc = ctx.u.reserve(sizeof(float) * 3 * (size_t)pl.ldk)) != EMSA_OK) return rc;  // P
    b.P = ctx.u.as<float>();
    if (!LOGV) {
        if ((rc = ctx.K.reserve(sizeof(float) * (size_t)N * pl.ldk)) != EMSA_OK) return rc;
        b.K = ctx.K.as<float>();
    }
    // u, v, bcol each start on a 256-byte boundary (v and bcol are read as float4)
    const size_t npad = ((size_t)N + 63) / 64 * 64;
    if ((rc = ctx.v.reserve(sizeof(float) * (npad + 2 * (size_t)pl.ldk))) != EMSA_OK) return rc;
    b.u = ctx.v.as<float>();
    b.v = b.u + npad;
    b.bcol = b.v + pl.ldk;
    if ((rc = ctx.orow.reserve(sizeof(float) * ((size_t)N + M))) != EMSA_OK) return rc;
    b.orow = ctx.orow.as<float>();
    b.ocol = b.orow + N;
    if ((rc = ctx.rowpart.reserve(sizeof(float) * 4 * kSaRowsPB * (size_t)pl.rowBlocks * pl.G)) !=
        EMSA_OK)
        return rc;
    b.rpart = ctx.rowpart.as<float>();
    if ((rc = ctx.colpart.reserve(sizeof(float) * (size_t)pl.H * pl.ldk)) != EMSA_OK) return rc;
    b.cpart = ctx.colpart.as<float>();
    if ((rc = ctx.samom.reserve(sizeof(double) *
                                ((size_t)pl.rowBlocks * kSaMomW + (size_t)pl.colBlocks * 3))) !=
        EMSA_OK)
        return rc;
    b.mom = ctx.samom.as<double>();
    b.cmom = b.mom + (size_t)pl.rowBlocks * kSaMomW;
    const size_t oldc = ctx.sacnt.bytes;
    if ((rc = ctx.sacnt.reserve(sizeof(unsigned) * ((size_t)pl.rowBlocks + pl.colBlocks + 2))) !=
        EMSA_OK)
        return rc;
    if (ctx.sacnt.bytes != oldc)
        EMSA_HIP_TRY(hipMemsetAsync(ctx.sacnt.ptr, 0, ctx.sacnt.bytes, s));
    b.rcnt = ctx.sacnt.as<unsigned>();
    b.ccnt = b.rcnt + pl.rowBlocks + 1;
    if ((rc = ctx.sastate.reserve(sizeof(double) * kStateW)) != EMSA_OK) return rc;
    b.state = ctx.sastate.as<double>();
    b.trace = trace;

    hipLaunchKernelGGL(sa_setup_kernel, dim3(1), dim3(1024), 0, s, b, N, M, pl.ldk, dRtIn,
                       p.moutlier);
    EMSA_HIP_TRY(hipGetLastError());
    const dim3 rgrid((unsigned)(pl.rowBlocks * pl.G)), cgrid((unsigned)(pl.colBlocks * pl.H));
    const dim3 blk(kWave * kSaWaves);
    float T = p.T_0;
    int it = 0;
    for (int Titer = 1; Titer <= p.JMAX; ++Titer) {
        const double k1d = kLog2e / (double)T;
        const float k1 = (float)k1d;
        const float k0 = (float)((double)p.alpha * k1d - 0.5 * std::log2((double)T));
        for (int i0 = 0; i0 < p.I0; ++i0, ++it) {
            for (int sw = 0; sw < p.I1; ++sw) {
                if (sw == 0) {
                    if (LOGV)
                        hipLaunchKernelGGL((sa_rows_kernel<true, 0>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
                    else
                        hipLaunchKernelGGL((sa_rows_kernel<false, 0>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
                } else {
                    if (LOGV)
                        hipLaunchKernelGGL((sa_rows_kernel<true, 1>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
                    else
                        hipLaunchKernelGGL((sa_rows_kernel<false, 1>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
                }
                EMSA_HIP_TRY(hipGetLastError());
                if (LOGV) {
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_cols_kernel<true, true>), cgrid, blk, 0, s, b, N, M,
                                           pl.ldk, pl.H, pl.chunkRows, k1, k0);
                    else
                        hipLaunchKernelGGL((sa_cols_kernel<true, false>), cgrid, blk, 0, s, b, N,
                                           M, pl.ldk, pl.H, pl.chunkRows, k1, k0);
                } else {
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_cols_kernel<false, true>), cgrid, blk, 0, s, b, N,
                                           M, pl.ldk, pl.H, pl.chunkRows, k1, k0);
                    else
                        hipLaunchKernelGGL((sa_cols_kernel<false, false>), cgrid, blk, 0, s, b, N,
                                           M, pl.ldk, pl.H, pl.chunkRows, k1, k0);
                }
                EMSA_HIP_TRY(hipGetLastError());
            }
            if (LOGV)
                hipLaunchKernelGGL((sa_rows_kernel<true, 2>), rgrid, blk, 0, s, b, N, M, pl.ldk,
                                   pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
            else
                hipLaunchKernelGGL((sa_rows_kernel<false, 2>), rgrid, blk, 0, s, b, N, M, pl.ldk,
                                   pl.G, pl.chunkCols, pl.colBlocks, k1, k0, T, it);
            EMSA_HIP_TRY(hipGetLastError());
        }
        T = T * p.TFACTOR;  // softassign.cu:820, float
    }
    hipLaunchKernelGGL(sa_output_kernel, dim3(1), dim3(64), 0, s, b.state, dRtOut);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // namespace emsa
