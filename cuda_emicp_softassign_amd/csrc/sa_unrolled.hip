// Differentiable unrolled Softassign, batched over B registrations of equal
// size (no counterpart in the reference; it differentiates the annealing loop
// of src/softassign.cu:486-822, one inner iteration = :494-818).
//
// The matrix is never stored: every pass recomputes
//   K_rc = exp(-(|x_r - P_c|^2 - alpha) / T) / sqrt(T) = exp2(k0 - k1 s),
// P_c = R y_c + t, and the current matrix is K_rc u_r v_c with linear scales
// (u^0 = v^0 = 1 at the start of every inner iteration). Sweep k:
//   A_r = u_r sum_c K_rc v_c + o_row_r;  u_r /= A_r;  o_row_r /= A_r
//   B_c = v_c sum_r K_rc u_r + o_col_c;  v_c /= B_c;  o_col_c /= B_c
// then s_r = sum_c M_rc, D'_r = sum_c M_rc B_c y_c and the moments of S
// (the same assembly as the plain Softassign), Horn on the device.
//
// Forward checkpoints per registration and inner iteration: the step record,
// per sweep A^k, the row mass u^{k-1} sum_c K v^{k-1}, B^k and the column
// mass, the slack vectors at the iteration start, and s_r, D'_r.
//
// Backward per inner iteration (reverse), with the gradient of (R', t') and
// of the slack vectors carried from the later iteration:
//  1. Horn backwards (as the EM-ICP adjoint) and the Softassign moment map
//     -> dL/d(Sigma, sum s x, sum s x D'^T, sum s^2 x, sum D', sum B y);
//  2. per row: dL/ds_r = alpha_r, dL/dD'_r = beta_r, dL/dU_r;
//  3. a column pass sum_r K_rc U_r (alpha_r, beta_r) gives dL/dV_c, dL/dB_c;
//  4. the sweeps backwards: sweep k's column half needs sum_c K_rc z_c (a row
//     pass), its row half sum_r K_rc z_r (a column pass);
//  5. every term above adds a rank-1 piece L_r Q_c to dL/dK; one pass over the
//     pairs recomputes K and accumulates dL/dx_r, dL/dP_c and dL/dT from
//     G = (sum_t L_r^t Q_c^t) K (4 + 2 I1 terms);
//  6. dL/dy_c = R^T dL/dP_c + direct terms, dL/d(R, t) of the iteration's
//     input, dL/dT_0 += TFACTOR^j dL/dT_j.
// All reductions have a fixed order; no workgroup waits on another.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace emsa {

namespace {

constexpr int kRW = 4;                 // rows per wave in row passes
constexpr int kWG = 4;                 // waves per workgroup
constexpr int kRowsPB = kRW * kWG;     // rows per row-pass workgroup
constexpr int kColsPB = 256;           // columns per column-pass workgroup
constexpr int kMomW = 22;              // per-wave moments of the final row pass
constexpr int kRecW = 40;              // record: trace layout + m13 (32..34), m16 (35..37)
constexpr int kMaxI1 = 4;
constexpr int kMaxT = 4 + 2 * kMaxI1;  // rank-1 terms of dL/dK
constexpr int kCoefW = 22;             // moment coefficients
constexpr int kAdjW = 48;              // per registration: gR(9) gt(3) gT0(1) coef(22 at 16)
constexpr int kFRPL = 4;               // final pass: rows per lane
constexpr int kFRows = kFRPL * kWave;  // 256 rows per tile
constexpr int kFCols = 1024;           // columns per final-pass workgroup (4 waves x 256)
constexpr int kStageW = 3 * kColGroup;
constexpr int kStageLd = kWave + 1;
constexpr float kSentinel = 1.0e15f;
// The row and column passes evaluate the exponent expanded
// (a_r + b_c + x'_r . p_c) at the temperatures where k1 max |x|^2 over the
// batch's target clouds is at most this (the expanded terms then round to
// about 1e-5 of the exponent at worst), the direct difference below.
constexpr float kExpandMax = 64.f;
typedef const __attribute__((address_space(4))) float* cfloat_p;

struct Sau {
    int B, N, M, ldk, H, chunkRows, I1, nIt, rowWaves, colBlocks, ckF, rowTiles, fchunks;
    float alpha;
    const float* X;    // [B][3][N]
    const float* Y;    // [B][3][M]
    const float* Rt0;  // [B][12]
    float* P;          // [B][3][ldk]
    double* state;     // [B][kStateW]
    float* u;          // [B][N]
    float* v;          // [B][ldk]
    float* orow;       // [B][N]
    float* ocol;       // [B][ldk]
    float* bcol;       // [B][ldk]
    float* wfin;       // [B][4][ldk] final-pass column weights V, V B y
    float* cpart;      // [B][H][4][ldk]
    double* rmom;      // [B][rowWaves][kMomW]
    double* cby;       // [B][colBlocks][3]
    double* rec;       // [B][nIt][kRecW]
    float* ck;         // [B][nIt][ckF]
    // backward
    float* gU;         // [B][N]
    float* gV;         // [B][ldk]
    float* gOr;        // [B][N]
    float* gOc;        // [B][ldk]
    float* L;          // [B][kMaxT][N]
    float* Q;          // [B][kMaxT][ldk]
    float* zr;         // [B][4][N]
    float* zc;         // [B][ldk]
    float* gX;         // [B][3][N]
    float* gY;         // [B][3][M]
    double* adj;       // [B][kAdjW]
    float* fcol;       // [B][rowTiles][ldk][3]
    float* frow;       // [B][rowTiles * fchunks][5][kFRows]
    double* fmomr;     // [B][rowTiles][2]
    double* fmomc;     // [B][colBlocks][12]
};

// ---- checkpoint layout of one registration and inner iteration -----------
__host__ __device__ inline size_t ck_A(const Sau& d, int k) { return (size_t)(k - 1) * d.N; }
__host__ __device__ inline size_t ck_rm(const Sau& d, int k) {
    return (size_t)d.I1 * d.N + (size_t)(k - 1) * d.N;
}
__host__ __device__ inline size_t ck_or0(const Sau& d) { return 2 * (size_t)d.I1 * d.N; }
__host__ __device__ inline size_t ck_sD(const Sau& d) { return ck_or0(d) + d.N; }  // 4N: s, D'
__host__ __device__ inline size_t ck_B(const Sau& d, int k) {
    return ck_sD(d) + 4 * (size_t)d.N + (size_t)(k - 1) * d.M;
}
__host__ __device__ inline size_t ck_cm(const Sau& d, int k) {
    return ck_sD(d) + 4 * (size_t)d.N + (size_t)d.I1 * d.M + (size_t)(k - 1) * d.M;
}
__host__ __device__ inline size_t ck_oc0(const Sau& d) {
    return ck_sD(d) + 4 * (size_t)d.N + 2 * (size_t)d.I1 * d.M;
}
__host__ __device__ inline size_t ck_floats(const Sau& d) { return ck_oc0(d) + d.M; }

__device__ inline const float* ckp(const Sau& d, int b, int it) {
    return d.ck + ((size_t)b * d.nIt + it) * d.ckF;
}
__device__ inline float* ckw(const Sau& d, int b, int it) {
    return d.ck + ((size_t)b * d.nIt + it) * d.ckF;
}

// term indices of dL/dK: 0..3 the final matrix, then per sweep k (I1 .. 1)
// the column half and the row half
__device__ inline int t_col(const Sau& d, int k) { return 4 + 2 * (d.I1 - k); }
__device__ inline int t_row(const Sau& d, int k) { return 5 + 2 * (d.I1 - k); }

// Input transform of inner iteration it (doubles).
__device__ inline void iter_input(const Sau& d, int b, int it, double* R, double* t) {
    if (it == 0) {
        const float* r0 = d.Rt0 + (size_t)b * 12;
        for (int i = 0; i < 9; ++i) R[i] = r0[i];
        for (int i = 0; i < 3; ++i) t[i] = r0[9 + i];
    } else {
        const double* o = d.rec + ((size_t)b * d.nIt + it - 1) * kRecW;
        for (int i = 0; i < 9; ++i) R[i] = o[17 + i];
        for (int i = 0; i < 3; ++i) t[i] = o[26 + i];
    }
}

// Row scale u^k and slack o_row^k of row r rebuilt from the checkpoint in the
// forward's float operation order (u^0 = 1, o_row^0 from the checkpoint).
__device__ inline void row_state(const float* ck, const Sau& d, int r, int k, float& u, float& o) {
    u = 1.f;
    o = ck[ck_or0(d) + r];
    for (int j = 1; j <= k; ++j) {
        const float A = ck[ck_A(d, j) + r];
        o = o / A;
        u = u / A;
    }
}
__device__ inline void col_state(const float* ck, const Sau& d, int c, int k, float& v, float& o) {
    v = 1.f;
    o = ck[ck_oc0(d) + c];
    for (int j = 1; j <= k; ++j) {
        const float Bk = ck[ck_B(d, j) + c];
        o = o / Bk;
        v = v / Bk;
    }
}

__device__ inline void load4(const float* p, float (&v)[4]) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}

// ---------------------------------------------------------------------------
// Setup and per-iteration transform
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) sau_init_kernel(Sau d, float moutlier) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.N) d.orow[(size_t)b * d.N + i] = moutlier;
    if (i < d.ldk) d.ocol[(size_t)b * d.ldk + i] = (i < d.M) ? moutlier : 0.f;
    if (blockIdx.x == 0 && threadIdx.x < kStateW) {
        double* st = d.state + (size_t)b * kStateW;
        st[threadIdx.x] = threadIdx.x < 12 ? (double)d.Rt0[(size_t)b * 12 + threadIdx.x] : 0.0;
    }
}

// P_c = R y_c + t for the input transform of iteration it.
__global__ void __launch_bounds__(256) sau_transform_kernel(Sau d, int it) {
    const int b = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ldk) return;
    double R[9], t[3];
    iter_input(d, b, it, R, t);
    const int M = d.M, ldk = d.ldk;
    float px = kSentinel, py = kSentinel, pz = kSentinel;
    if (c < M) {
        const float* Y = d.Y + (size_t)b * 3 * M;
        const double y0 = Y[c], y1 = Y[(size_t)M + c], y2 = Y[2 * (size_t)M + c];
        px = (float)(R[0] * y0 + R[1] * y1 + R[2] * y2 + t[0]);
        py = (float)(R[3] * y0 + R[4] * y1 + R[5] * y2 + t[1]);
        pz = (float)(R[6] * y0 + R[7] * y1 + R[8] * y2 + t[2]);
    }
    float* P = d.P + (size_t)b * 3 * ldk;
    P[c] = px;
    P[(size_t)ldk + c] = py;
    P[2 * (size_t)ldk + c] = pz;
}

// ---------------------------------------------------------------------------
// Row passes: a wave owns kRW whole rows, lanes take 4 consecutive columns.
// MODE 0: forward sweep k (weights v^{k-1}); MODE 1: forward final pass
// (weights V, V B y) with per-wave FP64 moments; MODE 2: backward of sweep k
// (weights z_c), then the row half of sweep k backwards.
// ---------------------------------------------------------------------------
enum { kRowFwd = 0, kRowFinal = 1, kRowBwd = 2 };

template <int MODE, bool EXP>
__global__ void __launch_bounds__(kWave * kWG)
sau_rowpass_kernel(Sau d, int k, int it, float k1, float k0) {
    constexpr int NW = (MODE == kRowFinal) ? 4 : 1;
    const int b = blockIdx.y;
    const int N = d.N, ldk = d.ldk;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int gw = blockIdx.x * kWG + wave;
    const int r0 = gw * kRW;
    const float* X = d.X + (size_t)b * 3 * N;
    const float* P = d.P + (size_t)b * 3 * ldk;
    const float* vb = d.v + (size_t)b * ldk;
    const float* wfb = d.wfin + (size_t)b * 4 * ldk;
    const float* zcb = d.zc + (size_t)b * ldk;
    const bool ones = (MODE == kRowFwd && k == 1);  // v^0 = 1

    float xr[kRW][3];
#pragma unroll
    for (int j = 0; j < kRW; ++j) {
        const int r = min(r0 + j, N - 1);
        xr[j][0] = X[r];
        xr[j][1] = X[(size_t)N + r];
        xr[j][2] = X[2 * (size_t)N + r];
    }
    // EXP: the exponent k0 - k1 |x_r - p_c|^2 as a_r + b_c + x'_r . p_c with
    // a_r = k0 - k1 |x_r|^2, x'_r = 2 k1 x_r, b_c = -k1 |p_c|^2 (one add and
    // 3 FMA per element)
    float ar[kRW], xq[kRW][3];
    if (EXP) {
#pragma unroll
        for (int j = 0; j < kRW; ++j) {
            ar[j] = fmaf(-k1,
                         fmaf(xr[j][0], xr[j][0], fmaf(xr[j][1], xr[j][1], xr[j][2] * xr[j][2])),
                         k0);
#pragma unroll
            for (int a = 0; a < 3; ++a) xq[j][a] = 2.f * k1 * xr[j][a];
        }
    }
    float acc[kRW][NW];
#pragma unroll
    for (int j = 0; j < kRW; ++j)
#pragma unroll
        for (int q = 0; q < NW; ++q) acc[j][q] = 0.f;

    for (int c = 4 * lane; c < ldk; c += 4 * kWave) {
        float p[3][4], w[4][NW], bc[4];
        load4(P + c, p[0]);
        load4(P + (size_t)ldk + c, p[1]);
        load4(P + 2 * (size_t)ldk + c, p[2]);
        if (EXP)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                bc[i] = -k1 * fmaf(p[0][i], p[0][i], fmaf(p[1][i], p[1][i], p[2][i] * p[2][i]));
        if (MODE == kRowFwd) {
            if (ones) {
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i][0] = 1.f;
            } else {
                float vv[4];
                load4(vb + c, vv);
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i][0] = vv[i];
            }
        } else if (MODE == kRowBwd) {
            float zz[4];
            load4(zcb + c, zz);
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i][0] = zz[i];
        } else {
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                float ww[4];
                load4(wfb + (size_t)q * ldk + c, ww);
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i][q] = ww[i];
            }
        }
#pragma unroll
        for (int j = 0; j < kRW; ++j) {
            if (r0 + j >= N) continue;  // wave-uniform
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float t;
                if (EXP) {
                    t = ar[j] + bc[i];
                    t = fmaf(xq[j][0], p[0][i], t);
                    t = fmaf(xq[j][1], p[1][i], t);
                    t = fmaf(xq[j][2], p[2][i], t);
                } else {
                    const float dx = xr[j][0] - p[0][i];
                    const float dy = xr[j][1] - p[1][i];
                    const float dz = xr[j][2] - p[2][i];
                    float s = dx * dx;
                    s = fmaf(dy, dy, s);
                    s = fmaf(dz, dz, s);
                    t = fmaf(s, -k1, k0);
                }
                const float K = __builtin_amdgcn_exp2f(t);
#pragma unroll
                for (int q = 0; q < NW; ++q) acc[j][q] = fmaf(K, w[i][q], acc[j][q]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kRW; ++j)
#pragma unroll
        for (int q = 0; q < NW; ++q) acc[j][q] = wave_sum(acc[j][q]);

    if (MODE == kRowFwd) {
        if (lane < kRW) {
            float tot = acc[0][0];
#pragma unroll
            for (int j = 1; j < kRW; ++j)
                if (lane == j) tot = acc[j][0];
            const int r = r0 + lane;
            if (r < N) {
                float* ck = ckw(d, b, it);
                float* ub = d.u + (size_t)b * N;
                float* ob = d.orow + (size_t)b * N;
                const float su = (k == 1) ? 1.f : ub[r];
                const float orv = ob[r];
                if (k == 1) ck[ck_or0(d) + r] = orv;
                const float mass = su * tot;
                const float A = mass + orv;
                ob[r] = orv / A;
                ub[r] = su / A;
                ck[ck_A(d, k) + r] = A;
                ck[ck_rm(d, k) + r] = mass;
            }
        }
        return;
    }
    if (MODE == kRowBwd) {
        if (lane < kRW) {
            float tot = acc[0][0];
#pragma unroll
            for (int j = 1; j < kRW; ++j)
                if (lane == j) tot = acc[j][0];
            const int r = r0 + lane;
            if (r < N) {
                const float* ck = ckp(d, b, it);
                float um, om, uk, ok;
                row_state(ck, d, r, k - 1, um, om);
                uk = um / ck[ck_A(d, k) + r];
                ok = om / ck[ck_A(d, k) + r];
                const double A = ck[ck_A(d, k) + r];
                const double mass = ck[ck_rm(d, k) + r];
                float* gUb = d.gU + (size_t)b * N;
                float* gOb = d.gOr + (size_t)b * N;
                const double gu = (double)gUb[r] + (double)tot;  // dL/du^k complete
                const double go = gOb[r];
                const double gA = -(gu * uk + go * ok) / A;
                gUb[r] = (float)(gu / A + gA * mass / (double)um);  // dL/du^{k-1}
                gOb[r] = (float)(go / A + gA);                      // dL/do_row^{k-1}
                const float z = (float)(gA * (double)um);
                d.zr[(size_t)b * 4 * N + r] = z;
                float* Lb = d.L + (size_t)b * kMaxT * N;
                Lb[(size_t)t_col(d, k) * N + r] = uk;
                Lb[(size_t)t_row(d, k) * N + r] = z;
            }
        }
        return;
    }

    // MODE kRowFinal: s_r = U sum K V, D'_r = U sum K V B y; FP64 moments
    //  [0] s  [1..3] s x  [4 + 3i + j] s x_i D'_j  [13..15] s^2 x  [16..18] D'
    double m[kMomW];
#pragma unroll
    for (int q = 0; q < kMomW; ++q) m[q] = 0.0;
    float* ck = ckw(d, b, it);
#pragma unroll
    for (int j = 0; j < kRW; ++j) {
        const int r = r0 + j;
        if (r >= N) continue;
        const double U = d.u[(size_t)b * N + r];
        const double s = U * acc[j][0];
        const double D[3] = {U * acc[j][1], U * acc[j][2], U * acc[j][3]};
        const double x[3] = {xr[j][0], xr[j][1], xr[j][2]};
        if (lane == 0) {
            ck[ck_sD(d) + r] = (float)s;
            ck[ck_sD(d) + N + r] = (float)D[0];
            ck[ck_sD(d) + 2 * (size_t)N + r] = (float)D[1];
            ck[ck_sD(d) + 3 * (size_t)N + r] = (float)D[2];
        }
        m[0] += s;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            m[1 + i] += s * x[i];
            m[13 + i] += s * s * x[i];
            m[16 + i] += D[i];
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) m[4 + 3 * i + jj] += s * x[i] * D[jj];
        }
    }
    if (lane < kMomW) {
        double val = m[0];
#pragma unroll
        for (int q = 1; q < kMomW; ++q)
            if (lane == q) val = m[q];
        d.rmom[((size_t)b * d.rowWaves + gw) * kMomW + lane] = val;
    }
}

// ---------------------------------------------------------------------------
// Column pass: a workgroup = 256 columns x one row chunk; lanes take 4
// consecutive columns, the 4 waves split the chunk's rows. NW row weights
// z[w][r] (w < NW) -> cpart[b][h][w][c].
// ---------------------------------------------------------------------------
template <int NW, bool EXP>
__global__ void __launch_bounds__(kWave * kWG)
sau_colpass_kernel(Sau d, const float* __restrict__ z, size_t zb, float k1, float k0) {
    const int b = blockIdx.y;
    const int H = d.H, N = d.N, ldk = d.ldk;
    const int cbk = blockIdx.x / H;
    const int h = blockIdx.x - cbk * H;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int c = cbk * kColsPB + 4 * lane;
    const int q = d.chunkRows / kWG;
    const int rs = h * d.chunkRows + wave * q;
    const int re = min(rs + q, N);
    const float* X = d.X + (size_t)b * 3 * N;
    const float* P = d.P + (size_t)b * 3 * ldk;
    const float* zz = z + (size_t)b * zb;
    float p[3][4], bc[4];
    load4(P + c, p[0]);
    load4(P + (size_t)ldk + c, p[1]);
    load4(P + 2 * (size_t)ldk + c, p[2]);
    // EXP: the exponent expanded as in the row passes:
    // (k0 - k1 |x_r|^2) + b_c + 2 k1 x_r . p_c
    if (EXP)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            bc[i] = -k1 * fmaf(p[0][i], p[0][i], fmaf(p[1][i], p[1][i], p[2][i] * p[2][i]));
    const float k2 = 2.f * k1;
    float acc[NW][4];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[w][i] = 0.f;
    // the row index is wave-uniform: rows and their weights come in through
    // scalar loads, four rows per batch so that the loads are issued together
    const cfloat_p Xs = (cfloat_p)X;
    const cfloat_p Zs = (cfloat_p)zz;
    auto pair4 = [&](float x0, float x1, float x2, const float (&zw)[NW]) {
        float a = 0.f, q0 = 0.f, q1 = 0.f, q2 = 0.f;
        if (EXP) {
            a = fmaf(-k1, fmaf(x0, x0, fmaf(x1, x1, x2 * x2)), k0);
            q0 = k2 * x0;
            q1 = k2 * x1;
            q2 = k2 * x2;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float t;
            if (EXP) {
                t = a + bc[i];
                t = fmaf(q0, p[0][i], t);
                t = fmaf(q1, p[1][i], t);
                t = fmaf(q2, p[2][i], t);
            } else {
                const float dx = x0 - p[0][i], dy = x1 - p[1][i], dz = x2 - p[2][i];
                float s = dx * dx;
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                t = fmaf(s, -k1, k0);
            }
            const float K = __builtin_amdgcn_exp2f(t);
#pragma unroll
            for (int w = 0; w < NW; ++w) acc[w][i] = fmaf(K, zw[w], acc[w][i]);
        }
    };
    int r = rs;
    for (; r + 4 <= re; r += 4) {
        float x0[4], x1[4], x2[4], zw[4][NW];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x0[j] = Xs[r + j];
            x1[j] = Xs[(size_t)N + r + j];
            x2[j] = Xs[2 * (size_t)N + r + j];
#pragma unroll
            for (int w = 0; w < NW; ++w) zw[j][w] = Zs[(size_t)w * N + r + j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) pair4(x0[j], x1[j], x2[j], zw[j]);
    }
    for (; r < re; ++r) {
        float zw[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) zw[w] = Zs[(size_t)w * N + r];
        pair4(Xs[r], Xs[(size_t)N + r], Xs[2 * (size_t)N + r], zw);
    }
    __shared__ float red[kWG][NW][kColsPB];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[wave][w][4 * lane + i] = acc[w][i];
    __syncthreads();
    const int tid = threadIdx.x;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        float s = red[0][w][tid];
#pragma unroll
        for (int g = 1; g < kWG; ++g) s += red[g][w][tid];
        d.cpart[(((size_t)b * H + h) * 4 + w) * ldk + cbk * kColsPB + tid] = s;
    }
}

// Column sweep k backwards for column c: given dL/dv^k, dL/do_col^k and a
// direct dL/dB^k, writes dL/dv^{k-1}, dL/do_col^{k-1}, z_c and the two Q terms.
__device__ inline void col_sweep_bwd(const Sau& d, int b, int it, int k, int c, double gv,
                                     double gBdir) {
    const float* ck = ckp(d, b, it);
    float vm, om;
    col_state(ck, d, c, k - 1, vm, om);
    const float Bk = ck[ck_B(d, k) + c];
    const float vk = vm / Bk, okk = om / Bk;
    const double Bd = Bk, mass = ck[ck_cm(d, k) + c];
    float* gOcb = d.gOc + (size_t)b * d.ldk;
    const double go = gOcb[c];
    const double gB = gBdir - (gv * vk + go * okk) / Bd;
    d.gV[(size_t)b * d.ldk + c] = (float)(gv / Bd + gB * mass / (double)vm);
    gOcb[c] = (float)(go / Bd + gB);
    const float z = (float)(gB * (double)vm);
    d.zc[(size_t)b * d.ldk + c] = z;
    float* Qb = d.Q + (size_t)b * kMaxT * d.ldk;
    Qb[(size_t)t_col(d, k) * d.ldk + c] = z;
    Qb[(size_t)t_row(d, k) * d.ldk + c] = vm;
}

// Column finalize: one thread per column sums the H row chunks (FP64, fixed
// order). MODE 0: forward sweep k; MODE 1: backward, final-matrix terms then
// sweep I1; MODE 2: backward sweep k (adds the row half's column sums).
enum { kColFwd = 0, kColBwdM = 1, kColBwd = 2 };

template <int MODE>
__global__ void __launch_bounds__(kWave * kWG) sau_colfin_kernel(Sau d, int k, int it) {
    constexpr int NW = (MODE == kColBwdM) ? 4 : 1;
    const int b = blockIdx.y;
    const int tid = threadIdx.x;
    const int c = blockIdx.x * kColsPB + tid;
    const int M = d.M, ldk = d.ldk, H = d.H;
    double tot[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        tot[w] = 0.0;
        for (int h = 0; h < H; ++h) tot[w] += d.cpart[(((size_t)b * H + h) * 4 + w) * ldk + c];
    }
    const float* Yb = d.Y + (size_t)b * 3 * M;
    if (MODE == kColFwd) {
        double by[3] = {0.0, 0.0, 0.0};
        float* vb = d.v + (size_t)b * ldk;
        if (c < M) {
            float* ck = ckw(d, b, it);
            float* ob = d.ocol + (size_t)b * ldk;
            const float sv = (k == 1) ? 1.f : vb[c];
            const float ov = ob[c];
            if (k == 1) ck[ck_oc0(d) + c] = ov;
            const float mass = (float)((double)sv * tot[0]);
            const float Bc = mass + ov;
            ob[c] = ov / Bc;
            vb[c] = sv / Bc;
            d.bcol[(size_t)b * ldk + c] = Bc;
            ck[ck_B(d, k) + c] = Bc;
            ck[ck_cm(d, k) + c] = mass;
            const float y0 = Yb[c], y1 = Yb[(size_t)M + c], y2 = Yb[2 * (size_t)M + c];
            by[0] = (double)Bc * y0;
            by[1] = (double)Bc * y1;
            by[2] = (double)Bc * y2;
            if (k == d.I1) {  // weights of the final row pass: V and V B y
                float* wf = d.wfin + (size_t)b * 4 * ldk;
                const float V = vb[c], base = V * Bc;
                wf[c] = V;
                wf[(size_t)ldk + c] = base * y0;
                wf[2 * (size_t)ldk + c] = base * y1;
                wf[3 * (size_t)ldk + c] = base * y2;
            }
        } else {
            vb[c] = 0.f;
            d.bcol[(size_t)b * ldk + c] = 0.f;
            if (k == d.I1) {
                float* wf = d.wfin + (size_t)b * 4 * ldk;
                for (int q = 0; q < 4; ++q) wf[(size_t)q * ldk + c] = 0.f;
            }
        }
        __shared__ double lds[kWG * 3];
        double total;
        block_sum<3>(by, lds, kWG, total);
        if (tid < 3) d.cby[((size_t)b * d.colBlocks + blockIdx.x) * 3 + tid] = total;
        return;
    }
    float* Qb = d.Q + (size_t)b * kMaxT * ldk;
    if (c >= M) {
        for (int t = 0; t < kMaxT; ++t) Qb[(size_t)t * ldk + c] = 0.f;
        d.zc[(size_t)b * ldk + c] = 0.f;
        return;
    }
    if (MODE == kColBwdM) {
        const float* ck = ckp(d, b, it);
        const double* coef = d.adj + (size_t)b * kAdjW + 16;
        float V, oc;
        col_state(ck, d, c, d.I1, V, oc);
        const double Bc = ck[ck_B(d, d.I1) + c];
        const double y[3] = {Yb[c], Yb[(size_t)M + c], Yb[2 * (size_t)M + c]};
        const double w[3] = {Bc * y[0], Bc * y[1], Bc * y[2]};
        const double Z0 = tot[0], Z[3] = {tot[1], tot[2], tot[3]};
        const double gV = Z0 + w[0] * Z[0] + w[1] * Z[1] + w[2] * Z[2];
        const double* gby = coef + 19;
        double gB = (double)V * (y[0] * Z[0] + y[1] * Z[1] + y[2] * Z[2]);
        for (int i = 0; i < 3; ++i) gB += gby[i] * y[i];
        float* gYb = d.gY + (size_t)b * 3 * M;
        for (int i = 0; i < 3; ++i)
            gYb[(size_t)i * M + c] += (float)(Bc * ((double)V * Z[i] + gby[i]));
        Qb[c] = V;
        for (int i = 0; i < 3; ++i) Qb[(size_t)(1 + i) * ldk + c] = (float)(w[i] * (double)V);
        col_sweep_bwd(d, b, it, d.I1, c, gV, gB);
        return;
    }
    // MODE kColBwd: dL/dv^k = (carried) + sum_r K z_r
    const double gv = (double)d.gV[(size_t)b * ldk + c] + tot[0];
    col_sweep_bwd(d, b, it, k, c, gv, 0.0);
}

// ---------------------------------------------------------------------------
// Forward solve: one workgroup per registration.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWave * kWG) sau_solve_kernel(Sau d, int it, float T) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    double m[kMomW];
#pragma unroll
    for (int q = 0; q < kMomW; ++q) m[q] = 0.0;
    const double* rm = d.rmom + (size_t)b * d.rowWaves * kMomW;
    for (int w = tid; w < d.rowWaves; w += blockDim.x)
#pragma unroll
        for (int q = 0; q < kMomW; ++q) m[q] += rm[(size_t)w * kMomW + q];
    double by[3] = {0.0, 0.0, 0.0};
    const double* cb = d.cby + (size_t)b * d.colBlocks * 3;
    for (int q = tid; q < d.colBlocks; q += blockDim.x)
#pragma unroll
        for (int i = 0; i < 3; ++i) by[i] += cb[(size_t)q * 3 + i];
    __shared__ double lds[kWG * kMomW];
    __shared__ double mt[kMomW + 3];
    double total;
    block_sum<kMomW>(m, lds, kWG, total);
    if (tid < kMomW) mt[tid] = total;
    block_sum<3>(by, lds, kWG, total);
    if (tid < 3) mt[kMomW + tid] = total;
    __syncthreads();
    if (tid != 0) return;
    double* st = d.state + (size_t)b * kStateW;
    const double Sg = mt[0];
    double Xc[3], Yc[3], S[9], Rn[9], tn[3];
    for (int i = 0; i < 3; ++i) {
        Xc[i] = mt[1 + i] / Sg;
        Yc[i] = mt[kMomW + i] / Sg;
    }
    // S[i + 3j], i over X (target), j over Y (source): softassign.cu:769-772
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[i + 3 * j] = mt[4 + 3 * i + j] - Yc[j] * mt[13 + i] - Xc[i] * mt[16 + j] +
                           Xc[i] * Yc[j] * Sg;
    const bool ok = horn_rt(Xc, Yc, S, Rn, tn);
    if (ok) {
        for (int i = 0; i < 9; ++i) st[kStR + i] = Rn[i];
        for (int i = 0; i < 3; ++i) st[kStT + i] = tn[i];
    }
    double* o = d.rec + ((size_t)b * d.nIt + it) * kRecW;
    o[0] = T;
    o[1] = Sg;
    for (int i = 0; i < 3; ++i) {
        o[2 + i] = Xc[i];
        o[5 + i] = Yc[i];
    }
    for (int i = 0; i < 9; ++i) o[8 + i] = S[i];
    for (int i = 0; i < 9; ++i) o[17 + i] = st[kStR + i];
    for (int i = 0; i < 3; ++i) o[26 + i] = st[kStT + i];
    o[29] = 0.0;
    o[30] = 0.0;
    o[31] = ok ? 0.0 : 1.0;
    for (int i = 0; i < 3; ++i) {
        o[32 + i] = mt[13 + i];
        o[35 + i] = mt[16 + i];
    }
    o[38] = 0.0;
    o[39] = 0.0;
}

__global__ void sau_output_kernel(Sau d, float* __restrict__ Rt) {
    const int b = blockIdx.x;
    if (threadIdx.x < 12)
        Rt[(size_t)b * 12 + threadIdx.x] = (float)d.state[(size_t)b * kStateW + threadIdx.x];
}

// ---------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------

// Horn backwards for one (Xc, Yc, S) with S[a*3+b] = sum (source)_a (target)_b:
// gR (row-major), gt -> dL/dS (same layout), dL/dXc, dL/dYc. Returns false when
// the step's eigenproblem has no usable gap (then the outputs are zero).
__device__ inline void horn_bwd_core(const double* Xc, const double* Yc, const double* S,
                                     const double* gR, const double* gt, double* gS, double* gXc,
                                     double* gYc) {
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
    double V[4][4], w[4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 4; ++i) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < 4; ++j) off += A[i][j] * A[i][j];
        }
        if (off <= 1e-36 * diag || off == 0.0) break;
        for (int p = 0; p < 3; ++p)
            for (int r = p + 1; r < 4; ++r) {
                const double apq = A[p][r];
                if (apq == 0.0) continue;
                const double tau = (A[r][r] - A[p][p]) / (2.0 * apq);
                const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = tt * cs;
                for (int q = 0; q < 4; ++q) {
                    const double akp = A[q][p], akq = A[q][r];
                    A[q][p] = cs * akp - sn * akq;
                    A[q][r] = sn * akp + cs * akq;
                }
                for (int q = 0; q < 4; ++q) {
                    const double apk = A[p][q], aqk = A[r][q];
                    A[p][q] = cs * apk - sn * aqk;
                    A[r][q] = sn * apk + cs * aqk;
                }
                for (int q = 0; q < 4; ++q) {
                    const double vkp = V[q][p], vkq = V[q][r];
                    V[q][p] = cs * vkp - sn * vkq;
                    V[q][r] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int q = 0; q < 4; ++q) w[q] = A[q][q];
    int top = 0;
    for (int q = 1; q < 4; ++q)
        if (w[q] > w[top]) top = q;
    double qv[4];
    for (int i = 0; i < 4; ++i) qv[i] = V[i][top];
    const double q0 = qv[0], qx = qv[1], qy = qv[2], qz = qv[3];
    const double R[9] = {q0 * q0 + qx * qx - qy * qy - qz * qz, 2 * (qx * qy - q0 * qz),
                         2 * (qx * qz + q0 * qy),           2 * (qy * qx + q0 * qz),
                         q0 * q0 - qx * qx + qy * qy - qz * qz, 2 * (qy * qz - q0 * qx),
                         2 * (qz * qx - q0 * qy),           2 * (qz * qy + q0 * qx),
                         q0 * q0 - qx * qx - qy * qy + qz * qz};
    double G[9];
    for (int i = 0; i < 3; ++i) gXc[i] = gt[i];
    for (int j = 0; j < 3; ++j) gYc[j] = -(R[j] * gt[0] + R[3 + j] * gt[1] + R[6 + j] * gt[2]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) G[3 * i + j] = gR[3 * i + j] - gt[i] * Yc[j];
    double gq[4];
    gq[0] = 2 * (q0 * G[0] - qz * G[1] + qy * G[2] + qz * G[3] + q0 * G[4] - qx * G[5] -
                 qy * G[6] + qx * G[7] + q0 * G[8]);
    gq[1] = 2 * (qx * G[0] + qy * G[1] + qz * G[2] + qy * G[3] - qx * G[4] - q0 * G[5] +
                 qz * G[6] + q0 * G[7] - qx * G[8]);
    gq[2] = 2 * (-qy * G[0] + qx * G[1] + q0 * G[2] + qx * G[3] + qy * G[4] + qz * G[5] -
                 q0 * G[6] + qz * G[7] - qy * G[8]);
    gq[3] = 2 * (-qz * G[0] - q0 * G[1] + qx * G[2] + q0 * G[3] - qz * G[4] + qy * G[5] +
                 qx * G[6] + qy * G[7] + qz * G[8]);
    double gA[4][4];
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) gA[a][c] = 0.0;
    for (int j = 0; j < 4; ++j) {
        if (j == top) continue;
        const double gap = w[top] - w[j];
        if (gap == 0.0) continue;
        double dot = 0.0;
        for (int i = 0; i < 4; ++i) dot += V[i][j] * gq[i];
        const double cf = dot / gap;
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 4; ++c) gA[a][c] += cf * V[a][j] * qv[c];
    }
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
}

// Horn backwards + the Softassign moment map, per registration (one thread).
// coef (at adj + 16): [0] Sigma [1..3] sum s x [4 + 3i + j] sum s x_i D'_j
// [13..15] sum s^2 x [16..18] sum D' [19..21] sum B y.
__global__ void sau_adj_solve_kernel(Sau d, int it) {
    if (threadIdx.x != 0) return;
    const int b = blockIdx.x;
    double* adj = d.adj + (size_t)b * kAdjW;
    double* coef = adj + 16;
    for (int i = 0; i < kCoefW; ++i) coef[i] = 0.0;
    const double* o = d.rec + ((size_t)b * d.nIt + it) * kRecW;
    const double Sg = o[1];
    if (o[31] != 0.0 || !(Sg > 0.0)) return;  // failed step: (R, t) passed through
    const double* Xc = o + 2;
    const double* Yc = o + 5;
    const double* S = o + 8;  // S[i + 3j]: i target, j source == h_S[3j + i]
    const double* m13 = o + 32;
    const double* m16 = o + 35;
    double gS[9], gXc[3], gYc[3];
    horn_bwd_core(Xc, Yc, S, adj, adj + 9, gS, gXc, gYc);
    // S_ij = m4_ij - Yc_j m13_i - Xc_i m16_j + Xc_i Yc_j Sigma
    double gSg = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double g = gS[i + 3 * j];
            coef[4 + 3 * i + j] = g;
            coef[13 + i] -= g * Yc[j];
            coef[16 + j] -= g * Xc[i];
            gXc[i] += g * (Yc[j] * Sg - m16[j]);
            gYc[j] += g * (Xc[i] * Sg - m13[i]);
            gSg += g * Xc[i] * Yc[j];
        }
    // Xc = sum s x / Sigma, Yc = sum B y / Sigma
    for (int i = 0; i < 3; ++i) {
        coef[1 + i] = gXc[i] / Sg;
        coef[19 + i] = gYc[i] / Sg;
        gSg -= (gXc[i] * Xc[i] + gYc[i] * Yc[i]) / Sg;
    }
    coef[0] = gSg;
}

// Per row: dL/ds_r, dL/dD'_r, the direct dL/dx_r, dL/dU_r, the final-matrix
// L terms and the four row weights U (alpha, beta) of the next column pass.
__global__ void __launch_bounds__(256) sau_adj_mrow_kernel(Sau d, int it) {
    const int b = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = d.N;
    if (r >= N) return;
    const float* ck = ckp(d, b, it);
    const double* coef = d.adj + (size_t)b * kAdjW + 16;
    const float* X = d.X + (size_t)b * 3 * N;
    const double x[3] = {X[r], X[(size_t)N + r], X[2 * (size_t)N + r]};
    const double s = ck[ck_sD(d) + r];
    const double D[3] = {ck[ck_sD(d) + N + r], ck[ck_sD(d) + 2 * (size_t)N + r],
                         ck[ck_sD(d) + 3 * (size_t)N + r]};
    double alpha = coef[0], beta[3], gx[3];
    for (int i = 0; i < 3; ++i) {
        alpha += coef[1 + i] * x[i] + 2.0 * s * coef[13 + i] * x[i];
        gx[i] = coef[1 + i] * s + coef[13 + i] * s * s;
    }
    for (int j = 0; j < 3; ++j) beta[j] = coef[16 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double g = coef[4 + 3 * i + j];
            alpha += g * x[i] * D[j];
            beta[j] += g * s * x[i];
            gx[i] += g * s * D[j];
        }
    float* gXb = d.gX + (size_t)b * 3 * N;
    for (int i = 0; i < 3; ++i) gXb[(size_t)i * N + r] += (float)gx[i];
    float U, o;
    row_state(ck, d, r, d.I1, U, o);
    const double Ud = U;
    d.gU[(size_t)b * N + r] =
        (float)((alpha * s + beta[0] * D[0] + beta[1] * D[1] + beta[2] * D[2]) / Ud);
    float* Lb = d.L + (size_t)b * kMaxT * N;
    float* zb = d.zr + (size_t)b * 4 * N;
    Lb[r] = (float)(alpha * Ud);
    zb[r] = (float)(alpha * Ud);
    for (int j = 0; j < 3; ++j) {
        Lb[(size_t)(1 + j) * N + r] = (float)(beta[j] * Ud);
        zb[(size_t)(1 + j) * N + r] = (float)(beta[j] * Ud);
    }
}

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Final pass over the pairs: G = (sum_t L_r^t Q_c^t) K. A workgroup = 256 rows
// x 1024 columns; lanes hold 4 rows, the 4 waves split the columns. Row sums
// (G d (3), G s, G) -> frow; column sums G d over the wave's 256 rows are
// staged through LDS -> fcol.
template <int NT>
__global__ void __launch_bounds__(kWave * kWG) sau_final_kernel(Sau d, float k1, float k0) {
    const int b = blockIdx.y;
    const int N = d.N, ldk = d.ldk, chunks = d.fchunks;
    const int rt = blockIdx.x / chunks;
    const int g = blockIdx.x - rt * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const float* X = d.X + (size_t)b * 3 * N;
    const float* Lb = d.L + (size_t)b * kMaxT * N;
    float x[kFRPL][3], Lr[kFRPL][NT];
#pragma unroll
    for (int j = 0; j < kFRPL; ++j) {
        const int r = rt * kFRows + j * kWave + lane;
        const bool in = r < N;
        const int rr = in ? r : 0;
        x[j][0] = X[rr];
        x[j][1] = X[(size_t)N + rr];
        x[j][2] = X[2 * (size_t)N + rr];
#pragma unroll
        for (int t = 0; t < NT; ++t) Lr[j][t] = in ? Lb[(size_t)t * N + rr] : 0.f;
    }
    const float* Pb = d.P + (size_t)b * 3 * ldk;
    const float* Qb = d.Q + (size_t)b * kMaxT * ldk;
    const cfloat_p Px = (cfloat_p)Pb;
    const cfloat_p Py = (cfloat_p)(Pb + ldk);
    const cfloat_p Pz = (cfloat_p)(Pb + 2 * (size_t)ldk);
    const cfloat_p Qc = (cfloat_p)Qb;
    const int cb = g * kFCols + wave * (kFCols / kWG);
    const int ce = min(cb + kFCols / kWG, ldk);
    float ax[kFRPL], ay[kFRPL], az[kFRPL], as[kFRPL], ag[kFRPL];
#pragma unroll
    for (int j = 0; j < kFRPL; ++j) ax[j] = ay[j] = az[j] = as[j] = ag[j] = 0.f;
    __shared__ float lds[kWG * kStageW * kStageLd];
    float* __restrict__ stage = lds + wave * kStageW * kStageLd;
    float* __restrict__ fcol = d.fcol + ((size_t)b * d.rowTiles + rt) * ldk * 3;
    for (int c = cb; c < ce; c += kColGroup) {
#pragma unroll 2
        for (int u = 0; u < kColGroup; ++u) {
            const int cc = c + u;
            const float px = Px[cc], py = Py[cc], pz = Pz[cc];
            float q[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) q[t] = Qc[(size_t)t * ldk + cc];
            float cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll
            for (int j = 0; j < kFRPL; ++j) {
                const float dx = x[j][0] - px, dy = x[j][1] - py, dz = x[j][2] - pz;
                float s = dx * dx;
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                const float K = __builtin_amdgcn_exp2f(fmaf(s, -k1, k0));
                float gk = Lr[j][0] * q[0];
#pragma unroll
                for (int t = 1; t < NT; ++t) gk = fmaf(Lr[j][t], q[t], gk);
                const float G = gk * K;
                ax[j] = fmaf(G, dx, ax[j]);
                ay[j] = fmaf(G, dy, ay[j]);
                az[j] = fmaf(G, dz, az[j]);
                as[j] = fmaf(G, s, as[j]);
                ag[j] += G;
                cx = fmaf(G, dx, cx);
                cy = fmaf(G, dy, cy);
                cz = fmaf(G, dz, cz);
            }
            stage[(3 * u + 0) * kStageLd + lane] = cx;
            stage[(3 * u + 1) * kStageLd + lane] = cy;
            stage[(3 * u + 2) * kStageLd + lane] = cz;
        }
        wave_lds_sync();
        if (lane < kStageW) {
            const float* __restrict__ row = stage + lane * kStageLd;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
            for (int i = 0; i < kWave; i += 4) {
                a0 += row[i];
                a1 += row[i + 1];
                a2 += row[i + 2];
                a3 += row[i + 3];
            }
            fcol[(size_t)c * 3 + lane] = (a0 + a1) + (a2 + a3);
        }
        wave_lds_sync();
    }
    __syncthreads();
    float* red = lds;  // [kWG][5][kFRows]
#pragma unroll
    for (int j = 0; j < kFRPL; ++j) {
        const int i = j * kWave + lane;
        red[(wave * 5 + 0) * kFRows + i] = ax[j];
        red[(wave * 5 + 1) * kFRows + i] = ay[j];
        red[(wave * 5 + 2) * kFRows + i] = az[j];
        red[(wave * 5 + 3) * kFRows + i] = as[j];
        red[(wave * 5 + 4) * kFRows + i] = ag[j];
    }
    __syncthreads();
    float* F = d.frow + (((size_t)b * d.rowTiles + rt) * chunks + g) * 5 * kFRows;
    for (int q = 0; q < 5; ++q) {
        float acc = red[q * kFRows + threadIdx.x];
#pragma unroll
        for (int w = 1; w < kWG; ++w) acc += red[(w * 5 + q) * kFRows + threadIdx.x];
        F[q * kFRows + threadIdx.x] = acc;
    }
}

// Rows of the final pass: dL/dx_r += -(2/T) sum G d; tile sums of G s, G.
__global__ void __launch_bounds__(kFRows) sau_final_rows_kernel(Sau d, float T) {
    const int b = blockIdx.y;
    const int rt = blockIdx.x;
    const int N = d.N, chunks = d.fchunks;
    const int r = rt * kFRows + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (r < N) {
        double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const float* F = d.frow + (((size_t)b * d.rowTiles + rt) * chunks) * 5 * kFRows;
        for (int g = 0; g < chunks; ++g)
            for (int q = 0; q < 5; ++q) a[q] += F[((size_t)g * 5 + q) * kFRows + threadIdx.x];
        float* gXb = d.gX + (size_t)b * 3 * N;
        const double f = -2.0 / (double)T;
        for (int i = 0; i < 3; ++i) gXb[(size_t)i * N + r] += (float)(f * a[i]);
        v[0] = a[3];
        v[1] = a[4];
    }
    __shared__ double lds[(kFRows / kWave) * 2];
    double tot;
    block_sum<2>(v, lds, kFRows / kWave, tot);
    if (threadIdx.x < 2) d.fmomr[((size_t)b * d.rowTiles + rt) * 2 + threadIdx.x] = tot;
}

// Columns of the final pass: dL/dP_c = (2/T) sum G d; dL/dy_c += R^T dL/dP_c;
// block sums of dL/dP y^T (9) and dL/dP (3) for dL/d(R, t) of the input.
__global__ void __launch_bounds__(256) sau_final_cols_kernel(Sau d, int it, float T) {
    const int b = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int M = d.M, ldk = d.ldk;
    double v[12];
    for (int i = 0; i < 12; ++i) v[i] = 0.0;
    if (c < M) {
        double gp[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < d.rowTiles; ++t) {
            const float* o = d.fcol + (((size_t)b * d.rowTiles + t) * ldk + c) * 3;
            gp[0] += o[0];
            gp[1] += o[1];
            gp[2] += o[2];
        }
        const double f = 2.0 / (double)T;
        for (int i = 0; i < 3; ++i) gp[i] *= f;
        double R[9], tv[3];
        iter_input(d, b, it, R, tv);
        const float* Yb = d.Y + (size_t)b * 3 * M;
        const double y[3] = {Yb[c], Yb[(size_t)M + c], Yb[2 * (size_t)M + c]};
        float* gYb = d.gY + (size_t)b * 3 * M;
        for (int j = 0; j < 3; ++j)
            gYb[(size_t)j * M + c] += (float)(R[j] * gp[0] + R[3 + j] * gp[1] + R[6 + j] * gp[2]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) v[3 * i + j] = gp[i] * y[j];
        for (int i = 0; i < 3; ++i) v[9 + i] = gp[i];
    }
    __shared__ double lds[4 * 12];
    double tot;
    block_sum<12>(v, lds, 4, tot);
    if (threadIdx.x < 12) d.fmomc[((size_t)b * d.colBlocks + blockIdx.x) * 12 + threadIdx.x] = tot;
}

// Per registration: dL/d(R, t) of the iteration's input and dL/dT_0.
__global__ void __launch_bounds__(64) sau_adj_finish_kernel(Sau d, int it, double fT, float T) {
    const int b = blockIdx.x;
    const int i = threadIdx.x;
    double* adj = d.adj + (size_t)b * kAdjW;
    const bool failed = d.rec[((size_t)b * d.nIt + it) * kRecW + 31] != 0.0;
    __shared__ double gT;
    if (i < 12) {
        double s = 0.0;
        for (int q = 0; q < d.colBlocks; ++q) s += d.fmomc[((size_t)b * d.colBlocks + q) * 12 + i];
        adj[i] = failed ? adj[i] + s : s;
    }
    if (i == 12) {
        double gs = 0.0, gg = 0.0;
        for (int t = 0; t < d.rowTiles; ++t) {
            gs += d.fmomr[((size_t)b * d.rowTiles + t) * 2 + 0];
            gg += d.fmomr[((size_t)b * d.rowTiles + t) * 2 + 1];
        }
        const double Td = T;
        // dK/dT = K ((s - alpha) / T^2 - 1 / (2T))
        gT = (gs - (double)d.alpha * gg) / (Td * Td) - gg / (2.0 * Td);
        adj[12] += fT * gT;
    }
}

__global__ void __launch_bounds__(256) sau_adj_setup_kernel(Sau d, const float* __restrict__ gRt) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = d.N, M = d.M, ldk = d.ldk;
    if (i < 3 * N) d.gX[(size_t)b * 3 * N + i] = 0.f;
    if (i < 3 * M) d.gY[(size_t)b * 3 * M + i] = 0.f;
    if (i < N) d.gOr[(size_t)b * N + i] = 0.f;
    if (i < ldk) d.gOc[(size_t)b * ldk + i] = 0.f;
    if (blockIdx.x == 0 && threadIdx.x < kAdjW)
        d.adj[(size_t)b * kAdjW + threadIdx.x] =
            threadIdx.x < 12 ? (double)gRt[(size_t)b * 12 + threadIdx.x] : 0.0;
}

// dL/dT_0 through the default slack value moutlier = exp(-1) / sqrt(T_0):
// d moutlier / dT_0 = -moutlier / (2 T_0), times the summed slack gradients.
__global__ void __launch_bounds__(256)
sau_adj_output_kernel(Sau d, float* __restrict__ gRt0, double* __restrict__ gT0, double dmo) {
    const int b = blockIdx.x;
    double v[1] = {0.0};
    for (int r = threadIdx.x; r < d.N; r += blockDim.x) v[0] += d.gOr[(size_t)b * d.N + r];
    for (int c = threadIdx.x; c < d.M; c += blockDim.x) v[0] += d.gOc[(size_t)b * d.ldk + c];
    __shared__ double lds[4];
    double tot;
    block_sum<1>(v, lds, 4, tot);
    const double* adj = d.adj + (size_t)b * kAdjW;
    if (gRt0 && threadIdx.x < 12) gRt0[(size_t)b * 12 + threadIdx.x] = (float)adj[threadIdx.x];
    if (gT0 && threadIdx.x == 0) gT0[b] = adj[12] + dmo * tot;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// max over the batch's target clouds of |x|^2 into out[0] (one workgroup;
// the values are non-negative, so their bit patterns order as the floats)
// max |x|^2 over one chunk of kXmaxChunk points (blockIdx.x) of target cloud
// blockIdx.y, written to out[blockIdx.y * gridDim.x + blockIdx.x]
constexpr int kXmaxThreads = 256;
constexpr int kXmaxChunk = 4096;
__global__ void __launch_bounds__(kXmaxThreads) sau_xmax_kernel(const float* __restrict__ X,
                                                                int N, float* __restrict__ out) {
    const float* __restrict__ x = X + (size_t)blockIdx.y * 3 * N;
    const int beg = blockIdx.x * kXmaxChunk, end = min(beg + kXmaxChunk, N);
    float m = 0.f;
    for (int r = beg + (int)threadIdx.x; r < end; r += kXmaxThreads) {
        const float x0 = x[r], x1 = x[(size_t)N + r], x2 = x[2 * (size_t)N + r];
        m = fmaxf(m, fmaf(x0, x0, fmaf(x1, x1, x2 * x2)));
    }
    m = wave_max(m);
    __shared__ float wm[kXmaxThreads / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = wm[0];
        for (int w = 1; w < kXmaxThreads / kWave; ++w) t = fmaxf(t, wm[w]);
        out[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

struct SauWork {
    DevBuf xmax;
    DevBuf P, state, u, v, orow, ocol, bcol, wfin, cpart, rmom, cby;
    DevBuf gU, gV, gOr, gOc, L, Q, zr, zc, adj, fcol, frow, fmomr, fmomc;
    WorkFence fence;  // the last forward or backward that used it, on any stream
};
constexpr int kMaxDevices = 64;
SauWork* g_sau[kMaxDevices];  // per device, created under that device's lock, never freed

// max |x|^2 over the batch's target clouds, read back to the host (it picks
// the exponent's form per temperature)
int sau_xmax2(SauWork& w, const float* dX, int B, int N, hipStream_t s, float& out) {
    int rc;
    const int chunks = (N + kXmaxChunk - 1) / kXmaxChunk;
    const size_t parts = (size_t)B * chunks;
    if ((rc = w.xmax.reserve(sizeof(float) * parts)) != EMSA_OK) return rc;
    hipLaunchKernelGGL(sau_xmax_kernel, dim3((unsigned)chunks, (unsigned)B), dim3(kXmaxThreads), 0,
                       s, dX, N, w.xmax.as<float>());
    EMSA_HIP_TRY(hipGetLastError());
    std::vector<float> h(parts);
    EMSA_HIP_TRY(hipMemcpyAsync(h.data(), w.xmax.ptr, sizeof(float) * parts,
                                hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    out = 0.f;
    for (float v : h) out = std::fmax(out, v);
    return EMSA_OK;
}

template <int MODE>
void sau_rows(bool ex, dim3 g, dim3 b, hipStream_t s, const Sau& d, int k, int it, float k1,
              float k0) {
    if (ex) hipLaunchKernelGGL((sau_rowpass_kernel<MODE, true>), g, b, 0, s, d, k, it, k1, k0);
    else hipLaunchKernelGGL((sau_rowpass_kernel<MODE, false>), g, b, 0, s, d, k, it, k1, k0);
}

template <int NW>
void sau_cols(bool ex, dim3 g, dim3 b, hipStream_t s, const Sau& d, const float* z, size_t zb,
              float k1, float k0) {
    if (ex) hipLaunchKernelGGL((sau_colpass_kernel<NW, true>), g, b, 0, s, d, z, zb, k1, k0);
    else hipLaunchKernelGGL((sau_colpass_kernel<NW, false>), g, b, 0, s, d, z, zb, k1, k0);
}

SauWork* sau_work(const DeviceCtx& ctx) {
    if (ctx.device < 0 || ctx.device >= kMaxDevices) return nullptr;
    if (!g_sau[ctx.device]) g_sau[ctx.device] = new SauWork();
    return g_sau[ctx.device];
}

struct SauArgs {
    int B, N, M, JMAX, I0, I1;
    float T_0, TFACTOR, alpha, moutlier;
};

int sau_check(const SauArgs& a) {
    if (a.B <= 0 || a.N <= 0 || a.M <= 0 || a.JMAX <= 0 || a.I0 <= 0)
        return set_error("softassign_unrolled: sizes and loop counts must be positive"), EMSA_EINVAL;
    if (a.I1 < 1 || a.I1 > kMaxI1)
        return set_error("softassign_unrolled: I1 must be in 1..4"), EMSA_EINVAL;
    if (!(a.T_0 > 0.f) || !(a.TFACTOR > 0.f) || !std::isfinite(a.T_0) ||
        !std::isfinite(a.TFACTOR) || !std::isfinite(a.alpha) || !std::isfinite(a.moutlier))
        return set_error("softassign_unrolled: T_0, TFACTOR must be positive and finite"),
               EMSA_EINVAL;
    if ((long long)a.B * a.N * 3 > 0x7fffffffLL || (long long)a.B * a.M * 3 > 0x7fffffffLL)
        return set_error("softassign_unrolled: batch too large"), EMSA_EINVAL;
    return EMSA_OK;
}

// Fills the sizes of d (no pointers).
void sau_dims(Sau& d, const SauArgs& a, int numCU) {
    d = Sau{};
    d.B = a.B;
    d.N = a.N;
    d.M = a.M;
    d.I1 = a.I1;
    d.nIt = a.JMAX * a.I0;
    d.alpha = a.alpha;
    d.ldk = (a.M + kColsPB - 1) / kColsPB * kColsPB;
    d.colBlocks = d.ldk / kColsPB;
    const int rowBlocks = (a.N + kRowsPB - 1) / kRowsPB;
    d.rowWaves = rowBlocks * kWG;
    const int target = std::max(1, numCU) * 4;
    int H = std::max(1, (target + d.colBlocks * a.B - 1) / (d.colBlocks * a.B));
    H = std::min(H, (a.N + kWG - 1) / kWG);
    d.chunkRows = ((a.N + H - 1) / H + kWG - 1) / kWG * kWG;
    d.H = (a.N + d.chunkRows - 1) / d.chunkRows;
    d.ckF = (int)ck_floats(d);
    d.rowTiles = (a.N + kFRows - 1) / kFRows;
    d.fchunks = (d.ldk + kFCols - 1) / kFCols;
}

size_t sau_ckpt_bytes(const Sau& d) {
    return sizeof(double) * (size_t)d.B * d.nIt * kRecW +
           sizeof(float) * (size_t)d.B * d.nIt * d.ckF;
}

std::vector<float> sau_temps(const SauArgs& a) {
    std::vector<float> T;
    float t = a.T_0;
    for (int j = 0; j < a.JMAX; ++j) {
        for (int i = 0; i < a.I0; ++i) T.push_back(t);
        t = t * a.TFACTOR;  // softassign.cu:820, float
    }
    return T;
}

void sau_consts(float T, float alpha, float& k1, float& k0) {
    const double k1d = kLog2e / (double)T;
    k1 = (float)k1d;
    k0 = (float)((double)alpha * k1d - 0.5 * std::log2((double)T));
}

int sau_bind_forward(Sau& d, SauWork& w, const float* X, const float* Y, const float* Rt0,
                     void* ckpt) {
    const size_t B = d.B, N = d.N, ldk = d.ldk;
    int rc;
    if ((rc = w.P.reserve(sizeof(float) * B * 3 * ldk)) != EMSA_OK) return rc;
    if ((rc = w.state.reserve(sizeof(double) * B * kStateW)) != EMSA_OK) return rc;
    if ((rc = w.u.reserve(sizeof(float) * B * N)) != EMSA_OK) return rc;
    if ((rc = w.v.reserve(sizeof(float) * B * ldk)) != EMSA_OK) return rc;
    if ((rc = w.orow.reserve(sizeof(float) * B * N)) != EMSA_OK) return rc;
    if ((rc = w.ocol.reserve(sizeof(float) * B * ldk)) != EMSA_OK) return rc;
    if ((rc = w.bcol.reserve(sizeof(float) * B * ldk)) != EMSA_OK) return rc;
    if ((rc = w.wfin.reserve(sizeof(float) * B * 4 * ldk)) != EMSA_OK) return rc;
    if ((rc = w.cpart.reserve(sizeof(float) * B * d.H * 4 * ldk)) != EMSA_OK) return rc;
    if ((rc = w.rmom.reserve(sizeof(double) * B * d.rowWaves * kMomW)) != EMSA_OK) return rc;
    if ((rc = w.cby.reserve(sizeof(double) * B * d.colBlocks * 3)) != EMSA_OK) return rc;
    d.X = X;
    d.Y = Y;
    d.Rt0 = Rt0;
    d.P = w.P.as<float>();
    d.state = w.state.as<double>();
    d.u = w.u.as<float>();
    d.v = w.v.as<float>();
    d.orow = w.orow.as<float>();
    d.ocol = w.ocol.as<float>();
    d.bcol = w.bcol.as<float>();
    d.wfin = w.wfin.as<float>();
    d.cpart = w.cpart.as<float>();
    d.rmom = w.rmom.as<double>();
    d.cby = w.cby.as<double>();
    d.rec = static_cast<double*>(ckpt);
    d.ck = reinterpret_cast<float*>(d.rec + B * d.nIt * kRecW);
    return EMSA_OK;
}

float default_moutlier(float T_0) { return 1.0f / sqrtf(T_0) * expf(-1.0f); }  // main.cpp default

}  // namespace

}  // namespace emsa

using namespace emsa;

extern "C" {

long long emsa_softassign_unrolled_ckpt_bytes(int B, int N, int M, int JMAX, int I0, int I1) {
    clear_error();
    SauArgs a{B, N, M, JMAX, I0, I1, 1.f, 0.5f, 0.f, 0.f};
    int rc;
    if ((rc = sau_check(a)) != EMSA_OK) return rc;
    Sau d;
    sau_dims(d, a, 1);
    return (long long)sau_ckpt_bytes(d);
}

int emsa_softassign_unrolled_fwd(const float* dX, const float* dY, int B, int N, int M,
                                 const float* dRt0, float T_0, float TFACTOR, float alpha,
                                 float moutlier, int JMAX, int I0, int I1, float* dRtOut,
                                 void* dckpt, void* stream) {
    clear_error();
    const float mo = moutlier < 0.f ? default_moutlier(T_0) : moutlier;
    SauArgs a{B, N, M, JMAX, I0, I1, T_0, TFACTOR, alpha, mo};
    int rc;
    if ((rc = sau_check(a)) != EMSA_OK) return rc;
    if (!dX || !dY || !dRt0 || !dRtOut || !dckpt)
        return set_error("softassign_unrolled_fwd: null buffer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    SauWork* wp = sau_work(*ctx);
    if (!wp) return set_error("softassign_unrolled: device index out of range"), EMSA_ENODEV;
    hipStream_t s = (hipStream_t)stream;
    // a call on another stream may still use this workspace
    if ((rc = wp->fence.wait()) != EMSA_OK) return rc;
    FenceMark mark(wp->fence, s);
    Sau d;
    sau_dims(d, a, ctx->numCU);
    if ((rc = sau_bind_forward(d, *wp, dX, dY, dRt0, dckpt)) != EMSA_OK) return rc;
    const dim3 blk(kWave * kWG);
    const unsigned ub = (unsigned)B;
    hipLaunchKernelGGL(sau_init_kernel, dim3((unsigned)((std::max(N, d.ldk) + 255) / 256), ub),
                       dim3(256), 0, s, d, mo);
    EMSA_HIP_TRY(hipGetLastError());
    const dim3 rgrid((unsigned)((N + kRowsPB - 1) / kRowsPB), ub);
    const dim3 cgrid((unsigned)(d.colBlocks * d.H), ub), fgrid((unsigned)d.colBlocks, ub);
    const dim3 tgrid((unsigned)(d.ldk / 256), ub);
    const std::vector<float> Ts = sau_temps(a);
    float xmax2 = 0.f;
    if ((rc = sau_xmax2(*wp, dX, B, N, s, xmax2)) != EMSA_OK) return rc;
    for (int it = 0; it < d.nIt; ++it) {
        float k1, k0;
        sau_consts(Ts[it], alpha, k1, k0);
        const bool ex = k1 * xmax2 <= kExpandMax;  // the exponent's form
        hipLaunchKernelGGL(sau_transform_kernel, tgrid, dim3(256), 0, s, d, it);
        for (int k = 1; k <= I1; ++k) {
            sau_rows<kRowFwd>(ex, rgrid, blk, s, d, k, it, k1, k0);
            sau_cols<1>(ex, cgrid, blk, s, d, (const float*)d.u, (size_t)N, k1, k0);
            hipLaunchKernelGGL(sau_colfin_kernel<kColFwd>, fgrid, blk, 0, s, d, k, it);
        }
        sau_rows<kRowFinal>(ex, rgrid, blk, s, d, I1, it, k1, k0);
        hipLaunchKernelGGL(sau_solve_kernel, dim3(ub), blk, 0, s, d, it, Ts[it]);
        EMSA_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(sau_output_kernel, dim3(ub), dim3(64), 0, s, d, dRtOut);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

int emsa_softassign_unrolled_bwd(const float* dX, const float* dY, int B, int N, int M,
                                 const float* dRt0, float T_0, float TFACTOR, float alpha,
                                 float moutlier, int JMAX, int I0, int I1, const void* dckpt,
                                 const float* dgRt, float* dgX, float* dgY, float* dgRt0,
                                 double* dgT0, void* stream) {
    clear_error();
    const bool defmo = moutlier < 0.f;
    const float mo = defmo ? default_moutlier(T_0) : moutlier;
    SauArgs a{B, N, M, JMAX, I0, I1, T_0, TFACTOR, alpha, mo};
    int rc;
    if ((rc = sau_check(a)) != EMSA_OK) return rc;
    if (!dX || !dY || !dRt0 || !dckpt || !dgRt || !dgX || !dgY)
        return set_error("softassign_unrolled_bwd: null buffer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    SauWork* wp = sau_work(*ctx);
    if (!wp) return set_error("softassign_unrolled: device index out of range"), EMSA_ENODEV;
    SauWork& w = *wp;
    hipStream_t s = (hipStream_t)stream;
    // a call on another stream may still use this workspace
    if ((rc = w.fence.wait()) != EMSA_OK) return rc;
    FenceMark mark(w.fence, s);
    Sau d;
    sau_dims(d, a, ctx->numCU);
    if ((rc = sau_bind_forward(d, w, dX, dY, dRt0, const_cast<void*>(dckpt))) != EMSA_OK)
        return rc;
    const size_t Bz = B, Nz = N, ldk = d.ldk;
    if ((rc = w.gU.reserve(sizeof(float) * Bz * Nz)) != EMSA_OK) return rc;
    if ((rc = w.gV.reserve(sizeof(float) * Bz * ldk)) != EMSA_OK) return rc;
    if ((rc = w.gOr.reserve(sizeof(float) * Bz * Nz)) != EMSA_OK) return rc;
    if ((rc = w.gOc.reserve(sizeof(float) * Bz * ldk)) != EMSA_OK) return rc;
    if ((rc = w.L.reserve(sizeof(float) * Bz * kMaxT * Nz)) != EMSA_OK) return rc;
    if ((rc = w.Q.reserve(sizeof(float) * Bz * kMaxT * ldk)) != EMSA_OK) return rc;
    if ((rc = w.zr.reserve(sizeof(float) * Bz * 4 * Nz)) != EMSA_OK) return rc;
    if ((rc = w.zc.reserve(sizeof(float) * Bz * ldk)) != EMSA_OK) return rc;
    if ((rc = w.adj.reserve(sizeof(double) * Bz * kAdjW)) != EMSA_OK) return rc;
    if ((rc = w.fcol.reserve(sizeof(float) * Bz * d.rowTiles * ldk * 3)) != EMSA_OK) return rc;
    if ((rc = w.frow.reserve(sizeof(float) * Bz * d.rowTiles * d.fchunks * 5 * kFRows)) !=
        EMSA_OK)
        return rc;
    if ((rc = w.fmomr.reserve(sizeof(double) * Bz * d.rowTiles * 2)) != EMSA_OK) return rc;
    if ((rc = w.fmomc.reserve(sizeof(double) * Bz * d.colBlocks * 12)) != EMSA_OK) return rc;
    d.gU = w.gU.as<float>();
    d.gV = w.gV.as<float>();
    d.gOr = w.gOr.as<float>();
    d.gOc = w.gOc.as<float>();
    d.L = w.L.as<float>();
    d.Q = w.Q.as<float>();
    d.zr = w.zr.as<float>();
    d.zc = w.zc.as<float>();
    d.gX = dgX;
    d.gY = dgY;
    d.adj = w.adj.as<double>();
    d.fcol = w.fcol.as<float>();
    d.frow = w.frow.as<float>();
    d.fmomr = w.fmomr.as<double>();
    d.fmomc = w.fmomc.as<double>();

    const dim3 blk(kWave * kWG);
    const unsigned ub = (unsigned)B;
    const int maxv = std::max({3 * N, 3 * M, d.ldk});
    hipLaunchKernelGGL(sau_adj_setup_kernel, dim3((unsigned)((maxv + 255) / 256), ub), dim3(256), 0,
                       s, d, dgRt);
    EMSA_HIP_TRY(hipGetLastError());
    const dim3 rgrid((unsigned)((N + kRowsPB - 1) / kRowsPB), ub);
    const dim3 cgrid((unsigned)(d.colBlocks * d.H), ub), fgrid((unsigned)d.colBlocks, ub);
    const dim3 tgrid((unsigned)(d.ldk / 256), ub);
    const dim3 mgrid((unsigned)((N + 255) / 256), ub);
    const dim3 pgrid((unsigned)(d.rowTiles * d.fchunks), ub);
    const dim3 frgrid((unsigned)d.rowTiles, ub), fcgrid((unsigned)d.colBlocks, ub);
    const std::vector<float> Ts = sau_temps(a);
    float xmax2 = 0.f;
    if ((rc = sau_xmax2(*wp, dX, B, N, s, xmax2)) != EMSA_OK) return rc;
    for (int it = d.nIt - 1; it >= 0; --it) {
        const float T = Ts[it];
        float k1, k0;
        sau_consts(T, alpha, k1, k0);
        const bool ex = k1 * xmax2 <= kExpandMax;  // the forward's form at this temperature
        const double fT = std::pow((double)TFACTOR, (double)(it / I0));
        hipLaunchKernelGGL(sau_transform_kernel, tgrid, dim3(256), 0, s, d, it);
        hipLaunchKernelGGL(sau_adj_solve_kernel, dim3(ub), dim3(64), 0, s, d, it);
        hipLaunchKernelGGL(sau_adj_mrow_kernel, mgrid, dim3(256), 0, s, d, it);
        sau_cols<4>(ex, cgrid, blk, s, d, (const float*)d.zr, 4 * Nz, k1, k0);
        hipLaunchKernelGGL(sau_colfin_kernel<kColBwdM>, fgrid, blk, 0, s, d, I1, it);
        for (int k = I1; k >= 1; --k) {
            sau_rows<kRowBwd>(ex, rgrid, blk, s, d, k, it, k1, k0);
            if (k > 1) {
                sau_cols<1>(ex, cgrid, blk, s, d, (const float*)d.zr, 4 * Nz, k1, k0);
                hipLaunchKernelGGL(sau_colfin_kernel<kColBwd>, fgrid, blk, 0, s, d, k - 1, it);
            }
        }
        switch (4 + 2 * I1) {
            case 6: hipLaunchKernelGGL(sau_final_kernel<6>, pgrid, blk, 0, s, d, k1, k0); break;
            case 8: hipLaunchKernelGGL(sau_final_kernel<8>, pgrid, blk, 0, s, d, k1, k0); break;
            case 10: hipLaunchKernelGGL(sau_final_kernel<10>, pgrid, blk, 0, s, d, k1, k0); break;
            default: hipLaunchKernelGGL(sau_final_kernel<12>, pgrid, blk, 0, s, d, k1, k0); break;
        }
        hipLaunchKernelGGL(sau_final_rows_kernel, frgrid, dim3(kFRows), 0, s, d, T);
        hipLaunchKernelGGL(sau_final_cols_kernel, fcgrid, dim3(256), 0, s, d, it, T);
        hipLaunchKernelGGL(sau_adj_finish_kernel, dim3(ub), dim3(64), 0, s, d, it, fT, T);
        EMSA_HIP_TRY(hipGetLastError());
    }
    const double dmo = defmo ? -(double)mo / (2.0 * (double)T_0) : 0.0;
    hipLaunchKernelGGL(sau_adj_output_kernel, dim3(ub), dim3(256), 0, s, d, dgRt0, dgT0, dmo);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // extern "C"
