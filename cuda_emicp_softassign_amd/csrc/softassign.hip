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
// Launches per inner iteration (I1 = 3: 11 log domain, 8 materialised), with
// no signalling between workgroups: row passes give each wave whole rows (its
// sums need no other workgroup); a column pass writes partial sums (over row
// chunks, or per workgroup of the one-pass materialised sweep) that a finalize
// kernel per 64 columns adds up; the final row pass writes per-workgroup
// FP64 moments that one solve workgroup reduces (fixed order) before Horn's
// step, and the next inner iteration's first row pass transforms the sources
// with the new (R, t) as it streams them.
// S is assembled from FP64 moments of the final pass:
//   D_r = D'_r - s_r Yc with D'_r = sum_c M_rc b_c y_c, and
//   S_ij = sum s x_i D'_j - Yc_j sum s^2 x_i - Xc_i sum D'_j + Xc_i Yc_j Sigma.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>

namespace emsa {

namespace {

constexpr int kSaWaves = 4;                  // waves per workgroup
constexpr int kSaRowsWG = 4;                 // rows per row-pass workgroup (4 measured faster than 8, 16)
constexpr int kSaFuseRows = 8;               // rows per workgroup of the one-pass materialised sweep
constexpr int kSaFuseWaves = 8;              // its waves (they split the columns)
constexpr int kSaFuseMaxSlabs = 3;           // column slabs of 256 kSaFuseWaves held in registers
constexpr int kSaFinalRows = 8;              // materialised final pass: rows per workgroup
constexpr int kSaFinalWaves = 8;             // and its waves
constexpr int kSaColsPB = 256;               // columns per column-pass workgroup (64 lanes x 4)
constexpr int kSaMomW = 22;                  // per-workgroup moments of the final pass
constexpr int kSaMomPad = 24;                // moment rows allocated (a multiple of kSaWaves)
constexpr float kSaSentinel = 1.0e15f;
constexpr float kSaMasked = -1.0e30f;        // exponent term of a padded row or column
constexpr int kSaMfItems = 4;                // rows (columns) per workgroup of a matrix-core pass
constexpr int kSaMfWaves = 4;                // its waves: they split the inner tiles
constexpr int kSaMfTile = 64;                // columns (rows) per matrix-core tile
constexpr int kSaMfU = 4;                    // tiles per wave and step (loads of the next step in flight)
constexpr int kSaMfStride = kSaMfWaves * kSaMfU;  // tiles per step of a workgroup
// the packs are padded (masked entries) to whole steps plus one step, so
// every load of a pass, the prefetch past the last step included, is in bounds
__host__ __device__ constexpr int sa_mf_tiles(int n) {
    return ((n + kSaMfTile - 1) / kSaMfTile + kSaMfStride - 1) / kSaMfStride * kSaMfStride;
}
__host__ __device__ constexpr int sa_mf_entries(int n) {
    return (sa_mf_tiles(n) + kSaMfStride) * kSaMfTile;
}
// The log domain evaluates the exponent expanded (a_r + b_c + x'_r . p_c)
// at the temperatures where k1 max_r |x_r|^2 is at most this: the expanded
// terms then round to about 1e-5 of the exponent at worst. Where the
// affinity is not negligible p_c lies near some x_r, so the target cloud's
// bound covers |p_c|^2 there. Colder steps use the direct difference.
constexpr float kSaExpandMax = 64.f;

struct SaPlan {
    int ldk = 0;        // padded column count (multiple of 256)
    int rowBlocks = 0;  // ceil(N / kSaRowsWG) row-pass workgroups
    int colBlocks = 0;  // ldk / 256
    int H = 0;          // row chunks per column block
    int chunkRows = 0;  // rows per chunk (multiple of kSaWaves)
    int fuseBlocks = 0; // ceil(N / kSaFuseRows) workgroups of the one-pass sweep
};

// Splits n rows into chunks so that colBlocks x chunks is about 4 workgroups
// per CU; chunk lengths are multiples of kSaWaves.
void split_rows(int n, int colBlocks, int numCU, int& H, int& chunkRows) {
    const int target = std::max(1, numCU) * 4;
    int h = std::max(1, (target + colBlocks - 1) / colBlocks);
    h = std::min(h, (n + kSaWaves - 1) / kSaWaves);
    chunkRows = ((n + h - 1) / h + kSaWaves - 1) / kSaWaves * kSaWaves;
    H = (n + chunkRows - 1) / chunkRows;
}

SaPlan plan_sa(int N, int M, int numCU) {
    SaPlan p;
    p.ldk = (M + kSaColsPB - 1) / kSaColsPB * kSaColsPB;
    p.rowBlocks = (N + kSaRowsWG - 1) / kSaRowsWG;
    p.colBlocks = p.ldk / kSaColsPB;
    split_rows(N, p.colBlocks, numCU, p.H, p.chunkRows);
    p.fuseBlocks = (N + kSaFuseRows - 1) / kSaFuseRows;
    return p;
}

struct SaBufs {
    const float* X;   // [3][N]
    const float* Y;   // [3][M]
    float* P;         // [3][ldk] transformed source points R y + t (written by the first row pass)
    float* Yp;        // [3][ldk] source points, zero past M (aligned vector loads)
    float* K;         // [N][ldk] (materialised only)
    float* u;         // [N] row scale (log: f = log2 u)
    float* v;         // [ldk] column scale (log: g = log2 v)
    float* orow;      // [N]
    float* ocol;      // [M]
    float* bcol;      // [ldk] last column sums b_c
    float* wfin;      // [4][ldk] final-pass column weights (scale, scale b y)
    float* cpart;     // [H][ldk] column partial sums (one-pass sweep: [fuseBlocks][ldk])
    double* rmom;     // [kSaMomPad][final-pass workgroups] moments of the final pass
    double* cby;      // [3][ldk / 64] sum_c b_c y_c per finalize block
    double* state;    // kStateW
    double* trace;    // nullable
    // nullable taps of one Sinkhorn sweep (emsa_softassign_step): the row sums
    // a_r [N] and column sums b_c [M] before normalisation, slack included, and
    // the post-Sinkhorn row sums s_r [N] of the final pass
    double* tapA;
    double* tapB;
    double* tapS;
    // log domain, expanded exponent on the matrix cores: per-column terms
    // {beta_c = k0 + g_c - k1 |p_c|^2, p_c} [ldk] for the row passes and
    // per-row terms {alpha'_r = k0 + f_r - k1 |x_r|^2, 2 k1 x_r} for the
    // column passes; both padded to whole steps of a pass (sa_mf_entries) with
    // masked entries (beta, alpha' = kSaMasked, the rest 0)
    float4* cpack;
    float4* rpack;
};

// The transformed sources p_c = R y_c + t (FP64, rounded) of the coming
// inner iteration with fresh column potentials (g = 0): P, and the column
// terms of the matrix-core row passes for exponent constants (k1, k0).
__device__ void sa_prep_columns(const SaBufs& b, int M, int ldk, float k1, float k0,
                                const double* Rm, const double* tm, int c0, int c1) {
    for (int c = c0 + (int)threadIdx.x; c < c1; c += blockDim.x) {
        float p[3];
        if (c < M) {
            const double y0 = b.Y[c], y1 = b.Y[(size_t)M + c], y2 = b.Y[2 * (size_t)M + c];
            for (int a = 0; a < 3; ++a)
                p[a] = (float)(Rm[3 * a] * y0 + Rm[3 * a + 1] * y1 + Rm[3 * a + 2] * y2 + tm[a]);
        } else {
            p[0] = p[1] = p[2] = kSaSentinel;
        }
        for (int a = 0; a < 3; ++a) b.P[(size_t)a * ldk + c] = p[a];
        if (b.cpack)
            b.cpack[c] = c < M ? make_float4(fmaf(-k1, fmaf(p[0], p[0], fmaf(p[1], p[1], p[2] * p[2])),
                                                  k0 + 0.f),
                                             p[0], p[1], p[2])
                               : make_float4(kSaMasked, 0.f, 0.f, 0.f);
    }
}

__global__ void __launch_bounds__(1024)
sa_setup_kernel(SaBufs b, int N, int M, int ldk, const float* __restrict__ Rt, float moutlier,
                int prep, float k1, float k0) {
    if (threadIdx.x < 12) b.state[threadIdx.x] = (double)Rt[threadIdx.x];
    if (threadIdx.x == 12) b.state[kStStatus] = 0.0;
    for (int r = threadIdx.x; r < N; r += blockDim.x) b.orow[r] = moutlier;
    for (int c = threadIdx.x; c < M; c += blockDim.x) b.ocol[c] = moutlier;
    for (int c = threadIdx.x; c < ldk; c += blockDim.x)
        for (int a = 0; a < 3; ++a) b.Yp[(size_t)a * ldk + c] = (c < M) ? b.Y[(size_t)a * M + c] : 0.f;
    if (b.rpack) {
        for (int r = N + threadIdx.x; r < sa_mf_entries(N); r += blockDim.x)
            b.rpack[r] = make_float4(kSaMasked, 0.f, 0.f, 0.f);
        for (int c = ldk + threadIdx.x; c < sa_mf_entries(M); c += blockDim.x)
            b.cpack[c] = make_float4(kSaMasked, 0.f, 0.f, 0.f);
    }
    if (prep) {
        double Rm[9], tm[3];
        for (int i = 0; i < 9; ++i) Rm[i] = (double)Rt[i];
        for (int i = 0; i < 3; ++i) tm[i] = (double)Rt[9 + i];
        sa_prep_columns(b, M, ldk, k1, k0, Rm, tm, 0, ldk);
    }
}

__global__ void sa_output_kernel(const double* __restrict__ state, float* __restrict__ Rt) {
    if (threadIdx.x < 12) Rt[threadIdx.x] = (float)state[threadIdx.x];
}

__device__ inline void load4(const float* p, float (&v)[4]) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}

// ---------------------------------------------------------------------------
// Row passes: a workgroup owns kSaRowsWG whole rows; its 4 waves take
// interleaved slabs of 256 columns (lanes: 4 consecutive columns each), so
// every load of the transformed sources P and the column scale serves all
// kSaRowsWG rows. Row sums: a DPP reduction per wave into lane 63, then the 4
// waves combine in LDS (fixed order). MODE 0: first row pass of an inner
// iteration (materialised: also writes K); MODE 1: Sinkhorn row pass; MODE 2:
// final pass (s_r, D'_r and the workgroup's FP64 moments).
// ---------------------------------------------------------------------------

template <bool LOG, int MODE, int R = kSaRowsWG, int NW = kSaWaves, bool EXP = false>
__global__ void __launch_bounds__(kWave * NW)
sa_rows_kernel(SaBufs b, int N, int M, int ldk, float k1, float k0) {
    constexpr int Q = (MODE == 2) ? 4 : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r0 = blockIdx.x * R;
    const int nr = min(R, N - r0);  // rows of this workgroup (uniform)

    float xr[R][3];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int r = r0 + min(j, nr - 1);
        xr[j][0] = b.X[r];
        xr[j][1] = b.X[(size_t)N + r];
        xr[j][2] = b.X[2 * (size_t)N + r];
    }
    // log domain, EXP: the exponent k0 + g_c - k1 |x_r - p_c|^2 as
    // a_r + b_c + x'_r . p_c with a_r = -k1 |x_r|^2, x'_r = 2 k1 x_r and
    // b_c = k0 + g_c - k1 |p_c|^2 (one add and 3 FMA per element)
    float ar[R], xq[R][3];
    if (LOG && EXP) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            ar[j] = -k1 * fmaf(xr[j][0], xr[j][0], fmaf(xr[j][1], xr[j][1], xr[j][2] * xr[j][2]));
#pragma unroll
            for (int a = 0; a < 3; ++a) xq[j][a] = 2.f * k1 * xr[j][a];
        }
    }
    float acc[R][Q];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[j][q] = 0.f;
    // MODE 0: the transformed sources P = R y + t of this inner iteration are
    // formed here (FP64, rounded to FP32) from the state the last solve left;
    // workgroup 0 stores them for the later passes
    double Rm[9], tm[3];
    if (MODE == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) Rm[i] = b.state[kStR + i];
#pragma unroll
        for (int i = 0; i < 3; ++i) tm[i] = b.state[kStT + i];
    }

    for (int c = 4 * lane + 4 * kWave * wave; c < ldk; c += 4 * kWave * NW) {
        float w[4][Q];  // per-column weights
        float p[3][4];
        float kg[4];
        if (MODE == 0) {
            float y[3][4];
            load4(b.Yp + c, y[0]);
            load4(b.Yp + (size_t)ldk + c, y[1]);
            load4(b.Yp + 2 * (size_t)ldk + c, y[2]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double y0 = y[0][i], y1 = y[1][i], y2 = y[2][i];
                const bool in = c + i < M;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    p[a][i] = in ? (float)(Rm[3 * a] * y0 + Rm[3 * a + 1] * y1 + Rm[3 * a + 2] * y2 +
                                           tm[a])
                                 : kSaSentinel;
            }
            if (blockIdx.x == 0)
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    *reinterpret_cast<float4*>(b.P + (size_t)a * ldk + c) =
                        make_float4(p[a][0], p[a][1], p[a][2], p[a][3]);
        } else if (LOG) {
            load4(b.P + c, p[0]);
            load4(b.P + (size_t)ldk + c, p[1]);
            load4(b.P + 2 * (size_t)ldk + c, p[2]);
        }
        // column scale of a fresh inner iteration: v = 1 (log: g = 0)
        float vv[4] = {LOG ? 0.f : 1.f, LOG ? 0.f : 1.f, LOG ? 0.f : 1.f, LOG ? 0.f : 1.f};
        if (MODE != 0) load4(b.v + c, vv);
#pragma unroll
        for (int i = 0; i < 4; ++i) kg[i] = LOG ? (k0 + vv[i]) : k0;  // log: v holds g_c
        float bc[4];
        if (LOG && EXP)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                bc[i] = fmaf(-k1, fmaf(p[0][i], p[0][i], fmaf(p[1][i], p[1][i], p[2][i] * p[2][i])),
                             kg[i]);
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i][0] = LOG ? 1.f : vv[i];
        if (MODE == 2) {  // precomputed by the last column finalize
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                float ww[4];
                load4(b.wfin + (size_t)q * ldk + c, ww);
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i][q] = ww[i];
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j >= nr) continue;  // uniform
            const int r = r0 + j;
            float val[4];
            if (!LOG && MODE != 0) {
                load4(b.K + (size_t)r * ldk + c, val);
            } else if (LOG && EXP) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float t = ar[j] + bc[i];
                    t = fmaf(xq[j][0], p[0][i], t);
                    t = fmaf(xq[j][1], p[1][i], t);
                    t = fmaf(xq[j][2], p[2][i], t);
                    val[i] = __builtin_amdgcn_exp2f(t);
                }
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
    // row sums: lane 63 of each wave, then the waves in a fixed order
    __shared__ float red[NW][R * Q];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float t = wave_sum_last(acc[j][q]);
            if (lane == kWave - 1) red[wave][j * Q + q] = t;
        }
    __syncthreads();
    __shared__ float tot[R * Q];
    if (threadIdx.x < R * Q) {
        float v = red[0][threadIdx.x];
#pragma unroll
        for (int w2 = 1; w2 < NW; ++w2) v += red[w2][threadIdx.x];
        tot[threadIdx.x] = v;
    }
    __syncthreads();

    if (MODE != 2) {
        // a_r = (row scale) * sum + o_row[r]; scale /= a_r; o_row /= a_r
        const int j = threadIdx.x;
        if (j < nr) {
            const int r = r0 + j;
            const float su = (MODE == 0) ? (LOG ? 0.f : 1.f) : b.u[r];
            const float scale = LOG ? exp2f(su) : su;
            const float a = scale * tot[j] + b.orow[r];
            b.orow[r] = b.orow[r] / a;
            b.u[r] = LOG ? (su - log2f(a)) : (su / a);
            if (b.tapA) b.tapA[r] = (double)a;
        }
        return;
    }

    // MODE 2: FP64 moments of this workgroup's rows: lane j < nr forms row
    // j's terms, then each moment is summed over the rows in row order
    //  [0] s  [1..3] s x  [4 + 3i + j] s x_i D'_j  [13..15] s^2 x  [16..18] D'
    __shared__ double mm[R][kSaMomW];
    if (threadIdx.x < nr) {
        const int j = threadIdx.x;
        const int r = r0 + j;
        const float su = b.u[r];
        const double scale = LOG ? (double)exp2f(su) : (double)su;
        const double s = scale * tot[j * Q];
        const double D[3] = {scale * tot[j * Q + 1], scale * tot[j * Q + 2],
                             scale * tot[j * Q + 3]};
        if (b.tapS) b.tapS[r] = s;
        const double x[3] = {b.X[r], b.X[(size_t)N + r], b.X[2 * (size_t)N + r]};
        mm[j][0] = s;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            mm[j][1 + i] = s * x[i];
            mm[j][13 + i] = s * s * x[i];
            mm[j][16 + i] = D[i];
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) mm[j][4 + 3 * i + jj] = s * x[i] * D[jj];
        }
    }
    __syncthreads();
    if (threadIdx.x < kSaMomW) {
        const int q = threadIdx.x;
        double acc = 0.0;
        for (int j = 0; j < nr; ++j) acc += mm[j][q];
        b.rmom[(size_t)q * gridDim.x + blockIdx.x] = acc;
    }
}

// ---------------------------------------------------------------------------
// Materialised variant, one Sinkhorn sweep in one pass over K: a workgroup
// of kSaFuseWaves waves owns kSaFuseRows whole rows and keeps their K values
// in registers (S slabs of 256 kSaFuseWaves columns: lane 4 columns per slab
// and row, ldk <= S * 256 kSaFuseWaves). The row
// sums with the current column scale give the new row scales (as the row
// pass does), then the column partial sums over the workgroup's rows with the
// new scales go to cpart[blockIdx.x] (as the column pass does over its row
// chunk); the column finalize adds these fuseBlocks partial rows up. K is
// read once per sweep
// instead of twice. FIRST: the first sweep of an inner
// iteration, which forms K from the transformed sources, stores it, and
// starts from a fresh column scale v = 1.
// ---------------------------------------------------------------------------
template <bool FIRST, int S>
__global__ void __launch_bounds__(kWave * kSaFuseWaves) __attribute__((amdgpu_waves_per_eu(4)))
sa_rowcol_kernel(SaBufs b, int N, int M, int ldk, float k1, float k0) {
    constexpr int R = kSaFuseRows;
    constexpr int NW = kSaFuseWaves;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r0 = blockIdx.x * R;
    const int nr = min(R, N - r0);  // rows of this workgroup (uniform)

    float xr[R][3];
    double Rm[9], tm[3];
    if (FIRST) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int r = r0 + min(j, nr - 1);
            xr[j][0] = b.X[r];
            xr[j][1] = b.X[(size_t)N + r];
            xr[j][2] = b.X[2 * (size_t)N + r];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) Rm[i] = b.state[kStR + i];
#pragma unroll
        for (int i = 0; i < 3; ++i) tm[i] = b.state[kStT + i];
    }
    float kv[S][R][4];
    float acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.f;
#pragma unroll
    for (int si = 0; si < S; ++si) {
        const int c = 4 * lane + 4 * kWave * (wave + NW * si);
        if (c >= ldk) {  // wave-uniform (ldk is a multiple of 256)
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) kv[si][j][i] = 0.f;
            continue;
        }
        float vv[4] = {1.f, 1.f, 1.f, 1.f};
        float p[3][4];
        if (FIRST) {
            float y[3][4];
            load4(b.Yp + c, y[0]);
            load4(b.Yp + (size_t)ldk + c, y[1]);
            load4(b.Yp + 2 * (size_t)ldk + c, y[2]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double y0 = y[0][i], y1 = y[1][i], y2 = y[2][i];
                const bool in = c + i < M;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    p[a][i] = in ? (float)(Rm[3 * a] * y0 + Rm[3 * a + 1] * y1 + Rm[3 * a + 2] * y2 +
                                           tm[a])
                                 : kSaSentinel;
            }
        } else {
            load4(b.v + c, vv);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            float val[4] = {0.f, 0.f, 0.f, 0.f};
            if (j < nr) {  // uniform
                const int r = r0 + j;
                if (FIRST) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float dx = xr[j][0] - p[0][i];
                        const float dy = xr[j][1] - p[1][i];
                        const float dz = xr[j][2] - p[2][i];
                        float sq = dx * dx;
                        sq = fmaf(dy, dy, sq);
                        sq = fmaf(dz, dz, sq);
                        val[i] = __builtin_amdgcn_exp2f(fmaf(sq, -k1, k0));
                    }
                    *reinterpret_cast<float4*>(b.K + (size_t)r * ldk + c) =
                        make_float4(val[0], val[1], val[2], val[3]);
                } else {
                    load4(b.K + (size_t)r * ldk + c, val);
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                kv[si][j][i] = val[i];
                acc[j] = fmaf(val[i], vv[i], acc[j]);
            }
        }
    }
    // row sums: lane 63 of each wave, then the waves in a fixed order;
    // a_r = (row scale) * sum + o_row[r]; scale /= a_r; o_row /= a_r
    __shared__ float red[NW][R];
    __shared__ float unew[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const float t = wave_sum_last(acc[j]);
        if (lane == kWave - 1) red[wave][j] = t;
    }
    __syncthreads();
    if (threadIdx.x < nr) {
        const int j = threadIdx.x;
        const int r = r0 + j;
        float tot = red[0][j];
#pragma unroll
        for (int w2 = 1; w2 < NW; ++w2) tot += red[w2][j];
        const float su = FIRST ? 1.f : b.u[r];
        const float a = su * tot + b.orow[r];
        if (b.tapA) b.tapA[r] = (double)a;
        b.orow[r] = b.orow[r] / a;
        const float un = su / a;
        b.u[r] = un;
        unew[j] = un;
    }
    __syncthreads();
    float uu[R];
#pragma unroll
    for (int j = 0; j < R; ++j) uu[j] = (j < nr) ? unew[j] : 0.f;
    // column partial sums over this workgroup's rows with the new row scales
    float* cp = b.cpart + (size_t)blockIdx.x * ldk;
#pragma unroll
    for (int si = 0; si < S; ++si) {
        const int c = 4 * lane + 4 * kWave * (wave + NW * si);
        if (c >= ldk) continue;
        float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < R; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) part[i] = fmaf(kv[si][j][i], uu[j], part[i]);
        *reinterpret_cast<float4*>(cp + c) = make_float4(part[0], part[1], part[2], part[3]);
    }
}

// ---------------------------------------------------------------------------
// Column pass, stage 1: a workgroup = 256 columns x one row chunk; lanes own
// 4 consecutive columns, the 4 waves split the chunk's rows.
// ---------------------------------------------------------------------------
template <bool LOG, bool EXP = false>
__global__ void __launch_bounds__(kWave * kSaWaves)
sa_cols_kernel(SaBufs b, int N, int ldk, int H, int chunkRows, float k1, float k0) {
    const int cbk = blockIdx.x / H;
    const int h = blockIdx.x - cbk * H;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int c = cbk * kSaColsPB + 4 * lane;
    const int q = chunkRows / kSaWaves;
    const int rs = h * chunkRows + wave * q;
    const int re = min(rs + q, N);

    float p[3][4], pc[4];
    if (LOG) {
        load4(b.P + c, p[0]);
        load4(b.P + (size_t)ldk + c, p[1]);
        load4(b.P + 2 * (size_t)ldk + c, p[2]);
        // EXP: the exponent k0 + f_r - k1 |x_r - p_c|^2 expanded as in the
        // row passes: (k0 + f_r - k1 |x_r|^2) + (-k1 |p_c|^2) + 2 k1 x_r . p_c
        if (EXP)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pc[i] = -k1 * fmaf(p[0][i], p[0][i], fmaf(p[1][i], p[1][i], p[2][i] * p[2][i]));
    }
    const float k2 = 2.f * k1;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    // 4 rows per trip: their scalar loads (x_r, f_r) are issued together
    // instead of one wait per row (log domain at 5032^2: 2% per registration)
#pragma unroll 4
    for (int r = rs; r < re; ++r) {
        if (LOG && EXP) {
            const float x0 = b.X[r], x1 = b.X[(size_t)N + r], x2 = b.X[2 * (size_t)N + r];
            const float a = fmaf(-k1, fmaf(x0, x0, fmaf(x1, x1, x2 * x2)), k0 + b.u[r]);
            const float q0 = k2 * x0, q1 = k2 * x1, q2 = k2 * x2;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float t = a + pc[i];
                t = fmaf(q0, p[0][i], t);
                t = fmaf(q1, p[1][i], t);
                t = fmaf(q2, p[2][i], t);
                acc[i] += __builtin_amdgcn_exp2f(t);
            }
        } else if (LOG) {
            const float x0 = b.X[r], x1 = b.X[(size_t)N + r], x2 = b.X[2 * (size_t)N + r];
            const float kf = k0 + b.u[r];  // f_r
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float dx = x0 - p[0][i], dy = x1 - p[1][i], dz = x2 - p[2][i];
                float s = dx * dx;
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                acc[i] += __builtin_amdgcn_exp2f(fmaf(s, -k1, kf));
            }
        } else {
            const float ur = b.u[r];
            float kv[4];
            load4(b.K + (size_t)r * ldk + c, kv);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] = fmaf(kv[i], ur, acc[i]);
        }
    }
    __shared__ float red[kSaWaves][kSaColsPB];
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][4 * lane + i] = acc[i];
    __syncthreads();
    const int tid = threadIdx.x;
    float s = red[0][tid];
#pragma unroll
    for (int w = 1; w < kSaWaves; ++w) s += red[w][tid];
    b.cpart[(size_t)h * ldk + cbk * kSaColsPB + tid] = s;
}

// Column pass, stage 2: a workgroup finalizes 64 columns; its NWF waves sum
// interleaved partial rows in FP64 (U partial sums per lane, so U loads are
// in flight, added pairwise) and combine in LDS (fixed order). Then
// b_c = (column scale) * sum + o_col[c]; scale /= b_c; o_col /= b_c; b_c kept;
// sum_c b_c y_c of the block for Yc. After the last sweep (last != 0) also the
// column weights of the final row pass. Log domain: the H row-chunk partials
// of the column pass (4 waves, U = 4); materialised one-pass sweep: the
// partial row of every fused-pass workgroup (N / 8 rows: 16 waves, U = 8).
constexpr int kSaFinCols = kWave;
constexpr int kSaFinWavesWide = 16;

template <bool LOG, bool FIRST, int NWF = kSaWaves, int U = 4>
__global__ void __launch_bounds__(kWave * NWF)
sa_colfin_kernel(SaBufs b, int M, int ldk, int H, int last) {
    static_assert(U == 4 || U == 8, "pairwise combine of 4 or 8 partial sums");
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int cc = blockIdx.x * kSaFinCols + lane;
    double pa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) pa[u] = 0.0;
    int k = wave;
    for (; k + (U - 1) * NWF < H; k += U * NWF)
#pragma unroll
        for (int u = 0; u < U; ++u) pa[u] += b.cpart[(size_t)(k + u * NWF) * ldk + cc];
    for (; k < H; k += NWF) pa[0] += b.cpart[(size_t)k * ldk + cc];
    double part = (pa[0] + pa[1]) + (pa[2] + pa[3]);
    if (U == 8) part += (pa[4 % U] + pa[5 % U]) + (pa[6 % U] + pa[7 % U]);
    __shared__ double red[NWF][kSaFinCols];
    red[wave][lane] = part;
    __syncthreads();
    if (wave != 0) return;
    double tot = red[0][lane];
#pragma unroll
    for (int w = 1; w < NWF; ++w) tot += red[w][lane];
    double by[3] = {0.0, 0.0, 0.0};
    float w0 = 0.f, wy[3] = {0.f, 0.f, 0.f};
    if (cc < M) {
        const float sv = FIRST ? (LOG ? 0.f : 1.f) : b.v[cc];
        const float scale = LOG ? exp2f(sv) : sv;
        const float bc = (float)(scale * tot) + b.ocol[cc];
        if (b.tapB) b.tapB[cc] = (double)bc;
        b.ocol[cc] = b.ocol[cc] / bc;
        const float vn = LOG ? (sv - log2f(bc)) : (sv / bc);
        b.v[cc] = vn;
        b.bcol[cc] = bc;
        const float y0 = b.Y[cc], y1 = b.Y[(size_t)M + cc], y2 = b.Y[2 * (size_t)M + cc];
        by[0] = (double)bc * y0;
        by[1] = (double)bc * y1;
        by[2] = (double)bc * y2;
        // final pass: M_rc b_c y_c = K u_r (v_c b_c) y_c; log-domain keeps v in the exponent
        w0 = LOG ? 1.f : vn;
        const float base = w0 * bc;
        wy[0] = base * y0;
        wy[1] = base * y1;
        wy[2] = base * y2;
    } else {
        b.v[cc] = 0.f;
        b.bcol[cc] = 0.f;
    }
    if (last) {
        b.wfin[cc] = w0;
        b.wfin[(size_t)ldk + cc] = wy[0];
        b.wfin[2 * (size_t)ldk + cc] = wy[1];
        b.wfin[3 * (size_t)ldk + cc] = wy[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) by[i] = wave_sum(by[i]);
    if (lane < 3)
        b.cby[(size_t)lane * gridDim.x + blockIdx.x] = (lane == 0) ? by[0] : (lane == 1 ? by[1] : by[2]);
}

// Totals of the per-workgroup moments and of sum b y, S, Horn and the new
// (R, t) in the state (workgroup 0). With prep, every workgroup of the launch
// computes the same totals and (R, t) (same loads, same order: the same bits)
// and prepares its own slice of kWave * kSaSolveWaves columns for the next
// inner iteration's matrix-core passes; without, one workgroup, and the next
// inner iteration's first row pass transforms the sources itself. Wave w sums quantities q = w % 4 + 4 k over
// the (w / 4)-th of kSaSolveWaves / 4 ranges of the partials (lanes striding,
// four partial sums per lane so the loads are in flight together), then a
// wave sum; the ranges are added in a fixed order in LDS.
constexpr int kSaSolveWaves = 16;

__global__ void __launch_bounds__(kWave * kSaSolveWaves)
sa_solve_kernel(SaBufs b, int nmom, int nfin, float T, int iter, int M, int ldk, int prep,
                float k1n, float k0n) {
    constexpr int kParts = kSaSolveWaves / kSaWaves;
    static_assert(kParts * kSaWaves == kSaSolveWaves, "whole quantity classes per range");
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int qc = wave % kSaWaves;
    const int part = wave / kSaWaves;
    const int len = (nmom + kParts - 1) / kParts;
    const int wbeg = part * len, wend = min(wbeg + len, nmom);
    // rows up to kSaMomPad - 1 exist (the ones past kSaMomW are not used):
    // no per-quantity branch. Each lane loads kU partials of each of its kPer
    // quantities in one round (predicated, all in flight together: one round
    // trip for up to kU * 64 partials per range), then adds them in a fixed
    // order.
    constexpr int kPer = kSaMomPad / kSaWaves;
    constexpr int kU = 5;
    double acc[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
    for (int w0 = wbeg; w0 < wend; w0 += kU * kWave) {
        double v[kU][kPer];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int w = w0 + u * kWave + lane;
#pragma unroll
            for (int k = 0; k < kPer; ++k)
                v[u][k] = (w < wend) ? b.rmom[(size_t)(qc + kSaWaves * k) * nmom + w] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            acc[k] += ((v[0][k] + v[1][k]) + (v[2][k] + v[3][k])) + v[4][k];
    }
    // component qc < 3 of sum b y: its partials split over the kParts ranges
    // like the moments
    double by = 0.0;
    if (qc < 3) {
        const int blen = (nfin + kParts - 1) / kParts;
        const int bbeg = part * blen, bend = min(bbeg + blen, nfin);
        for (int c0 = bbeg; c0 < bend; c0 += kU * kWave) {
            double v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int c = c0 + u * kWave + lane;
                v[u] = (c < bend) ? b.cby[(size_t)qc * nfin + c] : 0.0;
            }
            by += ((v[0] + v[1]) + (v[2] + v[3])) + v[4];
        }
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = wave_sum(acc[k]);
    by = wave_sum(by);
    __shared__ double red[kParts][kSaMomPad];
    __shared__ double rby[kParts][3];
    __shared__ double mt[kSaMomW + 3];
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) red[part][qc + kSaWaves * k] = acc[k];
        if (qc < 3) rby[part][qc] = by;
    }
    __syncthreads();
    if (tid < kSaMomW) {
        double t = red[0][tid];
#pragma unroll
        for (int pp = 1; pp < kParts; ++pp) t += red[pp][tid];
        mt[tid] = t;
    } else if (tid < kSaMomW + 3) {
        double t = rby[0][tid - kSaMomW];
#pragma unroll
        for (int pp = 1; pp < kParts; ++pp) t += rby[pp][tid - kSaMomW];
        mt[tid] = t;
    }
    __syncthreads();
    __shared__ double rtNew[12];  // (R, t) after this solve, for the column slice
    if (tid == 0) {
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
        const bool owner = blockIdx.x == 0;  // the one that writes the state and trace
        if (ok) {
            if (owner) {
                for (int i = 0; i < 9; ++i) b.state[kStR + i] = Rn[i];
                for (int i = 0; i < 3; ++i) b.state[kStT + i] = tn[i];
            }
        } else {  // (R, t) stay: no workgroup writes them on this path
            if (owner) b.state[kStStatus] += 1.0;
            for (int i = 0; i < 9; ++i) Rn[i] = b.state[kStR + i];
            for (int i = 0; i < 3; ++i) tn[i] = b.state[kStT + i];
        }
        for (int i = 0; i < 9; ++i) rtNew[i] = Rn[i];
        for (int i = 0; i < 3; ++i) rtNew[9 + i] = tn[i];
        if (b.trace && owner) {
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
    // the next inner iteration's transformed sources and column terms, for
    // the matrix-core passes (exponent constants of its temperature)
    if (prep) {
        __syncthreads();
        double Rm[9], tm[3];
        for (int i = 0; i < 9; ++i) Rm[i] = rtNew[i];
        for (int i = 0; i < 3; ++i) tm[i] = rtNew[9 + i];
        const int per = kWave * kSaSolveWaves;
        const int c0 = (int)blockIdx.x * per;
        sa_prep_columns(b, M, ldk, k1n, k0n, Rm, tm, c0, min(ldk, c0 + per));
    }
}

// ---------------------------------------------------------------------------
// Log domain with the expanded exponent on the matrix cores. The exponent of
// element (r, c) is a rank-5 product,
//   E_rc = alpha_r * 1 + 1 * beta_c + (2 k1 x_r) . p_c,
// evaluated with five v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4, contraction
// depth 1 each, accumulator chained), in the order the VALU form adds the
// terms. A wave owns kSaMfItems = 4 outer items (rows in a row pass, columns
// in a column pass) and streams every inner item (columns / rows) in tiles of
// 64, one inner item per lane: lane l supplies the inner item's terms (A
// operand), lane 4b + j the outer item j's (B operand, the same for every b),
// and lane 4b + j receives E for inner items 4b..4b+3 of the tile and outer
// item j. Each lane adds exp2(E) of its 4 elements (one FP64 sum per tile),
// and the 16 lanes of an outer item are added at the end, so a wave holds
// complete row (column) sums: the normalisation, the slack update and the
// terms the next pass needs are written by the wave itself, with no partial
// sums and no finalize kernel.
// ---------------------------------------------------------------------------
typedef float sa_f4 __attribute__((ext_vector_type(4)));

// E for one tile: inner terms (ia: alpha or 1, ib: 1 or beta, ip: 3), outer (oa, ob, op: 3)
__device__ inline sa_f4 sa_exponent(float ia, float ib, const float (&ip)[3], float oa, float ob,
                                    const float (&op)[3]) {
    sa_f4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_4x4x1f32(ia, oa, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x1f32(ib, ob, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x1f32(ip[0], op[0], d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x1f32(ip[1], op[1], d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x1f32(ip[2], op[2], d, 0, 0, 0);
    return d;
}

// Sum over the 16 lanes holding the same outer item (lane bits 2..5).
__device__ inline double sa_item_sum(double v) {
#pragma unroll
    for (int m = 4; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Row passes (MODE 0: first of an inner iteration, 1: Sinkhorn). Columns in
// tiles of 64 from cpack. The final pass (its four weighted sums per element
// and the moments) is the VALU row pass, which reads the same P, v and
// weights.
template <int MODE>
__global__ void __launch_bounds__(kWave * kSaMfWaves)
sa_mrows_kernel(SaBufs b, int N, int M, int ldk, float k1, float k0) {
    static_assert(MODE == 0 || MODE == 1, "Sinkhorn row passes; the final pass is the VALU one");
    constexpr int U = kSaMfU;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int j = lane & (kSaMfItems - 1);
    const int r0 = blockIdx.x * kSaMfItems;
    const int r = r0 + j;
    // outer terms of row r: alpha_r = -k1 |x_r|^2, 1, 2 k1 x_r (masked past N)
    float x[3] = {0.f, 0.f, 0.f};
    if (r < N) {
        x[0] = b.X[r];
        x[1] = b.X[(size_t)N + r];
        x[2] = b.X[2 * (size_t)N + r];
    }
    const float oa = (r < N) ? -k1 * fmaf(x[0], x[0], fmaf(x[1], x[1], x[2] * x[2])) : kSaMasked;
    const float op[3] = {2.f * k1 * x[0], 2.f * k1 * x[1], 2.f * k1 * x[2]};
    // the epilogue's operands, loaded now so that they arrive during the loop
    const bool fin = threadIdx.x < kSaMfItems && r < N;
    const float su = (fin && MODE != 0) ? b.u[r] : 0.f;
    const float orow = fin ? b.orow[r] : 0.f;
    double acc = 0.0;
    // wave w takes tiles w, w + W, ... in steps of U tiles; the next step's
    // loads (into the other buffer) are issued before the current step's
    // arithmetic. The packs are padded to whole steps, so no load is guarded.
    const int tiles = (r0 < N) ? sa_mf_tiles(M) : 0;  // block-uniform
    auto load = [&](int t0, float4 (&cp)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) cp[u] = b.cpack[(size_t)(t0 + u * kSaMfWaves) * kSaMfTile + lane];
    };
    auto step = [&](const float4 (&cp)[U]) {
        sa_f4 d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float ip[3] = {cp[u].y, cp[u].z, cp[u].w};
            d[u] = sa_exponent(1.f, cp[u].x, ip, oa, 1.f, op);
        }
        float sq = 0.f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float e[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) e[i] = __builtin_amdgcn_exp2f(d[u][i]);
            sq += (e[0] + e[1]) + (e[2] + e[3]);
        }
        acc += (double)sq;
    };
    float4 ca[U], cb[U];
    if (tiles > 0) load(wave, ca);
    for (int t0 = wave; t0 < tiles; t0 += 2 * kSaMfStride) {
        load(t0 + kSaMfStride, cb);
        step(ca);
        if (t0 + kSaMfStride >= tiles) break;  // uniform
        load(t0 + 2 * kSaMfStride, ca);
        step(cb);
    }
    // the 16 lanes of each row, then the waves in a fixed order
    __shared__ double red[kSaMfWaves][kSaMfItems];
    acc = sa_item_sum(acc);
    if (lane < kSaMfItems) red[wave][j] = acc;
    __syncthreads();
    if (threadIdx.x >= kSaMfItems) return;  // thread j < 4 (wave 0) finishes row r
    double tot = red[0][j];
#pragma unroll
    for (int w2 = 1; w2 < kSaMfWaves; ++w2) tot += red[w2][j];
    // a_r = (row scale) * sum + o_row[r]; scale /= a_r; o_row /= a_r; and the
    // row terms of the column passes
    if (r < N) {
        const float scale = exp2f(su);
        const float a = scale * (float)tot + orow;
        b.orow[r] = orow / a;
        const float un = su - log2f(a);
        b.u[r] = un;
        if (b.tapA) b.tapA[r] = (double)a;
        b.rpack[r] = make_float4(fmaf(-k1, fmaf(x[0], x[0], fmaf(x[1], x[1], x[2] * x[2])), k0 + un),
                                 op[0], op[1], op[2]);
    }
}

// Column passes (FIRST: the first sweep of an inner iteration, fresh column
// scale). Rows in tiles of 64 from rpack. The workgroup's 4 columns get
// b_c = (column scale) * sum + o_col[c], the new potential, the column terms
// of the next row pass, and after the last sweep the final pass's weights;
// the workgroup's sum_c b_c y_c (columns in order) goes to cby.
template <bool FIRST>
__global__ void __launch_bounds__(kWave * kSaMfWaves)
sa_mcols_kernel(SaBufs b, int N, int M, int ldk, float k1, float k0, int last) {
    constexpr int U = kSaMfU;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int j = lane & (kSaMfItems - 1);
    const int quad = lane >> 2;
    const int c0 = blockIdx.x * kSaMfItems;
    const int c = c0 + j;
    // outer terms of column c: 1, beta'_c = -k1 |p_c|^2, p_c (masked past M)
    const float p[3] = {b.P[c], b.P[(size_t)ldk + c], b.P[2 * (size_t)ldk + c]};
    const float ob =
        (c < M) ? -k1 * fmaf(p[0], p[0], fmaf(p[1], p[1], p[2] * p[2])) : kSaMasked;
    const float op[3] = {c < M ? p[0] : 0.f, c < M ? p[1] : 0.f, c < M ? p[2] : 0.f};
    // the epilogue's operands, loaded now so that they arrive during the loop
    const bool fin = threadIdx.x < kSaMfItems && c < M;
    const float sv = (fin && !FIRST) ? b.v[c] : 0.f;
    const float ocol = fin ? b.ocol[c] : 0.f;
    const float y[3] = {fin ? b.Y[c] : 0.f, fin ? b.Y[(size_t)M + c] : 0.f,
                        fin ? b.Y[2 * (size_t)M + c] : 0.f};
    double acc = 0.0;
    const int tiles = (c0 < M) ? sa_mf_tiles(N) : 0;  // block-uniform
    auto load = [&](int t0, float4 (&rp)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) rp[u] = b.rpack[(size_t)(t0 + u * kSaMfWaves) * kSaMfTile + lane];
    };
    auto step = [&](const float4 (&rp)[U]) {
        sa_f4 d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float ip[3] = {rp[u].y, rp[u].z, rp[u].w};
            // beta'_c + alpha'_r (equal to alpha'_r + beta'_c; the first
            // product is the same for every tile)
            d[u] = sa_exponent(1.f, rp[u].x, ip, ob, 1.f, op);
        }
        float sq = 0.f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float e[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) e[i] = __builtin_amdgcn_exp2f(d[u][i]);
            sq += (e[0] + e[1]) + (e[2] + e[3]);
        }
        acc += (double)sq;
    };
    float4 ra[U], rb[U];
    if (tiles > 0) load(wave, ra);
    for (int t0 = wave; t0 < tiles; t0 += 2 * kSaMfStride) {
        load(t0 + kSaMfStride, rb);
        step(ra);
        if (t0 + kSaMfStride >= tiles) break;  // uniform
        load(t0 + 2 * kSaMfStride, ra);
        step(rb);
    }
    __shared__ double red[kSaMfWaves][kSaMfItems];
    acc = sa_item_sum(acc);
    if (quad == 0) red[wave][j] = acc;
    __syncthreads();
    if (threadIdx.x >= kSaMfItems) return;  // thread j < 4 finishes column c
    double tot = red[0][j];
#pragma unroll
    for (int w2 = 1; w2 < kSaMfWaves; ++w2) tot += red[w2][j];
    double by[3] = {0.0, 0.0, 0.0};
    if (c < M) {
        const float scale = exp2f(sv);
        const float bc = (float)(scale * tot) + ocol;
        if (b.tapB) b.tapB[c] = (double)bc;
        b.ocol[c] = ocol / bc;
        const float vn = sv - log2f(bc);
        b.v[c] = vn;
        b.bcol[c] = bc;
        const float y0 = y[0], y1 = y[1], y2 = y[2];
        by[0] = (double)bc * y0;
        by[1] = (double)bc * y1;
        by[2] = (double)bc * y2;
        if (last) {
            b.wfin[c] = 1.f;
            b.wfin[(size_t)ldk + c] = bc * y0;
            b.wfin[2 * (size_t)ldk + c] = bc * y1;
            b.wfin[3 * (size_t)ldk + c] = bc * y2;
        }
        b.cpack[c] = make_float4(fmaf(-k1, fmaf(p[0], p[0], fmaf(p[1], p[1], p[2] * p[2])),
                                      k0 + vn),
                                 p[0], p[1], p[2]);
    } else {
        b.v[c] = 0.f;
        b.bcol[c] = 0.f;
        if (last)
            for (int q = 0; q < 4; ++q) b.wfin[(size_t)q * ldk + c] = 0.f;
        b.cpack[c] = make_float4(kSaMasked, 0.f, 0.f, 0.f);
    }
    // columns j = 0..3 are lanes 0..3 of wave 0: add in order
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double b1 = __shfl(by[i], 1, kWave), b2 = __shfl(by[i], 2, kWave),
                     b3 = __shfl(by[i], 3, kWave);
        if (j == 0) b.cby[(size_t)i * gridDim.x + blockIdx.x] = ((by[i] + b1) + b2) + b3;
    }
}

template <int S>
bool launch_rowcol_upto(int slabs, bool first, dim3 grid, dim3 blk, hipStream_t s,
                        const SaBufs& b, int N, int M, int ldk, float k1, float k0) {
    if (slabs == S) {
        if (first)
            hipLaunchKernelGGL((sa_rowcol_kernel<true, S>), grid, blk, 0, s, b, N, M, ldk, k1, k0);
        else
            hipLaunchKernelGGL((sa_rowcol_kernel<false, S>), grid, blk, 0, s, b, N, M, ldk, k1, k0);
        return true;
    }
    if constexpr (S < kSaFuseMaxSlabs)
        return launch_rowcol_upto<S + 1>(slabs, first, grid, blk, s, b, N, M, ldk, k1, k0);
    return false;
}

// Log-domain row pass of the given mode, expanded exponent or direct difference.
template <bool EXP>
void launch_log_rows(int mode, dim3 grid, dim3 blk, hipStream_t s, const SaBufs& b, int N, int M,
                     int ldk, float k1, float k0) {
    if (mode == 0)
        hipLaunchKernelGGL((sa_rows_kernel<true, 0, kSaRowsWG, kSaWaves, EXP>), grid, blk, 0, s, b,
                           N, M, ldk, k1, k0);
    else if (mode == 1)
        hipLaunchKernelGGL((sa_rows_kernel<true, 1, kSaRowsWG, kSaWaves, EXP>), grid, blk, 0, s, b,
                           N, M, ldk, k1, k0);
    else
        hipLaunchKernelGGL((sa_rows_kernel<true, 2, kSaRowsWG, kSaWaves, EXP>), grid, blk, 0, s, b,
                           N, M, ldk, k1, k0);
}

int launch_rowcol(bool first, int slabs, dim3 grid, dim3 blk, hipStream_t s, const SaBufs& b,
                  int N, int M, int ldk, float k1, float k0) {
    if (!launch_rowcol_upto<1>(slabs, first, grid, blk, s, b, N, M, ldk, k1, k0))
        return set_error("softassign: bad slab count"), EMSA_EINVAL;
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // namespace

int softassign_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                      const float* dRtIn, float* dRtOut, const emsa_params& p, int variant,
                      hipStream_t s, double* trace, float xmax2, const SaStep* step) {
    const bool LOGV = (variant == EMSA_SOFTASSIGN_LOGDOMAIN);
    const SaPlan pl = plan_sa(N, M, ctx.numCU);
    // moment rows of the final pass (one per final-pass workgroup)
    const int finalBlocks = LOGV ? pl.rowBlocks : (N + kSaFinalRows - 1) / kSaFinalRows;
    const int nmom = finalBlocks;
    const int finBlocks = pl.ldk / kSaFinCols;  // column-finalize workgroups
    // matrix-core passes (log domain, expanded exponent): workgroups of the
    // row and column passes; the final row pass writes mfRows moment rows and
    // each column pass mfCols partial sums of sum_c b_c y_c
    const int mfRows = (N + kSaMfItems - 1) / kSaMfItems;
    const int mfCols = pl.ldk / kSaMfItems;
    int rc;
    SaBufs b{};
    b.X = dX;
    b.Y = dY;
    if ((rc = ctx.u.reserve(sizeof(float) * 6 * (size_t)pl.ldk)) != EMSA_OK) return rc;  // P, Yp
    b.P = ctx.u.as<float>();
    b.Yp = b.P + 3 * (size_t)pl.ldk;
    if (!LOGV) {
        if ((rc = ctx.K.reserve(sizeof(float) * (size_t)N * pl.ldk)) != EMSA_OK) return rc;
        b.K = ctx.K.as<float>();
    }
    // u, v, bcol, wfin each start on a 256-byte boundary (read as float4)
    const size_t npad = ((size_t)N + 63) / 64 * 64;
    if ((rc = ctx.v.reserve(sizeof(float) * (npad + 6 * (size_t)pl.ldk))) != EMSA_OK) return rc;
    b.u = ctx.v.as<float>();
    b.v = b.u + npad;
    b.bcol = b.v + pl.ldk;
    b.wfin = b.bcol + pl.ldk;
    if ((rc = ctx.orow.reserve(sizeof(float) * ((size_t)N + M))) != EMSA_OK) return rc;
    b.orow = ctx.orow.as<float>();
    b.ocol = b.orow + N;
    // materialised sweeps in one pass over K when a lane's share of its rows
    // fits in registers (partials per row-pass workgroup); otherwise a row
    // pass and a column pass (partials per row chunk)
    constexpr int slabCols = kSaColsPB * kSaFuseWaves;
    const int slabs =
        (!LOGV && pl.ldk <= kSaFuseMaxSlabs * slabCols) ? (pl.ldk + slabCols - 1) / slabCols : 0;
    const size_t partRows = slabs > 0 ? (size_t)pl.fuseBlocks : (size_t)pl.H;
    if ((rc = ctx.colpart.reserve(sizeof(float) * partRows * pl.ldk)) != EMSA_OK) return rc;
    b.cpart = ctx.colpart.as<float>();
    if ((rc = ctx.samom.reserve(sizeof(double) *
                                ((size_t)nmom * kSaMomPad +
                                 (size_t)std::max(finBlocks, mfCols) * 3))) != EMSA_OK)
        return rc;
    b.rmom = ctx.samom.as<double>();
    b.cby = b.rmom + (size_t)nmom * kSaMomPad;
    if ((rc = ctx.sastate.reserve(sizeof(double) * kStateW)) != EMSA_OK) return rc;
    b.state = ctx.sastate.as<double>();
    b.trace = trace;
    if (LOGV) {
        const size_t ce = (size_t)sa_mf_entries(M), re = (size_t)sa_mf_entries(N);
        if ((rc = ctx.sapack.reserve(sizeof(float4) * (std::max(ce, (size_t)pl.ldk) + re))) !=
            EMSA_OK)
            return rc;
        b.cpack = ctx.sapack.as<float4>();
        b.rpack = b.cpack + std::max(ce, (size_t)pl.ldk);
    }
    float T = step ? step->T : p.T_0;
    // exponent constants and form at temperature T (log domain: expanded
    // while k1 max|x|^2 <= kSaExpandMax, on the matrix cores)
    auto consts = [&](float Tt, float& k1, float& k0, bool& expand) {
        const double k1d = kLog2e / (double)Tt;
        k1 = (float)k1d;
        k0 = (float)((double)p.alpha * k1d - 0.5 * std::log2((double)Tt));
        expand = LOGV && k1 * xmax2 <= kSaExpandMax;
    };
    {
        float k1, k0;
        bool expand;
        consts(T, k1, k0, expand);
        hipLaunchKernelGGL(sa_setup_kernel, dim3(1), dim3(1024), 0, s, b, N, M, pl.ldk, dRtIn,
                           p.moutlier, expand ? 1 : 0, k1, k0);
    }
    EMSA_HIP_TRY(hipGetLastError());
    if (step) {  // the caller's slack state instead of moutlier
        EMSA_HIP_TRY(hipMemcpyAsync(b.orow, step->orow, sizeof(float) * N,
                                    hipMemcpyDeviceToDevice, s));
        EMSA_HIP_TRY(hipMemcpyAsync(b.ocol, step->ocol, sizeof(float) * M,
                                    hipMemcpyDeviceToDevice, s));
    }
    const int JMAX = step ? 1 : p.JMAX, I0 = step ? 1 : p.I0;
    const dim3 rgrid((unsigned)pl.rowBlocks), cgrid((unsigned)(pl.colBlocks * pl.H)),
        fgrid((unsigned)finBlocks);
    const dim3 blk(kWave * kSaWaves), wblk(kWave * kSaFinWavesWide);
    const dim3 xgrid((unsigned)pl.fuseBlocks), xblk(kWave * kSaFuseWaves);
    const dim3 mrgrid((unsigned)mfRows), mcgrid((unsigned)mfCols), mblk(kWave * kSaMfWaves);
    int it = 0;
    for (int Titer = 1; Titer <= JMAX; ++Titer) {
        float k1, k0;
        bool expand;
        consts(T, k1, k0, expand);
        auto log_rows = [&](int mode) {
            if (expand) launch_log_rows<true>(mode, rgrid, blk, s, b, N, M, pl.ldk, k1, k0);
            else launch_log_rows<false>(mode, rgrid, blk, s, b, N, M, pl.ldk, k1, k0);
        };
        for (int i0 = 0; i0 < I0; ++i0, ++it) {
            for (int sw = 0; sw < p.I1; ++sw) {
                const int last = (sw + 1 == p.I1) ? 1 : 0;
                if (step) {  // this sweep's taps
                    b.tapA = step->a ? step->a + (size_t)sw * N : nullptr;
                    b.tapB = step->b ? step->b + (size_t)sw * M : nullptr;
                }
                if (LOGV && expand) {  // matrix-core passes, no finalize
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_mrows_kernel<0>), mrgrid, mblk, 0, s, b, N, M,
                                           pl.ldk, k1, k0);
                    else
                        hipLaunchKernelGGL((sa_mrows_kernel<1>), mrgrid, mblk, 0, s, b, N, M,
                                           pl.ldk, k1, k0);
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_mcols_kernel<true>), mcgrid, mblk, 0, s, b, N, M,
                                           pl.ldk, k1, k0, last);
                    else
                        hipLaunchKernelGGL((sa_mcols_kernel<false>), mcgrid, mblk, 0, s, b, N, M,
                                           pl.ldk, k1, k0, last);
                } else if (LOGV) {
                    log_rows(sw == 0 ? 0 : 1);
                    if (expand)
                        hipLaunchKernelGGL((sa_cols_kernel<true, true>), cgrid, blk, 0, s, b, N,
                                           pl.ldk, pl.H, pl.chunkRows, k1, k0);
                    else
                        hipLaunchKernelGGL((sa_cols_kernel<true, false>), cgrid, blk, 0, s, b, N,
                                           pl.ldk, pl.H, pl.chunkRows, k1, k0);
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_colfin_kernel<true, true>), fgrid, blk, 0, s, b, M,
                                           pl.ldk, pl.H, last);
                    else
                        hipLaunchKernelGGL((sa_colfin_kernel<true, false>), fgrid, blk, 0, s, b,
                                           M, pl.ldk, pl.H, last);
                } else if (slabs > 0) {
                    if ((rc = launch_rowcol(sw == 0, slabs, xgrid, xblk, s, b, N, M, pl.ldk, k1,
                                            k0)) != EMSA_OK)
                        return rc;
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_colfin_kernel<false, true, kSaFinWavesWide, 8>),
                                           fgrid, wblk, 0, s, b, M, pl.ldk, pl.fuseBlocks, last);
                    else
                        hipLaunchKernelGGL((sa_colfin_kernel<false, false, kSaFinWavesWide, 8>),
                                           fgrid, wblk, 0, s, b, M, pl.ldk, pl.fuseBlocks, last);
                } else {
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_rows_kernel<false, 0>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, k1, k0);
                    else
                        hipLaunchKernelGGL((sa_rows_kernel<false, 1>), rgrid, blk, 0, s, b, N, M,
                                           pl.ldk, k1, k0);
                    hipLaunchKernelGGL((sa_cols_kernel<false>), cgrid, blk, 0, s, b, N, pl.ldk,
                                       pl.H, pl.chunkRows, k1, k0);
                    if (sw == 0)
                        hipLaunchKernelGGL((sa_colfin_kernel<false, true>), fgrid, blk, 0, s, b,
                                           M, pl.ldk, pl.H, last);
                    else
                        hipLaunchKernelGGL((sa_colfin_kernel<false, false>), fgrid, blk, 0, s, b,
                                           M, pl.ldk, pl.H, last);
                }
                EMSA_HIP_TRY(hipGetLastError());
            }
            b.tapA = b.tapB = nullptr;
            b.tapS = step ? step->s : nullptr;
            if (LOGV)  // the final pass reads P, v and the weights the column passes left
                log_rows(2);
            else
                hipLaunchKernelGGL((sa_rows_kernel<false, 2, kSaFinalRows, kSaFinalWaves>),
                                   dim3((unsigned)finalBlocks), dim3(kWave * kSaFinalWaves), 0, s,
                                   b, N, M, pl.ldk, k1, k0);
            b.tapS = nullptr;
            // the next inner iteration: its temperature, and whether its
            // passes run on the matrix cores (then the solve prepares them)
            const bool more = (i0 + 1 < I0) || (Titer < JMAX);
            const float Tn = (i0 + 1 < I0) ? T : T * p.TFACTOR;
            float k1n = 0.f, k0n = 0.f;
            bool expn = false;
            if (more) consts(Tn, k1n, k0n, expn);
            // with the columns to prepare, one workgroup per kWave * kSaSolveWaves of them
            const unsigned sgrid = expn ? (unsigned)((pl.ldk + kWave * kSaSolveWaves - 1) /
                                                     (kWave * kSaSolveWaves))
                                        : 1u;
            hipLaunchKernelGGL(sa_solve_kernel, dim3(sgrid), dim3(kWave * kSaSolveWaves), 0, s, b,
                               (LOGV && expand) ? mfRows : nmom,
                               (LOGV && expand) ? mfCols : finBlocks, T, it, M, pl.ldk,
                               expn ? 1 : 0, k1n, k0n);
            EMSA_HIP_TRY(hipGetLastError());
        }
        T = T * p.TFACTOR;  // softassign.cu:820, float
    }
    if (step) {
        EMSA_HIP_TRY(hipMemcpyAsync(step->orow, b.orow, sizeof(float) * N,
                                    hipMemcpyDeviceToDevice, s));
        EMSA_HIP_TRY(hipMemcpyAsync(step->ocol, b.ocol, sizeof(float) * M,
                                    hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(sa_output_kernel, dim3(1), dim3(64), 0, s, b.state, dRtOut);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

}  // namespace emsa
