// EM-ICP on gfx950: three kernel launches per sigma^2 step, for one
// registration or for a whole batch of them (uneven sizes, multi-start).
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
// Target points are wave-uniform: the all-VALU kernel reads them with scalar
// loads (SGPR operands; one exp2 and 15 VALU ops per pair); the default
// kernel reads them from its LDS stage with broadcast reads and needs 10 VALU
// ops per pair (B and W summed on the matrix cores, the exponent scale carried
// by the coordinates).
//
// Work decomposition per sigma^2 step, three launches with no signalling
// between workgroups (kernel boundaries order them):
//  1. emicp_affinity_kernel: a table of registrations, each cut into tiles of
//     64 source rows x one column chunk; a workgroup is one tile, its 4 waves
//     split the chunk and combine in LDS; the tile's per-row partial sums
//     (C, B, W) go to HBM.
//  2. emicp_rows_kernel: one wave per row block reduces the chunk partials in
//     a fixed order in FP64, forms lambda_r and lambda_r X'_r (or the floored
//     row) and the 18 FP64 moments of its 64 rows. A registration whose tiles
//     span its whole target (one chunk) has this done in the affinity
//     kernel's epilogue instead, from the same sums (bit-identical), and the
//     launch is dropped when every registration is such.
//  3. emicp_solve_kernel: one workgroup per registration sums the row-block
//     moments in a fixed order, forms Xc, Yc, S and solves Horn's problem in
//     FP64, updating (R, t).
// The reduction order never depends on scheduling, so results are
// reproducible run to run.
#include "emsa_device.h"
#include "emsa_internal.h"
#include "emsa_order.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace emsa {

namespace {

constexpr float kSentinel = 1.0e15f;  // padding target: exp2 underflows to 0 for any sigma
typedef const __attribute__((address_space(4))) float* cfloat_p;

// Orders LDS writes of some lanes before reads by other lanes of the same
// wave (a wave's LDS instructions execute in order; this keeps the compiler
// from moving accesses across and waits for outstanding LDS operations).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Registration served by workgroup b: the last descriptor with tileBase <= b.
__device__ inline int find_reg(const EmRegDesc* __restrict__ regs, int nregs, int b) {
    int lo = 0, hi = nregs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (regs[mid].tileBase <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Working order of the clouds. Each cloud (the target and the source of every
// pair) is put in the order of a 3D Hilbert curve over its bounding box
// (kOrdBits bits per axis), so that consecutive points are close in space: a
// wave's 64 source rows and each group of kCullCols target columns then have
// small bounding boxes. The affinity kernel skips a column group for a wave
// when the boxes are so far apart that every pair of the group has
// e = exp2(-|d|^2 k) <= 2^-64, i.e. an affinity A = e^2 (the reference's
// expf(-|d|^2 / sigma^2)) below 2^-128, under FLT_MIN: the reference's
// fast-math build flushes such an A to exactly 0, and an FP32 sum of such
// terms changes a normaliser that is above the floor (C > 1e-6) by less than
// 1e-28 relative. A second mode skips only the groups whose e underflow to
// exactly 0 (2^-150 already rounds to 0), so that every sum stays bit for bit
// what the full loop over the same order gives. Per-row results are written
// back in the caller's row order.
// ---------------------------------------------------------------------------
// |exp2 argument| of e beyond which a group is skipped: mode 1 (the default)
// 64 (A < 2^-128 < FLT_MIN), mode 2 160 (e exactly 0 in FP32); the margins
// cover the rounding of the box distance.
constexpr float kCullArgUnderflow = 64.f;
constexpr float kCullArgExact = 160.f;

// One workgroup per cloud: blockIdx.x = 2 * registration + (0: target,
// 1: source); only the registration that owns a pair's buffers orders it.
// The permutation (working position -> caller's index), the cloud in working
// order (the target padded with far sentinels past N), and the target's
// column-group boxes.
__global__ void __launch_bounds__(kOrdThreads)
emicp_order_kernel(const EmRegDesc* __restrict__ regs) {
    const EmRegDesc& d = regs[blockIdx.x >> 1];
    if (!d.padOwner || !d.rowperm) return;  // workgroup-uniform
    const bool src = (blockIdx.x & 1) != 0;
    const int n = src ? d.M : d.N;
    order_cloud(src ? d.Y : d.X, n, src ? d.rowperm : d.colperm,
                d.sortTmp + (src ? 4 * (size_t)d.N : 0), src ? d.Ys : d.Xpad, src ? n : d.ldx,
                kSentinel);
    if (src || !d.gbox) return;
    __syncthreads();  // the boxes read the padded target written above
    group_boxes<kCullCols>(d.Xpad, d.ldx, n, d.gbox);
}

// One workgroup per registration: sum the target in FP64 (for floored rows,
// in the caller's order), pad it when the pair is not reordered (the ordering
// kernel writes the padded target otherwise; once per target cloud), and load
// the initial transform.
__global__ void __launch_bounds__(256)
emicp_setup_kernel(const EmRegDesc* __restrict__ regs, const float* __restrict__ Rt0) {
    const EmRegDesc& d = regs[blockIdx.x];
    const int N = d.N, ldx = d.ldx;
    const bool pad = d.padOwner && !d.colperm;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c = threadIdx.x; c < ldx; c += blockDim.x) {
        float x = kSentinel, y = kSentinel, z = kSentinel;
        if (c < N) {
            x = d.X[c];
            y = d.X[(size_t)N + c];
            z = d.X[2 * (size_t)N + c];
            sx += x;
            sy += y;
            sz += z;
        }
        if (pad) {
            d.Xpad[c] = x;
            d.Xpad[(size_t)ldx + c] = y;
            d.Xpad[2 * (size_t)ldx + c] = z;
        }
    }
    __shared__ double lds[4 * 3];
    double v[3] = {sx, sy, sz};
    double tot;
    block_sum<3>(v, lds, blockDim.x / kWave, tot);
    if (threadIdx.x < 3) d.state[kStXsum + threadIdx.x] = tot;
    if (threadIdx.x < 12) d.state[threadIdx.x] = (double)Rt0[(size_t)blockIdx.x * 12 + threadIdx.x];
    if (threadIdx.x == 12) d.state[kStStatus] = 0.0;
    if (threadIdx.x == 13) d.state[kStResid] = 0.0;
}

__global__ void emicp_output_kernel(const EmRegDesc* __restrict__ regs, float* __restrict__ Rt,
                                    double* __restrict__ resid) {
    const EmRegDesc& d = regs[blockIdx.x];
    if (threadIdx.x < 12) Rt[(size_t)d.slot * 12 + threadIdx.x] = (float)d.state[threadIdx.x];
    if (resid && threadIdx.x == 12) resid[d.slot] = d.state[kStResid];
}

// Registration owning global row block gr: the last descriptor with
// rowBase <= gr.
__device__ inline int find_reg_rows(const EmRegDesc* __restrict__ regs, int nregs, int gr) {
    int lo = 0, hi = nregs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (regs[mid].rowBase <= gr) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// p_r = R y_r + t in FP64 from the registration state (rounded by the caller).
__device__ inline void transform_row(const double* __restrict__ state, float yx, float yy,
                                     float yz, double& pxd, double& pyd, double& pzd) {
    double Rm[9], tv[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Rm[i] = state[kStR + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = state[kStT + i];
    pxd = Rm[0] * yx + Rm[1] * yy + Rm[2] * yz + tv[0];
    pyd = Rm[3] * yx + Rm[4] * yy + Rm[5] * yz + tv[1];
    pzd = Rm[6] * yx + Rm[7] * yy + Rm[8] * yz + tv[2];
}

// Stage 1. COMP: subtract the rounding residual of p_r in every difference,
// d = (x - p_hi) - p_lo, where x - p_hi is exact for the pairs that matter
// (Sterbenz), so d carries no error from rounding p_r to FP32. Without it the
// relative error of e is about 1.44 |d| eps |p| / sigma^2 (eps = 6e-8): below
// 2e-6 for the pairs that carry the normalisers once sigma^2 >= 2e-3, where
// the 3 subtractions are skipped.
constexpr float kCompBelowSigma2 = 2e-3f;
template <bool COMP>
__global__ void __launch_bounds__(kEmRows * kEmWaves)
emicp_affinity_kernel(const EmRegDesc* __restrict__ regs, int nregs, float negk) {
    const int ri = find_reg(regs, nregs, blockIdx.x);
    const EmRegDesc& d = regs[ri];
    const int M = d.M, ldx = d.ldx, chunks = d.chunks, waveCols = d.waveCols;
    const int local = blockIdx.x - d.tileBase;
    const int rb = local / chunks;
    const int g = local - rb * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    // wave index made provably uniform so that target loads become scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r = rb * kEmRows + lane;

    float yx = 0.f, yy = 0.f, yz = 0.f;
    if (r < M) {
        yx = d.Ys[r];
        yy = d.Ys[(size_t)M + r];
        yz = d.Ys[2 * (size_t)M + r];
    }
    double pxd, pyd, pzd;
    transform_row(d.state, yx, yy, yz, pxd, pyd, pzd);
    const float px = (float)pxd, py = (float)pyd, pz = (float)pzd;
    const float plx = (float)(pxd - (double)px), ply = (float)(pyd - (double)py),
                plz = (float)(pzd - (double)pz);

    // The padded target is read-only during the launch: reading it through
    // the constant address space with a wave-uniform index makes the loads
    // scalar (s_load_dwordx16), and the values enter the VALU as SGPRs.
    const cfloat_p Xx = (cfloat_p)d.Xpad;
    const cfloat_p Xy = (cfloat_p)(d.Xpad + ldx);
    const cfloat_p Xz = (cfloat_p)(d.Xpad + 2 * (size_t)ldx);
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

    // Combine the 4 waves of the chunk in LDS, fixed order; wave 0 stores.
    __shared__ float red[kEmWaves][5][kEmRows];
    red[wave][0][lane] = C;
    red[wave][1][lane] = B;
    red[wave][2][lane] = Wx;
    red[wave][3][lane] = Wy;
    red[wave][4][lane] = Wz;
    __syncthreads();
    if (wave == 0) {
        float* P = d.partial + (size_t)local * 5 * kEmRows;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            float acc = red[0][q][lane];
#pragma unroll
            for (int w = 1; w < kEmWaves; ++w) acc += red[w][q][lane];
            P[q * kEmRows + lane] = acc;
        }
    }
}

// Stage 2, one wave and one row block (64 source rows) of registration d:
// the wave-summed 18 FP64 moments of the block, valid in every lane.
// Row finalize of row rb * 64 + lane from its sums over all target columns
// (C without the outlier term, B, W): the floor, lambda, lambda X', the
// per-row outputs and checkpoint; then the 18 FP64 moments of the row block,
// each summed over the 64 rows in row order by lane q from the rows' terms in
// LDS (t: this wave's kRowTerms x 64 doubles), and stored by lane q
// (moment-major). Moment q is the product of terms kMomA[q] and kMomB[q]:
//  [0] lambda  [1..3] lambda X'  [4..6] lambda y  [7 + 3i + j] lambda X'_i y_j
//  [16] lambda |X' - (R y + t)|^2  [17] floored rows
// terms: 0 lambda, 1..3 lambda X', 4..6 y, 7 residual, 8 floored, 9 one
constexpr int kRowTerms = 10;
__device__ __attribute__((always_inline)) inline void row_block_finalize(
    const EmRegDesc& d, int rb, int lane, double Cs, double Bs, double Xs, double Ys, double Zs,
    double outlier, int last, int iter, double* __restrict__ t) {    const int N = d.N, M = d.M;
    const int r = rb * kEmRows + lane;
    const bool valid = r < M;
    const double* __restrict__ state = d.state;
    Cs += outlier;
    double lam, Vx, Vy, Vz, fl = 0.0;
    if ((float)Cs > 10e-7f) {  // emicp.cu:155, compared in float
        const double rs = 1.0 / sqrt(Cs);
        lam = Bs * rs;
        Vx = Xs * rs;
        Vy = Ys * rs;
        Vz = Zs * rs;
    } else {  // emicp.cu:159: A' = 1/N for the whole row
        lam = (double)N * d.floorA;
        Vx = d.floorA * state[kStXsum + 0];
        Vy = d.floorA * state[kStXsum + 1];
        Vz = d.floorA * state[kStXsum + 2];
        fl = 1.0;
    }
    float yx = 0.f, yy = 0.f, yz = 0.f;
    int ro = r;  // the caller's index of this row
    if (valid) {
        yx = d.Ys[r];
        yy = d.Ys[(size_t)M + r];
        yz = d.Ys[2 * (size_t)M + r];
        if (d.rowperm) ro = d.rowperm[r];
    } else {
        lam = Vx = Vy = Vz = fl = 0.0;
    }
    if (last && d.rows && valid) {
        d.rows[ro] = Cs;
        d.rows[(size_t)M + ro] = lam;
    }
    if (d.ckpt && valid) {
        double* ck = d.ckpt + ((size_t)iter * M + ro) * 5;
        ck[0] = Cs;
        ck[1] = lam;
        ck[2] = Vx;
        ck[3] = Vy;
        ck[4] = Vz;
    }
    double pxd, pyd, pzd;
    transform_row(state, yx, yy, yz, pxd, pyd, pzd);
    double res = 0.0;
    if (lam > 0.0) {
        const double ex = Vx / lam - pxd, ey = Vy / lam - pyd, ez = Vz / lam - pzd;
        res = lam * (ex * ex + ey * ey + ez * ez);
    }
    t[0 * kEmRows + lane] = lam;
    t[1 * kEmRows + lane] = Vx;
    t[2 * kEmRows + lane] = Vy;
    t[3 * kEmRows + lane] = Vz;
    t[4 * kEmRows + lane] = (double)yx;
    t[5 * kEmRows + lane] = (double)yy;
    t[6 * kEmRows + lane] = (double)yz;
    t[7 * kEmRows + lane] = res;
    t[8 * kEmRows + lane] = fl;
    t[9 * kEmRows + lane] = 1.0;
    wave_lds_sync();
    static_assert(kMomW == 18, "moment layout above");
    if (lane < kMomW) {
        const int q = lane;
        int ia, ib;
        if (q < 4) { ia = q; ib = 9; }
        else if (q < 7) { ia = 0; ib = q; }
        else if (q < 16) { ia = 1 + (q - 7) / 3; ib = 4 + (q - 7) % 3; }
        else { ia = q - 9; ib = 9; }
        const double* __restrict__ ta = t + ia * kEmRows;
        const double* __restrict__ tb = t + ib * kEmRows;
        double sum = 0.0;
#pragma unroll 8
        for (int rr = 0; rr < kEmRows; ++rr) sum = fma(ta[rr], tb[rr], sum);
        d.moments[(size_t)q * d.rowBlocks + rb] = sum;
    }
}

// Stage 2: one workgroup per row block. Wave w sums the chunk partials
// k = w, w + 4, ... in FP64 (in chunk order; the loads of 4 chunks are issued
// before their adds), the 4 waves' sums are added in a fixed order in LDS,
// then wave 0 finalizes the rows and stores the moments (moment-major). With
// one chunk only wave 0 has a partial and the others add +0.0, so the sums
// equal the affinity kernel's own (the fused finalize) bit for bit.
__global__ void __launch_bounds__(kEmRows * kEmWaves)
emicp_rows_kernel(const EmRegDesc* __restrict__ regs, int nregs, long long totalRowBlocks,
                  double outlier, int last, int iter, int skipSingle) {
    const long long gr = blockIdx.x;
    if (gr >= totalRowBlocks) return;  // workgroup-uniform
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int ri = find_reg_rows(regs, nregs, (int)gr);
    const EmRegDesc& d = regs[ri];
    if (skipSingle && d.chunks == 1) return;  // finalized by the affinity kernel
    const int rb = (int)gr - d.rowBase;
    const int chunks = d.chunks;
    double v5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const float* Pr = d.partial + (size_t)rb * chunks * 5 * kEmRows + lane;
    int k = wave;
    for (; k + 3 * kEmWaves < chunks; k += 4 * kEmWaves) {
        float v[4][5];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 5; ++q) v[i][q] = Pr[((size_t)(k + i * kEmWaves) * 5 + q) * kEmRows];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 5; ++q) v5[q] += v[i][q];
    }
    for (; k < chunks; k += kEmWaves) {
#pragma unroll
        for (int q = 0; q < 5; ++q) v5[q] += Pr[((size_t)k * 5 + q) * kEmRows];
    }
    __shared__ double part[kEmWaves][5][kEmRows];
    __shared__ double terms[kRowTerms * kEmRows];
#pragma unroll
    for (int q = 0; q < 5; ++q) part[wave][q][lane] = v5[q];
    __syncthreads();
    if (wave != 0) return;
    double sm[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        sm[q] = ((part[0][q][lane] + part[1][q][lane]) + part[2][q][lane]) + part[3][q][lane];
    row_block_finalize(d, rb, lane, sm[0], sm[1], sm[2], sm[3], sm[4], outlier, last, iter, terms);
}

// Stage 3 for registration d from its 18 summed moments (one thread):
// Xc, Yc, S, Horn, the state update and the trace record.
__device__ inline void solve_registration(const EmRegDesc& d, const double* mom, float sigma_p2,
                                          int iter) {
    double* __restrict__ state = d.state;
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
    if (d.trace) {
        double* o = d.trace + (size_t)iter * EMSA_TRACE_W;
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

// Stage 1 on the matrix cores: B_r = sum_c e_rc and W_r = sum_c e_rc x_c are
// a product of the e tile (rows x columns) with [x | y | z | 1] (columns x 4),
// with the column dimension as the contraction. v_mfma_f32_4x4x1f32 works on
// 16 independent 4x4 blocks of lanes: lane 4b + j, accumulator i receives
// a(lane 4b + i) * b(lane 4b + j). With a = e (lane = source row, as in the
// VALU kernel) and b = component j of the current column, lane 4b + j sums
// component j (x, y, z, 1) for rows 4b .. 4b + 3. The columns' components are
// staged per wave in LDS as four rows, so one ds_read_b128 gives a lane its
// component for 4 columns. Per pair on the VALU: the scaled difference
// d' = ka x - p' (one FMA per coordinate, plus 3 subtractions for the
// compensation), |d'|^2 and C += e^2 (10 ops with the compensation, 7
// without) + 1 exp of -|d'|^2 (the sign is a source modifier), and
// one matrix instruction per column and wave. UNR columns per loop iteration,
// NACC independent accumulators (measured on MI355X, 8 x 10k batched
// sequence: with the difference form only and 2048-column tiles NACC 4 was 4%
// faster than 2 and 1, 8 was 8% slower; with the expanded form, which has its
// own accumulators, 2 and 4 run the same; UNR is the column group of the
// skipping, 8). One matrix instruction per column and
// wave costs about 16 SIMD cycles here: moving C = sum e^2 onto
// the matrix cores as well (the diagonal of e e^T) made the launch 30% slower.
// cull != 0 (and boxes present): per piece of up to 32 groups of kCullCols
// columns, lane jj tests group jj's box against the box of the wave's valid
// rows, and the loop visits only the groups with a pair whose exp2 argument
// can be above -kCullArgUnderflow (cull 1) or -kCullArgExact (cull 2) (a
// ballot mask, walked lowest bit first); a piece with no such group is not
// even staged. visits (nullable, measurement): visits[0]
// receives the number of (valid row, valid column) pairs the launch evaluates,
// visits[1] those evaluated in the expanded form. perm (nullable): the tile of
// each workgroup slot (else tile = slot); work (nullable): receives each
// tile's visited columns, 3 per column of a wave in the expanded form and 4
// in the difference form (about their relative issue costs), the key a later
// order is sorted by. permNext (nullable): workgroup 0 writes the next
// launch's order there from workPrev (the previous launch's work) and the
// tiles take slots 1.. (balance_tiles).
constexpr int kMfPiece = 256;  // columns staged per wave at a time
// row stride of a wave's stage: 4 floats of padding put the four component
// rows, which lanes 4b + j read together (row j), in different LDS banks
constexpr int kMfStageLd = kMfPiece + 4;
// stage rows: x, y, z (or w), the constant 1, and (expanded form) -|w|^2
constexpr int kMfStageRows = 5;
// EXP stage value of a slot the piece does not fill: exp2 of its exponent is 0
// (-|w|^2 overflows to -inf, and the products with |u|^2 <= xmax stay finite)
constexpr float kExpFar = 1.0e30f;

// -|w|^2 of 4 staged columns (the signs on the FMA operands), formed once per
// piece: the accumulator init of the expanded exponent
__device__ inline float4 neg_sq_norm4(float4 x, float4 y, float4 z) {
    return make_float4(fmaf(-x.x, x.x, fmaf(-y.x, y.x, -z.x * z.x)),
                       fmaf(-x.y, x.y, fmaf(-y.y, y.y, -z.y * z.y)),
                       fmaf(-x.z, x.z, fmaf(-y.z, y.z, -z.z * z.z)),
                       fmaf(-x.w, x.w, fmaf(-y.w, y.w, -z.w * z.w)));
}

// Affinity workgroup order, computed by one more workgroup (kEmRows * kEmWaves
// threads) of an affinity launch, for the launch after it, in `lds` (the
// launch's stage, kBalanceLdsBytes or more): the n <= kBalanceMaxTiles tiles
// in order of their work (one step older than the launch's own), heaviest
// first, in 64 levels (a counting sort: per level and per group of 64 tiles,
// the count and each tile's rank among the lanes of its level, found with one
// ballot per level bit, then prefix sums; equal levels keep the tile order),
// then dealt to the workgroup slots in rounds of `width` (the compute units;
// slots b, b + width, ... share one). Up to kGreedyMaxTiles tiles (and width
// kGreedyMaxWidth) each round gives its heaviest tile to the compute unit with
// the least work so far (ties: the lower unit), the next to the next, ...;
// above, a round in even position goes ascending and a full round in odd
// position descending. perm[b] = the tile of slot b, a permutation of 0..n-1
// for any work values.
constexpr int kGreedyMaxTiles = 1280;
constexpr int kGreedyMaxWidth = 512;
constexpr int kBalanceLevels = 64;
// LDS of the sort for n tiles: counts (2 bytes per group of 64 tiles and
// level), the tiles' levels (1 byte each), the prefix sums, and for the
// greedy deal the sorted tiles and the units' keys and picks
constexpr int balance_lds_bytes(int n, bool greedy) {
    return ((n + kEmRows * kEmWaves - 1) / (kEmRows * kEmWaves)) * kEmWaves * kBalanceLevels * 2 +
           (n + 15) / 16 * 16 + 4 * (kEmWaves * kBalanceLevels + kBalanceLevels + kEmWaves) +
           (greedy ? 4 * kGreedyMaxTiles + 8 * kGreedyMaxWidth : 0);
}
constexpr int kBalanceLdsBytes = balance_lds_bytes(kBalanceMaxTiles, false) >
                                         balance_lds_bytes(kGreedyMaxTiles, true)
                                     ? balance_lds_bytes(kBalanceMaxTiles, false)
                                     : balance_lds_bytes(kGreedyMaxTiles, true);
__device__ void balance_tiles(const int* __restrict__ work, int* __restrict__ perm, int n,
                              int width, unsigned char* __restrict__ lds) {
    constexpr int kLevels = kBalanceLevels;
    constexpr int kThreads = kEmRows * kEmWaves;
    static_assert(kLevels == kWave && kThreads == kWave * kEmWaves, "one level per lane");
    static_assert(kBalanceMaxTiles <= 4096, "tile index in 12 bits");
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int rounds = (n + kThreads - 1) / kThreads;  // groups of 64 tiles: rounds * kEmWaves
    unsigned short* cnt = reinterpret_cast<unsigned short*>(lds);  // [group][level]
    unsigned char* lvl = lds + (size_t)rounds * kEmWaves * kLevels * 2;  // [tile] its level
    int* segTot = reinterpret_cast<int*>(lvl + (n + 15) / 16 * 16);     // [wave][level]
    int* lvBase = segTot + kEmWaves * kLevels;  // [level]
    int* wmax = lvBase + kLevels;               // [wave]
    int* order = wmax + kEmWaves;               // greedy: (work >> shift) << 12 | tile, heaviest first
    int* ckey = order + kGreedyMaxTiles;        // greedy: (work so far) << 10 | unit
    int* pick = ckey + kGreedyMaxWidth;         // greedy: this round's position of each unit
    const bool greedy = n <= kGreedyMaxTiles && width <= kGreedyMaxWidth;
    int m = 0;
#pragma unroll 4
    for (int i = tid; i < n; i += kThreads) m = max(m, work[i]);
    for (int o = 1; o < kWave; o <<= 1) m = max(m, __shfl_xor(m, o, kWave));
    if (lane == 0) wmax[wave] = m;
    for (int e = tid; e < rounds * kEmWaves * kLevels; e += kThreads) cnt[e] = 0;
    __syncthreads();
    const int mw = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    // level of work w: 63 - floor(w * 64 / (max + 1)) (0: heaviest), in FP32
    const float scale = (float)kLevels / ((float)mw + 1.f);
    int shift = 0;  // greedy loads in 16 bits per tile
    while ((mw >> shift) >= (1 << 16)) ++shift;
#pragma unroll 4
    for (int i = tid; i < n; i += kThreads)  // each tile's level, into LDS
        lvl[i] = (unsigned char)(kLevels - 1 -
                                 min(kLevels - 1, (int)((float)max(work[i], 0) * scale)));
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    // level, and lanes of the wave with the same level (one ballot per bit)
    auto classify = [&](int r, bool& valid, int& lv, unsigned long long& same) {
        const int i = r * kThreads + tid;
        valid = i < n;
        lv = valid ? (int)lvl[i] : 0;
        same = __ballot(valid);
        for (int k = 0; k < 6; ++k) {
            const unsigned long long b = __ballot(valid && ((lv >> k) & 1));
            same &= ((lv >> k) & 1) ? b : ~b;
        }
    };
    for (int r = 0; r < rounds; ++r) {  // tiles per (group, level)
        bool valid;
        int lv;
        unsigned long long same;
        classify(r, valid, lv, same);
        if (valid && (same & below) == 0)
            cnt[(r * kEmWaves + wave) * kLevels + lv] = (unsigned short)__popcll(same);
    }
    __syncthreads();
    {  // per level, exclusive prefix over the groups: wave s takes groups [s rounds, (s+1) rounds)
        int run = 0;
        for (int k = 0; k < rounds; ++k) {
            unsigned short* c = cnt + (wave * rounds + k) * kLevels + lane;
            const int v = *c;
            *c = (unsigned short)run;
            run += v;
        }
        segTot[wave * kLevels + lane] = run;
    }
    __syncthreads();
    if (wave == 0) {  // the waves' offsets, then the levels' exclusive prefix
        int t = 0;
        for (int w2 = 0; w2 < kEmWaves; ++w2) {
            const int v = segTot[w2 * kLevels + lane];
            segTot[w2 * kLevels + lane] = t;
            t += v;
        }
        int inc = t;
        for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += v;
        }
        lvBase[lane] = inc - t;
    }
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {  // each tile's position
        bool valid;
        int lv;
        unsigned long long same;
        classify(r, valid, lv, same);
        if (!valid) continue;
        const int g = r * kEmWaves + wave;
        const int pos = lvBase[lv] + segTot[(g / rounds) * kLevels + lv] + cnt[g * kLevels + lv] +
                        __popcll(same & below);
        const int i = r * kThreads + tid;
        if (greedy) {
            order[pos] = ((max(work[i], 0) >> shift) << 12) | i;
        } else {
            const int rd = pos / width, p = pos - rd * width;
            const bool full = (rd + 1) * width <= n;
            perm[rd * width + (((rd & 1) && full) ? width - 1 - p : p)] = i;
        }
    }
    if (!greedy) return;
    __syncthreads();
    for (int c = tid; c < min(width, n); c += kThreads) {  // round 0: unit c, the c-th heaviest
        const int o = order[c];
        perm[c] = o & 4095;
        ckey[c] = ((o >> 12) << 10) | c;
    }
    __syncthreads();
    for (int base = width; base < n; base += width) {
        const int avail = min(width, n - base);
        for (int c = tid; c < avail; c += kThreads) {  // rank of unit c by its work so far
            const int kc = ckey[c];
            int rank = 0, c2 = 0;
            for (; c2 + 16 <= avail; c2 += 16) {  // 4 reads in flight
                int4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const int4*>(ckey + c2 + 4 * u);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    rank += (v[u].x < kc) + (v[u].y < kc) + (v[u].z < kc) + (v[u].w < kc);
            }
            for (; c2 < avail; ++c2) rank += (ckey[c2] < kc) ? 1 : 0;
            pick[c] = rank;
        }
        __syncthreads();
        for (int c = tid; c < avail; c += kThreads) {
            const int o = order[base + pick[c]];
            perm[base + c] = o & 4095;
            ckey[c] = (((ckey[c] >> 10) + (o >> 12)) << 10) | c;
        }
        __syncthreads();
    }
}

template <bool COMP, int UNR, int NACC, bool FUSE>
__global__ void __launch_bounds__(kEmRows * kEmWaves) __attribute__((amdgpu_waves_per_eu(8)))
emicp_affinity_mfma_kernel(const EmRegDesc* __restrict__ regs, int nregs, float ka, int cull,
                           float xmax, unsigned long long* __restrict__ visits, double outlier,
                           int last, int iter, const int* __restrict__ perm,
                           int* __restrict__ work, const int* __restrict__ workPrev,
                           int* __restrict__ permNext, int ntiles, int width) {
    static_assert(UNR == kCullCols, "one loop iteration per column group");
    static_assert(kMfPiece / kCullCols <= 32, "group mask of a piece fits 32 bits");
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ __attribute__((aligned(16))) float stage[kEmWaves][kMfStageRows][kMfStageLd];
    static_assert(sizeof(stage) >= kBalanceLdsBytes, "the order's sort fits the stage");
    if (permNext) {  // workgroup 0 orders the next launch's workgroups
        if (blockIdx.x == 0) {
            balance_tiles(workPrev, permNext, ntiles, width,
                          reinterpret_cast<unsigned char*>(&stage[0][0][0]));
            return;
        }
    }
    const int slot = (int)blockIdx.x - (permNext ? 1 : 0);
    const int tile = perm ? perm[slot] : slot;
    const int ri = find_reg(regs, nregs, tile);
    const EmRegDesc& d = regs[ri];
    const int M = d.M, ldx = d.ldx, chunks = d.chunks, waveCols = d.waveCols;
    const int local = tile - d.tileBase;
    const int rb = local / chunks;
    const int g = local - rb * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int r = rb * kEmRows + lane;

    float yx = 0.f, yy = 0.f, yz = 0.f;
    if (r < M) {
        yx = d.Ys[r];
        yy = d.Ys[(size_t)M + r];
        yz = d.Ys[2 * (size_t)M + r];
    }
    double pxd, pyd, pzd;
    transform_row(d.state, yx, yy, yz, pxd, pyd, pzd);
    const float px = (float)pxd, py = (float)pyd, pz = (float)pzd;
    // the row in units of the exponent scale: p' = ka p (FP64, then split
    // into a rounded part and its residual for the compensated difference),
    // so that d' = ka x - p' comes from one FMA per coordinate and the
    // exponent is -|d'|^2 with no scaling multiply
    const double kad = (double)ka;
    const double qxd = kad * pxd, qyd = kad * pyd, qzd = kad * pzd;
    const float qx = (float)qxd, qy = (float)qyd, qz = (float)qzd;
    const float qlx = (float)(qxd - (double)qx), qly = (float)(qyd - (double)qy),
                qlz = (float)(qzd - (double)qz);

    const cfloat_p Xx = (cfloat_p)d.Xpad;
    const cfloat_p Xy = (cfloat_p)(d.Xpad + ldx);
    const cfloat_p Xz = (cfloat_p)(d.Xpad + 2 * (size_t)ldx);

    // box of the wave's valid rows (for the skipping and the expanded form's
    // test); a non-finite row (NaN affinities in every column) turns both off
    // for the wave
    const bool rv = r < M;
    const float nk = ka * ka;
    bool doCull = cull && d.gbox;
    const float cullArg = (cull == 2) ? kCullArgExact : kCullArgUnderflow;
    float rlo[3] = {0.f, 0.f, 0.f}, rhi[3] = {0.f, 0.f, 0.f};
    bool finiteRows = false;
    if (doCull || xmax > 0.f) {
        const bool bad = rv && !(fabsf(px) <= FLT_MAX && fabsf(py) <= FLT_MAX && fabsf(pz) <= FLT_MAX);
        finiteRows = __ballot(bad) == 0;
        const float pv[3] = {px, py, pz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            rlo[a] = lane63(wave_min_last(rv ? pv[a] : INFINITY));
            rhi[a] = lane63(wave_max_last(rv ? pv[a] : -INFINITY));
        }
    }
    doCull = doCull && finiteRows;

    // Expanded exponent on the matrix cores (EXP), per wave about the centre
    // c of its rows' box. With w = ka x - ka c (the column, one FMA per
    // coordinate, staged) and u = ka p - ka c (the row, FP64 then rounded),
    //   -|d'|^2 = -|u|^2 + 2 w.u - |w|^2,
    // a v_mfma_f32_16x16x4f32 per 16 rows (s) and 16 columns: A[i][k] (lane
    // i + 16k) = (w_x, w_y, w_z, 1) of column i, read from the stage, B[k][jr]
    // (lane jr + 16k) = (2u_x, 2u_y, 2u_z, -|u|^2) of row 16s + jr, and the
    // accumulator init -|w|^2 of the lane's 4 columns (shared by the four
    // row groups), so that lane jr + 16q receives the exponents of row
    // 16s + jr and columns 4q .. 4q + 3. B and W then go through v_mfma_f32_4x4x1f32 as in the
    // difference form, with (w, 1) as the column operand: lane 4b + j sums
    // component j over rows 16s + 4(b mod 4) .. + 3 for the columns 4q + v of
    // phase q = b / 4; the four phases are added at the end, which leaves the
    // same layout as the difference form, and W = (sum e w + ka c B) / ka.
    // Per 64 pairs: 1/4 of a 16x16x4 product (instead of 6 VALU ops of the
    // difference, 9 with the compensation). The rounding of the expanded
    // terms grows with ka^2 (|w| + |u|)^2, so a wave takes this form only
    // when all its valid rows lie within a of c with ka^2 a^2 <= xmax
    // (xmax = 0: never); a wave with a non-finite row keeps the difference.
    const int j = lane & 3;  // this lane's component in the matrix product
    bool expd = false;
    float cwx = 0.f, cwy = 0.f, cwz = 0.f;  // ka c, FP32
    float bop[4] = {0.f, 0.f, 0.f, 0.f};    // B operand of rows 16s + (lane & 15)
    if (xmax > 0.f && finiteRows) {
        float cw[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) cw[a] = (rlo[a] <= rhi[a]) ? 0.5f * (rlo[a] + rhi[a]) : 0.f;
        const float ax = px - cw[0], ay = py - cw[1], az = pz - cw[2];
        const float a2 = fmaf(ax, ax, fmaf(ay, ay, az * az));
        const bool bad = rv && !(nk * a2 <= xmax);
        expd = __ballot(bad) == 0;
        if (expd) {  // wave-uniform
            cwx = (float)(kad * (double)cw[0]);
            cwy = (float)(kad * (double)cw[1]);
            cwz = (float)(kad * (double)cw[2]);
            float ux = 0.f, uy = 0.f, uz = 0.f, n2 = 0.f;
            if (rv) {
                ux = (float)(qxd - (double)cwx);
                uy = (float)(qyd - (double)cwy);
                uz = (float)(qzd - (double)cwz);
                n2 = -(float)((double)ux * ux + (double)uy * uy + (double)uz * uz);
            }
            const int k = lane >> 4;
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) {
                const int src = 16 * sb + (lane & 15);
                const float vx = __shfl(2.f * ux, src, kWave), vy = __shfl(2.f * uy, src, kWave),
                            vz = __shfl(2.f * uz, src, kWave), vn = __shfl(n2, src, kWave);
                bop[sb] = (k == 0) ? vx : (k == 1) ? vy : (k == 2) ? vz : vn;
            }
        }
    }

    float (*st)[kMfStageLd] = stage[wave];
    // the operand's constant fourth component (B = sum e * 1): the compensated
    // instance writes it once (restaging it per piece kept a constant float4
    // live across the loop and spilled it at 8 waves per SIMD); the plain one
    // restages it with each piece (same-process A/B: writing it once changed
    // its register allocation and made its launches 7% slower)
    static_assert(kMfPiece == 4 * kWave, "one float4 of ones per lane");
    *reinterpret_cast<float4*>(&st[3][4 * lane]) = make_float4(1.f, 1.f, 1.f, 1.f);
    float C = 0.f;
    f4 D[NACC];
#pragma unroll
    for (int a = 0; a < NACC; ++a) D[a] = f4{0.f, 0.f, 0.f, 0.f};
    // EXP: per 16 rows s, lane jr + 16q sums e^2 of row 16s + jr (Cx) and
    // the B/W matrix accumulators of phase q (Dx)
    float Cx[4] = {0.f, 0.f, 0.f, 0.f};
    f4 Dx[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) Dx[a] = f4{0.f, 0.f, 0.f, 0.f};
    int nvis = 0;  // valid columns visited by this wave
    // The chunk's groups of kCullCols columns are dealt to the 4 waves in
    // turn (group G0 + wave + 4k), so that the columns near the row block,
    // which are consecutive in working order, are shared by all waves.
    constexpr int kPieceGroups = kMfPiece / kCullCols;
    const int gpw = waveCols / kCullCols;  // groups per wave
    const int G0 = g * kEmWaves * gpw;
    const int Gend = min(G0 + kEmWaves * gpw, ldx / kCullCols);
    const int nw = (Gend - G0 - wave + kEmWaves - 1) / kEmWaves;  // this wave's groups
    // the column loop, once per form of the exponent (a wave-uniform choice)
    auto columns = [&](auto expTag) {
    constexpr bool EXP = decltype(expTag)::value;
    // skipping masks of the current batch of 4 pieces (wave group k tested
    // by lane k - kb and k - kb - 64: bit k - kb of cm0, of cm1 past 64)
    unsigned long long cm0 = 0, cm1 = 0;
    auto group_live = [&](int k) {  // group G0 + wave + 4k has a pair above the threshold
        const float* bx = d.gbox + (size_t)(G0 + wave + kEmWaves * k) * 8;
        const float4 lo = *reinterpret_cast<const float4*>(bx);
        const float4 hi = *reinterpret_cast<const float4*>(bx + 4);
        const float gx = fmaxf(fmaxf(rlo[0] - hi.x, lo.x - rhi[0]), 0.f);
        const float gy = fmaxf(fmaxf(rlo[1] - hi.y, lo.y - rhi[1]), 0.f);
        const float gz = fmaxf(fmaxf(rlo[2] - hi.z, lo.z - rhi[2]), 0.f);
        const float g2 = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
        return !(g2 * nk > cullArg);
    };
    for (int k0 = 0; k0 < nw; k0 += kPieceGroups) {
        const int ng = min(kPieceGroups, nw - k0);
        const int gbase = G0 + wave + kEmWaves * k0;  // group of piece slot jj: gbase + 4 jj
        unsigned keep = (ng >= 32) ? 0xFFFFFFFFu : ((1u << ng) - 1u);
        if (doCull) {
            // the first piece of a batch tests the batch's 128 groups at once
            // (two per lane, their box loads in flight together): one memory
            // round trip per 4 pieces instead of one per piece
            const int kb = k0 & ~(4 * kPieceGroups - 1);
            if (k0 == kb) {
                const int ka0 = kb + lane, ka1 = kb + kWave + lane;
                const bool l0 = (ka0 < nw) && group_live(ka0);
                const bool l1 = (ka1 < nw) && group_live(ka1);
                cm0 = __ballot(l0);
                cm1 = __ballot(l1);
            }
            const int pi = (k0 - kb) / kPieceGroups;  // 0..3
            keep &= (unsigned)(((pi < 2) ? cm0 : cm1) >> (32 * (pi & 1)));
            if (keep == 0) continue;
        }
        // stage [x; y; z; 1] of the piece (EXP: [w; 1], and slots past the
        // piece's groups far away, since a 16-column step may read one): lane
        // 2 jj + h holds columns 4h .. 4h+3 of slot jj
        {
            const int jl = lane >> 1, h = lane & 1;
            const int sc = jl * kCullCols + 4 * h;
            if (jl < ng) {
                const int col = (gbase + kEmWaves * jl) * kCullCols + 4 * h;
                float4 vx = *reinterpret_cast<const float4*>(d.Xpad + col);
                float4 vy = *reinterpret_cast<const float4*>(d.Xpad + ldx + col);
                float4 vz = *reinterpret_cast<const float4*>(d.Xpad + 2 * (size_t)ldx + col);
                if constexpr (EXP) {
                    vx = make_float4(fmaf(vx.x, ka, -cwx), fmaf(vx.y, ka, -cwx),
                                     fmaf(vx.z, ka, -cwx), fmaf(vx.w, ka, -cwx));
                    vy = make_float4(fmaf(vy.x, ka, -cwy), fmaf(vy.y, ka, -cwy),
                                     fmaf(vy.z, ka, -cwy), fmaf(vy.w, ka, -cwy));
                    vz = make_float4(fmaf(vz.x, ka, -cwz), fmaf(vz.y, ka, -cwz),
                                     fmaf(vz.z, ka, -cwz), fmaf(vz.w, ka, -cwz));
                }
                *reinterpret_cast<float4*>(&st[0][sc]) = vx;
                *reinterpret_cast<float4*>(&st[1][sc]) = vy;
                *reinterpret_cast<float4*>(&st[2][sc]) = vz;
                if constexpr (EXP)
                    *reinterpret_cast<float4*>(&st[4][sc]) = neg_sq_norm4(vx, vy, vz);
            } else if (EXP) {
                const float4 far = make_float4(kExpFar, kExpFar, kExpFar, kExpFar);
                *reinterpret_cast<float4*>(&st[0][sc]) = far;
                *reinterpret_cast<float4*>(&st[1][sc]) = far;
                *reinterpret_cast<float4*>(&st[2][sc]) = far;
                *reinterpret_cast<float4*>(&st[4][sc]) = neg_sq_norm4(far, far, far);
            }
        }
        wave_lds_sync();
        if constexpr (EXP) {
            // 16-column steps: slots 2m and 2m + 1 together (a slot the box
            // test rejected has every e below the threshold in this form too:
            // with |u|^2 <= xmax its exponent's rounding is far below 1)
            unsigned km = (keep | (keep >> 1)) & 0x55555555u;
            while (km) {  // wave-uniform
                const int jj = __builtin_ctz(km);
                km &= km - 1u;
                const int c = jj * kCullCols;  // in the stage
                if (visits || work) {
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int cc0 = (gbase + kEmWaves * (jj + t)) * kCullCols;
                        if (jj + t < ng) nvis += max(0, min(kCullCols, d.N - cc0));
                    }
                }
                // A: component (lane >> 4) of column c + (lane & 15); the
                // lane's 4 columns of phase q = lane >> 4: their (w, 1) for
                // the B/W operand and -|w|^2 (staged) for the accumulator init
                const float aop = st[lane >> 4][c + (lane & 15)];
                const int cq = c + 4 * (lane >> 4);
                const f4 xq = *reinterpret_cast<const f4*>(&st[j][cq]);
                const f4 nw2 = *reinterpret_cast<const f4*>(&st[4][cq]);
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) {
                    const f4 ex = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, bop[sb], nw2, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float e = __builtin_amdgcn_exp2f(ex[v]);
                        Cx[sb] = fmaf(e, e, Cx[sb]);
                        Dx[sb] = __builtin_amdgcn_mfma_f32_4x4x1f32(e, xq[v], Dx[sb], 0, 0, 0);
                    }
                }
            }
            wave_lds_sync();
            continue;
        }
        while (keep) {  // wave-uniform: the slots still to visit
            const int jj = __builtin_ctz(keep);
            keep &= keep - 1u;
            const int c = jj * UNR;                               // in the stage
            const int cc0 = (gbase + kEmWaves * jj) * kCullCols;  // in the target
            if (visits || work) nvis += max(0, min(UNR, d.N - cc0));
            f4 bq[UNR / 4];
#pragma unroll
            for (int v = 0; v < UNR / 4; ++v)
                bq[v] = *reinterpret_cast<const f4*>(&st[j][c + 4 * v]);
#pragma unroll
            for (int h = 0; h < UNR / 4; ++h) {
                // the coordinates of 4 columns, read by every lane from the
                // stage at one address (an LDS broadcast): no scalar-memory
                // round trip per visited group
                const f4 xs = *reinterpret_cast<const f4*>(&st[0][c + 4 * h]);
                const f4 ys = *reinterpret_cast<const f4*>(&st[1][c + 4 * h]);
                const f4 zs = *reinterpret_cast<const f4*>(&st[2][c + 4 * h]);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int u = 4 * h + i;
                    float dx = fmaf(xs[i], ka, -qx), dy = fmaf(ys[i], ka, -qy),
                          dz = fmaf(zs[i], ka, -qz);
                    if (COMP) {
                        dx -= qlx;
                        dy -= qly;
                        dz -= qlz;
                    }
                    float s = dx * dx;
                    s = fmaf(dy, dy, s);
                    s = fmaf(dz, dz, s);
                    const float e = __builtin_amdgcn_exp2f(-s);
                    C = fmaf(e, e, C);
                    D[u % NACC] = __builtin_amdgcn_mfma_f32_4x4x1f32(e, bq[h][i], D[u % NACC],
                                                                     0, 0, 0);
                }
            }
        }
        wave_lds_sync();
    }
    };
    if (expd) {
        columns(std::true_type{});
        // add the four phases (lanes l, l ^ 16, l ^ 32, l ^ 48, fixed order)
        // and keep the rows of this lane's 16: C of row `lane`, D of rows
        // (lane & ~3) .. + 3 as in the difference form
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            Cx[sb] += __shfl_xor(Cx[sb], 16, kWave);
            Cx[sb] += __shfl_xor(Cx[sb], 32, kWave);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                Dx[sb][i] += __shfl_xor(Dx[sb][i], 16, kWave);
                Dx[sb][i] += __shfl_xor(Dx[sb][i], 32, kWave);
            }
        }
        const int sl = lane >> 4;
        C = (sl == 0) ? Cx[0] : (sl == 1) ? Cx[1] : (sl == 2) ? Cx[2] : Cx[3];
        D[0] = (sl == 0) ? Dx[0] : (sl == 1) ? Dx[1] : (sl == 2) ? Dx[2] : Dx[3];
    } else {
        columns(std::false_type{});
    }
#pragma unroll
    for (int a = 1; a < NACC; ++a) D[0] += D[a];
    if (visits && lane == 0 && nvis > 0) {
        const unsigned long long np =
            (unsigned long long)nvis * (unsigned long long)min(kEmRows, M - rb * kEmRows);
        atomicAdd(visits, np);
        if (expd) atomicAdd(visits + 1, np);
    }

    // Combine the 4 waves in LDS, fixed order; wave 0 stores. Lane 4b + j
    // holds component j of rows 4b .. 4b + 3: j = 0, 1, 2 -> W_x, W_y, W_z
    // (slots 2, 3, 4), j = 3 -> B (slot 1). Each wave's 5 x 64 partials go in
    // its own stage, which it no longer reads (its LDS accesses are in order),
    // so the workgroup needs 16 KB of LDS and 8 waves per SIMD fit.
    static_assert(5 * kEmRows <= 4 * kMfStageLd, "partials fit a wave's stage");
    __shared__ int waveWork[kEmWaves];
    if (work && lane == 0) waveWork[wave] = nvis * (expd ? 3 : 4);
    float* red = &stage[wave][0][0];
    red[lane] = C;
    const int q = (j == 3) ? 1 : 2 + j;
    const int rb4 = lane & ~3;
#pragma unroll
    for (int i = 0; i < 4; ++i) red[q * kEmRows + rb4 + i] = D[0][i];
    if (expd) {  // sums of e w -> W = sum e x = (sum e w + ka c B) / ka
        wave_lds_sync();
        const float Bs = red[1 * kEmRows + lane];
        const float cv[3] = {cwx, cwy, cwz};
#pragma unroll
        for (int a = 0; a < 3; ++a)
            red[(2 + a) * kEmRows + lane] = fmaf(cv[a], Bs, red[(2 + a) * kEmRows + lane]) / ka;
    }
    __syncthreads();
    if (wave == 0) {
        static_assert(kEmWaves == 4, "four wave weights");
        if (work && lane == 0)
            work[tile] = (waveWork[0] + waveWork[1]) + (waveWork[2] + waveWork[3]);
        float acc[5];
#pragma unroll
        for (int qq = 0; qq < 5; ++qq) {
            acc[qq] = (&stage[0][0][0])[qq * kEmRows + lane];
#pragma unroll
            for (int w = 1; w < kEmWaves; ++w) acc[qq] += (&stage[w][0][0])[qq * kEmRows + lane];
        }
        if (FUSE && d.chunks == 1) {
            // the tile spans the whole target (one chunk): these are the
            // row sums the row finalize would read back, so it runs here
            // the stage is free again: wave 0 has read it and the other
            // waves are done with it (the barrier above)
            static_assert(kRowTerms * kEmRows * sizeof(double) <= sizeof(stage),
                          "row terms fit the stage");
            wave_lds_sync();
            row_block_finalize(d, rb, lane, (double)acc[0], (double)acc[1], (double)acc[2],
                               (double)acc[3], (double)acc[4], outlier, last, iter,
                               reinterpret_cast<double*>(&stage[0][0][0]));
        } else {
            float* P = d.partial + (size_t)local * 5 * kEmRows;
#pragma unroll
            for (int qq = 0; qq < 5; ++qq) P[qq * kEmRows + lane] = acc[qq];
        }
    }
}


// Stage 3: one workgroup per registration.
__global__ void __launch_bounds__(kEmRows * kEmWaves)
emicp_solve_kernel(const EmRegDesc* __restrict__ regs, float sigma_p2, int iter) {
    const EmRegDesc& d = regs[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int rowBlocks = d.rowBlocks;
    __shared__ double mom[kMomW];
    // moment q = wave + 4 k: lanes stride over the row blocks, then a
    // fixed-order wave sum; the wave's loads are issued together (rows past
    // kMomW exist and are ignored)
    constexpr int kPer = kMomPad / kEmWaves;
    double acc[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
    for (int b = lane; b < rowBlocks; b += kWave) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            acc[k] += d.moments[(size_t)(wave + kEmWaves * k) * rowBlocks + b];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = wave + kEmWaves * k;
            if (q < kMomW) mom[q] = acc[k];
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    solve_registration(d, mom, sigma_p2, iter);
}


}  // namespace

// The affinity launch of one sigma^2 step: compensated difference below
// kCompBelowSigma2; B and W on the matrix cores unless the context selects
// the all-VALU kernel.
void launch_affinity(const DeviceCtx& ctx, const EmRegDesc* desc, int nregs, unsigned tiles,
                     float sigma_p2, hipStream_t s, unsigned long long* visits,
                     const EmRowFuse* fuse, const int* perm = nullptr, int* work = nullptr,
                     const int* workPrev = nullptr, int* permNext = nullptr) {
    const dim3 blk(kEmRows * kEmWaves);
    const bool comp = sigma_p2 < kCompBelowSigma2;
    if (ctx.emAffinityMfma) {
        // exponent scale sqrt(log2 e / (2 sigma^2)), applied to the coordinates
        const float ka = (float)std::sqrt(kLog2e / (2.0 * (double)sigma_p2));
        const double outlier = fuse ? fuse->outlier : 0.0;
        const int last = fuse ? fuse->last : 0, iter = fuse ? fuse->iter : 0;
        const unsigned grid = tiles + (permNext ? 1u : 0u);
// 2 matrix accumulators: with 4, the instances that finalize rows in their
// epilogue held one register pair in scratch between the row setup and the
// epilogue (2.6 MB written per batched 8 x 10k launch); 2 give the same launch
// times for the batched (whole-row and 2048-column tiles) and plain runs
#define EMSA_AFF(C_, F_)                                                                     \
    hipLaunchKernelGGL((emicp_affinity_mfma_kernel<C_, 8, 2, F_>), dim3(grid), blk, 0, s, desc, \
                       nregs, ka, ctx.emCull, ctx.emExpanded, visits, outlier, last, iter, perm, \
                       work, workPrev, permNext, (int)tiles, std::max(1, ctx.numCU))
        if (fuse) {
            if (comp) EMSA_AFF(true, true);
            else EMSA_AFF(false, true);
        } else {
            if (comp) EMSA_AFF(true, false);
            else EMSA_AFF(false, false);
        }
#undef EMSA_AFF
    } else {
        const float negk = (float)(-kLog2e / (2.0 * (double)sigma_p2));
        if (comp)
            hipLaunchKernelGGL(emicp_affinity_kernel<true>, dim3(tiles), blk, 0, s, desc, nregs,
                               negk);
        else
            hipLaunchKernelGGL(emicp_affinity_kernel<false>, dim3(tiles), blk, 0, s, desc, nregs,
                               negk);
    }
}

EmPlan plan_emicp(int N, int M, int numCU, int blocks_per_cu) {
    EmPlan pl;
    pl.rowBlocks = (M + kEmRows - 1) / kEmRows;
    pl.ldx = (N + kColGroup - 1) / kColGroup * kColGroup;
    const int groups = pl.ldx / kColGroup;  // column groups of 16
    // workgroups per CU: 7 where the launches run in the balanced order (at
    // least kBalanceMinPairs pairs; plain 10k registration 2.08 ms against
    // 2.24 ms with 12), else 12 (N = M = 1007: 0.92 ms against 0.93 with 6
    // to 8): more chunks speed up the affinity but lengthen the row finalize
    const int dflt = ((double)N * (double)M >= kBalanceMinPairs) ? 7 : 12;
    const int target = std::max(1, numCU) * (blocks_per_cu > 0 ? blocks_per_cu : dflt);
    int chunks = (target + pl.rowBlocks - 1) / pl.rowBlocks;
    chunks = std::max(1, std::min(chunks, (groups + kEmWaves - 1) / kEmWaves));
    const int gpw = (groups + chunks * kEmWaves - 1) / (chunks * kEmWaves);  // groups per wave
    pl.waveCols = gpw * kColGroup;
    pl.chunks = (pl.ldx + pl.waveCols * kEmWaves - 1) / (pl.waveCols * kEmWaves);
    return pl;
}

std::vector<float> sigma_schedule(const emsa_params& p) {
    std::vector<float> out;
    if (check_schedule(p, "") != EMSA_OK) return out;
    float s = p.sigma_p2;
    while (s > p.sigma_inf) {  // emicp.cu:443
        out.push_back(s);
        s *= p.sigma_factor;  // emicp.cu:668
        if (out.size() > 1000000) break;
    }
    return out;
}

int check_schedule(const emsa_params& p, const char* fn) {
    if (!(p.sigma_p2 > p.sigma_inf)) return EMSA_OK;  // the loop body never runs
    if (p.sigma_factor > 0.f && p.sigma_factor < 1.f && p.sigma_inf >= 0.f) return EMSA_OK;
    set_error(std::string(fn) +
              ": sigma_factor must lie in (0, 1) and sigma_inf be >= 0 when sigma_p2 > "
              "sigma_inf (the sigma^2 loop would not end)");
    return EMSA_EINVAL;
}

int em_layout(const DeviceCtx& ctx, EmWork& w, const EmPair* pairs, int count, int n_starts,
              int tileCols, double* trace, double* rows, const EmAux* aux, EmSizes& sz) {
    if (count <= 0 || n_starts <= 0) return set_error("emicp: empty batch"), EMSA_EINVAL;
    w.regs.clear();
    w.ordOff.clear();
    sz = EmSizes{};
    // working order (Hilbert) and column-group boxes, except for the unrolled
    // forward, whose per-step checkpoints follow the caller's row order
    const bool order = (aux == nullptr);
    long long tiles = 0, rowBlocks = 0;
    auto al64 = [](size_t v) { return (v + 63) / 64 * 64; };
    for (int k = 0; k < count; ++k) {
        const EmPair& pr = pairs[k];
        if (!pr.X || !pr.Y || pr.N <= 0 || pr.M <= 0)
            return set_error("emicp: null or empty cloud in batch"), EMSA_EINVAL;
        EmPlan pl;
        if (tileCols > 0) {
            pl.rowBlocks = (pr.M + kEmRows - 1) / kEmRows;
            pl.ldx = (pr.N + kColGroup - 1) / kColGroup * kColGroup;
            pl.waveCols = std::max(kColGroup, tileCols / kEmWaves / kColGroup * kColGroup);
            pl.chunks = (pl.ldx + pl.waveCols * kEmWaves - 1) / (pl.waveCols * kEmWaves);
        } else {
            pl = plan_emicp(pr.N, pr.M, ctx.numCU, ctx.emBlocksPerCU);
        }
        const size_t xoff = sz.xpadF;
        sz.xpadF += 3 * (size_t)pl.ldx;
        // per pair, in 4-byte words: boxes [ldx], Ys [3M], rowperm [M],
        // colperm [N], ordering scratch [4N + 4M]
        const size_t ooff = sz.ordW;
        if (order)
            sz.ordW += al64(pl.ldx) + al64(3 * (size_t)pr.M) + al64(pr.M) + al64(pr.N) +
                       al64(4 * ((size_t)pr.N + pr.M));
        for (int st = 0; st < n_starts; ++st) {
            EmRegDesc d{};
            d.X = pr.X;
            d.Y = pr.Y;
            d.N = pr.N;
            d.M = pr.M;
            d.ldx = pl.ldx;
            d.rowBlocks = pl.rowBlocks;
            d.chunks = pl.chunks;
            d.waveCols = pl.waveCols;
            d.floorA = (double)(1.0f / (float)pr.N);
            d.padOwner = (st == 0);
            d.slot = k * n_starts + st;
            d.tileBase = (int)tiles;
            d.rowBase = (int)rowBlocks;
            const long long t = (long long)pl.rowBlocks * pl.chunks;
            tiles += t;
            rowBlocks += pl.rowBlocks;
            // offsets for now (em_bind turns them into pointers)
            d.Xpad = reinterpret_cast<float*>(xoff);
            if (order) w.ordOff.push_back(ooff);
            d.partial = reinterpret_cast<float*>(sz.partF);
            sz.partF += (size_t)t * 5 * kEmRows;
            d.moments = reinterpret_cast<double*>(sz.momD);
            sz.momD += (size_t)pl.rowBlocks * kMomPad;
            d.trace = (count == 1 && n_starts == 1) ? trace : nullptr;
            d.rows = (count == 1 && n_starts == 1) ? rows : nullptr;
            d.ckpt = nullptr;
            if (aux) {
                d.trace = aux[k * n_starts + st].trace;
                d.ckpt = aux[k * n_starts + st].ckpt;
            }
            w.regs.push_back(d);
        }
    }
    if (tiles > 0x7fffffffLL) return set_error("emicp: batch too large for one launch"), EMSA_EINVAL;
    w.tiles = tiles;
    w.rowBlocks = rowBlocks;
    w.singleChunk = true;
    w.someSingleChunk = false;
    for (const EmRegDesc& d : w.regs) {
        w.singleChunk = w.singleChunk && d.chunks == 1;
        w.someSingleChunk = w.someSingleChunk || d.chunks == 1;
    }
    sz.nregs = (int)w.regs.size();
    return EMSA_OK;
}

void em_bind(EmWork& w, float* xpad, float* partial, double* moments, double* state,
             unsigned* order) {
    auto al64 = [](size_t v) { return (v + 63) / 64 * 64; };
    for (size_t i = 0; i < w.regs.size(); ++i) {
        EmRegDesc& d = w.regs[i];
        d.Xpad = xpad + reinterpret_cast<size_t>(d.Xpad);
        if (!w.ordOff.empty()) {
            float* o = reinterpret_cast<float*>(order) + w.ordOff[i];
            d.gbox = o;
            o += al64(d.ldx);
            d.Ys = o;
            o += al64(3 * (size_t)d.M);
            d.rowperm = reinterpret_cast<int*>(o);
            o += al64(d.M);
            d.colperm = reinterpret_cast<int*>(o);
            o += al64(d.N);
            d.sortTmp = reinterpret_cast<unsigned*>(o);
        } else {
            d.Ys = const_cast<float*>(d.Y);
            d.gbox = nullptr;
            d.rowperm = d.colperm = nullptr;
            d.sortTmp = nullptr;
        }
        d.partial = partial + reinterpret_cast<size_t>(d.partial);
        d.moments = moments + reinterpret_cast<size_t>(d.moments);
        d.state = state + i * kStateW;
    }
}

int em_prepare(DeviceCtx& ctx, EmWork& w, const EmPair* pairs, int count, int n_starts,
               int tileCols, double* trace, double* rows, hipStream_t s, const EmAux* aux,
               bool sync_upload) {
    if (count <= 0 || n_starts <= 0) return set_error("emicp: empty batch"), EMSA_EINVAL;
    if (&w == &ctx.em && ctx.emPending) {  // a sequence on another stream may still read it
        EMSA_HIP_TRY(hipEventSynchronize(ctx.emDone));
        ctx.emPending = false;
    }
    EmSizes sz;
    int rc = em_layout(ctx, w, pairs, count, n_starts, tileCols, trace, rows, aux, sz);
    if (rc != EMSA_OK) return rc;
    const int nregs = sz.nregs;
    if ((rc = w.xpad.reserve(sizeof(float) * sz.xpadF)) != EMSA_OK) return rc;
    if ((rc = w.partial.reserve(sizeof(float) * sz.partF)) != EMSA_OK) return rc;
    if ((rc = w.moments.reserve(sizeof(double) * sz.momD)) != EMSA_OK) return rc;
    if ((rc = w.state.reserve(sizeof(double) * kStateW * nregs)) != EMSA_OK) return rc;
    if ((rc = w.desc.reserve(sizeof(EmRegDesc) * nregs)) != EMSA_OK) return rc;
    if (sz.ordW > 0 && (rc = w.order.reserve(4 * sz.ordW)) != EMSA_OK) return rc;
    double launchPairs = 0.0;
    for (const EmRegDesc& d : w.regs) launchPairs += (double)d.N * (double)d.M;
    w.balance = ctx.emBalance && ctx.emAffinityMfma && ctx.emCull && !w.ordOff.empty() &&
                w.tiles <= kBalanceMaxTiles && launchPairs >= kBalanceMinPairs;
    if (w.balance && (rc = w.sched.reserve(4 * sizeof(int) * (size_t)w.tiles)) != EMSA_OK)
        return rc;
    em_bind(w, w.xpad.as<float>(), w.partial.as<float>(), w.moments.as<double>(),
            w.state.as<double>(), sz.ordW > 0 ? w.order.as<unsigned>() : nullptr);
    EMSA_HIP_TRY(hipMemcpyAsync(w.desc.ptr, w.regs.data(), sizeof(EmRegDesc) * nregs,
                                hipMemcpyHostToDevice, s));
    if (sync_upload) EMSA_HIP_TRY(hipStreamSynchronize(s));  // w.regs may be rewritten next
    return EMSA_OK;
}

int em_enqueue(DeviceCtx& ctx, EmWork& w, const float* dRt0, float* dRtOut, double* dResid,
               const emsa_params& p, const std::vector<float>& sigmas, hipStream_t s,
               const EmStepProbe* probe) {
    const int nregs = (int)w.regs.size();
    const EmRegDesc* desc = w.desc.as<EmRegDesc>();
    int rc = em_launch_setup(w, dRt0, s);
    if (rc != EMSA_OK) return rc;
    const unsigned rowGrid = (unsigned)w.rowBlocks;  // one workgroup per row block
    // registrations whose tiles span their whole target (one chunk) have
    // their rows finalized by the affinity kernel; the row kernel serves the
    // others and is not launched when there are none
    const bool fuse = ctx.emRowFuse && ctx.emAffinityMfma && w.someSingleChunk;
    const bool rowsKernel = !(fuse && w.singleChunk);
    // balanced workgroup order: steps 0 and 1 in tile order; the affinity
    // launch of step k >= 1 also orders step k + 1 from step k - 1's work
    // (work and orders double-buffered: [work even, work odd, order even,
    // order odd] steps)
    int* const sched = w.balance ? w.sched.as<int>() : nullptr;
    const size_t T = (size_t)w.tiles;
    for (size_t it = 0; it < sigmas.size(); ++it) {
        const float sig = sigmas[it];
        const double outlier = (double)expf(-p.d_02 / sig);  // emicp.cu:489, host float
        const int last = (it + 1 == sigmas.size()) ? 1 : 0;
        if (probe && probe->events) EMSA_HIP_TRY(hipEventRecord(probe->events[2 * it], s));
        const EmRowFuse fz{outlier, last, (int)it};
        const bool orderNext = sched && it >= 1 && it + 1 < sigmas.size();
        launch_affinity(ctx, desc, nregs, (unsigned)w.tiles, sig, s,
                        (probe && probe->visits) ? probe->visits + 2 * it : nullptr,
                        fuse ? &fz : nullptr,
                        (sched && it >= 2) ? sched + (2 + it % 2) * T : nullptr,
                        sched ? sched + (it % 2) * T : nullptr,
                        orderNext ? sched + ((it - 1) % 2) * T : nullptr,
                        orderNext ? sched + (2 + (it + 1) % 2) * T : nullptr);
        EMSA_HIP_TRY(hipGetLastError());
        if (probe && probe->events) EMSA_HIP_TRY(hipEventRecord(probe->events[2 * it + 1], s));
        if (rowsKernel) {
            hipLaunchKernelGGL(emicp_rows_kernel, dim3(rowGrid), dim3(kEmRows * kEmWaves), 0, s,
                               desc, nregs, w.rowBlocks, outlier, last, (int)it, fuse ? 1 : 0);
            EMSA_HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(emicp_solve_kernel, dim3(nregs), dim3(kEmRows * kEmWaves), 0, s, desc,
                           sig, (int)it);
        EMSA_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(emicp_output_kernel, dim3(nregs), dim3(64), 0, s, desc, dRtOut, dResid);
    EMSA_HIP_TRY(hipGetLastError());
    if (&w == &ctx.em) {
        EMSA_HIP_TRY(hipEventRecord(ctx.emDone, s));
        ctx.emPending = true;
    }
    return EMSA_OK;
}

int em_launch_affinity(const DeviceCtx& ctx, EmWork& w, float sigma_p2, hipStream_t s,
                       unsigned long long* visits) {
    launch_affinity(ctx, w.desc.as<EmRegDesc>(), (int)w.regs.size(), (unsigned)w.tiles,
                    sigma_p2, s, visits, nullptr);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

int em_launch_setup(EmWork& w, const float* dRt0, hipStream_t s) {
    const unsigned nregs = (unsigned)w.regs.size();
    if (nregs > 0 && w.regs[0].rowperm) {
        hipLaunchKernelGGL(emicp_order_kernel, dim3(2 * nregs), dim3(kOrdThreads), 0, s,
                           w.desc.as<EmRegDesc>());
        EMSA_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(emicp_setup_kernel, dim3(nregs), dim3(256), 0, s,
                       w.desc.as<EmRegDesc>(), dRt0);
    EMSA_HIP_TRY(hipGetLastError());
    return EMSA_OK;
}

int emicp_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                 const float* dRtIn, float* dRtOut, const emsa_params& p,
                 const std::vector<float>& sigmas, hipStream_t s, double* trace, double* rows) {
    EmPair pr{dX, N, dY, M};
    int rc = em_prepare(ctx, ctx.em, &pr, 1, 1, 0, trace, rows, s);
    if (rc != EMSA_OK) return rc;
    return em_enqueue(ctx, ctx.em, dRtIn, dRtOut, nullptr, p, sigmas, s);
}

}  // namespace emsa
