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
// Both clouds are put in 3D Hilbert order once per call (the target padded,
// with the bounding box of every group of 16 consecutive target points), so
// that a wave's 64 source rows are close together in space. The search is a
// branch and bound that stays exact: each row starts from a valid candidate
// (its previous iteration's neighbour, or at the first iteration the target
// point at the same rank along the curve), and a wave skips a target group
// when for every one of its rows the group box is farther than the row's
// best distance so far. The box distance is computed with the same rounded
// operations as a point's distance, and rounding is monotone, so no point of
// a skipped group can have a smaller or equal computed distance.
// Per iteration three launches: nearest candidates per (64-row block, column
// chunk) tile, each folded into its row's 64-bit (distance, index) minimum
// with an atomic min (order-independent; O(M) memory), then per row block
// the row's neighbour and the block's FP64 moments (caller's row order),
// then one workgroup sums the moments in a fixed order and runs Horn.
#include "emsa_device.h"
#include "emsa_internal.h"
#include "emsa_order.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace emsa {

namespace {

constexpr int kIcpRows = 64;         // source rows per tile (one per lane)
constexpr int kIcpWaves = 4;         // waves per tile, taking the chunk's super-groups in turn
constexpr int kIcpGroup = 16;        // target points per group box
constexpr int kIcpSuper = 16;        // groups per super-group box
constexpr int kIcpSuperCols = kIcpGroup * kIcpSuper;  // target points per super-group
constexpr int kIcpHyper = 16;        // super-groups per hyper-group box
constexpr unsigned long long kIcpNone = ~0ull;      // no candidate yet
constexpr int kIcpMom = 16;          // moments per row block (see icp_rows_kernel)
constexpr int kIcpSolveWaves = 4;

struct IcpPlan {
    int rowBlocks = 0;    // ceil(M / 64)
    int supers = 0;       // super-groups of the padded target (ldx / kIcpSuperCols)
    int chunks = 0;       // column chunks per row block
    int chunkSupers = 0;  // super-groups per chunk
};

// Column chunks while there are fewer than about 8 tiles per CU (the search
// is latency-bound: scalar loads, then a ballot).
IcpPlan plan_icp(int N, int M, int numCU) {
    IcpPlan p;
    p.rowBlocks = (M + kIcpRows - 1) / kIcpRows;
    p.supers = (N + kIcpSuperCols - 1) / kIcpSuperCols;
    int chunks = (8 * std::max(1, numCU) + p.rowBlocks - 1) / p.rowBlocks;
    chunks = std::min(std::max(chunks, 1), p.supers);
    p.chunkSupers = (p.supers + chunks - 1) / chunks;
    p.chunks = (p.supers + p.chunkSupers - 1) / p.chunkSupers;
    return p;
}

// The call's working buffers (all O(N + M)).
struct IcpBufs {
    float* Xs;                // [3][ldx] target in working order, +inf past N
    int* Xperm;               // [ldx] working position -> caller's target index (N past N)
    float* gbox;              // [ldx / kIcpGroup][8] group boxes
    float* sbox;              // [ldx / kIcpSuperCols][8] super-group boxes
    float* hbox;              // [ceil(supers / kIcpHyper)][8] hyper-group boxes
    float* Ys;                // [3][M] source in working order
    int* Yperm;               // [M] working position -> caller's row
    unsigned* tmp;            // [4 N + 4 M] sort scratch
    unsigned long long* key;  // [M] per caller's row: (distance bits << 32) | target index
    int* prev;                // [M] per caller's row: the last neighbour (-1: none)
    double* mom;
    double* state;
    int ldx;
};

__global__ void icp_setup_kernel(const float* __restrict__ Rt, double* __restrict__ state) {
    if (threadIdx.x < 12) state[threadIdx.x] = (double)Rt[threadIdx.x];
    if (threadIdx.x == 12) state[kStStatus] = 0.0;
}

// out[k] = the union of boxes in[k * per .. min(n, (k + 1) * per) - 1], for
// all k < ceil(n / per) (an empty box's lo > hi drops out; a non-finite
// group's infinite box is kept).
__device__ void union_boxes(const float* __restrict__ in, int n, int per, float* __restrict__ out) {
    for (int k = threadIdx.x; k < (n + per - 1) / per; k += blockDim.x) {
        float bb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int i = k * per; i < min(n, (k + 1) * per); ++i)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bb[a] = fminf(bb[a], in[(size_t)i * 8 + a]);
                bb[3 + a] = fmaxf(bb[3 + a], in[(size_t)i * 8 + 4 + a]);
            }
        float4* o = reinterpret_cast<float4*>(out + (size_t)k * 8);
        o[0] = make_float4(bb[0], bb[1], bb[2], 0.f);
        o[1] = make_float4(bb[3], bb[4], bb[5], 0.f);
    }
}

// Workgroup 0 orders the target (and its group boxes), workgroup 1 the
// source; workgroup 1 also clears the per-row candidates.
__global__ void __launch_bounds__(kOrdThreads)
icp_order_kernel(const float* __restrict__ X, int N, const float* __restrict__ Y, int M,
                 IcpBufs b) {
    // one inlined sort for both clouds (workgroup-uniform arguments)
    const bool tgt = blockIdx.x == 0;
    order_cloud(tgt ? X : Y, tgt ? N : M, tgt ? b.Xperm : b.Yperm,
                b.tmp + (tgt ? 0 : 4 * (size_t)N), tgt ? b.Xs : b.Ys, tgt ? b.ldx : M,
                tgt ? INFINITY : 0.f);
    if (tgt) {
        for (int i = N + threadIdx.x; i < b.ldx; i += blockDim.x) b.Xperm[i] = N;
        __syncthreads();  // the boxes read the ordered target written above
        group_boxes<kIcpGroup>(b.Xs, b.ldx, N, b.gbox);
        const int supers = b.ldx / kIcpSuperCols;
        __syncthreads();
        union_boxes(b.gbox, supers * kIcpSuper, kIcpSuper, b.sbox);
        __syncthreads();
        union_boxes(b.sbox, supers, kIcpHyper, b.hbox);
    } else {
        for (int r = threadIdx.x; r < M; r += blockDim.x) {
            b.key[r] = kIcpNone;
            b.prev[r] = -1;
        }
    }
}

__global__ void icp_output_kernel(const double* __restrict__ state, float* __restrict__ Rt) {
    if (threadIdx.x < 12) Rt[threadIdx.x] = (float)state[threadIdx.x];
}

// The nearest of the chunk's target points for each of the tile's 64 source
// rows (working order): the row's starting candidate, then a three-level
// descent over the chunk's boxes (hyper-groups of 4096 points, super-groups
// of 256, groups of 16), each box skipped when it is farther than every
// row's best so far; in a visited group the points in working order,
// keeping a strictly smaller distance, or an equal one with a lower caller's
// index. Each box contains the boxes below it, and the group box its points,
// so by monotone rounding a computed bound never exceeds what it bounds.
// Boxes and target points are wave-uniform (scalar loads). The 4 waves are
// combined by the same rule, and each row's result is folded into its
// 64-bit key with an atomic min.
__global__ void __launch_bounds__(kIcpRows * kIcpWaves)
icp_nn_kernel(const float* __restrict__ X, int N, int M, IcpBufs b, int chunks, int chunkSupers) {
    typedef const __attribute__((address_space(4))) float* cf_p;
    typedef const __attribute__((address_space(4))) int* ci_p;
    const int rb = blockIdx.x / chunks;
    const int g = blockIdx.x - rb * chunks;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int rs = rb * kIcpRows + lane;  // working position of the row
    const bool valid = rs < M;
    float yx = 0.f, yy = 0.f, yz = 0.f;
    int orow = 0;
    if (valid) {
        yx = b.Ys[rs];
        yy = b.Ys[(size_t)M + rs];
        yz = b.Ys[2 * (size_t)M + rs];
        orow = b.Yperm[rs];
    }
    double Rm[9], tv[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Rm[i] = b.state[kStR + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = b.state[kStT + i];
    const double pxd = Rm[0] * yx + Rm[1] * yy + Rm[2] * yz + tv[0];
    const double pyd = Rm[3] * yx + Rm[4] * yy + Rm[5] * yz + tv[1];
    const double pzd = Rm[6] * yx + Rm[7] * yy + Rm[8] * yz + tv[2];
    const float px = (float)pxd, py = (float)pyd, pz = (float)pzd;
    const float plx = (float)(pxd - (double)px), ply = (float)(pyd - (double)py),
                plz = (float)(pzd - (double)pz);
    auto dist2 = [&](float cx, float cy, float cz) {
        const float dx = (cx - px) - plx;
        const float dy = (cy - py) - ply;
        const float dz = (cz - pz) - plz;
        float q = dx * dx;
        q = fmaf(dy, dy, q);
        q = fmaf(dz, dz, q);
        return q;
    };
    // a box's distance, by the point distance's own operations
    auto box2 = [&](cf_p bx) {
        const float dlx = (bx[0] - px) - plx, dhx = (bx[4] - px) - plx;
        const float dly = (bx[1] - py) - ply, dhy = (bx[5] - py) - ply;
        const float dlz = (bx[2] - pz) - plz, dhz = (bx[6] - pz) - plz;
        const float gx = fmaxf(fmaxf(dlx, -dhx), 0.f);
        const float gy = fmaxf(fmaxf(dly, -dhy), 0.f);
        const float gz = fmaxf(fmaxf(dlz, -dhz), 0.f);
        float q = gx * gx;
        q = fmaf(gy, gy, q);
        q = fmaf(gz, gz, q);
        return q;
    };
    // starting candidate: the previous neighbour, or the target point at the
    // same rank along the curve
    float best = INFINITY;
    int bi = N;
    if (valid) {
        int c = b.prev[orow];
        if (c < 0) c = b.Xperm[min(N - 1, (int)((long long)rs * N / M))];
        best = dist2(X[c], X[(size_t)N + c], X[2 * (size_t)N + c]);
        bi = c;
    }
    const cf_p Xx = (cf_p)b.Xs;
    const cf_p Xy = (cf_p)(b.Xs + b.ldx);
    const cf_p Xz = (cf_p)(b.Xs + 2 * (size_t)b.ldx);
    const ci_p Xp = (ci_p)b.Xperm;
    const cf_p Gb = (cf_p)b.gbox;
    const cf_p Sb = (cf_p)b.sbox;
    const cf_p Hb = (cf_p)b.hbox;
    const int S0 = g * chunkSupers;
    const int Send = min(S0 + chunkSupers, b.ldx / kIcpSuperCols);
    // Every wave tests every hyper-group box of the chunk and, in a visited
    // one, every super-group box in the chunk, 4 boxes at a time (loaded
    // together; all 4 tested against the best before any is searched: a
    // larger best only visits more); in a visited super-group wave w takes
    // groups w, w + 4, w + 8, w + 12, so that the neighbouring groups near
    // the rows are searched by different waves.
    static_assert(kIcpSuper == 4 * kIcpWaves, "one batch of 4 groups per wave and super-group");
    auto near = [&](cf_p boxes, int k) {
        return __ballot(valid && box2(boxes + (size_t)k * 8) <= best) != 0;
    };
    auto search_super = [&](int S) {
        const int Gs = S * kIcpSuper + wave;
        bool hit[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) hit[h] = near(Gb, Gs + 4 * h);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            if (!hit[h]) continue;
            const int c0 = (Gs + 4 * h) * kIcpGroup;
#pragma unroll
            for (int u = 0; u < kIcpGroup; ++u) {
                const float q = dist2(Xx[c0 + u], Xy[c0 + u], Xz[c0 + u]);
                const int oi = Xp[c0 + u];
                if (q < best || (q == best && oi < bi)) {
                    best = q;
                    bi = oi;
                }
            }
        }
    };
    // boxes k0 .. k1 - 1 of a level, 4 at a time (indexes clamped in bounds,
    // the ones past k1 never visited)
    auto sweep = [&](cf_p boxes, int k0, int k1, auto&& visit) {
        for (int k4 = k0; k4 < k1; k4 += 4) {
            bool hit[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) hit[j] = near(boxes, min(k4 + j, k1 - 1)) && k4 + j < k1;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (hit[j]) visit(k4 + j);
        }
    };
    sweep(Hb, S0 / kIcpHyper, (Send - 1) / kIcpHyper + 1, [&](int H) {
        sweep(Sb, max(S0, H * kIcpHyper), min(Send, (H + 1) * kIcpHyper), search_super);
    });
    __shared__ float rsd[kIcpWaves][kIcpRows];
    __shared__ int rsi[kIcpWaves][kIcpRows];
    rsd[wave][lane] = best;
    rsi[wave][lane] = bi;
    __syncthreads();
    if (wave != 0 || !valid) return;
#pragma unroll
    for (int w = 1; w < kIcpWaves; ++w) {
        const float q = rsd[w][lane];
        const int k = rsi[w][lane];
        if (q < best || (q == best && k < bi)) {
            best = q;
            bi = k;
        }
    }
    // distances are >= 0 (or +inf), so their bits order as the values do
    const unsigned long long key =
        ((unsigned long long)__float_as_uint(best) << 32) | (unsigned)bi;
    atomicMin(b.key + orow, key);
}

// One wave per row block (caller's row order): each row's neighbour from its
// key (the key is cleared for the next iteration and the neighbour kept as
// its starting candidate), then the block's FP64 moments, each a wave sum in
// a fixed order:
//   0..2 sum x_n(r), 3..5 sum y_r, 6 + 3a + b sum y_a x_b, 15 sum |d|^2,
// stored moment-major (mom[q * rowBlocks + rb]). rowIdx / rowDist
// (nullable): each row's neighbour and squared distance.
__global__ void __launch_bounds__(kWave)
icp_rows_kernel(const float* __restrict__ X, int N, const float* __restrict__ Y, int M, IcpBufs b,
                int rowBlocks, int* __restrict__ rowIdx, float* __restrict__ rowDist) {
    const int rb = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = rb * kIcpRows + lane;
    double m[kIcpMom];
#pragma unroll
    for (int q = 0; q < kIcpMom; ++q) m[q] = 0.0;
    if (r < M) {
        const unsigned long long key = b.key[r];
        b.key[r] = kIcpNone;
        const float best = __uint_as_float((unsigned)(key >> 32));
        const int bi = min(max((int)(unsigned)(key & 0xffffffffu), 0), N - 1);
        b.prev[r] = bi;
        const double x[3] = {X[bi], X[(size_t)N + bi], X[2 * (size_t)N + bi]};
        const double y[3] = {Y[r], Y[(size_t)M + r], Y[2 * (size_t)M + r]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            m[a] = x[a];
            m[3 + a] = y[a];
#pragma unroll
            for (int c = 0; c < 3; ++c) m[6 + 3 * a + c] = y[a] * x[c];
        }
        m[15] = best;
        if (rowIdx) rowIdx[r] = bi;
        if (rowDist) rowDist[r] = best;
    }
#pragma unroll
    for (int q = 0; q < kIcpMom; ++q) {
        const double v = wave_sum(m[q]);
        if (lane == 0) b.mom[(size_t)q * rowBlocks + rb] = v;
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

int icp_reserve(DeviceCtx& ctx, const IcpPlan& pl, int N, int M, bool rows, IcpBufs& b) {
    int rc;
    b.ldx = pl.supers * kIcpSuperCols;
    const size_t ldx = (size_t)b.ldx;
    // one allocation, carved: the ordered target and its boxes first (16-byte
    // rows: ldx is a multiple of 256), then the 8-byte keys, then 4-byte arrays
    const size_t hypers = ((size_t)pl.supers + kIcpHyper - 1) / kIcpHyper;
    const size_t bytes = sizeof(float) * (3 * ldx + (ldx / kIcpGroup) * 8 +
                                          (ldx / kIcpSuperCols) * 8 + hypers * 8) +
                         sizeof(unsigned long long) * (size_t)M +
                         sizeof(float) * (ldx + 3 * (size_t)M + (size_t)M + 4 * ((size_t)N + M) +
                                          (size_t)M);
    if ((rc = ctx.icpWork.reserve(bytes)) != EMSA_OK) return rc;
    if ((rc = ctx.icpMom.reserve(sizeof(double) * kIcpMom * (size_t)pl.rowBlocks)) != EMSA_OK)
        return rc;
    if ((rc = ctx.icpState.reserve(sizeof(double) * kStateW)) != EMSA_OK) return rc;
    if (rows) {
        if ((rc = ctx.icpRowIdx.reserve(sizeof(int) * (size_t)M)) != EMSA_OK) return rc;
        if ((rc = ctx.icpRowDist.reserve(sizeof(float) * (size_t)M)) != EMSA_OK) return rc;
    }
    float* f = ctx.icpWork.as<float>();
    b.Xs = f;
    f += 3 * ldx;
    b.gbox = f;
    f += (ldx / kIcpGroup) * 8;
    b.sbox = f;
    f += (ldx / kIcpSuperCols) * 8;
    b.hbox = f;
    f += hypers * 8;
    b.key = reinterpret_cast<unsigned long long*>(f);
    f = reinterpret_cast<float*>(b.key + M);
    b.Xperm = reinterpret_cast<int*>(f);
    f += ldx;
    b.Ys = f;
    f += 3 * (size_t)M;
    b.Yperm = reinterpret_cast<int*>(f);
    f += M;
    b.tmp = reinterpret_cast<unsigned*>(f);
    f += 4 * ((size_t)N + M);
    b.prev = reinterpret_cast<int*>(f);
    b.mom = ctx.icpMom.as<double>();
    b.state = ctx.icpState.as<double>();
    return EMSA_OK;
}

// Enqueues setup, the ordering of both clouds, `iters` iterations (iters == 0
// with rowIdx: one correspondence query, no solve) and the output kernel.
int icp_enqueue(const IcpPlan& pl, const IcpBufs& b, const float* dX, int N, const float* dY, int M,
                const float* dRtIn, float* dRtOut, int iters, double* trace, int* rowIdx,
                float* rowDist, hipStream_t s) {
    const unsigned tiles = (unsigned)((size_t)pl.rowBlocks * pl.chunks);
    hipLaunchKernelGGL(icp_setup_kernel, dim3(1), dim3(64), 0, s, dRtIn, b.state);
    hipLaunchKernelGGL(icp_order_kernel, dim3(2), dim3(kOrdThreads), 0, s, dX, N, dY, M, b);
    const int rounds = iters > 0 ? iters : 1;
    for (int it = 0; it < rounds; ++it) {
        hipLaunchKernelGGL(icp_nn_kernel, dim3(tiles), dim3(kIcpRows * kIcpWaves), 0, s, dX, N, M,
                           b, pl.chunks, pl.chunkSupers);
        hipLaunchKernelGGL(icp_rows_kernel, dim3((unsigned)pl.rowBlocks), dim3(kWave), 0, s, dX, N,
                           dY, M, b, pl.rowBlocks, iters > 0 ? nullptr : rowIdx,
                           iters > 0 ? nullptr : rowDist);
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
    if ((long long)pl.rowBlocks * pl.chunks >= (1LL << 24))  // 256 threads per tile, < 2^32
        return set_error(std::string(who) + ": clouds too large for one nearest-neighbour launch"),
               EMSA_EINVAL;
    IcpBufs b;
    if ((rc = icp_reserve(*ctx, pl, N, M, query, b)) != EMSA_OK) return rc;
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
