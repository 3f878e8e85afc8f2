// Device helpers shared by the EM-ICP and Softassign kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace emsa {

constexpr int kWave = 64;
constexpr double kLog2e = 1.4426950408889634;

// Sum over the 64 lanes of a wave; every lane receives the total.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Minimum / maximum over the 64 lanes of a wave; every lane receives it.
__device__ inline float wave_min(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, kWave));
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
    return v;
}

// Sum over the 64 lanes of a wave with DPP-modified adds only (no LDS
// permutes); the total is valid in lane 63. Within each 16-lane row a
// butterfly (quad_perm xor 1, xor 2, row_half_mirror, row_mirror) leaves the
// row sum in every lane; row_bcast:15 into rows 1 and 3, then row_bcast:31
// into rows 2 and 3, accumulate the four row sums into row 3.
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK,
                                                          0xF, false));
}
__device__ inline float wave_sum_last(float v) {
    v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_add<0x140, 0xF>(v);  // row_mirror
    v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 -> rows 1, 3
    v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 -> rows 2, 3
    return v;
}

// Minimum / maximum over the 64 lanes of a wave with DPP-modified operations
// only, in the pattern of wave_sum_last; valid in lane 63. Lanes a DPP
// operation does not write keep their value (the `old` operand is v).
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_src(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL,
                                                      ROW_MASK, 0xF, false));
}
__device__ inline float wave_min_last(float v) {
    v = fminf(v, dpp_src<0xB1, 0xF>(v));
    v = fminf(v, dpp_src<0x4E, 0xF>(v));
    v = fminf(v, dpp_src<0x141, 0xF>(v));
    v = fminf(v, dpp_src<0x140, 0xF>(v));
    v = fminf(v, dpp_src<0x142, 0xA>(v));
    v = fminf(v, dpp_src<0x143, 0xC>(v));
    return v;
}
__device__ inline float wave_max_last(float v) {
    v = fmaxf(v, dpp_src<0xB1, 0xF>(v));
    v = fmaxf(v, dpp_src<0x4E, 0xF>(v));
    v = fmaxf(v, dpp_src<0x141, 0xF>(v));
    v = fmaxf(v, dpp_src<0x140, 0xF>(v));
    v = fmaxf(v, dpp_src<0x142, 0xA>(v));
    v = fmaxf(v, dpp_src<0x143, 0xC>(v));
    return v;
}
// The value of lane 63, wave-uniform (an SGPR).
__device__ inline float lane63(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Block-wide sum of K doubles per thread for a block of `nwaves` waves.
// lds must hold nwaves*K doubles. Thread k < K of the block receives total k
// in out (other threads: unspecified). Fixed reduction order: deterministic.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* lds, int nwaves, double& out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    }
    __syncthreads();
    out = 0.0;
    if (threadIdx.x < K) {
        for (int w = 0; w < nwaves; ++w) out += lds[w * K + threadIdx.x];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Horn's absolute orientation in FP64 (the maths of src/findRTfromS.cpp:33-159).
// S[3a+b] = sum (source)_a (target)_b. Writes R (row-major) and t = Xc - R Yc.
// Returns false when the input is not finite (R, t then unspecified).
//
// The top eigenvector of Horn's symmetric 4x4 matrix N replaces LAPACK dsyev_
// (findRTfromS.cpp:118-155). Fast path, after the quaternion characteristic
// polynomial method (QCP: Theobald 2005; Liu, Agrafiotis & Theobald 2010):
// the largest root of det(N - l I) by Newton's method from a Gershgorin upper
// bound (monotone, since all roots are real), then the eigenvector as the
// largest column of adj(N - l I). When that column is tiny or the residual
// check fails (near-degenerate top eigenvalue), a cyclic Jacobi solve is used.
// ---------------------------------------------------------------------------
__host__ __device__ inline double det3(double a, double b, double c, double d, double e, double f,
                              double g, double h, double i) {
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

__host__ __device__ inline void jacobi_top4(double A[4][4], double q[4]) {
    double V[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = 0.0, diag = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            diag += A[i][i] * A[i][i];
#pragma unroll
            for (int j = i + 1; j < 4; ++j) off += A[i][j] * A[i][j];
        }
        if (off <= 1e-36 * diag || off == 0.0) break;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int r = p + 1; r < 4; ++r) {
                const double apq = A[p][r];
                if (apq == 0.0) continue;
                const double tau = (A[r][r] - A[p][p]) / (2.0 * apq);
                const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + tt * tt);
                const double s = tt * c;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double akp = A[k][p], akq = A[k][r];
                    A[k][p] = c * akp - s * akq;
                    A[k][r] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double apk = A[p][k], aqk = A[r][k];
                    A[p][k] = c * apk - s * aqk;
                    A[r][k] = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double vkp = V[k][p], vkq = V[k][r];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][r] = s * vkp + c * vkq;
                }
            }
        }
    }
    // eigenvector of the largest eigenvalue, selected without dynamic indexing
    double best = A[0][0];
    q[0] = V[0][0]; q[1] = V[1][0]; q[2] = V[2][0]; q[3] = V[3][0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        if (A[k][k] > best) {
            best = A[k][k];
            q[0] = V[0][k]; q[1] = V[1][k]; q[2] = V[2][k]; q[3] = V[3][k];
        }
    }
}

// Fast path; false when the caller must fall back to Jacobi.
__host__ __device__ inline bool newton_top4(const double A[4][4], double q[4]) {
    // Faddeev-LeVerrier: det(lI - A) = l^4 + c3 l^3 + c2 l^2 + c1 l + c0
    double Mk[4][4], T[4][4];
    const double c3 = -(A[0][0] + A[1][1] + A[2][2] + A[3][3]);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Mk[i][j] = A[i][j] + (i == j ? c3 : 0.0);
    double cs[3];
#pragma unroll
    for (int k = 2; k <= 4; ++k) {
        double tr = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double acc = 0.0;
                for (int l = 0; l < 4; ++l) acc += A[i][l] * Mk[l][j];
                T[i][j] = acc;
            }
        for (int i = 0; i < 4; ++i) tr += T[i][i];
        const double c = -tr / k;
        cs[k - 2] = c;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) Mk[i][j] = T[i][j] + (i == j ? c : 0.0);
    }
    const double c2 = cs[0], c1 = cs[1], c0 = cs[2];
    double scale = 0.0, lam = -1e300;
    for (int i = 0; i < 4; ++i) {
        double r = A[i][i];
        for (int j = 0; j < 4; ++j) {
            scale = fmax(scale, fabs(A[i][j]));
            if (j != i) r += fabs(A[i][j]);
        }
        lam = fmax(lam, r);  // Gershgorin: an upper bound of the largest eigenvalue
    }
    if (!(scale > 0.0)) return false;
    for (int it = 0; it < 80; ++it) {
        const double p = (((lam + c3) * lam + c2) * lam + c1) * lam + c0;
        const double dp = ((4.0 * lam + 3.0 * c3) * lam + 2.0 * c2) * lam + c1;
        if (!(dp > 0.0)) break;
        const double step = p / dp;
        lam -= step;
        if (fabs(step) <= 1e-15 * scale) break;
    }
    // adjugate of B = A - lam I: row i of the cofactor matrix (B symmetric)
    double B[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) B[i][j] = A[i][j] - (i == j ? lam : 0.0);
    double best = -1.0;
    for (int i = 0; i < 4; ++i) {
        double v[4];
        int rows[3], n = 0;
        for (int k = 0; k < 4; ++k)
            if (k != i) rows[n++] = k;
        for (int j = 0; j < 4; ++j) {
            int cols[3], m = 0;
            for (int k = 0; k < 4; ++k)
                if (k != j) cols[m++] = k;
            const double dd = det3(B[rows[0]][cols[0]], B[rows[0]][cols[1]], B[rows[0]][cols[2]],
                                   B[rows[1]][cols[0]], B[rows[1]][cols[1]], B[rows[1]][cols[2]],
                                   B[rows[2]][cols[0]], B[rows[2]][cols[1]], B[rows[2]][cols[2]]);
            v[j] = ((i + j) & 1) ? -dd : dd;
        }
        const double nn = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
        if (nn > best) {
            best = nn;
            for (int k = 0; k < 4; ++k) q[k] = v[k];
        }
    }
    if (!(best > 1e-60 * scale * scale * scale * scale * scale * scale)) return false;
    const double inv = 1.0 / sqrt(best);
    for (int k = 0; k < 4; ++k) q[k] *= inv;
    // residual check |A q - lam q| small against the matrix scale
    double res = 0.0;
    for (int i = 0; i < 4; ++i) {
        double r = -lam * q[i];
        for (int j = 0; j < 4; ++j) r += A[i][j] * q[j];
        res = fmax(res, fabs(r));
    }
    return res <= 1e-10 * scale;
}

__host__ __device__ inline bool horn_rt(const double* Xc, const double* Yc, const double* S, double* R,
                               double* t) {
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
    double fro = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) fro += A[i][j] * A[i][j];
    if (!__builtin_isfinite(fro)) return false;
    double q[4];
    if (!newton_top4(A, q)) jacobi_top4(A, q);
    double q0 = q[0], qx = q[1], qy = q[2], qz = q[3];
    const double nrm = 1.0 / sqrt(q0 * q0 + qx * qx + qy * qy + qz * qz);
    q0 *= nrm; qx *= nrm; qy *= nrm; qz *= nrm;
    R[0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[1] = 2.0 * (qx * qy - q0 * qz);
    R[2] = 2.0 * (qx * qz + q0 * qy);
    R[3] = 2.0 * (qy * qx + q0 * qz);
    R[4] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[5] = 2.0 * (qy * qz - q0 * qx);
    R[6] = 2.0 * (qz * qx - q0 * qy);
    R[7] = 2.0 * (qz * qy + q0 * qx);
    R[8] = q0 * q0 - qx * qx - qy * qy + qz * qz;
    for (int i = 0; i < 3; ++i)
        t[i] = Xc[i] - (R[3 * i] * Yc[0] + R[3 * i + 1] * Yc[1] + R[3 * i + 2] * Yc[2]);
    for (int i = 0; i < 9; ++i)
        if (!__builtin_isfinite(R[i])) return false;
    return true;
}

}  // namespace emsa
