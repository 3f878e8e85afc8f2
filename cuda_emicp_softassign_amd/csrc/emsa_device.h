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
// Horn's absolute orientation in FP64 (the maths of src/findRTfromS.cpp:33-159,
// eigensolver: cyclic Jacobi in place of LAPACK dsyev_).
// S[3a+b] = sum (source)_a (target)_b. Writes R (row-major) and t = Xc - R Yc.
// Returns false when the input is not finite (R, t untouched).
// ---------------------------------------------------------------------------
__device__ inline bool horn_rt(const double* Xc, const double* Yc, const double* S, double* R,
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
    double V[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    double diag = 0.0, off = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        diag += A[i][i] * A[i][i];
#pragma unroll
        for (int j = i + 1; j < 4; ++j) off += A[i][j] * A[i][j];
    }
    if (!isfinite(diag) || !isfinite(off)) return false;
    for (int sweep = 0; sweep < 32; ++sweep) {
        off = 0.0;
        diag = 0.0;
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
            for (int q = p + 1; q < 4; ++q) {
                const double apq = A[p][q];
                if (apq == 0.0) continue;
                const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = rsqrt(1.0 + tt * tt);
                const double s = tt * c;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
        }
    }
    // eigenvector of the largest eigenvalue, selected without dynamic indexing
    double best = A[0][0], q0 = V[0][0], qx = V[1][0], qy = V[2][0], qz = V[3][0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        if (A[k][k] > best) {
            best = A[k][k];
            q0 = V[0][k];
            qx = V[1][k];
            qy = V[2][k];
            qz = V[3][k];
        }
    }
    const double nrm = rsqrt(q0 * q0 + qx * qx + qy * qy + qz * qz);
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
        if (!isfinite(R[i])) return false;
    return true;
}

}  // namespace emsa
