/*
 * emsa oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference's soft-correspondence registration
 * iterations (tttamaki/cuda_emicp_softassign), used as the checker for the HIP
 * product path. Only tests/, the smoke check and the benchmark's cpu_baseline
 * leg may load this library; the product library never links or calls it.
 *
 * Contents
 *   - Horn's closed-form absolute orientation (src/findRTfromS.cpp:33-159),
 *     with a cyclic Jacobi eigensolver in place of LAPACK dsyev_
 *     (findRTfromS.cpp:118-155): or_find_rt_from_s_f64 / _f32.
 *   - EM-ICP (src/emicp_cpu.cpp:60-527, src/emicp.cu:443-668) in FP32
 *     (the restatement of emicp_cpu, timed as the CPU baseline) and in FP64
 *     (the oracle of the identical iteration): or_emicp_f32 / _f64 and the
 *     single-step or_emicp_step_f32 / _f64.
 *   - Softassign (src/softassign.cu:486-822) in FP32 and FP64:
 *     or_softassign_f32 / _f64 and or_softassign_inner_f32 / _f64.
 *
 * Pinning (how the tests fix this restatement, since the reference ships no
 * golden outputs and its CPU path needs PCL and LAPACK, absent here):
 *   - Jacobi eigensolver against numpy.linalg.eigh;
 *   - Horn against SVD-Kabsch and against the exact rigid transform relating
 *     the reference's data/P*.txt and data/Qnew*.txt pairs;
 *   - EM-ICP against a literal, materialised-matrix restatement written as
 *     plain Python loops at tiny sizes, against closed forms (uniform
 *     affinities at huge sigma, the floored-row values, exact recovery of the
 *     P/Qnew transform), and FP32 against FP64;
 *   - Softassign against the same literal-loop restatement at tiny sizes and
 *     the Sinkhorn invariants (row sums plus slack equal 1 after a row pass,
 *     column sums plus slack equal 1 after a column pass).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_TRACE_EMICP 32
#define OR_TRACE_SOFTASSIGN 32

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* ------------------------------------------------------------------------- */
/* Cyclic Jacobi for a symmetric 4x4 matrix (row-major, destroyed).           */
/* On return w[k] = eigenvalues, column k of V (V[4*i+k]) = eigenvector k.    */
/* ------------------------------------------------------------------------- */
int or_jacobi4(double *A, double *V, double *w)
{
    for (int i = 0; i < 16; i++) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
    int sweep;
    for (sweep = 0; sweep < 64; sweep++) {
        double off = 0, diag = 0;
        for (int p = 0; p < 4; p++) {
            diag += A[5 * p] * A[5 * p];
            for (int q = p + 1; q < 4; q++) off += A[4 * p + q] * A[4 * p + q];
        }
        if (!(off == off) || !(diag == diag)) return -1; /* NaN input */
        if (off <= 1e-36 * diag || off == 0.0) break;
        for (int p = 0; p < 3; p++)
            for (int q = p + 1; q < 4; q++) {
                double apq = A[4 * p + q];
                if (apq == 0.0) continue;
                double tau = (A[5 * q] - A[5 * p]) / (2.0 * apq);
                double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
                for (int k = 0; k < 4; k++) { /* columns p, q */
                    double akp = A[4 * k + p], akq = A[4 * k + q];
                    A[4 * k + p] = c * akp - s * akq;
                    A[4 * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; k++) { /* rows p, q */
                    double apk = A[4 * p + k], aqk = A[4 * q + k];
                    A[4 * p + k] = c * apk - s * aqk;
                    A[4 * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; k++) {
                    double vkp = V[4 * k + p], vkq = V[4 * k + q];
                    V[4 * k + p] = c * vkp - s * vkq;
                    V[4 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int k = 0; k < 4; k++) w[k] = A[5 * k];
    return sweep;
}

/* Horn's 4x4 matrix from S (findRTfromS.cpp:38-74): S_ab = h_S[3a+b]. */
#define HORN_N(Nm, S)                                                                             \
    do {                                                                                          \
        Nm[0] = S[0] + S[4] + S[8];                                                               \
        Nm[1] = S[5] - S[7];                                                                      \
        Nm[2] = S[6] - S[2];                                                                      \
        Nm[3] = S[1] - S[3];                                                                      \
        Nm[4] = S[5] - S[7];                                                                      \
        Nm[5] = S[0] - S[4] - S[8];                                                               \
        Nm[6] = S[1] + S[3];                                                                      \
        Nm[7] = S[6] + S[2];                                                                      \
        Nm[8] = S[6] - S[2];                                                                      \
        Nm[9] = S[1] + S[3];                                                                      \
        Nm[10] = S[4] - S[0] - S[8];                                                              \
        Nm[11] = S[5] + S[7];                                                                     \
        Nm[12] = S[1] - S[3];                                                                     \
        Nm[13] = S[6] + S[2];                                                                     \
        Nm[14] = S[5] + S[7];                                                                     \
        Nm[15] = S[8] - S[0] - S[4];                                                              \
    } while (0)

/* Top eigenvector of Horn's N; returns 0 on success, -1 on non-finite input.
 * Mirrors eigenvectorOfN (findRTfromS.cpp:108-159) without the static q_pre:
 * on failure q is left untouched (the caller keeps its previous estimate). */
int or_horn_quaternion(const double *Nin, double *q)
{
    double A[16], V[16], w[4];
    memcpy(A, Nin, sizeof(A));
    if (or_jacobi4(A, V, w) < 0) return -1;
    int k = 0;
    for (int j = 1; j < 4; j++)
        if (w[j] > w[k]) k = j;
    for (int i = 0; i < 4; i++) q[i] = V[4 * i + k];
    return 0;
}

/* FP64 Procrustes: R row-major, t = Xc - R Yc (findRTfromS.cpp:80-96). */
int or_find_rt_from_s_f64(const double *Xc, const double *Yc, const double *S, double *R,
                          double *t)
{
    double Nm[16], q[4];
    HORN_N(Nm, S);
    if (or_horn_quaternion(Nm, q) < 0) return -1;
    double q0 = q[0], qx = q[1], qy = q[2], qz = q[3];
    R[0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[1] = 2 * (qx * qy - q0 * qz);
    R[2] = 2 * (qx * qz + q0 * qy);
    R[3] = 2 * (qy * qx + q0 * qz);
    R[4] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[5] = 2 * (qy * qz - q0 * qx);
    R[6] = 2 * (qz * qx - q0 * qy);
    R[7] = 2 * (qz * qy + q0 * qx);
    R[8] = q0 * q0 - qx * qx - qy * qy + qz * qz;
    t[0] = Xc[0] - (R[0] * Yc[0] + R[1] * Yc[1] + R[2] * Yc[2]);
    t[1] = Xc[1] - (R[3] * Yc[0] + R[4] * Yc[1] + R[5] * Yc[2]);
    t[2] = Xc[2] - (R[6] * Yc[0] + R[7] * Yc[1] + R[8] * Yc[2]);
    return 0;
}

/* FP32 Procrustes with the reference's mixed precision: N entries formed in
 * float and widened (findRTfromS.cpp:59-74), eigensolve in double, q narrowed
 * to float, R and t in float (findRTfromS.cpp:80-96). */
int or_find_rt_from_s_f32(const float *Xc, const float *Yc, const float *S, float *R, float *t)
{
    double Nm[16], qd[4];
    HORN_N(Nm, S); /* each right-hand side is float arithmetic */
    if (or_horn_quaternion(Nm, qd) < 0) return -1;
    float q0 = (float)qd[0], qx = (float)qd[1], qy = (float)qd[2], qz = (float)qd[3];
    R[0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[1] = 2 * (qx * qy - q0 * qz);
    R[2] = 2 * (qx * qz + q0 * qy);
    R[3] = 2 * (qy * qx + q0 * qz);
    R[4] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[5] = 2 * (qy * qz - q0 * qx);
    R[6] = 2 * (qz * qx - q0 * qy);
    R[7] = 2 * (qz * qy + q0 * qx);
    R[8] = q0 * q0 - qx * qx - qy * qy + qz * qz;
    t[0] = Xc[0] - (R[0] * Yc[0] + R[1] * Yc[1] + R[2] * Yc[2]);
    t[1] = Xc[1] - (R[3] * Yc[0] + R[4] * Yc[1] + R[5] * Yc[2]);
    t[2] = Xc[2] - (R[6] * Yc[0] + R[7] * Yc[1] + R[8] * Yc[2]);
    return 0;
}

/* The sigma schedule of the EM-ICP loop, in float (emicp_cpu.cpp:305, :527). */
int or_sigma_schedule(float sigma_p2, float sigma_inf, float sigma_factor, float *out, int cap)
{
    int n = 0;
    while (sigma_p2 > sigma_inf) {
        if (out && n < cap) out[n] = sigma_p2;
        n++;
        sigma_p2 *= sigma_factor;
        if (n > 100000000) return -1;
    }
    return n;
}

/* ---- FP32 restatement (emicp_cpu) ---- */
#define REAL float
#define SFX f32
#define EXPR expf
#define SQRTR sqrtf
#define FABSR fabsf
#define FIND_RT(Xc, Yc, S, R, t) (void)or_find_rt_from_s_f32(Xc, Yc, S, R, t)
#include "emsa_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXPR
#undef SQRTR
#undef FABSR
#undef FIND_RT

/* ---- FP64 oracle of the identical iteration ---- */
#define REAL double
#define SFX f64
#define EXPR exp
#define SQRTR sqrt
#define FABSR fabs
#define FIND_RT(Xc, Yc, S, R, t) (void)or_find_rt_from_s_f64(Xc, Yc, S, R, t)
#include "emsa_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXPR
#undef SQRTR
#undef FABSR
#undef FIND_RT

int or_trace_width(void) { return OR_TRACE_EMICP; }
