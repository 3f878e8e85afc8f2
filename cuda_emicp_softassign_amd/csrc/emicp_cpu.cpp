// EM-ICP on the host CPU in FP32: the library's counterpart of the
// reference's emicp_cpu (src/emicp_cpu.cpp:186-547), for callers without a
// GPU or that compare the two. Same iteration and operation order as the
// reference, in FP32 with libm expf/sqrtf:
//   A_rc  = expf(-(|x_c - (R y_r + t)|^2) / sigma^2), the difference taken per
//           component and the squares added x, y, z (emicp_cpu.cpp:82-101)
//   C_r   = sum_c A_rc + expf(-d_02 / sigma^2)                       (:342-357)
//   A'_rc = sqrtf(A_rc / C_r), or 1/N when C_r <= 1e-6               (:113-126)
//   lambda_r = sum_c A'_rc, Lambda = sum_r |lambda_r|                (:380-391)
//   X'_r  = (sum_c A'_rc x_c) / lambda_r                             (:412-419)
//   Xc    = (sum_r lambda_r X'_r) / Lambda, Yc likewise              (:435-459)
//   S[i+3j] = sum_r lambda_r (X'_r - Xc)_i (y_r - Yc)_j             (:469-495)
//   (R, t) <- Horn's step (findRTfromS, :505); sigma^2 *= factor.
// No N x M matrix is stored: each source row's affinities live in a per-thread
// row buffer. Rows are split over host threads; every sum over the rows is then
// taken in row order on one thread, so results do not depend on the thread
// count. Horn's step is the library's FP64 solver (the one the GPU kernels
// run), rounded to FP32 as the reference's float R and t; a failed solve keeps
// (R, t) for that step (the reference reuses a process-global quaternion).
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace emsa {

namespace {

int host_threads(int rows) {
    int n = 0;
    for (const char* v : {"EMSA_CPU_THREADS", "OMP_NUM_THREADS"}) {
        const char* e = std::getenv(v);
        if (e && std::atoi(e) > 0) {
            n = std::atoi(e);
            break;
        }
    }
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min({n, 256, rows}));
}

// Rows [r0, r1) of one sigma^2 step: C_r, lambda_r and X'_r.
void step_rows(const float* X, int N, const float* Y, int M, const float* R, const float* t,
               float sigma_p2, float outlier, int r0, int r1, float* C, float* lam, float* Xp) {
    std::vector<float> A((size_t)N);
    const float* Xx = X;
    const float* Xy = X + N;
    const float* Xz = X + 2 * (size_t)N;
    const float floorA = 1.0f / N;  // emicp_cpu.cpp:125
    for (int r = r0; r < r1; ++r) {
        const float Yx = Y[r], Yy = Y[(size_t)M + r], Yz = Y[2 * (size_t)M + r];
        float Cr = 0.f;
        for (int c = 0; c < N; ++c) {
            float dx = Xx[c] - (R[0] * Yx + R[1] * Yy + R[2] * Yz + t[0]);
            float dy = Xy[c] - (R[3] * Yx + R[4] * Yy + R[5] * Yz + t[1]);
            float dz = Xz[c] - (R[6] * Yx + R[7] * Yy + R[8] * Yz + t[2]);
            dx *= dx;
            dy *= dy;
            dz *= dz;
            dx += dy;
            dx += dz;
            dx /= sigma_p2;
            const float a = expf(-dx);
            A[c] = a;
            Cr += a;
        }
        Cr += outlier;
        float l = 0.f, wx = 0.f, wy = 0.f, wz = 0.f;
        if (Cr > 10e-7f) {
            for (int c = 0; c < N; ++c) {
                const float a = sqrtf(A[c] / Cr);
                l += a;
                wx += a * Xx[c];
                wy += a * Xy[c];
                wz += a * Xz[c];
            }
        } else {
            for (int c = 0; c < N; ++c) {
                l += floorA;
                wx += floorA * Xx[c];
                wy += floorA * Xy[c];
                wz += floorA * Xz[c];
            }
        }
        C[r] = Cr;
        lam[r] = l;
        Xp[r] = wx / l;  // elementwiseDivision, emicp_cpu.cpp:419
        Xp[(size_t)M + r] = wy / l;
        Xp[2 * (size_t)M + r] = wz / l;
    }
}

}  // namespace

int emicp_cpu_run(const float* X, int N, const float* Y, int M, float* Rio, float* tio,
                  const emsa_params& p, double* trace, int trace_cap) {
    const std::vector<float> sig = sigma_schedule(p);
    float R[9], t[3];
    std::memcpy(R, Rio, sizeof(R));
    std::memcpy(t, tio, sizeof(t));
    std::vector<float> C((size_t)M), lam((size_t)M), Xp(3 * (size_t)M);
    const int T = host_threads(M);
    for (size_t it = 0; it < sig.size(); ++it) {
        const float s2 = sig[it];
        const float outlier = expf(-p.d_02 / s2);  // emicp_cpu.cpp:356
        std::vector<std::thread> pool;
        const int per = (M + T - 1) / T;
        for (int k = 0; k < T; ++k) {
            const int r0 = k * per, r1 = std::min(M, r0 + per);
            if (r0 >= r1) break;
            pool.emplace_back(step_rows, X, N, Y, M, R, t, s2, outlier, r0, r1, C.data(),
                              lam.data(), Xp.data());
        }
        for (auto& th : pool) th.join();
        // sums over the rows in row order (sasum, sgemv 't', sgemm 't', 'n')
        float Lam = 0.f, Xc[3] = {0.f, 0.f, 0.f}, Yc[3] = {0.f, 0.f, 0.f};
        for (int r = 0; r < M; ++r) {
            Lam += std::fabs(lam[r]);
            for (int i = 0; i < 3; ++i) {
                Xc[i] += Xp[(size_t)i * M + r] * lam[r];
                Yc[i] += Y[(size_t)i * M + r] * lam[r];
            }
        }
        const float inv = 1.0f / Lam;  // emicp_cpu.cpp:457
        for (int i = 0; i < 3; ++i) {
            Xc[i] *= inv;
            Yc[i] *= inv;
        }
        float S[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int r = 0; r < M; ++r) {
            float xc[3], yc[3];
            for (int i = 0; i < 3; ++i) {
                xc[i] = (Xp[(size_t)i * M + r] - Xc[i]) * lam[r];  // centering, then weighting
                yc[i] = Y[(size_t)i * M + r] - Yc[i];
            }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) S[i + 3 * j] += xc[i] * yc[j];
        }
        double Xcd[3], Ycd[3], Sd[9], Rd[9], td[3];
        for (int i = 0; i < 3; ++i) {
            Xcd[i] = Xc[i];
            Ycd[i] = Yc[i];
        }
        for (int i = 0; i < 9; ++i) Sd[i] = S[i];
        const bool ok = horn_rt(Xcd, Ycd, Sd, Rd, td);
        if (ok) {
            for (int i = 0; i < 9; ++i) R[i] = (float)Rd[i];
            for (int i = 0; i < 3; ++i) t[i] = (float)td[i];
        }
        if (trace && (int)it < trace_cap) {
            double* o = trace + it * EMSA_TRACE_W;
            std::fill(o, o + EMSA_TRACE_W, 0.0);
            o[0] = s2;
            o[1] = Lam;
            for (int i = 0; i < 3; ++i) {
                o[2 + i] = Xc[i];
                o[5 + i] = Yc[i];
                o[26 + i] = t[i];
            }
            for (int i = 0; i < 9; ++i) {
                o[8 + i] = S[i];
                o[17 + i] = R[i];
            }
            o[31] = ok ? 0.0 : 1.0;
        }
    }
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error("emicp_cpu: non-finite result"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error("emicp_cpu: non-finite result"), EMSA_ENONFINITE;
    std::memcpy(Rio, R, sizeof(R));
    std::memcpy(tio, t, sizeof(t));
    return (int)sig.size();
}

}  // namespace emsa

using namespace emsa;

extern "C" {

int emsa_emicp_cpu(const float* X, int N, const float* Y, int M, float* R, float* t,
                   const emsa_params* p, double* trace, int trace_cap) {
    clear_error();
    if (!X || !Y || !R || !t || !p || N <= 0 || M <= 0 || trace_cap < 0)
        return set_error("emicp_cpu: null pointer or empty cloud"), EMSA_EINVAL;
    int rc = check_schedule(*p, "emicp_cpu");
    if (rc != EMSA_OK) return rc;
    if (!finite_params(*p) || !finite_cloud(X, N) || !finite_cloud(Y, M))
        return set_error("emicp_cpu: non-finite input"), EMSA_ENONFINITE;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error("emicp_cpu: non-finite R"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error("emicp_cpu: non-finite t"), EMSA_ENONFINITE;
    if (!trace && iteration_hook_set()) {
        std::vector<double> own((size_t)std::max<size_t>(sigma_schedule(*p).size(), 1) *
                                EMSA_TRACE_W);
        rc = emicp_cpu_run(X, N, Y, M, R, t, *p, own.data(), (int)(own.size() / EMSA_TRACE_W));
        if (rc >= 0) report_iterations(EMSA_ALG_EMICP_CPU, own.data(), rc);
        return rc;
    }
    rc = emicp_cpu_run(X, N, Y, M, R, t, *p, trace, trace ? trace_cap : 0);
    if (rc >= 0 && trace) report_iterations(EMSA_ALG_EMICP_CPU, trace, std::min(rc, trace_cap));
    return rc < 0 ? rc : std::min(rc, trace ? trace_cap : rc);
}

}  // extern "C"
