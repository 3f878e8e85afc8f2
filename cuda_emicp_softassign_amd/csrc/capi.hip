// The C ABI of libemsa (include/emsa.h): argument checks, per-device context
// and workspace, host<->device staging around the device entry points.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <cmath>
#include <cstring>
#include <memory>

namespace emsa {

namespace {
thread_local std::string g_err;
thread_local emsa_iteration_fn g_hook = nullptr;
thread_local void* g_hook_user = nullptr;
std::mutex g_ctx_mu;
// never destroyed: contexts outlive every static destructor that may still call in
std::vector<std::unique_ptr<DeviceCtx>>& g_ctx = *new std::vector<std::unique_ptr<DeviceCtx>>();

}  // namespace

bool finite_params(const emsa_params& p) {
    const float f[] = {p.alpha, p.T_0, p.TFACTOR, p.moutlier, p.sigma_p2, p.sigma_inf,
                       p.sigma_factor, p.d_02};
    for (float v : f)
        if (!std::isfinite(v)) return false;
    return true;
}

bool finite_cloud(const float* P, int n) {
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        if (!std::isfinite(P[i])) return false;
    return true;
}

namespace {

__global__ void find_rt_kernel(const float* __restrict__ in, float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double Xc[3], Yc[3], S[9], R[9], t[3];
    for (int i = 0; i < 3; ++i) {
        Xc[i] = in[i];
        Yc[i] = in[3 + i];
    }
    for (int i = 0; i < 9; ++i) S[i] = in[6 + i];
    const bool ok = horn_rt(Xc, Yc, S, R, t);
    for (int i = 0; i < 9; ++i) out[i] = (float)R[i];
    for (int i = 0; i < 3; ++i) out[9 + i] = (float)t[i];
    out[12] = ok ? 0.f : 1.f;
}

}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

bool iteration_hook_set() { return g_hook != nullptr; }

void report_iterations(int algorithm, const double* trace, int n) {
    const emsa_iteration_fn fn = g_hook;
    void* const user = g_hook_user;
    if (!fn) return;
    for (int i = 0; i < n; ++i) {
        const double* o = trace + (size_t)i * EMSA_TRACE_W;
        float R[9], t[3];
        for (int k = 0; k < 9; ++k) R[k] = (float)o[17 + k];
        for (int k = 0; k < 3; ++k) t[k] = (float)o[26 + k];
        fn(algorithm, i, R, t, user);
    }
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int DevBuf::reserve(size_t want) {
    if (want <= bytes) return EMSA_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    const size_t cap = want + want / 4 + 256;
    if (hipMalloc(&ptr, cap) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        set_error("hipMalloc of " + std::to_string(cap) + " bytes failed");
        return EMSA_ENOMEM;
    }
    bytes = cap;
    return EMSA_OK;
}

int WorkFence::wait() {
    if (!ev) EMSA_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (pending) {
        EMSA_HIP_TRY(hipEventSynchronize(ev));
        pending = false;
    }
    return EMSA_OK;
}

DeviceCtx* current_ctx() {
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return nullptr;
    }
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if ((int)g_ctx.size() < ndev) g_ctx.resize(ndev);
    if (!g_ctx[dev]) {
        auto c = std::make_unique<DeviceCtx>();
        c->device = dev;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            set_error("hipGetDeviceProperties failed");
            return nullptr;
        }
        c->numCU = prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreateWithFlags failed");
            return nullptr;
        }
        if (hipEventCreateWithFlags(&c->emDone, hipEventDisableTiming) != hipSuccess) {
            set_error("hipEventCreateWithFlags failed");
            return nullptr;
        }
        g_ctx[dev] = std::move(c);
    }
    return g_ctx[dev].get();
}

DeviceCtx* find_ctx(int dev) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (dev < 0 || dev >= (int)g_ctx.size()) return nullptr;
    return g_ctx[dev].get();
}

int find_rt_device(DeviceCtx& ctx, const float* hXc, const float* hYc, const float* hS, float* hR,
                   float* ht) {
    int rc;
    if ((rc = ctx.misc.reserve(sizeof(float) * 64)) != EMSA_OK) return rc;
    float in[15], out[13];
    std::memcpy(in, hXc, 12);
    std::memcpy(in + 3, hYc, 12);
    std::memcpy(in + 6, hS, 36);
    float* d = ctx.misc.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(d, in, sizeof(in), hipMemcpyHostToDevice, ctx.stream));
    hipLaunchKernelGGL(find_rt_kernel, dim3(1), dim3(64), 0, ctx.stream, d, d + 16);
    EMSA_HIP_TRY(hipGetLastError());
    EMSA_HIP_TRY(hipMemcpyAsync(out, d + 16, sizeof(out), hipMemcpyDeviceToHost, ctx.stream));
    EMSA_HIP_TRY(hipStreamSynchronize(ctx.stream));
    if (out[12] != 0.f) {
        set_error("findRTfromS: non-finite input");
        return EMSA_ENONFINITE;
    }
    std::memcpy(hR, out, 36);
    std::memcpy(ht, out + 9, 12);
    return EMSA_OK;
}

// Host-buffer EM-ICP shared by emsa_emicp / _traced / _step.
static int emicp_host(const float* X, int N, const float* Y, int M, float* R, float* t,
                      const emsa_params& p, const std::vector<float>& sigmas, double* htrace,
                      int trace_cap, double* hC, double* hLam) {
    if (!X || !Y || !R || !t || N <= 0 || M <= 0) {
        set_error("emicp: null pointer or empty cloud");
        return EMSA_EINVAL;
    }
    if (!finite_params(p) || !finite_cloud(X, N) || !finite_cloud(Y, M)) {
        set_error("emicp: non-finite input");
        return EMSA_ENONFINITE;
    }
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error("emicp: non-finite R"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error("emicp: non-finite t"), EMSA_ENONFINITE;
    if (sigmas.empty()) return EMSA_OK;  // loop body never runs: R, t unchanged
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    int rc;
    const int nrec = (int)sigmas.size();
    if ((rc = ctx->X.reserve(sizeof(float) * 3 * (size_t)N)) != EMSA_OK) return rc;
    if ((rc = ctx->Y.reserve(sizeof(float) * 3 * (size_t)M)) != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 32)) != EMSA_OK) return rc;
    if (htrace && (rc = ctx->trace.reserve(sizeof(double) * EMSA_TRACE_W * nrec)) != EMSA_OK)
        return rc;
    if ((hC || hLam) && (rc = ctx->rows.reserve(sizeof(double) * 2 * (size_t)M)) != EMSA_OK)
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
    rc = emicp_device(*ctx, ctx->X.as<float>(), N, ctx->Y.as<float>(), M, drt, drt + 16, p,
                      sigmas, s, htrace ? ctx->trace.as<double>() : nullptr,
                      (hC || hLam) ? ctx->rows.as<double>() : nullptr);
    if (rc != EMSA_OK) return rc;
    float out[12];
    EMSA_HIP_TRY(hipMemcpyAsync(out, drt + 16, sizeof(out), hipMemcpyDeviceToHost, s));
    std::vector<double> rows;
    if (hC || hLam) {
        rows.resize(2 * (size_t)M);
        EMSA_HIP_TRY(hipMemcpyAsync(rows.data(), ctx->rows.ptr, sizeof(double) * 2 * (size_t)M,
                                    hipMemcpyDeviceToHost, s));
    }
    std::vector<double> tr;
    if (htrace) {
        tr.resize((size_t)EMSA_TRACE_W * nrec);
        EMSA_HIP_TRY(hipMemcpyAsync(tr.data(), ctx->trace.ptr, sizeof(double) * tr.size(),
                                    hipMemcpyDeviceToHost, s));
    }
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(out[i])) return set_error("emicp: non-finite result"), EMSA_ENONFINITE;
    std::memcpy(R, out, 36);
    std::memcpy(t, out + 9, 12);
    if (htrace) {
        const int n = std::min(nrec, trace_cap);
        std::memcpy(htrace, tr.data(), sizeof(double) * EMSA_TRACE_W * (size_t)n);
    }
    if (hC) for (int i = 0; i < M; ++i) hC[i] = rows[i];
    if (hLam) for (int i = 0; i < M; ++i) hLam[i] = rows[(size_t)M + i];
    return EMSA_OK;
}

// Host-buffer Softassign shared by emsa_softassign / _traced.
static int softassign_host(const float* X, int N, const float* Y, int M, float* R, float* t,
                           const emsa_params& p, int variant, double* htrace, int trace_cap) {
    if (!X || !Y || !R || !t || N <= 0 || M <= 0) {
        set_error("softassign: null pointer or empty cloud");
        return EMSA_EINVAL;
    }
    if (p.I1 < 1) {  // softassign.cu:653-666 would read uninitialised column sums
        set_error("softassign: I1 must be >= 1");
        return EMSA_EINVAL;
    }
    if (variant != EMSA_SOFTASSIGN_MATERIALISED && variant != EMSA_SOFTASSIGN_LOGDOMAIN) {
        set_error("softassign: unknown variant");
        return EMSA_EINVAL;
    }
    if (!finite_params(p) || !finite_cloud(X, N) || !finite_cloud(Y, M) || !(p.T_0 > 0.f)) {
        set_error("softassign: non-finite input or T_0 <= 0");
        return EMSA_ENONFINITE;
    }
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error("softassign: non-finite R"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error("softassign: non-finite t"), EMSA_ENONFINITE;
    const int nrec = std::max(0, p.JMAX) * std::max(0, p.I0);
    if (nrec == 0) return EMSA_OK;  // no inner iteration: R, t unchanged
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = ctx->X.reserve(sizeof(float) * 3 * (size_t)N)) != EMSA_OK) return rc;
    if ((rc = ctx->Y.reserve(sizeof(float) * 3 * (size_t)M)) != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 32)) != EMSA_OK) return rc;
    if (htrace && (rc = ctx->trace.reserve(sizeof(double) * EMSA_TRACE_W * nrec)) != EMSA_OK)
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
    double xmax2 = 0.0;  // max |x|^2 of the target
    for (int r = 0; r < N; ++r) {
        const double a = X[r], b = X[(size_t)N + r], c = X[2 * (size_t)N + r];
        xmax2 = std::max(xmax2, a * a + b * b + c * c);
    }
    rc = softassign_device(*ctx, ctx->X.as<float>(), N, ctx->Y.as<float>(), M, drt, drt + 16, p,
                           variant, s, htrace ? ctx->trace.as<double>() : nullptr, (float)xmax2);
    if (rc != EMSA_OK) return rc;
    float out[12];
    EMSA_HIP_TRY(hipMemcpyAsync(out, drt + 16, sizeof(out), hipMemcpyDeviceToHost, s));
    std::vector<double> tr;
    if (htrace) {
        tr.resize((size_t)EMSA_TRACE_W * nrec);
        EMSA_HIP_TRY(hipMemcpyAsync(tr.data(), ctx->trace.ptr, sizeof(double) * tr.size(),
                                    hipMemcpyDeviceToHost, s));
    }
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(out[i]))
            return set_error("softassign: non-finite result"), EMSA_ENONFINITE;
    std::memcpy(R, out, 36);
    std::memcpy(t, out + 9, 12);
    if (htrace) {
        const int n = std::min(nrec, trace_cap);
        std::memcpy(htrace, tr.data(), sizeof(double) * EMSA_TRACE_W * (size_t)n);
    }
    return EMSA_OK;
}

// One inner iteration of Softassign on host buffers (emsa_softassign_step).
static int softassign_step_host(const float* X, int N, const float* Y, int M, float* R, float* t,
                                float T, float alpha, int I1, float* o_row, float* o_col,
                                int variant, double* ha, double* hb, double* hs, double* record) {
    if (!X || !Y || !R || !t || !o_row || !o_col || N <= 0 || M <= 0) {
        set_error("softassign_step: null pointer or empty cloud");
        return EMSA_EINVAL;
    }
    if (I1 < 1) return set_error("softassign_step: I1 must be >= 1"), EMSA_EINVAL;
    if (variant != EMSA_SOFTASSIGN_MATERIALISED && variant != EMSA_SOFTASSIGN_LOGDOMAIN)
        return set_error("softassign_step: unknown variant"), EMSA_EINVAL;
    if (!(T > 0.f) || !std::isfinite(T) || !std::isfinite(alpha) || !finite_cloud(X, N) ||
        !finite_cloud(Y, M))
        return set_error("softassign_step: non-finite input or T <= 0"), EMSA_ENONFINITE;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(R[i])) return set_error("softassign_step: non-finite R"), EMSA_ENONFINITE;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(t[i])) return set_error("softassign_step: non-finite t"), EMSA_ENONFINITE;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    int rc;
    const size_t nTap = (size_t)I1 * N + (size_t)I1 * M + (size_t)N;
    if ((rc = ctx->X.reserve(sizeof(float) * 3 * (size_t)N)) != EMSA_OK) return rc;
    if ((rc = ctx->Y.reserve(sizeof(float) * 3 * (size_t)M)) != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 32)) != EMSA_OK) return rc;
    if ((rc = ctx->trace.reserve(sizeof(double) * EMSA_TRACE_W)) != EMSA_OK) return rc;
    if ((rc = ctx->sastep.reserve(sizeof(double) * nTap + sizeof(float) * ((size_t)N + M))) !=
        EMSA_OK)
        return rc;
    double* dA = ctx->sastep.as<double>();
    double* dB = dA + (size_t)I1 * N;
    double* dS = dB + (size_t)I1 * M;
    float* dOrow = reinterpret_cast<float*>(dA + nTap);
    float* dOcol = dOrow + N;
    float rt[12];
    std::memcpy(rt, R, 36);
    std::memcpy(rt + 9, t, 12);
    float* drt = ctx->rt.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->X.ptr, X, sizeof(float) * 3 * (size_t)N,
                                hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(ctx->Y.ptr, Y, sizeof(float) * 3 * (size_t)M,
                                hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(drt, rt, sizeof(rt), hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(dOrow, o_row, sizeof(float) * N, hipMemcpyHostToDevice, s));
    EMSA_HIP_TRY(hipMemcpyAsync(dOcol, o_col, sizeof(float) * M, hipMemcpyHostToDevice, s));
    double xmax2 = 0.0;
    for (int r = 0; r < N; ++r) {
        const double a = X[r], b = X[(size_t)N + r], c = X[2 * (size_t)N + r];
        xmax2 = std::max(xmax2, a * a + b * b + c * c);
    }
    emsa_params p;
    emsa_default_params(&p);
    p.I1 = I1;
    p.alpha = alpha;
    const SaStep st{T, dOrow, dOcol, dA, dB, dS};
    rc = softassign_device(*ctx, ctx->X.as<float>(), N, ctx->Y.as<float>(), M, drt, drt + 16, p,
                           variant, s, ctx->trace.as<double>(), (float)xmax2, &st);
    if (rc != EMSA_OK) return rc;
    float out[12];
    std::vector<double> taps(nTap), rec(EMSA_TRACE_W);
    std::vector<float> orow(N), ocol(M);
    EMSA_HIP_TRY(hipMemcpyAsync(out, drt + 16, sizeof(out), hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipMemcpyAsync(taps.data(), dA, sizeof(double) * nTap, hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipMemcpyAsync(rec.data(), ctx->trace.ptr, sizeof(double) * EMSA_TRACE_W,
                                hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipMemcpyAsync(orow.data(), dOrow, sizeof(float) * N, hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipMemcpyAsync(ocol.data(), dOcol, sizeof(float) * M, hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(out[i]))
            return set_error("softassign_step: non-finite result"), EMSA_ENONFINITE;
    std::memcpy(R, out, 36);
    std::memcpy(t, out + 9, 12);
    std::memcpy(o_row, orow.data(), sizeof(float) * N);
    std::memcpy(o_col, ocol.data(), sizeof(float) * M);
    if (ha) std::memcpy(ha, taps.data(), sizeof(double) * I1 * (size_t)N);
    if (hb) std::memcpy(hb, taps.data() + (size_t)I1 * N, sizeof(double) * I1 * (size_t)M);
    if (hs) std::memcpy(hs, taps.data() + (size_t)I1 * (N + M), sizeof(double) * N);
    if (record) std::memcpy(record, rec.data(), sizeof(double) * EMSA_TRACE_W);
    return EMSA_OK;
}

}  // namespace emsa

using namespace emsa;

extern "C" {

void emsa_default_params(emsa_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->JMAX = 100;
    p->I0 = 5;
    p->I1 = 3;
    p->alpha = 3.0f;
    p->T_0 = 100.0f;
    p->TFACTOR = 0.95f;
    p->moutlier = (float)(1 / sqrtf(p->T_0) * expf(-1.0f));  // main.cpp:286
    p->sigma_p2 = 0.01f;
    p->sigma_inf = 0.00001f;
    p->sigma_factor = 0.9f;
    p->d_02 = 0.01f;
    p->maxIteration = 30;
}

const char* emsa_last_error(void) { return g_err.c_str(); }

int emsa_emicp_iterations(const emsa_params* p) {
    clear_error();
    if (!p) return set_error("emicp_iterations: null params"), EMSA_EINVAL;
    const int rc = check_schedule(*p, "emicp_iterations");
    if (rc != EMSA_OK) return rc;
    return (int)sigma_schedule(*p).size();
}

void emsa_set_iteration_callback(emsa_iteration_fn fn, void* user) {
    g_hook = fn;
    g_hook_user = fn ? user : nullptr;
}

int emsa_emicp(const float* X, int N, const float* Y, int M, float* R, float* t,
               const emsa_params* p) {
    clear_error();
    if (!p) return set_error("emicp: null params"), EMSA_EINVAL;
    const int rc = check_schedule(*p, "emicp");
    if (rc != EMSA_OK) return rc;
    const std::vector<float> sig = sigma_schedule(*p);
    if (!iteration_hook_set())
        return emicp_host(X, N, Y, M, R, t, *p, sig, nullptr, 0, nullptr, nullptr);
    std::vector<double> trace(std::max<size_t>(sig.size(), 1) * EMSA_TRACE_W);
    const int hrc = emicp_host(X, N, Y, M, R, t, *p, sig, trace.data(), (int)sig.size(), nullptr,
                               nullptr);
    if (hrc == EMSA_OK) report_iterations(EMSA_ALG_EMICP, trace.data(), (int)sig.size());
    return hrc;
}

int emsa_emicp_traced(const float* X, int N, const float* Y, int M, float* R, float* t,
                      const emsa_params* p, double* trace, int trace_cap) {
    clear_error();
    if (!p || !trace || trace_cap < 0) return set_error("emicp: bad trace args"), EMSA_EINVAL;
    const int crc = check_schedule(*p, "emicp_traced");
    if (crc != EMSA_OK) return crc;
    const std::vector<float> sig = sigma_schedule(*p);
    const int rc = emicp_host(X, N, Y, M, R, t, *p, sig, trace, trace_cap, nullptr, nullptr);
    return rc == EMSA_OK ? (int)std::min<size_t>(sig.size(), (size_t)trace_cap) : rc;
}

int emsa_emicp_step(const float* X, int N, const float* Y, int M, float* R, float* t,
                    float sigma_p2, float d_02, double* C, double* lambda, double* record) {
    clear_error();
    emsa_params p;
    emsa_default_params(&p);
    p.sigma_p2 = sigma_p2;
    p.d_02 = d_02;
    if (!(sigma_p2 > 0.f)) return set_error("emicp_step: sigma_p2 must be > 0"), EMSA_EINVAL;
    std::vector<float> sig(1, sigma_p2);
    double rec[EMSA_TRACE_W];
    const int rc = emicp_host(X, N, Y, M, R, t, p, sig, rec, 1, C, lambda);
    if (rc == EMSA_OK && record) std::memcpy(record, rec, sizeof(rec));
    return rc;
}

int emsa_softassign(const float* X, int N, const float* Y, int M, float* R, float* t,
                    const emsa_params* p, int variant) {
    clear_error();
    if (!p) return set_error("softassign: null params"), EMSA_EINVAL;
    if (!iteration_hook_set()) return softassign_host(X, N, Y, M, R, t, *p, variant, nullptr, 0);
    const int nrec = std::max(0, p->JMAX) * std::max(0, p->I0);
    std::vector<double> trace(std::max(nrec, 1) * (size_t)EMSA_TRACE_W);
    const int rc = softassign_host(X, N, Y, M, R, t, *p, variant, trace.data(), nrec);
    if (rc == EMSA_OK) report_iterations(EMSA_ALG_SOFTASSIGN, trace.data(), nrec);
    return rc;
}

int emsa_softassign_traced(const float* X, int N, const float* Y, int M, float* R, float* t,
                           const emsa_params* p, int variant, double* trace, int trace_cap) {
    clear_error();
    if (!p || !trace || trace_cap < 0) return set_error("softassign: bad trace args"), EMSA_EINVAL;
    const int rc = softassign_host(X, N, Y, M, R, t, *p, variant, trace, trace_cap);
    const int nrec = std::max(0, p->JMAX) * std::max(0, p->I0);
    return rc == EMSA_OK ? std::min(nrec, trace_cap) : rc;
}

int emsa_softassign_step(const float* X, int N, const float* Y, int M, float* R, float* t,
                         float T, float alpha, int I1, float* o_row, float* o_col, int variant,
                         double* a, double* b, double* s, double* record) {
    clear_error();
    return softassign_step_host(X, N, Y, M, R, t, T, alpha, I1, o_row, o_col, variant, a, b, s,
                                record);
}

int emsa_find_rt_from_s(const float* Xc, const float* Yc, const float* S, float* R, float* t) {
    clear_error();
    if (!Xc || !Yc || !S || !R || !t) return set_error("findRTfromS: null pointer"), EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return find_rt_device(*ctx, Xc, Yc, S, R, t);
}

}  // extern "C"
