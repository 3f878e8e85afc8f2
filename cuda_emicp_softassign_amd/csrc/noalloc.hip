// EM-ICP on device buffers without any device allocation: the variant of
// emsa_emicp_dev for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc). Its
// workspace (descriptor, state, row-block moments, chunk partial sums, padded
// target, ordering buffers) lives in one file-scope device array of fixed
// size; the kernels and the work decomposition are those of emsa_emicp_dev,
// so the results are the same bit for bit.
//
// Each device holds its own copy of the array until the process ends. Every
// call on a device uses that device's copy and leaves the last call's
// contents in it; calls that overlap on one device (on different streams or
// from different threads without ordering) race on it.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <cstdio>

namespace emsa {

namespace {

constexpr size_t kScratchBytes = (size_t)16 << 20;  // 16 MiB per device
constexpr int kMaxDevices = 64;

// The workspace. In host code `scratch` names a host-side stand-in, not the
// device array: host code uses only the address hipGetSymbolAddress returns
// for the current device (g_scratch below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

size_t al256(size_t v) { return (v + 255) / 256 * 256; }

// Writes the registration's descriptor into the workspace; the descriptor
// travels as a kernel argument, so the host copy need not outlive the call.
__global__ void noalloc_store_desc_kernel(EmRegDesc d, EmRegDesc* __restrict__ out) {
    if (threadIdx.x == 0) *out = d;
}

// The device address of this device's workspace, looked up once per device
// (with that device current) on its first call.
int scratch_ptr(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("emicp_dev_noalloc: device index beyond " + std::to_string(kMaxDevices));
        return EMSA_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("emicp_dev_noalloc: hipGetSymbolAddress(scratch): ") +
                      hipGetErrorString(e));
            return EMSA_EHIP;
        }
        g_scratch[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch[dev]);
    return EMSA_OK;
}

int emicp_noalloc(DeviceCtx& c, unsigned char* base, const float* dX, int N, const float* dY,
                  int M, float* dRt, const emsa_params& p, const std::vector<float>& sig,
                  hipStream_t s) {
    EmWork w;  // host-side table only; its DevBufs are pointed into the workspace, never sized
    const EmPair pr{dX, N, dY, M};
    EmSizes sz;
    int rc = em_layout(c, w, &pr, 1, 1, 0, nullptr, nullptr, nullptr, sz);
    if (rc != EMSA_OK) return rc;
    size_t off = 0;
    const size_t oDesc = off;
    off += al256(sizeof(EmRegDesc));
    const size_t oState = off;
    off += al256(sizeof(double) * kStateW * (size_t)sz.nregs);
    const size_t oMom = off;
    off += al256(sizeof(double) * sz.momD);
    const size_t oPart = off;
    off += al256(sizeof(float) * sz.partF);
    const size_t oXpad = off;
    off += al256(sizeof(float) * sz.xpadF);
    const size_t oOrd = off;
    off += al256(4 * sz.ordW);
    if (off > kScratchBytes) {
        char msg[256];
        std::snprintf(msg, sizeof(msg),
                      "emicp_dev_noalloc: N = %d, M = %d need %zu bytes of workspace, more than "
                      "its fixed %zu bytes (use emsa_emicp_dev)",
                      N, M, off, kScratchBytes);
        set_error(msg);
        return EMSA_EINVAL;
    }
    em_bind(w, reinterpret_cast<float*>(base + oXpad), reinterpret_cast<float*>(base + oPart),
            reinterpret_cast<double*>(base + oMom), reinterpret_cast<double*>(base + oState),
            reinterpret_cast<unsigned*>(base + oOrd));
    EmRegDesc* dDesc = reinterpret_cast<EmRegDesc*>(base + oDesc);
    w.desc.ptr = dDesc;
    w.desc.bytes = 0;
    hipLaunchKernelGGL(noalloc_store_desc_kernel, dim3(1), dim3(64), 0, s, w.regs[0], dDesc);
    EMSA_HIP_TRY(hipGetLastError());
    rc = em_enqueue(c, w, dRt, dRt, nullptr, p, sig, s);
    w.desc.ptr = nullptr;  // the workspace is not this table's to free
    return rc;
}

}  // namespace

}  // namespace emsa

using namespace emsa;

extern "C" {

int emsa_emicp_dev_noalloc(const float* dX, int N, const float* dY, int M, float* dRt,
                           const emsa_params* p, void* stream) {
    clear_error();
    if (!dX || !dY || !dRt || !p || N <= 0 || M <= 0)
        return set_error("emicp_dev_noalloc: bad arguments"), EMSA_EINVAL;
    if (!finite_params(*p))
        return set_error("emicp_dev_noalloc: non-finite params"), EMSA_ENONFINITE;
    int rc = check_schedule(*p, "emicp_dev_noalloc");
    if (rc != EMSA_OK) return rc;
    const std::vector<float> sig = sigma_schedule(*p);
    if (sig.empty()) return EMSA_OK;
    int dev = 0;
    EMSA_HIP_TRY(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    if ((rc = scratch_ptr(dev, &base)) != EMSA_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    // the decomposition and kernel selection of the device's context when one
    // exists (so results equal emsa_emicp_dev's), else the defaults; no
    // context is created here (that would create a stream)
    DeviceCtx* ctx = find_ctx(dev);
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        return emicp_noalloc(*ctx, base, dX, N, dY, M, dRt, *p, sig, s);
    }
    DeviceCtx local;
    local.device = dev;
    EMSA_HIP_TRY(hipDeviceGetAttribute(&local.numCU, hipDeviceAttributeMultiprocessorCount, dev));
    return emicp_noalloc(local, base, dX, N, dY, M, dRt, *p, sig, s);
}

}  // extern "C"
