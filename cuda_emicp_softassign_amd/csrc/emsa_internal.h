// Host-side internals shared by the translation units of libemsa.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emsa.h"

namespace emsa {

void set_error(const std::string& msg);
void clear_error();

#define EMSA_HIP_TRY(call)                                                               \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::emsa::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return EMSA_EHIP;                                                            \
        }                                                                                \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t want);  // EMSA_OK / EMSA_ENOMEM
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// State of the EM-ICP loop, in doubles, one slot of kStateW per registration.
constexpr int kStateW = 32;    // R[9] t[3] Xsum[3] status ...
constexpr int kStR = 0, kStT = 9, kStXsum = 12, kStStatus = 15, kStResid = 16, kStLambda = 17;
constexpr int kMomW = 18;      // moments per row block (see emicp.hip)
constexpr int kEmRows = 64;    // source rows per EM-ICP workgroup (one per lane)
constexpr int kEmWaves = 4;    // waves per workgroup, splitting the column chunk
constexpr int kColGroup = 16;  // columns per unrolled group of scalar loads

struct EmPlan {
    int rowBlocks = 0;   // ceil(M / kEmRows)
    int chunks = 0;      // column chunks per row block
    int waveCols = 0;    // columns per wave (multiple of kColGroup)
    int ldx = 0;         // padded column count = roundup(N, kColGroup)
};

struct DeviceCtx {
    int device = 0;
    int numCU = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // workspaces
    DevBuf X, Y, Xpad, partial, moments, counters, state, rt, trace, rows, misc;
    // softassign workspaces
    DevBuf K, u, v, orow, ocol, rowsum, colsum, colpart, rowpart, sacnt, samom;
};

// Returns the context of the current HIP device (created on first use) or
// nullptr with the error set.
DeviceCtx* current_ctx();

// The sigma^2 schedule in float, exactly as the reference loop runs it.
std::vector<float> sigma_schedule(const emsa_params& p);

EmPlan plan_emicp(int N, int M, int numCU, int chunks_override = 0);

// Runs the EM-ICP loop on device buffers. dX, dY: SoA device arrays.
// dRtIn/dRtOut: 12 floats (R row-major, t). trace: device array of
// n*EMSA_TRACE_W doubles or null. rows: device array of 2*M doubles (C, lambda
// of the last step) or null.
int emicp_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                 const float* dRtIn, float* dRtOut, const emsa_params& p,
                 const std::vector<float>& sigmas, hipStream_t s, double* trace, double* rows);

// Softassign on device buffers (see softassign.hip).
int softassign_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                      const float* dRtIn, float* dRtOut, const emsa_params& p, int variant,
                      hipStream_t s, double* trace);

// Horn Procrustes on the device for one (Xc, Yc, S) (all float in, float out).
int find_rt_device(DeviceCtx& ctx, const float* hXc, const float* hYc, const float* hS,
                   float* hR, float* ht);

}  // namespace emsa
