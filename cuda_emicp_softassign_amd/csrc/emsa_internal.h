// Host-side internals shared by the translation units of libemsa.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emsa.h"

namespace emsa {

void set_error(const std::string& msg);
void clear_error();
// The calling thread's per-iteration hook (emsa_set_iteration_callback), and
// its replay of n trace records (R at record[17..25], t at [26..28]).
bool iteration_hook_set();
void report_iterations(int algorithm, const double* trace, int n);
bool finite_params(const emsa_params& p);
bool finite_cloud(const float* P, int n);

#define EMSA_HIP_TRY(call)                                                               \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::emsa::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return EMSA_EHIP;                                                            \
        }                                                                                \
    } while (0)

// Grow-only device buffer. Not freed by a destructor: contexts live until the
// process ends (freeing after the runtime's own teardown is unsafe); graph
// cache entries call release() when they are evicted.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    int reserve(size_t want);  // EMSA_OK / EMSA_ENOMEM
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Completion of the last work, on whatever stream it ran, that used a
// workspace shared by calls on one device. A call that rewrites the
// workspace first waits for it (wait), then marks its own work (a FenceMark
// records the event on the call's stream when it goes out of scope, so every
// return path marks what was already enqueued).
struct WorkFence {
    hipEvent_t ev = nullptr;
    bool pending = false;
    int wait();  // EMSA_OK / EMSA_EHIP
};
struct FenceMark {
    WorkFence& f;
    hipStream_t s;
    FenceMark(WorkFence& f_, hipStream_t s_) : f(f_), s(s_) {}
    FenceMark(const FenceMark&) = delete;
    FenceMark& operator=(const FenceMark&) = delete;
    ~FenceMark() {
        if (f.ev && hipEventRecord(f.ev, s) == hipSuccess) f.pending = true;
        else (void)hipGetLastError();
    }
};

// State of the EM-ICP loop, in doubles, one slot of kStateW per registration.
constexpr int kStateW = 32;    // R[9] t[3] Xsum[3] status ...
constexpr int kStR = 0, kStT = 9, kStXsum = 12, kStStatus = 15, kStResid = 16, kStLambda = 17;
constexpr int kMomW = 18;      // moments per row block (see emicp.hip)
constexpr int kMomPad = 20;    // moment rows allocated (a multiple of kEmWaves)
constexpr int kEmRows = 64;    // source rows per EM-ICP workgroup (one per lane)
constexpr int kEmWaves = 4;    // waves per workgroup, splitting the column chunk
constexpr int kColGroup = 16;  // columns per unrolled group of scalar loads
constexpr int kCullCols = 8;   // columns per bounding box of the target

struct EmPlan {
    int rowBlocks = 0;   // ceil(M / kEmRows)
    int chunks = 0;      // column chunks per row block
    int waveCols = 0;    // columns per wave (multiple of kColGroup)
    int ldx = 0;         // padded column count = roundup(N, kColGroup)
};

// Device-side descriptor of one EM-ICP registration (one source/target pair
// and one initial transform). A launch covers a table of them; workgroup b
// serves the registration whose [tileBase, tileBase + rowBlocks*chunks)
// contains b.
struct EmRegDesc {
    const float* X;      // target, SoA [3][N] as given
    float* Xpad;         // target, SoA [3][ldx] in working order, padded with far sentinels
    const float* Y;      // source, SoA [3][M] as given
    float* Ys;           // source in working order (an ordered copy, or Y itself)
    int* rowperm;        // working source row -> caller's row (null: identity)
    int* colperm;        // working target column -> caller's column (null: identity)
    float* gbox;         // [ldx / 8][8]: lo xyz, -, hi xyz, - of each 8-column group (null: none)
    unsigned* sortTmp;   // ordering scratch: 4 N words (target), then 4 M words (source)
    double* state;       // kStateW doubles
    float* partial;      // rowBlocks*chunks*5*kEmRows chunk partial sums
    double* moments;     // kMomPad x rowBlocks row-block moments (moment-major; rows
                         // kMomW.. are not written or used)
    double* trace;       // nullable: EMSA_TRACE_W doubles per sigma step
    double* rows;        // nullable: C_r, lambda_r (2*M) of the last step
    double* ckpt;        // nullable: [steps][M][5] C_r, lambda_r, lambda_r X'_r per step
    double floorA;       // (double)(1.0f / N), the floored affinity
    int N, M, ldx, rowBlocks, chunks, waveCols;
    int tileBase;        // first affinity workgroup of this registration
    int rowBase;         // first row block (over the whole table) of this registration
    int padOwner;        // 1: this registration writes Xpad (once per target cloud)
    int slot;            // output slot
};

// One source/target pair of an EM-ICP batch (device or host pointers,
// depending on the entry point).
struct EmPair {
    const float* X;
    int N;
    const float* Y;
    int M;
};

// Affinity workgroup order across sigma^2 steps: each step's affinity launch
// writes the work of every tile (its visited columns, weighted by the form of
// the exponent) and one more workgroup of the next launch sorts the tiles by
// it for the launch after that (so the sort runs beside the tiles, off the
// critical path), heaviest first, dealt to the workgroup slots so that the
// ones that share a compute unit (the hardware places workgroups b and
// b + numCU on the same one while all are resident) add up to about the same
// work. Used when the launches skip column groups (ordered clouds), so that
// the work per tile differs, for at most kBalanceMaxTiles tiles and at least
// kBalanceMinPairs pairs per launch (below, a launch is too short for the
// extra load of the order to pay). The order changes which workgroup
// computes a tile, never what it computes.
constexpr int kBalanceMaxTiles = 4096;
constexpr double kBalanceMinPairs = 16777216.0;

// Workspace and descriptor table of one EM-ICP batch.
struct EmWork {
    DevBuf xpad, partial, moments, state, desc, order;
    DevBuf sched;  // balanced order: work [2][tiles], then the tile of each slot [2][tiles]
    bool balance = false;  // sched is sized and the affinity launches use it
    std::vector<EmRegDesc> regs;
    std::vector<size_t> ordOff;  // per registration: its pair's offset in `order` (words);
                                 // empty when the clouds keep the caller's order
    long long tiles = 0;       // affinity workgroups per step
    long long rowBlocks = 0;   // row blocks per step over the whole table
    bool singleChunk = false;      // every registration's tile spans its whole target
    bool someSingleChunk = false;  // some registration's tile spans its whole target
};

// What the affinity kernel needs to finalize the rows itself (one chunk per
// row block): the step's outlier term, whether it is the last step, its index.
struct EmRowFuse {
    double outlier;
    int last;
    int iter;
};

struct GraphEntry {
    std::vector<uint64_t> key;
    EmWork work;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // recorded after each replay: the entry's last use
    uint64_t lastUse = 0;
};

struct DeviceCtx {
    int device = 0;
    int numCU = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // workspaces of the eager (non-graph) entry points
    DevBuf X, Y, rt, trace, rows, misc, resid;
    EmWork em;
    // softassign workspaces (sastep: staging of emsa_softassign_step's slack
    // vectors and taps)
    DevBuf K, u, v, orow, ocol, colpart, samom, sastate, sastep, sapack;
    // ICP workspaces: the ordered clouds, group boxes and per-row candidates
    // (O(N + M)), row-block moments, state, per-row neighbour and distance
    // (the correspondence query)
    DevBuf icpWork, icpMom, icpState, icpRowIdx, icpRowDist;
    // captured EM-ICP launch sequences of the device entry points, replayed
    // when a call repeats the same buffers, sizes and parameters; keys seen
    // once (a sequence is recorded on its second call)
    std::vector<std::unique_ptr<GraphEntry>> graphs;
    std::vector<std::vector<uint64_t>> seen;
    // workspaces of calls launched directly (not yet recorded), used in turn
    // so that consecutive calls do not wait for each other; ringDone[i] marks
    // the completion of the last sequence that used ring[i]
    static constexpr int kRing = 4;
    EmWork ring[kRing];
    hipEvent_t ringDone[kRing] = {};
    bool ringPending[kRing] = {};
    int ringNext = 0;
    // completion of the last launch sequence that used `em` on any stream;
    // em_prepare waits for it before rewriting that workspace
    hipEvent_t emDone = nullptr;
    bool emPending = false;
    int emTileCols = 0;        // batched column tile (0: heuristic); set by the autotuner
    int emAffinityMfma = 1;    // 1: B and W of the affinity on the matrix cores; 0: on the VALU
    int emCull = 1;            // skip column groups: 1 all A < 2^-128 (< FLT_MIN), 2 all e == 0, 0 none
    int emRowFuse = 1;         // 1: single-chunk tiles finalize their rows in the affinity kernel
    float emExpanded = 32.f;   // affinity exponent on the matrix cores while ka^2 a^2 <= this (0: never)
    int emBalance = 1;         // 1: affinity workgroups in the balanced order (kBalanceMaxTiles)
    int emBlocksPerCU = 0;     // plain plan occupancy target (0: heuristic)
};

// Returns the context of the current HIP device (created on first use) or
// nullptr with the error set.
DeviceCtx* current_ctx();
// The context of device dev if one was created, else nullptr (creates nothing).
DeviceCtx* find_ctx(int dev);

// The sigma^2 schedule in float, exactly as the reference loop runs it.
std::vector<float> sigma_schedule(const emsa_params& p);
// EMSA_EINVAL (error set, prefixed by fn) when that loop would not end:
// sigma_p2 > sigma_inf with sigma_factor outside (0, 1) or sigma_inf < 0 (the
// reference loops forever there, src/emicp.cu:443, :668); else EMSA_OK.
int check_schedule(const emsa_params& p, const char* fn);

EmPlan plan_emicp(int N, int M, int numCU, int chunks_override = 0);

// Lays out a batch of count pairs x n_starts initial transforms in w
// (allocating its buffers and uploading the descriptors on stream s).
// tileCols > 0: fixed column tile per workgroup (batched); 0: the single-
// registration plan that spreads one registration over the whole device.
// aux (nullable): per registration, a trace base ([steps][EMSA_TRACE_W]) and a
// checkpoint base ([steps][M][5]) for the unrolled (differentiable) forward.
struct EmAux {
    double* trace;
    double* ckpt;
};
// sync_upload: wait for the descriptor upload before returning (callers that
// rewrite w.regs before the stream has consumed it need this).
// em_prepare = em_layout (descriptors with buffer offsets, and the buffer
// sizes) + buffers of the sizes + em_bind (offsets to pointers) + upload.
struct EmSizes {
    size_t xpadF = 0;  // floats of the padded targets
    size_t partF = 0;  // floats of the chunk partial sums
    size_t momD = 0;   // doubles of the row-block moments
    size_t ordW = 0;   // 4-byte words of the ordering buffers (0: caller's order)
    int nregs = 0;     // registrations (kStateW doubles of state each)
};
int em_layout(const DeviceCtx& ctx, EmWork& w, const EmPair* pairs, int count, int n_starts,
              int tileCols, double* trace, double* rows, const EmAux* aux, EmSizes& sz);
void em_bind(EmWork& w, float* xpad, float* partial, double* moments, double* state,
             unsigned* order);
int em_prepare(DeviceCtx& ctx, EmWork& w, const EmPair* pairs, int count, int n_starts,
               int tileCols, double* trace, double* rows, hipStream_t s,
               const EmAux* aux = nullptr, bool sync_upload = true);

// Enqueues setup, one launch per sigma^2 step, and the output kernel.
// dRt0: [nregs][12] initial (R, t); dRtOut: [nregs][12]; dResid: [nregs] or null.
// Measurement hooks of one enqueued sequence (eager launches only): events
// (nullable, 2 per sigma^2 step) recorded around each affinity launch, and
// device counters (nullable, two per sigma^2 step) that the affinity launches
// add their evaluated pairs to (all, and those of the expanded exponent).
struct EmStepProbe {
    hipEvent_t* events;
    unsigned long long* visits;
};
int em_enqueue(DeviceCtx& ctx, EmWork& w, const float* dRt0, float* dRtOut, double* dResid,
               const emsa_params& p, const std::vector<float>& sigmas, hipStream_t s,
               const EmStepProbe* probe = nullptr);

// Single launches of the setup and affinity kernels (measurement probes).
int em_launch_setup(EmWork& w, const float* dRt0, hipStream_t s);
// visits (nullable, two counters): receive the number of pairs the launch
// evaluates (the matrix-core kernel; pairs in skipped column groups are not
// counted) and those evaluated in the expanded form.
int em_launch_affinity(const DeviceCtx& ctx, EmWork& w, float sigma_p2, hipStream_t s,
                       unsigned long long* visits = nullptr);
int default_tile_cols(const DeviceCtx& ctx);

// Runs the EM-ICP loop on device buffers for one registration (eager).
int emicp_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                 const float* dRtIn, float* dRtOut, const emsa_params& p,
                 const std::vector<float>& sigmas, hipStream_t s, double* trace, double* rows);

// One inner iteration of Softassign instead of the whole schedule
// (emsa_softassign_step): at temperature T, from the slack vectors orow [N]
// and ocol [M] (device, in/out: read after set-up, written back at the end),
// with the per-sweep taps a [I1][N], b [I1][M] and the post-Sinkhorn row sums
// s [N] (device, each nullable).
struct SaStep {
    float T;
    float* orow;
    float* ocol;
    double* a;
    double* b;
    double* s;
};

// Softassign on device buffers (see softassign.hip). xmax2: max |x|^2 over
// the target cloud (chooses the log domain's form of the exponent).
int softassign_device(DeviceCtx& ctx, const float* dX, int N, const float* dY, int M,
                      const float* dRtIn, float* dRtOut, const emsa_params& p, int variant,
                      hipStream_t s, double* trace, float xmax2, const SaStep* step = nullptr);

// Horn Procrustes on the device for one (Xc, Yc, S) (all float in, float out).
int find_rt_device(DeviceCtx& ctx, const float* hXc, const float* hYc, const float* hS,
                   float* hR, float* ht);

}  // namespace emsa
