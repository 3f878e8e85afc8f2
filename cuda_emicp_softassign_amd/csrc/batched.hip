// Device-buffer EM-ICP entry points: single registrations, uneven-size
// multi-start batches, the hipGraph cache that replays their launch sequence,
// and the autotuner that picks their work decomposition.
#include "emsa_device.h"
#include "emsa_internal.h"

#include <algorithm>
#include <cstring>

namespace emsa {

namespace {

constexpr size_t kMaxGraphs = 32;
constexpr size_t kMaxSeen = 64;
constexpr int kDefaultTileCols = 2048;
uint64_t g_clock = 0;

uint64_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

void push_params(std::vector<uint64_t>& k, const emsa_params& p) {
    k.push_back(fbits(p.sigma_p2));
    k.push_back(fbits(p.sigma_inf));
    k.push_back(fbits(p.sigma_factor));
    k.push_back(fbits(p.d_02));
}

// Waits for the entry's last replay (on whatever stream it ran), then frees
// it; other work on the device is not waited for.
void drop_entry(GraphEntry& e) {
    if (e.done) {
        (void)hipEventSynchronize(e.done);
        (void)hipEventDestroy(e.done);
        e.done = nullptr;
    }
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
    e.work.xpad.release();
    e.work.partial.release();
    e.work.moments.release();
    e.work.state.release();
    e.work.desc.release();
}

int tile_cols(const DeviceCtx& ctx) { return default_tile_cols(ctx); }

// Runs this call's launch sequence on `user`: replayed from the cache when
// the call was recorded before; launched directly (with the context's own
// workspace) the first time a call is seen; recorded and replayed when it is
// seen again. A sequence is recorded on the library's own stream (the null
// stream cannot be captured) and only kernels are recorded: the workspace is
// set up (descriptors uploaded) before capture.
int run_graph(DeviceCtx& ctx, const std::vector<uint64_t>& key, const EmPair* pairs, int count,
              int n_starts, int tileCols, const float* dRt0, float* dRtOut, double* dResid,
              const emsa_params& p, const std::vector<float>& sig, hipStream_t user) {
    ++g_clock;
    for (auto& e : ctx.graphs) {
        if (e->key == key) {
            e->lastUse = g_clock;
            // the entry's workspace is shared by its replays: a replay on
            // another stream starts after the previous one (a device-side wait)
            EMSA_HIP_TRY(hipStreamWaitEvent(user, e->done, 0));
            EMSA_HIP_TRY(hipGraphLaunch(e->exec, user));
            EMSA_HIP_TRY(hipEventRecord(e->done, user));
            return EMSA_OK;
        }
    }
    auto seen = std::find(ctx.seen.begin(), ctx.seen.end(), key);
    if (seen == ctx.seen.end()) {
        if (ctx.seen.size() >= kMaxSeen) ctx.seen.erase(ctx.seen.begin());
        ctx.seen.push_back(key);
        // direct launch with the next ring workspace: waits only for the call
        // that used this slot kRing calls ago (its descriptors and host table
        // are then free to rewrite)
        const int slot = ctx.ringNext;
        ctx.ringNext = (slot + 1) % DeviceCtx::kRing;
        EmWork& w = ctx.ring[slot];
        if (!ctx.ringDone[slot])
            EMSA_HIP_TRY(hipEventCreateWithFlags(&ctx.ringDone[slot], hipEventDisableTiming));
        if (ctx.ringPending[slot]) {
            EMSA_HIP_TRY(hipEventSynchronize(ctx.ringDone[slot]));
            ctx.ringPending[slot] = false;
        }
        int rc = em_prepare(ctx, w, pairs, count, n_starts, tileCols, nullptr, nullptr, user,
                            nullptr, false);
        if (rc != EMSA_OK) return rc;
        if ((rc = em_enqueue(ctx, w, dRt0, dRtOut, dResid, p, sig, user)) != EMSA_OK) return rc;
        EMSA_HIP_TRY(hipEventRecord(ctx.ringDone[slot], user));
        ctx.ringPending[slot] = true;
        return EMSA_OK;
    }
    ctx.seen.erase(seen);
    if (ctx.graphs.size() >= kMaxGraphs) {  // evict the least recently used
        auto it = std::min_element(ctx.graphs.begin(), ctx.graphs.end(),
                                   [](const std::unique_ptr<GraphEntry>& a,
                                      const std::unique_ptr<GraphEntry>& b) {
                                       return a->lastUse < b->lastUse;
                                   });
        drop_entry(**it);
        ctx.graphs.erase(it);
    }
    auto e = std::make_unique<GraphEntry>();
    e->key = key;
    EMSA_HIP_TRY(hipEventCreateWithFlags(&e->done, hipEventDisableTiming));
    int rc = em_prepare(ctx, e->work, pairs, count, n_starts, tileCols, nullptr, nullptr,
                        ctx.stream);
    if (rc != EMSA_OK) {
        drop_entry(*e);
        return rc;
    }
    EMSA_HIP_TRY(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
    rc = em_enqueue(ctx, e->work, dRt0, dRtOut, dResid, p, sig, ctx.stream);
    const hipError_t ec = hipStreamEndCapture(ctx.stream, &e->graph);
    if (rc != EMSA_OK || ec != hipSuccess) {
        if (rc == EMSA_OK) set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        drop_entry(*e);
        return rc != EMSA_OK ? rc : EMSA_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        drop_entry(*e);
        return EMSA_EHIP;
    }
    e->lastUse = g_clock;
    EMSA_HIP_TRY(hipGraphLaunch(e->exec, user));
    EMSA_HIP_TRY(hipEventRecord(e->done, user));
    ctx.graphs.push_back(std::move(e));
    return EMSA_OK;
}

int check_pairs(const emsa_pair* pairs, int count) {
    if (!pairs || count <= 0) return set_error("emicp: empty batch"), EMSA_EINVAL;
    for (int k = 0; k < count; ++k)
        if (!pairs[k].X || !pairs[k].Y || pairs[k].N <= 0 || pairs[k].M <= 0)
            return set_error("emicp: null or empty cloud in batch"), EMSA_EINVAL;
    return EMSA_OK;
}

}  // namespace

int default_tile_cols(const DeviceCtx& ctx) {
    return ctx.emTileCols > 0 ? ctx.emTileCols : kDefaultTileCols;
}

}  // namespace emsa

using namespace emsa;

extern "C" {

int emsa_emicp_dev(const float* dX, int N, const float* dY, int M, float* dRt,
                   const emsa_params* p, void* stream) {
    clear_error();
    if (!dX || !dY || !dRt || !p || N <= 0 || M <= 0)
        return set_error("emicp_dev: bad arguments"), EMSA_EINVAL;
    if (!finite_params(*p)) return set_error("emicp_dev: non-finite params"), EMSA_ENONFINITE;
    int rc = check_schedule(*p, "emicp_dev");
    if (rc != EMSA_OK) return rc;
    const std::vector<float> sig = sigma_schedule(*p);
    if (sig.empty()) return EMSA_OK;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<uint64_t> key = {1, (uint64_t)dX, (uint64_t)N, (uint64_t)dY, (uint64_t)M,
                                 (uint64_t)dRt, (uint64_t)ctx->emBlocksPerCU};
    push_params(key, *p);
    const EmPair pr{dX, N, dY, M};
    return run_graph(*ctx, key, &pr, 1, 1, 0, dRt, dRt, nullptr, *p, sig, (hipStream_t)stream);
}

int emsa_emicp_batched_dev(const emsa_pair* pairs, int count, const float* dRt0, int n_starts,
                           const emsa_params* p, float* dRtOut, double* dResid, void* stream) {
    clear_error();
    int rc;
    if ((rc = check_pairs(pairs, count)) != EMSA_OK) return rc;
    if (!dRt0 || !dRtOut || !p || n_starts <= 0)
        return set_error("emicp_batched_dev: bad arguments"), EMSA_EINVAL;
    if (!finite_params(*p)) return set_error("emicp_batched_dev: non-finite params"), EMSA_ENONFINITE;
    if ((rc = check_schedule(*p, "emicp_batched_dev")) != EMSA_OK) return rc;
    const std::vector<float> sig = sigma_schedule(*p);
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (sig.empty()) {  // no step runs: the result is the initial transform
        EMSA_HIP_TRY(hipMemcpyAsync(dRtOut, dRt0, sizeof(float) * 12 * (size_t)count * n_starts,
                                    hipMemcpyDeviceToDevice, (hipStream_t)stream));
        return EMSA_OK;
    }
    std::vector<uint64_t> key = {2, (uint64_t)count, (uint64_t)n_starts, (uint64_t)dRt0,
                                 (uint64_t)dRtOut, (uint64_t)dResid, (uint64_t)tile_cols(*ctx)};
    push_params(key, *p);
    for (int k = 0; k < count; ++k) {
        key.push_back((uint64_t)pairs[k].X);
        key.push_back((uint64_t)pairs[k].N);
        key.push_back((uint64_t)pairs[k].Y);
        key.push_back((uint64_t)pairs[k].M);
    }
    std::vector<EmPair> pv(count);
    for (int k = 0; k < count; ++k) pv[k] = EmPair{pairs[k].X, pairs[k].N, pairs[k].Y, pairs[k].M};
    return run_graph(*ctx, key, pv.data(), count, n_starts, tile_cols(*ctx), dRt0, dRtOut, dResid,
                     *p, sig, (hipStream_t)stream);
}

int emsa_emicp_batched(const emsa_pair* pairs, int count, const float* Rt0, int n_starts,
                       const emsa_params* p, float* Rt_out, double* residual) {
    clear_error();
    int rc;
    if ((rc = check_pairs(pairs, count)) != EMSA_OK) return rc;
    if (!Rt0 || !Rt_out || !p || n_starts <= 0)
        return set_error("emicp_batched: bad arguments"), EMSA_EINVAL;
    if (!finite_params(*p)) return set_error("emicp_batched: non-finite params"), EMSA_ENONFINITE;
    if ((rc = check_schedule(*p, "emicp_batched")) != EMSA_OK) return rc;
    const size_t nregs = (size_t)count * n_starts;
    for (size_t i = 0; i < 12 * nregs; ++i)
        if (!std::isfinite(Rt0[i])) return set_error("emicp_batched: non-finite Rt0"), EMSA_ENONFINITE;
    size_t total = 0;
    for (int k = 0; k < count; ++k) {
        if (!finite_cloud(pairs[k].X, pairs[k].N) || !finite_cloud(pairs[k].Y, pairs[k].M))
            return set_error("emicp_batched: non-finite input"), EMSA_ENONFINITE;
        total += 3 * ((size_t)pairs[k].N + pairs[k].M);
    }
    const std::vector<float> sig = sigma_schedule(*p);
    if (sig.empty()) {
        std::memcpy(Rt_out, Rt0, sizeof(float) * 12 * nregs);
        if (residual) std::fill(residual, residual + nregs, 0.0);
        return EMSA_OK;
    }
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    if ((rc = ctx->X.reserve(sizeof(float) * total)) != EMSA_OK) return rc;
    if ((rc = ctx->rt.reserve(sizeof(float) * 24 * nregs)) != EMSA_OK) return rc;
    if ((rc = ctx->resid.reserve(sizeof(double) * nregs)) != EMSA_OK) return rc;
    std::vector<EmPair> pv(count);
    float* base = ctx->X.as<float>();
    size_t off = 0;
    for (int k = 0; k < count; ++k) {
        const size_t nx = 3 * (size_t)pairs[k].N, ny = 3 * (size_t)pairs[k].M;
        EMSA_HIP_TRY(hipMemcpyAsync(base + off, pairs[k].X, sizeof(float) * nx,
                                    hipMemcpyHostToDevice, s));
        EMSA_HIP_TRY(hipMemcpyAsync(base + off + nx, pairs[k].Y, sizeof(float) * ny,
                                    hipMemcpyHostToDevice, s));
        pv[k] = EmPair{base + off, pairs[k].N, base + off + nx, pairs[k].M};
        off += nx + ny;
    }
    float* drt = ctx->rt.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(drt, Rt0, sizeof(float) * 12 * nregs, hipMemcpyHostToDevice, s));
    rc = em_prepare(*ctx, ctx->em, pv.data(), count, n_starts, tile_cols(*ctx), nullptr, nullptr, s);
    if (rc != EMSA_OK) return rc;
    rc = em_enqueue(*ctx, ctx->em, drt, drt + 12 * nregs, ctx->resid.as<double>(), *p, sig, s);
    if (rc != EMSA_OK) return rc;
    std::vector<float> out(12 * nregs);
    std::vector<double> res(nregs);
    EMSA_HIP_TRY(hipMemcpyAsync(out.data(), drt + 12 * nregs, sizeof(float) * out.size(),
                                hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipMemcpyAsync(res.data(), ctx->resid.ptr, sizeof(double) * nregs,
                                hipMemcpyDeviceToHost, s));
    EMSA_HIP_TRY(hipStreamSynchronize(s));
    for (float v : out)
        if (!std::isfinite(v)) return set_error("emicp_batched: non-finite result"), EMSA_ENONFINITE;
    std::memcpy(Rt_out, out.data(), sizeof(float) * out.size());
    if (residual) std::memcpy(residual, res.data(), sizeof(double) * nregs);
    return EMSA_OK;
}

int emsa_autotune_emicp(const emsa_pair* pairs, int count, int n_starts, const emsa_params* p,
                        void* stream, int* choice) {
    clear_error();
    int rc;
    if ((rc = check_pairs(pairs, count)) != EMSA_OK) return rc;
    if (!p || n_starts <= 0) return set_error("autotune: bad arguments"), EMSA_EINVAL;
    if ((rc = check_schedule(*p, "autotune")) != EMSA_OK) return rc;
    std::vector<float> sig = sigma_schedule(*p);
    if (sig.empty()) return set_error("autotune: the schedule runs no sigma^2 step"), EMSA_EINVAL;
    // 6 steps spread over the schedule: the large-sigma steps visit nearly
    // every column, the small-sigma ones skip most (different best tiles)
    if (sig.size() > 6) {
        std::vector<float> sub;
        for (int i = 0; i < 6; ++i) sub.push_back(sig[(size_t)i * (sig.size() - 1) / 5]);
        sig.swap(sub);
    }
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    const bool single = (count == 1 && n_starts == 1);
    std::vector<int> cands = single ? std::vector<int>{2, 4, 8, 16, 32}
                                    : std::vector<int>{512, 1024, 2048, 4096, 8192};
    if (!single) {  // whole targets: one chunk each, rows finalized in the affinity kernel
        int wmax = 0;
        for (int k = 0; k < count; ++k)
            wmax = std::max(wmax, (pairs[k].N + kEmWaves * kColGroup - 1) /
                                      (kEmWaves * kColGroup) * (kEmWaves * kColGroup));
        if (wmax > cands.back() && wmax <= 65536) cands.push_back(wmax);
    }
    const size_t nregs = (size_t)count * n_starts;
    if ((rc = ctx->rt.reserve(sizeof(float) * 24 * nregs)) != EMSA_OK) return rc;
    std::vector<float> rt0(12 * nregs, 0.f);
    for (size_t i = 0; i < nregs; ++i) rt0[12 * i] = rt0[12 * i + 4] = rt0[12 * i + 8] = 1.f;
    float* drt = ctx->rt.as<float>();
    EMSA_HIP_TRY(hipMemcpyAsync(drt, rt0.data(), sizeof(float) * rt0.size(),
                                hipMemcpyHostToDevice, s));
    std::vector<EmPair> pv(count);
    for (int k = 0; k < count; ++k) pv[k] = EmPair{pairs[k].X, pairs[k].N, pairs[k].Y, pairs[k].M};
    hipEvent_t e0, e1;
    EMSA_HIP_TRY(hipEventCreate(&e0));
    EMSA_HIP_TRY(hipEventCreate(&e1));
    const int saveB = ctx->emBlocksPerCU;
    int best = cands[0];
    float bestMs = 1e30f;
    for (int c : cands) {
        if (single) ctx->emBlocksPerCU = c;
        rc = em_prepare(*ctx, ctx->em, pv.data(), count, n_starts, single ? 0 : c, nullptr,
                        nullptr, s);
        if (rc != EMSA_OK) break;
        if ((rc = em_enqueue(*ctx, ctx->em, drt, drt + 12 * nregs, nullptr, *p, sig, s)) !=
            EMSA_OK)
            break;  // warm-up
        float ms = 1e30f;
        for (int rep = 0; rep < 2 && rc == EMSA_OK; ++rep) {
            EMSA_HIP_TRY(hipEventRecord(e0, s));
            rc = em_enqueue(*ctx, ctx->em, drt, drt + 12 * nregs, nullptr, *p, sig, s);
            EMSA_HIP_TRY(hipEventRecord(e1, s));
            EMSA_HIP_TRY(hipEventSynchronize(e1));
            float m = 0.f;
            EMSA_HIP_TRY(hipEventElapsedTime(&m, e0, e1));
            ms = std::min(ms, m);
        }
        if (ms < bestMs) {
            bestMs = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (single) ctx->emBlocksPerCU = saveB;
    if (rc != EMSA_OK) return rc;
    if (single) ctx->emBlocksPerCU = best;
    else ctx->emTileCols = best;
    if (choice) *choice = best;
    return EMSA_OK;
}

int emsa_set_tiling(int blocks_per_cu, int tile_cols) {
    clear_error();
    if (blocks_per_cu < 0 || blocks_per_cu > 64 || tile_cols < 0 || tile_cols > 65536 ||
        (tile_cols % (kEmWaves * kColGroup)) != 0)
        return set_error("set_tiling: blocks_per_cu in [0, 64], tile_cols a multiple of 64"),
               EMSA_EINVAL;
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->emBlocksPerCU = blocks_per_cu;
    ctx->emTileCols = tile_cols;
    return EMSA_OK;
}

int emsa_graph_cache_clear(void) {
    clear_error();
    DeviceCtx* ctx = current_ctx();
    if (!ctx) return EMSA_ENODEV;
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto& e : ctx->graphs) drop_entry(*e);  // each waits for its own last replay
    ctx->graphs.clear();
    ctx->seen.clear();
    return EMSA_OK;
}

}  // extern "C"
