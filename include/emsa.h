/*
 * emsa -- MI355X-native soft-correspondence rigid registration (EM-ICP and
 * Softassign), C ABI.
 *
 * This is the thin boundary under the C++ drop-in header 3dregistration.h.
 * Plain pointers and sizes only. Point clouds are SoA float arrays
 * [x0..x(n-1), y0.., z0..], exactly the layout cloud2data builds
 * (src/cloud2data.cpp:29-47). R is 3x3 row-major (r11 r12 r13 r21 ...),
 * t is (tx, ty, tz); both are in/out: on entry the initial estimate, on exit
 * the result (src/3dregistration.h:82-99, README.md:206-214). Every algorithm
 * solves min ||X - (R Y + t)|| with X the target (cloud_target) and Y the
 * source (cloud_source).
 *
 * Every entry point returns EMSA_OK (0) or a negative code. On failure R and
 * t are left unchanged and emsa_last_error() describes the failure. Calls
 * are serialised per device; the library never tears down the caller's HIP
 * context (unlike cudaThreadExit in src/emicp.cu:713).
 */
#ifndef EMSA_H_
#define EMSA_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EMSA_OK = 0,
    EMSA_EINVAL = -1,    /* bad argument (sizes, NULL pointers, I1 < 1, ...) */
    EMSA_ENOMEM = -2,    /* device allocation failed */
    EMSA_EHIP = -3,      /* HIP runtime error */
    EMSA_ENONFINITE = -4,/* non-finite input or result */
    EMSA_ENODEV = -5     /* no HIP device */
};

enum {
    EMSA_SOFTASSIGN_MATERIALISED = 0, /* FP32 (N x M) kernel matrix in HBM */
    EMSA_SOFTASSIGN_LOGDOMAIN = 1     /* potentials, affinities recomputed per pass */
};

/* Numeric fields of registrationParameters (src/3dregistration.h:39-65), same
 * order and meaning. Viewer members, argc/argv are host-side only. */
typedef struct emsa_params {
    int JMAX;         /* annealing temperatures (softassign), default 100 */
    int I0;           /* inner iterations per temperature, default 5 */
    int I1;           /* Sinkhorn sweeps per inner iteration, default 3 */
    float alpha;      /* outlier parameter, default 3 */
    float T_0;        /* initial temperature, default 100 */
    float TFACTOR;    /* temperature factor, default 0.95 */
    float moutlier;   /* slack row/column value, default expf(-1)/sqrtf(T_0) */
    float sigma_p2;   /* EM-ICP initial sigma^2, default 0.01 */
    float sigma_inf;  /* EM-ICP final sigma^2, default 1e-5 */
    float sigma_factor; /* EM-ICP sigma^2 factor, default 0.9 */
    float d_02;       /* EM-ICP outlier distance^2, default 0.01 */
    int maxIteration; /* ICP iterations (emsa_icp), default 30 */
    int noviewer;
    int nostop;
    int notimer;
} emsa_params;

/* Fills the defaults of src/main.cpp:280-315. */
void emsa_default_params(emsa_params *p);

/* Message of the last failure on the calling thread ("" if none). */
const char *emsa_last_error(void);

/* Number of sigma^2 steps the EM-ICP loop runs for these parameters, computed
 * in float exactly as the reference loop (src/emicp.cu:443, :668). */
int emsa_emicp_iterations(const emsa_params *p);

/* Per-iteration hook: the counterpart of the reference's viewer update inside
 * its loops (src/emicp.cu:652-664, src/softassign.cu:805-815,
 * src/icp.cpp:272-282), which transforms the source cloud with the current
 * (R, t) and redraws it. Once set on a thread, emsa_emicp, emsa_softassign,
 * emsa_icp and emsa_emicp_cpu called from that thread call fn once per
 * iteration (sigma^2 step, inner iteration, ICP iteration), in order, with
 * that iteration's R (9, row-major) and t (3). The GPU loops have no host
 * round trip per iteration, so the calls come after the registration has
 * finished and before the entry point returns; none come when it fails.
 * fn = NULL removes the hook. algorithm: EMSA_ALG_*. */
enum { EMSA_ALG_EMICP = 1, EMSA_ALG_SOFTASSIGN = 2, EMSA_ALG_ICP = 3, EMSA_ALG_EMICP_CPU = 4 };
typedef void (*emsa_iteration_fn)(int algorithm, int iteration, const float *R, const float *t,
                                  void *user);
void emsa_set_iteration_callback(emsa_iteration_fn fn, void *user);

/* ---- Host-buffer entry points (the drop-in boundary) -------------------- */

/* EM-ICP on the GPU. Replaces emicp() (src/emicp.cu:247-716; declared at
 * src/3dregistration.h:87-90). X: N target points, Y: M source points. */
int emsa_emicp(const float *X, int N, const float *Y, int M, float *R, float *t,
               const emsa_params *p);

/* Same as emsa_emicp, also writing one record of EMSA_TRACE_W doubles per
 * sigma step into trace (capacity trace_cap records): sigma_p2, Lambda,
 * Xc[3], Yc[3], S[9] (S[i+3j], i target side, j source side), R[9], t[3],
 * residual. Returns the number of steps written, or a negative code. */
#define EMSA_TRACE_W 32
int emsa_emicp_traced(const float *X, int N, const float *Y, int M, float *R, float *t,
                      const emsa_params *p, double *trace, int trace_cap);

/* One sigma^2 step from (R, t): the per-source-row normalisers
 * C_r = sum_c A_rc + exp(-d_02/sigma_p2) and lambda_r = sum_c A'_rc
 * (src/emicp.cu:477-519), the trace record of the step, and R, t updated.
 * C and lambda are double arrays of M entries (either may be NULL). */
int emsa_emicp_step(const float *X, int N, const float *Y, int M, float *R, float *t,
                    float sigma_p2, float d_02, double *C, double *lambda, double *record);

/* EM-ICP on the host CPU in FP32 (no GPU needed). Replaces emicp_cpu()
 * (src/emicp_cpu.cpp:186-547; declared at src/3dregistration.h:91-95): the
 * same iteration and FP32 operation order, libm expf/sqrtf, source rows split
 * over host threads (EMSA_CPU_THREADS or OMP_NUM_THREADS, else all cores;
 * results do not depend on the count), Horn's step in FP64. trace (nullable,
 * trace_cap records): one EMSA_TRACE_W record per sigma^2 step as in
 * emsa_emicp_traced. Returns the number of sigma^2 steps run (records written
 * when trace is given) or a negative code. */
int emsa_emicp_cpu(const float *X, int N, const float *Y, int M, float *R, float *t,
                   const emsa_params *p, double *trace, int trace_cap);

/* Softassign on the GPU. Replaces softassign() (src/softassign.cu:256-886;
 * declared at src/3dregistration.h:82-86). variant: EMSA_SOFTASSIGN_*. */
int emsa_softassign(const float *X, int N, const float *Y, int M, float *R, float *t,
                    const emsa_params *p, int variant);

/* As emsa_softassign with one EMSA_TRACE_W record per inner iteration:
 * T, sumM, Xc[3], Yc[3], S[9], R[9], t[3]. */
int emsa_softassign_traced(const float *X, int N, const float *Y, int M, float *R, float *t,
                           const emsa_params *p, int variant, double *trace, int trace_cap);

/* One inner iteration of Softassign (src/softassign.cu:494-818) from (R, t)
 * at temperature T: the affinities (:106-131), I1 Sinkhorn sweeps with the
 * slack vectors o_row [N] and o_col [M] (in/out: they persist across the
 * reference's whole run, :385-390, and are divided by every sweep's sums),
 * the post-Sinkhorn sums, the weighted centroids, S and Horn's step; R, t
 * updated. Per-sweep views for parity checks (each nullable):
 * a [I1][N]: row sums a_r = sum_c M_rc + o_row[r] before the row division
 * (:538-552); b [I1][M]: column sums b_c = sum_r M_rc + o_col[c] before the
 * column division (:579-590); s [N]: post-Sinkhorn row sums s_r (:626-639);
 * record: T, Sigma = sum s_r, Xc[3], Yc[3], S[9], R[9], t[3] as in
 * emsa_softassign_traced. variant: EMSA_SOFTASSIGN_*. */
int emsa_softassign_step(const float *X, int N, const float *Y, int M, float *R, float *t,
                         float T, float alpha, int I1, float *o_row, float *o_col, int variant,
                         double *a, double *b, double *s, double *record);

/* Horn's closed-form Procrustes step, computed on the device in FP64.
 * Replaces findRTfromS (src/findRTfromS.cpp:33-97; declared at
 * src/3dregistration.h:104-107). S is h_S of the reference: S[3a+b] =
 * sum (source)_a (target)_b. */
int emsa_find_rt_from_s(const float *Xc, const float *Yc, const float *S, float *R, float *t);

/* ICP with hard correspondences on the GPU. Replaces icp() (src/icp.cpp:122-289;
 * declared at src/3dregistration.h:96-99): p->maxIteration iterations of
 * nearest source-to-target correspondences, centroids, S and Horn's step.
 * The reference's FLANN kd-tree search (src/icp.cpp:175, :218-222) is
 * approximate; here every target point is compared (exact nearest neighbour
 * of the FP32 distance with the transformed source point compensated for its
 * FP32 rounding; equal distances resolve to the lowest target index).
 * R, t in/out as in emsa_emicp; maxIteration <= 0 leaves them unchanged. */
int emsa_icp(const float *X, int N, const float *Y, int M, float *R, float *t,
             const emsa_params *p);

/* As emsa_icp with one EMSA_TRACE_W record per iteration: iteration, mean
 * squared correspondence distance, Xc(3), Yc(3), h_S(9), R(9), t(3), -, -,
 * status. Returns the number of records written or a negative code. */
int emsa_icp_traced(const float *X, int N, const float *Y, int M, float *R, float *t,
                    const emsa_params *p, double *trace, int trace_cap);

/* One correspondence query of emsa_icp (no counterpart function in the
 * reference; it is the knnSearch of src/icp.cpp:218-222 with k = 1): for
 * every source point y_r, idx[r] receives the target point nearest to
 * R y_r + t and dist2[r] its squared distance (FP32), as emsa_icp finds them.
 * R, t are read only. */
int emsa_icp_nearest(const float *X, int N, const float *Y, int M, const float *R,
                     const float *t, int *idx, float *dist2);

/* ---- Device-buffer entry points (inputs resident in HBM) ---------------- */

/* EM-ICP with X, Y in device memory (SoA, unpadded) and Rt a device array of
 * 12 floats (R row-major, then t), in/out. stream: a hipStream_t or NULL for
 * the null stream. Asynchronous with respect to the host; errors detected at
 * launch are returned. */
int emsa_emicp_dev(const float *dX, int N, const float *dY, int M, float *dRt,
                   const emsa_params *p, void *stream);

/* EM-ICP on device buffers that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, kernels and results as emsa_emicp_dev, launched directly on
 * `stream` (never recorded into the library's own graphs). The workspace is a
 * fixed 16 MiB device array per GPU, held until the process ends; clouds
 * whose workspace does not fit (about N, M up to 65,536 points at the default
 * decomposition) are rejected with EMSA_EINVAL. Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU race. */
int emsa_emicp_dev_noalloc(const float *dX, int N, const float *dY, int M, float *dRt,
                           const emsa_params *p, void *stream);

/* ---- Batched EM-ICP (no counterpart in the reference) -------------------
 * count source/target pairs of uneven sizes, each registered from n_starts
 * initial transforms, all in one kernel launch per sigma^2 step (the same
 * schedule for every registration). Rt0: [count][n_starts][12] initial
 * (R row-major, t); Rt_out: [count][n_starts][12]; residual (nullable):
 * [count][n_starts], sum_r lambda_r |X'_r - (R y_r + t)|^2 / Lambda of the
 * last sigma^2 step (with the R, t that step started from). */
typedef struct emsa_pair {
    const float *X; /* target, SoA, N points */
    int N;
    const float *Y; /* source, SoA, M points */
    int M;
} emsa_pair;

/* Host buffers (pairs point to host SoA arrays). */
int emsa_emicp_batched(const emsa_pair *pairs, int count, const float *Rt0, int n_starts,
                       const emsa_params *p, float *Rt_out, double *residual);

/* Device buffers (pairs point to device SoA arrays; dRt0, dRt_out, dResidual
 * are device arrays). The first call with given buffers, sizes and parameters
 * launches its kernels directly; when the same call comes again its launch
 * sequence is recorded in a hipGraph and replayed on `stream` from then on.
 * Asynchronous; errors detected at launch are returned. */
int emsa_emicp_batched_dev(const emsa_pair *pairs, int count, const float *dRt0, int n_starts,
                           const emsa_params *p, float *dRt_out, double *dResidual,
                           void *stream);

/* Autotuner: times the candidate work decompositions on these device buffers
 * (6 sigma^2 steps spread over the schedule of p) and keeps the fastest for
 * later calls on the
 * current device. count == 1 && n_starts == 1 tunes the single-registration
 * plan (workgroups per CU, from 2..32), otherwise the batched column tile width
 * (512..8192). The
 * choice is written to *choice (nullable). Results of different choices
 * differ only by FP32 summation order. */
int emsa_autotune_emicp(const emsa_pair *pairs, int count, int n_starts, const emsa_params *p,
                        void *stream, int *choice);

/* Manual override of the decomposition the autotuner chooses, for the
 * current device: workgroups per CU of the single-registration plan and the
 * batched column tile width (0 keeps the built-in heuristic). */
int emsa_set_tiling(int blocks_per_cu, int tile_cols);

/* Drops every captured launch sequence of the current device and its
 * workspace. */
int emsa_graph_cache_clear(void);

/* ---- Differentiable unrolled EM-ICP (no counterpart in the reference) ----
 * For callers that train through the registration: `steps` iterations of the
 * EM-ICP step of src/emicp.cu:443-668 with sigma_k^2 = sigma2_0 * factor^k
 * (computed in double, rounded to float), outlier term exp(-d_02 / sigma_k^2).
 * All buffers are device buffers; every pair is one registration from one
 * initial transform. Asynchronous on `stream`. */

/* Number of doubles of the checkpoint buffer the forward fills: per pair and
 * step one EMSA_TRACE_W record and (C_r, lambda_r, lambda_r X'_r) per source
 * row. No affinity is stored; the backward recomputes them. */
long long emsa_emicp_unrolled_ckpt_size(const emsa_pair *pairs, int count, int steps);

/* Forward. dRt0, dRt_out: [count][12] (R row-major, t); dckpt: the size above. */
int emsa_emicp_unrolled_fwd(const emsa_pair *pairs, int count, const float *dRt0,
                            float sigma2_0, float factor, float d_02, int steps, float *dRt_out,
                            double *dckpt, void *stream);

/* Gradient buffers of one pair: dL/dX [3][N] and dL/dY [3][M], SoA, overwritten. */
typedef struct emsa_pair_grad {
    float *gX;
    float *gY;
} emsa_pair_grad;

/* Backward (the hand-written adjoint of the forward above, same arguments).
 * dgRt: [count][12] dL/d(R, t) of the output. Writes grads[k], dgRt0 (nullable,
 * [count][12] dL/d(R_0, t_0)) and dgsigma2 (nullable, [count] dL/dsigma2_0,
 * doubles). A step whose Horn solve failed in the forward passes the gradient
 * of its output through to its input. */
int emsa_emicp_unrolled_bwd(const emsa_pair *pairs, int count, const float *dRt0,
                            float sigma2_0, float factor, float d_02, int steps,
                            const double *dckpt, const float *dgRt, const emsa_pair_grad *grads,
                            float *dgRt0, double *dgsigma2, void *stream);

/* ---- Differentiable unrolled Softassign (no counterpart in the reference) --
 * For callers that train through the registration: JMAX temperatures
 * (T <- T * TFACTOR in float, as src/softassign.cu:820) of I0 inner iterations
 * (src/softassign.cu:494-818) with I1 (1..4) Sinkhorn sweeps each, for B
 * registrations of equal sizes at once. dX: [B][3][N], dY: [B][3][M] (SoA per
 * registration), dRt0, dRt_out: [B][12]; moutlier < 0 selects the default
 * exp(-1) / sqrt(T_0) (and its T_0 gradient). All buffers are device buffers;
 * asynchronous on `stream`. */

/* Bytes of the checkpoint buffer the forward fills (per registration and
 * inner iteration: a 40-double record, per sweep the row/column sums and
 * masses, the slack vectors at the iteration start, s_r and D'_r). */
long long emsa_softassign_unrolled_ckpt_bytes(int B, int N, int M, int JMAX, int I0, int I1);

int emsa_softassign_unrolled_fwd(const float *dX, const float *dY, int B, int N, int M,
                                 const float *dRt0, float T_0, float TFACTOR, float alpha,
                                 float moutlier, int JMAX, int I0, int I1, float *dRt_out,
                                 void *dckpt, void *stream);

/* Backward (hand-written adjoint, same arguments as the forward). dgRt:
 * [B][12] dL/d(R, t) of the output. Writes dgX [B][3][N], dgY [B][3][M],
 * dgRt0 (nullable, [B][12]) and dgT0 (nullable, [B] doubles, dL/dT_0). */
int emsa_softassign_unrolled_bwd(const float *dX, const float *dY, int B, int N, int M,
                                 const float *dRt0, float T_0, float TFACTOR, float alpha,
                                 float moutlier, int JMAX, int I0, int I1, const void *dckpt,
                                 const float *dgRt, float *dgX, float *dgY, float *dgRt0,
                                 double *dgT0, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EMSA_H_ */
