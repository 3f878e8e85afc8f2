/*
 * emsa measurement probe (no counterpart in the reference): device-wide issue
 * rates that bound the fused registration kernels, used to report their
 * achieved fraction of the binding compute ceiling.
 */
#ifndef EMSA_PROBE_H_
#define EMSA_PROBE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Measures, on the current device, the v_exp_f32 rate (exponentials per
 * second, all lanes of all CUs) and the FP32 FMA rate (FMA instructions per
 * lane per second, i.e. 2 flops each). Returns EMSA_OK or a negative code. */
int emsa_probe_rates(double *exp_per_s, double *fma_per_s);

/* Issue cost, in SIMD cycles per wave instruction (an FP32 FMA of a wave of
 * 64 lanes taking 4), of v_exp_f32, v_mfma_f32_4x4x1f32 and
 * v_mfma_f32_16x16x4f32 on the current device, measured in loops where each
 * instruction is mixed with independent FMAs (as in the fused kernels: the
 * matrix cores and the VALU do not run at the same time, so their costs add).
 * Returns EMSA_OK or a negative code. */
int emsa_probe_issue_cycles(double *exp_cycles, double *mfma4_cycles, double *mfma16_cycles);

/* Average duration in microseconds of one launch of the fused EM-ICP
 * affinity kernel over the batch `pairs` (device SoA pointers) x n_starts
 * identity starts, laid out exactly as emsa_emicp_batched_dev lays it out
 * (count == 1 && n_starts == 1: the single-registration plan of
 * emsa_emicp_dev), timed with HIP events over `reps` back-to-back launches on
 * the library's stream at the given sigma^2. *pairs_per_launch receives the
 * number of (target, source) pairs one launch evaluates: counted on the device
 * for the matrix-core kernel (column groups skipped because all their
 * affinities underflow to 0 are not evaluated), N * M per registration for
 * the all-VALU kernel. */
struct emsa_pair;
struct emsa_params;
int emsa_probe_emicp_affinity(const struct emsa_pair *pairs, int count, int n_starts,
                              float sigma_p2, int reps, double *us_per_launch,
                              double *pairs_per_launch);

/* Selects, on the current device, where the EM-ICP affinity kernel forms its
 * row sums B and weighted sums W: 1 (default) on the matrix cores
 * (v_mfma_f32_4x4x1f32 over the column dimension), 0 on the VALU. Both give
 * the same sums up to FP32 rounding; for A/B measurement. Drops the recorded
 * launch sequences of the device. */
int emsa_probe_set_affinity_path(int use_matrix_cores);

/* Runs the whole EM-ICP sequence (identity starts, the schedule of *params)
 * of a batch laid out as emsa_emicp_batched_dev lays it out (count == 1 &&
 * n_starts == 1: as emsa_emicp_dev), launched directly on the library's stream,
 * twice: once counting on the device the pairs every affinity launch
 * evaluates, once with HIP events around each affinity launch. Receives the
 * summed affinity time in microseconds, the summed evaluated pairs and the
 * number of affinity launches (sigma^2 steps); per_step (nullable) receives,
 * for the first per_step_cap steps, (sigma^2, microseconds, pairs, pairs of
 * waves that took the expanded exponent) each. */
int emsa_probe_emicp_sequence(const struct emsa_pair *pairs, int count, int n_starts,
                              const struct emsa_params *params, double *affinity_us,
                              double *pairs_evaluated, int *launches, double *per_step,
                              int per_step_cap);

/* The working order EM-ICP gives a cloud: dP is a device SoA cloud [3][n];
 * dPerm (device, n ints) receives the permutation (working position ->
 * caller's index): points sorted stably by their 30-bit Hilbert key over the
 * cloud's bounding box (10 bits per axis), non-finite points last. For
 * tests of the ordering. */
int emsa_probe_hilbert_order(const float *dP, int n, int *dPerm);

/* Matrix-core EM-ICP affinity of the current device: which groups of 8 target
 * columns a wave of 64 source rows skips (bounding boxes of the clouds in
 * Hilbert order). 1 (the default): the groups whose every affinity
 * A = exp(-|d|^2 / sigma^2) is below 2^-128, under FLT_MIN (the reference's
 * fast-math build flushes such an A to 0; the sums change by far less than
 * FP32 resolves); 2: only the groups whose every e = sqrt(A) underflows to
 * exactly 0 in FP32, which gives results bit-identical to 0; 0 visits every
 * column. Any other nonzero value is 1. For A/B measurement; drops the
 * recorded launch sequences of the device. */
int emsa_probe_set_affinity_cull(int mode);

/* Matrix-core EM-ICP affinity of the current device: a wave of 64 source rows
 * forms its exponents -|x - p|^2 / (2 sigma^2) on the matrix cores, expanded
 * about the centre of each block of 4 rows, when every block's rows lie
 * within a of their centre with (log2 e / (2 sigma^2)) a^2 <= max_ka2, and as
 * a difference on the VALU otherwise. 0 always takes the difference; the
 * default is 32. The two forms differ by FP32 rounding only. For A/B
 * measurement; drops the recorded launch sequences of the device. */
int emsa_probe_set_affinity_expanded(float max_ka2);

/* Selects, on the current device, where the EM-ICP row finalize runs when
 * every tile of a launch spans its registration's whole target (one column
 * chunk, e.g. emsa_set_tiling(0, tile_cols) with tile_cols >= N): 1 (default)
 * in the affinity kernel's epilogue, no row-finalize launch; 0 in its own
 * kernel after the partial sums went through HBM. Both give bit-identical
 * results. For A/B measurement; drops the recorded launch sequences. */
int emsa_probe_set_row_fuse(int enable);

/* Selects, on the current device, the order of the EM-ICP affinity
 * workgroups of launches that skip column groups (ordered clouds) and have at
 * most 4096 tiles and at least 2^24 pairs: 1 (default) from the third
 * sigma^2 step on, sorted by an earlier step's work per tile and dealt so
 * that the workgroups sharing a compute unit add up to about the same work
 * (one more workgroup of the previous launch sorts them); 0 in tile order. Which
 * workgroup computes a tile changes, what it computes does not: results are
 * bit-identical. For A/B measurement; drops the recorded launch sequences. */
int emsa_probe_set_affinity_balance(int enable);

#ifdef __cplusplus
}
#endif

#endif /* EMSA_PROBE_H_ */
