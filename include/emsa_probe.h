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

#ifdef __cplusplus
}
#endif

#endif /* EMSA_PROBE_H_ */
