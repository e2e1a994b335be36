/*
 * ebsd_probe.h -- measurement-only C ABI (libebsd_ldsprobe.so): the LDS read ceiling and the
 * vector-instruction issue ceiling of the current GPU, which the Radon gather's and the peak
 * kernel's rooflines are quoted against.  No counterpart in the
 * reference; nothing on the band-detection path calls it.
 */
#ifndef EBSD_PROBE_H
#define EBSD_PROBE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Conflict-free LDS reads of width_bytes (4, 8 or 16) by `blocks` work-groups of 1,024
 * threads, 16 reads per lane per step for `iters` steps, timed with HIP events on `stream`
 * (a hipStream_t) after one untimed launch.  *bytes_per_s = bytes read / elapsed time.
 * Returns 0, -1 on bad arguments, -3 on a HIP error. */
int ebsd_probe_lds_read(int width_bytes, int blocks, int iters, double *bytes_per_s, void *stream);

/* Vector-instruction issue ceiling: `blocks` work-groups of 256 threads each run 16
 * independent chains (kind 0: one v_fma_f32 each; 1: a v_xor_b32 and a v_add_u32 each; 2: a
 * v_max_i32 with a DPP lane shift and a v_sub_u32 each; 3: one v_fma_f64 each) for `iters`
 * steps, timed with HIP events on `stream` after one untimed launch.
 * *instr_per_s = wave-instructions issued / elapsed time.  Returns 0, -1 on bad arguments,
 * -3 on a HIP error. */
int ebsd_probe_valu_issue(int kind, int blocks, int iters, double *instr_per_s, void *stream);

#ifdef __cplusplus
}
#endif
#endif
