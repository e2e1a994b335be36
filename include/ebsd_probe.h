/*
 * ebsd_probe.h -- measurement-only C ABI (libebsd_ldsprobe.so): the LDS read ceiling of the
 * current GPU, which the Radon gather's roofline is quoted against.  No counterpart in the
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

#ifdef __cplusplus
}
#endif
#endif
