/*
 * saras_probes.h -- C ABI of the field probes of SARAS (lib/io/probes.cc) on device fields
 * (gfx950 HIP kernels).
 *
 * A probe is a global index triple (i, j, k); its record holds, for every probed field, the value
 * of that field's array at (i, j, k) (probes::getData, probes.cc:145-157).  Fields are device
 * arrays in the reference's padded layout (k fastest, logical index -1..n along each direction,
 * n the field's core size); the values are gathered on the device in one launch and copied to the
 * host in one transfer.  The fields must be complete when the call is made: every entry point that
 * writes them has returned (the time-step calls return with their work finished; operators of
 * saras_field.h after sr_fgrid_sync).  Status codes are those of saras_mg.h.
 */
#ifndef SARAS_PROBES_H
#define SARAS_PROBES_H

#ifdef __cplusplus
extern "C" {
#endif

#define SR_PROBE_MAX_FIELDS 10   /* dataStruct carries up to 10 field values (probes.h:64-72) */

/* probes::getData (probes.cc:145-157): out[p * nfields + f] = fields[f] at idx[3p..3p+2].  dims holds
   the core sizes nx, ny, nz of each field; every probe index must lie in each field's padded range
   -1..n (SR_ERR_ARG otherwise, where the reference reads past the array).  The entry point moves
   the probe list and the values through device buffers it allocates for the call; it returns
   with `out` (host memory) filled. */
int sr_probe_gather(int nfields, const double *const *fields, const int *dims, int nprobes, const int *idx,
                    double *out);

/* No counterpart in the reference.  sr_probe_gather for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
   gather and results, with the probe list and the values staged in a static device buffer of
   SR_PROBE_NOALLOC_MAX probes (368 KiB, held by every GPU the library runs on until the process
   ends) instead of buffers allocated per call.  More probes than that is SR_ERR_ARG.  Every call
   on a GPU shares that GPU's buffer: calls that overlap on one GPU (several host threads) race. */
#define SR_PROBE_NOALLOC_MAX 4096
int sr_probe_gather_noalloc(int nfields, const double *const *fields, const int *dims, int nprobes, const int *idx,
                            double *out);

#ifdef __cplusplus
}
#endif

#endif
