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

#ifdef __cplusplus
}
#endif

#endif
