/*
 * saras_restart.h -- C ABI of the restart file of SARAS (libsaras_io.so, host code over HDF5).
 *
 * The file is the reference's output/restartFile.h5: a scalar dataset "Time" and one dataset per
 * field, named by the field (Vx, Vy, Vz, P and, for scalar runs, T; hydro_d3.cc:140-143,
 * scalar_d3.cc:128-132), holding the field's core (no pads) as a 3D array of doubles in C order,
 * x slowest (writer.cc:455-519).  Arrays at this boundary are host arrays in the reference's padded
 * layout, (nx+2)(ny+2)(nz+2) with k fastest and index -1 first; nx, ny, nz are the core sizes.
 * Status codes are those of saras_mg.h (SR_ERR_IO for a file that cannot be created, opened, read or
 * written); the message of the last failure is sr_restart_last_error().
 */
#ifndef SARAS_RESTART_H
#define SARAS_RESTART_H

#ifdef __cplusplus
extern "C" {
#endif

/* writer::writeRestart (writer.cc:455-519): creates or truncates `path`, writes Time and the core
   of each field.  dims holds nx, ny, nz per field. */
int sr_restart_write(const char *path, double time, int nfields, const char *const *names, const int *dims,
                     const double *const *data);

/* reader::readData (reader.cc:213-296) with restartCheck (reader.cc:338-400): opens `path`, checks
   that dataset P, when it is among `names`, is 3D with the grid's node counts, reads Time and the
   core of each field into data (pads left as they are).  A missing file, a missing dataset or a
   dataset of other dimensions is an error (the reference prints and exits). */
int sr_restart_read(const char *path, int nfields, const char *const *names, const int *dims, double *const *data,
                    double *time);

const char *sr_restart_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
