// Restart file of SARAS over serial HDF5 (include/saras_restart.h): writer::writeRestart
// (lib/io/writer.cc:455-519) and reader::readData with restartCheck (lib/io/reader.cc:213-400) for
// one process that owns the whole domain.  The reference writes through MPI-IO hyperslabs of each
// rank's core; with one rank the file dataspace and the core coincide, so each field's core is
// gathered into a contiguous buffer and written whole.
#include <hdf5.h>

#include <string>
#include <vector>

#include "../../include/saras_mg.h"
#include "../../include/saras_restart.h"

namespace {

constexpr int kOk = SR_OK, kErrArg = SR_ERR_ARG, kErrIo = SR_ERR_IO;

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// core of a padded field <-> contiguous buffer
void pack(const double *padded, int nx, int ny, int nz, std::vector<double> &core) {
    core.resize((size_t)nx * ny * nz);
    size_t q = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const double *row = padded + ((size_t)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + 1;
            for (int k = 0; k < nz; k++) core[q++] = row[k];
        }
}

void unpack(const std::vector<double> &core, int nx, int ny, int nz, double *padded) {
    size_t q = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double *row = padded + ((size_t)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + 1;
            for (int k = 0; k < nz; k++) row[k] = core[q++];
        }
}

// closes an HDF5 identifier on scope exit
struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
};

int check_args(int nfields, const char *const *names, const int *dims, const void *data) {
    if (nfields < 1 || !names || !dims || !data) return fail(kErrArg, "bad arguments");
    for (int f = 0; f < nfields; f++) {
        if (!names[f] || !names[f][0]) return fail(kErrArg, "a field has no name");
        if (dims[3 * f] < 1 || dims[3 * f + 1] < 1 || dims[3 * f + 2] < 1) return fail(kErrArg, "bad field sizes");
    }
    return kOk;
}

}  // namespace

extern "C" const char *sr_restart_last_error(void) { return g_err.c_str(); }

extern "C" int sr_restart_write(const char *path, double time, int nfields, const char *const *names, const int *dims,
                                const double *const *data) {
    if (!path) return fail(kErrArg, "null path");
    int r = check_args(nfields, names, dims, data);
    if (r != kOk) return r;
    for (int f = 0; f < nfields; f++)
        if (!data[f]) return fail(kErrArg, "null field array");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(kErrIo, std::string("cannot create ") + path);
    {
        Hid space(H5Screate(H5S_SCALAR), H5Sclose);
        Hid ds(H5Dcreate2(file.id, "Time", H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               H5Dclose);
        if (ds.id < 0 || H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, space.id, space.id, H5P_DEFAULT, &time) < 0)
            return fail(kErrIo, "Error in writing output to HDF file");
    }
    std::vector<double> core;
    for (int f = 0; f < nfields; f++) {
        int nx = dims[3 * f], ny = dims[3 * f + 1], nz = dims[3 * f + 2];
        hsize_t d[3] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
        pack(data[f], nx, ny, nz, core);
        Hid space(H5Screate_simple(3, d, nullptr), H5Sclose);
        Hid ds(H5Dcreate2(file.id, names[f], H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               H5Dclose);
        if (ds.id < 0 || H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, core.data()) < 0)
            return fail(kErrIo, std::string("Error in writing output to HDF file (") + names[f] + ")");
    }
    if (H5Fflush(file.id, H5F_SCOPE_GLOBAL) < 0) return fail(kErrIo, "Error in writing output to HDF file");
    return kOk;
}

extern "C" int sr_restart_read(const char *path, int nfields, const char *const *names, const int *dims,
                               double *const *data, double *time) {
    if (!path || !time) return fail(kErrArg, "null argument");
    int r = check_args(nfields, names, dims, data);
    if (r != kOk) return r;
    for (int f = 0; f < nfields; f++)
        if (!data[f]) return fail(kErrArg, "null field array");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(kErrIo, "ERROR: Restart flag is true, but could not open restart file");
    // restartCheck: dimensionality and sizes of P against the grid
    for (int f = 0; f < nfields; f++) {
        if (std::string(names[f]) != "P") continue;
        Hid ds(H5Dopen2(file.id, "P", H5P_DEFAULT), H5Dclose);
        if (ds.id < 0) return fail(kErrIo, "restart file has no dataset P");
        Hid space(H5Dget_space(ds.id), H5Sclose);
        if (H5Sget_simple_extent_ndims(space.id) != 3)
            return fail(kErrArg, "ERROR: Dimensionality of restart file conflicts with solver parameters");
        hsize_t d[3];
        H5Sget_simple_extent_dims(space.id, d, nullptr);
        for (int a = 0; a < 3; a++)
            if ((int)d[a] != dims[3 * f + a])
                return fail(kErrArg, "ERROR: Array limits in restart file conflicts with solver parameters");
    }
    {
        Hid ds(H5Dopen2(file.id, "Time", H5P_DEFAULT), H5Dclose);
        if (ds.id < 0) return fail(kErrIo, "restart file has no dataset Time");
        Hid space(H5Screate(H5S_SCALAR), H5Sclose);
        if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, space.id, space.id, H5P_DEFAULT, time) < 0)
            return fail(kErrIo, "Error in reading input from HDF file");
    }
    std::vector<double> core;
    for (int f = 0; f < nfields; f++) {
        int nx = dims[3 * f], ny = dims[3 * f + 1], nz = dims[3 * f + 2];
        Hid ds(H5Dopen2(file.id, names[f], H5P_DEFAULT), H5Dclose);
        if (ds.id < 0) return fail(kErrIo, std::string("restart file has no dataset ") + names[f]);
        Hid space(H5Dget_space(ds.id), H5Sclose);
        hsize_t d[3] = {0, 0, 0};
        if (H5Sget_simple_extent_ndims(space.id) != 3) return fail(kErrArg, std::string("dataset ") + names[f] + " is not 3D");
        H5Sget_simple_extent_dims(space.id, d, nullptr);
        if ((int)d[0] != nx || (int)d[1] != ny || (int)d[2] != nz)
            return fail(kErrArg, std::string("dataset ") + names[f] + " has other sizes than the field");
        core.assign((size_t)nx * ny * nz, 0.0);
        if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, core.data()) < 0)
            return fail(kErrIo, "Error in reading input from HDF file");
        unpack(core, nx, ny, nz, data[f]);
    }
    return kOk;
}
