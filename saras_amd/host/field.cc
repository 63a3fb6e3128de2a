// Field classes over include/saras_field.h.  Each method enqueues its operator on the grid's
// field-operator stream and waits for it, so host copies (DArray3::download) and the pressure
// solve (which runs on its own stream) see finished results.
#include "field.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>

static void check(int status, const char *what) {
    if (status != SR_OK) {
        std::cout << what << " failed with status " << status << ": " << sr_last_error() << ". ABORTING" << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

static void finish(const grid &g) { check(sr_fgrid_sync(g.fieldOps()), "field operator"); }

static int locOf(bool xs, bool ys, bool zs) {
    if (xs && ys && zs) return SR_LOC_NODE;
    if (!xs && ys && zs) return SR_LOC_VX;
    if (xs && !ys && zs) return SR_LOC_VY;
    if (xs && ys && !zs) return SR_LOC_VZ;
    std::cout << "Field staggering not supported (one face direction at most). ABORTING" << std::endl;
    std::exit(EXIT_FAILURE);
}

static void dimsOf(const grid &g, int loc, int &nx, int &ny, int &nz) {
    check(sr_fgrid_dims(g.fieldOps(), loc, &nx, &ny, &nz), "field dimensions");
}

static sr_bc syncOnly() {
    sr_bc b{};
    for (int w = 0; w < 6; w++) b.kind[w] = SR_BC_NONE;
    return b;
}

// ---------------------------------------------------------------------------------------------
field::field(const grid &gridData, bool xStag, bool yStag, bool zStag, std::string fieldName)
    : gridData(gridData), xStag(xStag), yStag(yStag), zStag(zStag), fieldName(std::move(fieldName)),
      loc(locOf(xStag, yStag, zStag)) {
    int nx, ny, nz;
    dimsOf(gridData, loc, nx, ny, nz);
    F.resize(nx, ny, nz);
}

void field::syncData() {
    sr_bc b = syncOnly();
    check(sr_bc_fill(gridData.fieldOps(), loc, &b, F.data()), "syncData");
    finish(gridData);
}

void field::imposeBC(const sr_bc &bc) {
    check(sr_bc_fill(gridData.fieldOps(), loc, &bc, F.data()), "imposeBC");
    finish(gridData);
}

// ---------------------------------------------------------------------------------------------
derivative::derivative(const grid &gridData, const field &F) : gridData(gridData), F(F) {}

void derivative::run(int order, int dim, DArray3 &out) {
    sr_fgrid *g = gridData.fieldOps();
    check(order == 1 ? sr_deriv1(g, F.loc, dim, F.F.data(), out.data()) : sr_deriv2(g, F.loc, dim, F.F.data(), out.data()),
          "derivative");
    finish(gridData);
}

void derivative::calcDerivative1_x(DArray3 &outArray) { run(1, 0, outArray); }
void derivative::calcDerivative1_y(DArray3 &outArray) { run(1, 1, outArray); }
void derivative::calcDerivative1_z(DArray3 &outArray) { run(1, 2, outArray); }
void derivative::calcDerivative2xx(DArray3 &outArray) { run(2, 0, outArray); }
void derivative::calcDerivative2yy(DArray3 &outArray) { run(2, 1, outArray); }
void derivative::calcDerivative2zz(DArray3 &outArray) { run(2, 2, outArray); }

// ---------------------------------------------------------------------------------------------
plainsf::plainsf(const grid &gridData) : gridData(gridData) {
    int nx, ny, nz;
    dimsOf(gridData, SR_LOC_NODE, nx, ny, nz);
    F.resize(nx, ny, nz);
}

void plainsf::syncData() {
    sr_bc b = syncOnly();
    check(sr_bc_fill(gridData.fieldOps(), SR_LOC_NODE, &b, F.data()), "syncData");
    finish(gridData);
}

void plainsf::gradient(plainvf &gradF, const vfield &) {
    check(sr_gradient(gridData.fieldOps(), F.data(), gradF.Vx.data(), gradF.Vy.data(), gradF.Vz.data()), "gradient");
    finish(gridData);
}

plainvf::plainvf(const grid &gridData) : gridData(gridData) {
    int n[3];
    dimsOf(gridData, SR_LOC_VX, n[0], n[1], n[2]);
    Vx.resize(n[0], n[1], n[2]);
    dimsOf(gridData, SR_LOC_VY, n[0], n[1], n[2]);
    Vy.resize(n[0], n[1], n[2]);
    dimsOf(gridData, SR_LOC_VZ, n[0], n[1], n[2]);
    Vz.resize(n[0], n[1], n[2]);
}

void plainvf::syncData() {
    sr_bc b = syncOnly();
    sr_fgrid *g = gridData.fieldOps();
    check(sr_bc_fill(g, SR_LOC_VX, &b, Vx.data()), "syncData");
    check(sr_bc_fill(g, SR_LOC_VY, &b, Vy.data()), "syncData");
    check(sr_bc_fill(g, SR_LOC_VZ, &b, Vz.data()), "syncData");
    finish(gridData);
}

// ---------------------------------------------------------------------------------------------
sfield::sfield(const grid &gridData, std::string fieldName)
    : F(gridData, true, true, true, fieldName), derS(gridData, F), fieldName(fieldName) {}

void sfield::computeDiff(plainsf &H) {
    check(sr_sf_diff(F.gridData.fieldOps(), F.F.data(), H.F.data()), "sfield::computeDiff");
    finish(F.gridData);
}

void sfield::computeNLin(const vfield &V, plainsf &H) {
    check(sr_sf_nlin(F.gridData.fieldOps(), V.Vx.F.data(), V.Vy.F.data(), V.Vz.F.data(), F.F.data(), H.F.data()),
          "sfield::computeNLin");
    finish(F.gridData);
}

void sfield::gradient(plainvf &gradF, const vfield &) {
    check(sr_gradient(F.gridData.fieldOps(), F.F.data(), gradF.Vx.data(), gradF.Vy.data(), gradF.Vz.data()),
          "sfield::gradient");
    finish(F.gridData);
}

void sfield::syncData() { F.syncData(); }

vfield::vfield(const grid &gridData, std::string fieldName)
    : Vx(gridData, false, true, true, fieldName + "_Vx"), Vy(gridData, true, false, true, fieldName + "_Vy"),
      Vz(gridData, true, true, false, fieldName + "_Vz"), derVx(gridData, Vx), derVy(gridData, Vy),
      derVz(gridData, Vz), fieldName(fieldName) {}

void vfield::computeDiff(plainvf &H) {
    check(sr_vf_diff(Vx.gridData.fieldOps(), Vx.F.data(), Vy.F.data(), Vz.F.data(), H.Vx.data(), H.Vy.data(),
                     H.Vz.data()),
          "vfield::computeDiff");
    finish(Vx.gridData);
}

void vfield::computeNLin(const vfield &V, plainvf &H) {
    check(sr_vf_nlin(Vx.gridData.fieldOps(), Vx.F.data(), Vy.F.data(), Vz.F.data(), V.Vx.F.data(), V.Vy.F.data(),
                     V.Vz.F.data(), H.Vx.data(), H.Vy.data(), H.Vz.data()),
          "vfield::computeNLin");
    finish(Vx.gridData);
}

void vfield::divergence(plainsf &divV, const sfield &) {
    check(sr_divergence(Vx.gridData.fieldOps(), Vx.F.data(), Vy.F.data(), Vz.F.data(), divV.F.data()),
          "vfield::divergence");
    finish(Vx.gridData);
}

void vfield::syncData() {
    Vx.syncData();
    Vy.syncData();
    Vz.syncData();
}

int jacobiSolve(field &F, const DArray3 &R, DArray3 &work, const sr_jacobi_desc &jd, const sr_bc &bc) {
    int iters = 0;
    int s = sr_jacobi_visc(F.gridData.fieldOps(), F.loc, &jd, &bc, R.data(), F.F.data(), work.data(), &iters);
    if (s == SR_ERR_NOT_CONVERGED) {
        std::cout << "Jacobi iterations for solution of " << F.fieldName << " not converging. Aborting" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    check(s, "Jacobi solve");
    return iters;
}
