// Field classes of lib/field over the C ABI in include/saras_field.h: field, derivative, plainsf,
// plainvf, sfield and vfield with the reference's names and the methods a time step calls on them
// (derivative.h:78-84, sfield.h:73-76, vfield.h:76-80, plainsf.h:92-98).  Every array is a DArray3
// in HBM; each method has finished on the device when it returns, as the reference's host loops
// have.  Boundary conditions are given per field as an sr_bc (the walls of lib/boundary/*.cc).
#pragma once

#include <string>

#include "grid.h"
#include "saras_field.h"
#include "sr_darray.h"

class field {
  public:
    // staggering per direction as field.h:70-75: true = nodes ("staggered"), false = faces
    field(const grid &gridData, bool xStag, bool yStag, bool zStag, std::string fieldName);

    const grid &gridData;
    bool xStag, yStag, zStag;
    std::string fieldName;
    int loc;                 // SR_LOC_*
    DArray3 F;

    void syncData();                    // mpidata::syncData (mpidata.cc:278-306)
    void imposeBC(const sr_bc &bc);     // syncData, then the walls in the reference's order
};

class derivative {
  public:
    derivative(const grid &gridData, const field &F);

    void calcDerivative1_x(DArray3 &outArray);
    void calcDerivative1_y(DArray3 &outArray);
    void calcDerivative1_z(DArray3 &outArray);

    void calcDerivative2xx(DArray3 &outArray);
    void calcDerivative2yy(DArray3 &outArray);
    void calcDerivative2zz(DArray3 &outArray);

  private:
    void run(int order, int dim, DArray3 &out);
    const grid &gridData;
    const field &F;
};

class plainvf;
class vfield;

class plainsf {
  public:
    explicit plainsf(const grid &gridData);
    const grid &gridData;
    DArray3 F;
    void operator=(real a) { F = a; }
    void syncData();
    // plainsf::gradient (plainsf.h:92-98): the same expression as sfield::gradient
    void gradient(plainvf &gradF, const vfield &V);
};

class plainvf {
  public:
    explicit plainvf(const grid &gridData);
    const grid &gridData;
    DArray3 Vx, Vy, Vz;
    void operator=(real a) {
        Vx = a;
        Vy = a;
        Vz = a;
    }
    void syncData();
};

class sfield {
  public:
    sfield(const grid &gridData, std::string fieldName);

    field F;
    derivative derS;
    std::string fieldName;

    void computeDiff(plainsf &H);
    void computeNLin(const vfield &V, plainsf &H);
    void gradient(plainvf &gradF, const vfield &V);
    void syncData();
};

class vfield {
  public:
    vfield(const grid &gridData, std::string fieldName);

    field Vx, Vy, Vz;
    derivative derVx, derVy, derVz;
    std::string fieldName;

    void computeDiff(plainvf &H);
    void computeNLin(const vfield &V, plainvf &H);
    void divergence(plainsf &divV, const sfield &P);
    void syncData();
};

// hydro_d3::solveVx/Vy/Vz and scalar_d3::solveVx/Vy/Vz/solveT (hydro_d3.cc:370-533,
// scalar_d3.cc:373-594): Jacobi sweeps of (1 - dt coef lap / 2) F = R with the walls bc after each
// sweep, to the tolerance of jd; returns the sweep count (the reference aborts without convergence:
// so does this, printing the reason)
int jacobiSolve(field &F, const DArray3 &R, DArray3 &work, const sr_jacobi_desc &jd, const sr_bc &bc);
