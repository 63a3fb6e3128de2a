// grid of lib/grid/grid.h for a single sub-domain (one GPU owns the whole domain): node
// ("staggered") and face ("collocated") coordinates and metrics, and the device handle of the
// field operators (include/saras_field.h) built from them.
#pragma once

#include "parser.h"
#include "saras_field.h"
#include "sr_array.h"

class grid {
  public:
    explicit grid(const parser &solParam);
    ~grid();
    grid(const grid &) = delete;
    grid &operator=(const grid &) = delete;

    const parser &inputParams;
    int sizeIndex[3];
    int globalSize[3];
    real dXi, dEt, dZt;
    real xLen, yLen, zLen;

    Array1 xStaggr, yStaggr, zStaggr;
    Array1 xi_xStaggr, xixxStaggr, xix2Staggr;
    Array1 et_yStaggr, etyyStaggr, ety2Staggr;
    Array1 zt_zStaggr, ztzzStaggr, ztz2Staggr;

    Array1 xColloc, yColloc, zColloc;
    Array1 xi_xColloc, xixxColloc, xix2Colloc;
    Array1 et_yColloc, etyyColloc, ety2Colloc;
    Array1 zt_zColloc, ztzzColloc, ztz2Colloc;

    // field-operator handle on this grid (created on first use)
    sr_fgrid *fieldOps() const;

  private:
    mutable sr_fgrid *fops = nullptr;
};
