// Node ("staggered") grid of lib/grid/grid.h restricted to what the pressure solve reads, for a
// single sub-domain (one GPU owns the whole domain).
#pragma once

#include "parser.h"
#include "sr_array.h"

class grid {
  public:
    explicit grid(const parser &solParam);

    const parser &inputParams;
    int sizeIndex[3];
    int globalSize[3];
    real dXi, dEt, dZt;
    real xLen, yLen, zLen;

    Array1 xStaggr, yStaggr, zStaggr;
    Array1 xi_xStaggr, xixxStaggr, xix2Staggr;
    Array1 et_yStaggr, etyyStaggr, ety2Staggr;
    Array1 zt_zStaggr, ztzzStaggr, ztz2Staggr;
};
