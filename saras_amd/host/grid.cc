#include "grid.h"

#include <cstdio>
#include <cstdlib>

#include "saras_mg.h"

static void axis(int n, real len, int type, real beta, Array1 &x, Array1 &xi_x, Array1 &xixx, Array1 &xix2) {
    x = Array1(n); xi_x = Array1(n); xixx = Array1(n); xix2 = Array1(n);
    if (sr_grid_axis(n, len, type, beta, 0, x.data(), xi_x.data(), xixx.data(), xix2.data()) != SR_OK) {
        std::printf("Invalid grid parameters. ABORTING\n");
        std::exit(EXIT_FAILURE);
    }
}

// grid::grid (lib/grid/grid.cc:59-136) for one sub-domain: sizes 2^ind + 1, spacing 1/(N-1)
grid::grid(const parser &solParam) : inputParams(solParam) {
    sizeIndex[0] = solParam.xInd;
    sizeIndex[1] = solParam.yInd;
    sizeIndex[2] = solParam.zInd;
    for (int d = 0; d < 3; d++) globalSize[d] = (1 << sizeIndex[d]) + 1;
    xLen = solParam.Lx; yLen = solParam.Ly; zLen = solParam.Lz;
    dXi = 1.0 / real(globalSize[0] - 1);
    dEt = 1.0 / real(globalSize[1] - 1);
    dZt = 1.0 / real(globalSize[2] - 1);
    axis(globalSize[0], xLen, solParam.xGrid, solParam.betaX, xStaggr, xi_xStaggr, xixxStaggr, xix2Staggr);
    axis(globalSize[1], yLen, solParam.yGrid, solParam.betaY, yStaggr, et_yStaggr, etyyStaggr, ety2Staggr);
    axis(globalSize[2], zLen, solParam.zGrid, solParam.betaZ, zStaggr, zt_zStaggr, ztzzStaggr, ztz2Staggr);
}
