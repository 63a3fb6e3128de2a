#include "grid.h"

#include <cstdio>
#include <cstdlib>

#include "saras_mg.h"

// one direction's coordinates and metrics: nodes (length n, index -1..n) or faces (faces = 1:
// length n - 1, index -1..n - 1)
static void axis(int n, real len, int type, real beta, Array1 &x, Array1 &xi_x, Array1 &xixx, Array1 &xix2,
                 int faces = 0) {
    int m = faces ? n - 1 : n;
    x = Array1(m); xi_x = Array1(m); xixx = Array1(m); xix2 = Array1(m);
    if (sr_grid_axis(n, len, type, beta, faces, x.data(), xi_x.data(), xixx.data(), xix2.data()) != SR_OK) {
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
    axis(globalSize[0], xLen, solParam.xGrid, solParam.betaX, xColloc, xi_xColloc, xixxColloc, xix2Colloc, 1);
    axis(globalSize[1], yLen, solParam.yGrid, solParam.betaY, yColloc, et_yColloc, etyyColloc, ety2Colloc, 1);
    axis(globalSize[2], zLen, solParam.zGrid, solParam.betaZ, zColloc, zt_zColloc, ztzzColloc, ztz2Colloc, 1);
}

grid::~grid() {
    if (fops) sr_fgrid_destroy(fops);
}

sr_fgrid *grid::fieldOps() const {
    if (fops) return fops;
    const parser &p = inputParams;
    sr_fgrid_desc d{};
    d.xInd = p.xInd; d.yInd = p.yInd; d.zInd = p.zInd;
    d.Lx = p.Lx; d.Ly = p.Ly; d.Lz = p.Lz;
    d.xGrid = p.xGrid; d.yGrid = p.yGrid; d.zGrid = p.zGrid;
    d.betaX = p.betaX; d.betaY = p.betaY; d.betaZ = p.betaZ;
    d.xPer = p.xPer; d.yPer = p.yPer; d.zPer = p.zPer;
    d.iScheme = p.iScheme;
    if (sr_fgrid_create(&d, &fops) != SR_OK) {
        std::printf("Field operator set-up failed: %s. ABORTING\n", sr_last_error());
        std::exit(EXIT_FAILURE);
    }
    return fops;
}
