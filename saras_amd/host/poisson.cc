// poisson / multigrid_d3 over the C ABI (include/saras_mg.h).  Public behaviour follows
// lib/poisson/poisson.cc and poisson_d3.cc; fatal conditions print and exit (the reference exits
// with status 0, here the status is non-zero).
#include "poisson.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <vector>

#include "saras_mg.h"

poisson::poisson(const grid &mesh, const parser &solParam)
    : vLevel(0), zeroBC(false), mesh(mesh), inputParams(solParam), dev(nullptr) {}

poisson::~poisson() {
    if (dev) sr_mg_destroy(dev);
}

void poisson::check(int status) const {
    if (status != SR_OK && status != SR_ERR_NOT_CONVERGED) {
        std::cout << "Multigrid solver error " << status << ": " << sr_last_error() << ". ABORTING" << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

// TEST_POISSON Dirichlet planes r^2/6 (poisson_d3.cc:513-572), single sub-domain
static void r2Walls(const grid &mesh, std::vector<std::vector<real>> &w) {
    int n[3] = {mesh.globalSize[0], mesh.globalSize[1], mesh.globalSize[2]};
    real h[3] = {mesh.dXi, mesh.dEt, mesh.dZt};
    real L[3] = {mesh.xLen, mesh.yLen, mesh.zLen};
    int M[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    int half[3] = {M[0] / 2, M[1] / 2, M[2] / 2};
    w.assign(6, {});
    for (int d = 0; d < 3; d++) {
        int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
        std::vector<real> plane((size_t)(n[a] + 2) * (n[b] + 2), 0.0);
        real dd = h[d] + L[d] / 2.0;
        for (int p = 0; p < n[a]; p++)
            for (int q = 0; q < n[b]; q++) {
                real da = h[a] * (p - half[a]), db = h[b] * (q - half[b]);
                plane[(size_t)(p + 1) * (n[b] + 2) + (q + 1)] = (dd * dd + da * da + db * db) / 6.0;
            }
        w[2 * d] = plane;
        w[2 * d + 1] = plane;
    }
}

multigrid_d3::multigrid_d3(const grid &mesh, const parser &solParam) : poisson(mesh, solParam) {
    const parser &p = solParam;
    sr_mg_desc d{};
    d.xInd = p.xInd; d.yInd = p.yInd; d.zInd = p.zInd;
    d.vcDepth = p.vcDepth; d.vcCount = p.vcCount;
    d.preSmooth = p.preSmooth; d.postSmooth = p.postSmooth;
    d.gsSmooth = p.gsSmooth; d.resType = p.resType;
    d.xPer = p.xPer; d.yPer = p.yPer; d.zPer = p.zPer;
    d.testPoisson = p.testPoisson; d.semantics = p.semantics; d.precision = p.precision;
    d.padMode = p.padMode; d.coarseProject = p.coarseProject;
    d.mgTolerance = p.mgTolerance; d.relTolerance = p.relTolerance;
    d.dXi = mesh.dXi; d.dEt = mesh.dEt; d.dZt = mesh.dZt;
    d.xix2 = mesh.xix2Staggr.data(); d.xixx = mesh.xixxStaggr.data();
    d.ety2 = mesh.ety2Staggr.data(); d.etyy = mesh.etyyStaggr.data();
    d.ztz2 = mesh.ztz2Staggr.data(); d.ztzz = mesh.ztzzStaggr.data();
    std::vector<std::vector<real>> walls;
    if (p.testPoisson) {
        r2Walls(mesh, walls);
        for (int q = 0; q < 6; q++) d.wall[q] = walls[q].data();
    }
    check(sr_mg_create(&d, &dev));
}

// poisson.cc:100-205
void poisson::mgSolve(plainsf &inFn, const plainsf &rhs) {
    sr_mg_stats st{};
    // both fields live in HBM: the solve reads and writes them in place
    check(sr_mg_solve(dev, inFn.F.data(), rhs.F.data(), 1, &st));
    if (inputParams.printResidual) {
        std::vector<double> h(st.cycles > 0 ? st.cycles : 1);
        int n = sr_mg_history(dev, h.data(), (int)h.size());
        for (int i = 0; i < n; i++)
            std::cout << std::endl << "Residual after V Cycle " << i << " is " << std::setprecision(16) << h[i] << std::endl;
    }
    if (inputParams.testPoisson) {
        int n[3] = {mesh.globalSize[0], mesh.globalSize[1], mesh.globalSize[2]};
        Array3 sol = inFn.F.host();
        real gloMax = 0.0;
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++) {
                    real x = mesh.dXi * (i - (n[0] - 1) / 2), y = mesh.dEt * (j - (n[1] - 1) / 2),
                         z = mesh.dZt * (k - (n[2] - 1) / 2);
                    real a = (x * x + y * y + z * z) / 6.0;
                    gloMax = std::max(gloMax, std::fabs(a - sol(i, j, k)));
                }
        std::cout << std::endl << "Maximum absolute deviation from analytic solution is: " << gloMax << std::endl
                  << std::endl;
    }
}

void poisson::coarsen() { check(sr_mg_op_coarsen(dev)); vLevel += 1; }
void poisson::prolong() { check(sr_mg_op_prolong(dev)); vLevel -= 1; }
void poisson::computeResidual() { check(sr_mg_op_compute_residual(dev)); }
void poisson::smooth(const int smoothCount) {
    check(sr_mg_set_zero_bc(dev, zeroBC));
    check(sr_mg_op_smooth(dev, smoothCount));
}
real poisson::computeError(const int normOrder) {
    double r = 0.0;
    check(sr_mg_op_compute_error(dev, normOrder, &r));
    return r;
}
void poisson::solve() {
    check(sr_mg_set_zero_bc(dev, zeroBC));
    check(sr_mg_op_solve_coarsest(dev));
}
void poisson::imposeBC() {
    check(sr_mg_set_zero_bc(dev, zeroBC));
    check(sr_mg_op_impose_bc(dev));
}
void poisson::updatePads(int which) { check(sr_mg_op_update_pads(dev, which)); }
void poisson::vCycle() { check(sr_mg_op_vcycle(dev)); vLevel = 0; }

// ---- known-answer tests of poisson_d3.cc:689-857, level by level --------------------------------
static std::vector<real> padded(int nx, int ny, int nz) { return std::vector<real>((size_t)(nx + 2) * (ny + 2) * (nz + 2), 0.0); }
static size_t at(int ny, int nz, int i, int j, int k) { return ((size_t)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + (k + 1); }

// testProlong: 1000 + 100 i + 10 j + k (the fine-index formula, coarse node I on fine node 2I)
real poisson::testProlong() {
    real maxErr = 0.0;
    for (int l = 1; l <= inputParams.vcDepth; l++) {
        int cx, cy, cz, fx, fy, fz;
        sr_mg_level_dims(dev, l, &cx, &cy, &cz);
        sr_mg_level_dims(dev, l - 1, &fx, &fy, &fz);
        std::vector<real> c = padded(cx, cy, cz), f = padded(fx, fy, fz);
        for (int i = 0; i < cx; i++)
            for (int j = 0; j < cy; j++)
                for (int k = 0; k < cz; k++) c[at(cy, cz, i, j, k)] = 1000.0 + 200.0 * i + 20.0 * j + 2.0 * k;
        check(sr_mg_upload(dev, SR_PRESSURE, l, c.data()));
        check(sr_mg_set_level(dev, l));
        check(sr_mg_op_prolong(dev));
        check(sr_mg_download(dev, SR_PRESSURE, l - 1, f.data()));
        for (int i = 0; i < fx; i++)
            for (int j = 0; j < fy; j++)
                for (int k = 0; k < fz; k++)
                    maxErr = std::max(maxErr, std::fabs(f[at(fy, fz, i, j, k)] - (1000.0 + 100.0 * i + 10.0 * j + k)));
    }
    check(sr_mg_set_level(dev, 0));
    return maxErr;
}

// testTransfer: pads after updatePads hold the periodic partner planes (self exchange, one rank)
real poisson::testTransfer() {
    real maxErr = 0.0;
    for (int l = 0; l <= inputParams.vcDepth; l++) {
        int nx, ny, nz;
        sr_mg_level_dims(dev, l, &nx, &ny, &nz);
        std::vector<real> a = padded(nx, ny, nz);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) a[at(ny, nz, i, j, k)] = 1000.0 + 100.0 * i + 10.0 * j + k;
        check(sr_mg_upload(dev, SR_PRESSURE, l, a.data()));
        check(sr_mg_set_level(dev, l));
        check(sr_mg_op_update_pads(dev, SR_PRESSURE));
        check(sr_mg_download(dev, SR_PRESSURE, l, a.data()));
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                if (!inputParams.xPer) continue;
                maxErr = std::max(maxErr, std::fabs(a[at(ny, nz, -1, j, k)] - (1000.0 + (nx - 2) * 100.0 + 10.0 * j + k)));
                maxErr = std::max(maxErr, std::fabs(a[at(ny, nz, nx, j, k)] - (1000.0 + 100.0 + 10.0 * j + k)));
            }
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) {
                if (!inputParams.yPer) continue;
                maxErr = std::max(maxErr, std::fabs(a[at(ny, nz, i, -1, k)] - (1000.0 + 100.0 * i + (ny - 2) * 10.0 + k)));
                maxErr = std::max(maxErr, std::fabs(a[at(ny, nz, i, ny, k)] - (1000.0 + 100.0 * i + 10.0 + k)));
            }
    }
    check(sr_mg_set_level(dev, 0));
    return maxErr;
}

// testPeriodic: imposeBC ghosts of sin(2 pi x/Lx) cos(2 pi y/Ly) cos(2 pi z/Lz) along periodic directions
real poisson::testPeriodic() {
    real maxErr = 0.0;
    const Array1 *xs[3] = {&mesh.xStaggr, &mesh.yStaggr, &mesh.zStaggr};
    real len[3] = {mesh.xLen, mesh.yLen, mesh.zLen};
    bool per[3] = {inputParams.xPer, inputParams.yPer, inputParams.zPer};
    for (int l = 0; l <= inputParams.vcDepth; l++) {
        int n[3];
        sr_mg_level_dims(dev, l, &n[0], &n[1], &n[2]);
        int s = 1 << l;
        auto coord = [&](int d, int i) {
            const Array1 &x = *xs[d];
            if (i == -1) return x(0) - (x(s) - x(0));
            if (i == n[d]) return x((n[d] - 1) * s) + (x((n[d] - 1) * s) - x((n[d] - 2) * s));
            return x(i * s);
        };
        auto fn = [&](int i, int j, int k) {
            return std::sin(2.0 * M_PI * coord(0, i) / len[0]) * std::cos(2.0 * M_PI * coord(1, j) / len[1]) *
                   std::cos(2.0 * M_PI * coord(2, k) / len[2]);
        };
        std::vector<real> a = padded(n[0], n[1], n[2]);
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++) a[at(n[1], n[2], i, j, k)] = fn(i, j, k);
        check(sr_mg_upload(dev, SR_PRESSURE, l, a.data()));
        check(sr_mg_set_level(dev, l));
        check(sr_mg_set_zero_bc(dev, 1));
        check(sr_mg_op_impose_bc(dev));
        check(sr_mg_download(dev, SR_PRESSURE, l, a.data()));
        for (int i = 0; i < n[0]; i++)
            for (int j = 0; j < n[1]; j++)
                for (int k = 0; k < n[2]; k++) {
                    int idx[3] = {i, j, k};
                    for (int d = 0; d < 3; d++) {
                        if (!per[d]) continue;
                        for (int g : {-1, n[d]}) {
                            int q[3] = {idx[0], idx[1], idx[2]};
                            q[d] = g;
                            maxErr = std::max(maxErr, std::fabs(a[at(n[1], n[2], q[0], q[1], q[2])] - fn(q[0], q[1], q[2])));
                        }
                    }
                }
    }
    check(sr_mg_set_level(dev, 0));
    return maxErr;
}
