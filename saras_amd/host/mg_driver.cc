// Command-line driver of the C++ drop-in classes (saras_amd/host), in the manner of the reference's
// test programs:
//   saras_mg_driver kat <ind> <domain>      testProlong / testTransfer / testPeriodic, printed like
//                                           src/tests/global/unittest.cc printResult (tol 1e-10)
//   saras_mg_driver mgtest <ind> <depth> <cycles>
//                                           TEST_POISSON r^2/6 case of tests/testPoisson.sh on one GPU
//   saras_mg_driver fieldtest <ind>         the field classes (field.h) on a periodic unit cube: the
//                                           unit tests of src/tests/units (differ_test.cc,
//                                           field_test.cc, nlin_test.cc) and the ghost values of
//                                           hydro_d3::testPeriodic (hydro_d3.cc:536-679)
// Exit status is non-zero on a failed check.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include <cmath>
#include <vector>

#include "field.h"
#include "grid.h"
#include "parser.h"
#include "plainsf.h"
#include "poisson.h"

static int printResult(const char *what, real err, real tol) {
    bool ok = err < tol;
    std::printf("%-50s %s (%.3e)\n", what, ok ? "[PASSED]" : "[FAILED]", err);
    return ok ? 0 : 1;
}

// testError (src/tests/global/unittest.cc:45-79), errorMom 1: mean absolute difference over the
// array without its outermost layer
static real meanAbsInterior(const Array3 &A, const Array3 &B) {
    real err = 0.0;
    long cnt = 0;
    for (int i = A.lbound(0) + 1; i <= A.ubound(0) - 1; i++)
        for (int j = A.lbound(1) + 1; j <= A.ubound(1) - 1; j++)
            for (int k = A.lbound(2) + 1; k <= A.ubound(2) - 1; k++) {
                err += std::fabs(A(i, j, k) - B(i, j, k));
                cnt += 1;
            }
    return err / (real)cnt;
}

// Taylor-Green velocity on the whole padded arrays (field_test.cc / nlin_test.cc taylorGreen)
static void taylorGreenV(vfield &V, const grid &m, Array3 *Hx) {
    const real tp = 2.0 * M_PI;
    Array3 vx(V.Vx.F.ubound(0), V.Vx.F.ubound(1), V.Vx.F.ubound(2));
    Array3 vy(V.Vy.F.ubound(0), V.Vy.F.ubound(1), V.Vy.F.ubound(2));
    if (Hx) Hx->resize(V.Vx.F.ubound(0), V.Vx.F.ubound(1), V.Vx.F.ubound(2));
    for (int i = -1; i <= vx.ubound(0); i++)
        for (int j = -1; j <= vx.ubound(1); j++)
            for (int k = -1; k <= vx.ubound(2); k++) {
                vx(i, j, k) = std::sin(tp * m.xColloc(i) / m.xLen) * std::cos(tp * m.yStaggr(j) / m.yLen) *
                              std::cos(tp * m.zStaggr(k) / m.zLen);
                if (Hx)
                    (*Hx)(i, j, k) = M_PI * std::sin(4.0 * M_PI * m.xColloc(i) / m.xLen) *
                                     std::pow(std::cos(tp * m.zStaggr(k) / m.zLen), 2.0);
            }
    for (int i = -1; i <= vy.ubound(0); i++)
        for (int j = -1; j <= vy.ubound(1); j++)
            for (int k = -1; k <= vy.ubound(2); k++)
                vy(i, j, k) = -std::cos(tp * m.xStaggr(i) / m.xLen) * std::sin(tp * m.yColloc(j) / m.yLen) *
                              std::cos(tp * m.zStaggr(k) / m.zLen);
    V.Vx.F.upload(vx);
    V.Vy.F.upload(vy);
    V.Vz.F = 0.0;
}

static int fieldTests(int ind) {
    parser p;
    p.xInd = p.yInd = p.zInd = ind;
    p.vcDepth = ind - 1;
    p.xPer = p.yPer = p.zPer = true;
    grid g(p);
    const real tp = 2.0 * M_PI;
    const real hmax = std::max(std::max(g.dXi, g.dEt), g.dZt);
    int bad = 0;
    // differ_test.cc: d/dx of sin(2 pi x) cos(2 pi y) cos(2 pi z) on the nodes
    {
        sfield F(g, "F");
        Array3 f(g.globalSize[0], g.globalSize[1], g.globalSize[2]), df = f;
        for (int i = -1; i <= f.ubound(0); i++)
            for (int j = -1; j <= f.ubound(1); j++)
                for (int k = -1; k <= f.ubound(2); k++) {
                    real cy = std::cos(tp * g.yStaggr(j) / g.yLen), cz = std::cos(tp * g.zStaggr(k) / g.zLen);
                    f(i, j, k) = std::sin(tp * g.xStaggr(i) / g.xLen) * cy * cz;
                    df(i, j, k) = tp * std::cos(tp * g.xStaggr(i) / g.xLen) * cy * cz;
                }
        F.F.F.upload(f);
        DArray3 out(g.globalSize[0], g.globalSize[1], g.globalSize[2]);
        F.derS.calcDerivative1_x(out);
        bad += printResult("Testing computation of derivatives in sfield", meanAbsInterior(df, out.host()),
                           g.dXi * g.dXi * 4.0 * M_PI * M_PI);
    }
    // field_test.cc: the divergence of the Taylor-Green velocity
    {
        vfield V(g, "V");
        sfield P(g, "P");
        plainsf div(g);
        taylorGreenV(V, g, nullptr);
        V.divergence(div, P);
        Array3 d = div.F.host();
        real m = 0.0;
        for (int i = 0; i < g.globalSize[0]; i++)
            for (int j = 0; j < g.globalSize[1]; j++)
                for (int k = 0; k < g.globalSize[2]; k++) m = std::max(m, std::fabs(d(i, j, k)));
        bad += printResult("Testing divergence calculation of vfield", m, hmax);
    }
    // nlin_test.cc: the advective term of the Taylor-Green velocity; computeNLin subtracts it from
    // H (vfield.cc:172), so the analytic (u . grad) u is compared with -H
    {
        vfield V(g, "V");
        plainvf H(g);
        Array3 hx;
        taylorGreenV(V, g, &hx);
        H = 0.0;
        V.computeNLin(V, H);
        Array3 h = H.Vx.host();
        for (int i = -1; i <= h.ubound(0); i++)
            for (int j = -1; j <= h.ubound(1); j++)
                for (int k = -1; k <= h.ubound(2); k++) h(i, j, k) = -h(i, j, k);
        bad += printResult("Testing non-linear term calculation", meanAbsInterior(hx, h), 2.0 * M_PI * hmax);
    }
    // hydro_d3::testPeriodic: core set, pads from imposeUBCs, compared with the analytic field at the
    // mirrored ghost coordinates of the x, y and z pads
    {
        vfield V(g, "V");
        taylorGreenV(V, g, nullptr);
        Array3 before = V.Vx.F.host();
        Array3 z0(before.ubound(0), before.ubound(1), before.ubound(2));
        for (int i = -1; i <= z0.ubound(0); i++)
            for (int j = -1; j <= z0.ubound(1); j++)
                for (int k = -1; k <= z0.ubound(2); k++) {
                    bool core = i >= 0 && j >= 0 && k >= 0 && i < z0.ubound(0) && j < z0.ubound(1) && k < z0.ubound(2);
                    z0(i, j, k) = core ? before(i, j, k) : 0.0;
                }
        V.Vx.F.upload(z0);
        sr_bc per{};
        per.kind[4] = per.kind[5] = SR_BC_PERIODIC;   // z walls: periodicFC / CC; x, y: syncData
        V.Vx.imposeBC(per);
        Array3 after = V.Vx.F.host();
        int nx = after.ubound(0), ny = after.ubound(1), nz = after.ubound(2);
        real err = 0.0;
        auto val = [&](real x, real y, real z) {
            return std::sin(tp * x / g.xLen) * std::cos(tp * y / g.yLen) * std::cos(tp * z / g.zLen);
        };
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                real xl = g.xColloc(0) - (g.xColloc(1) - g.xColloc(0));
                real xr = g.xColloc(nx - 1) + (g.xColloc(nx - 1) - g.xColloc(nx - 2));
                err = std::max(err, std::fabs(after(-1, j, k) - val(xl, g.yStaggr(j), g.zStaggr(k))));
                err = std::max(err, std::fabs(after(nx, j, k) - val(xr, g.yStaggr(j), g.zStaggr(k))));
            }
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) {
                real yl = g.yStaggr(0) - (g.yStaggr(1) - g.yStaggr(0));
                real yr = g.yStaggr(ny - 1) + (g.yStaggr(ny - 1) - g.yStaggr(ny - 2));
                err = std::max(err, std::fabs(after(i, -1, k) - val(g.xColloc(i), yl, g.zStaggr(k))));
                err = std::max(err, std::fabs(after(i, ny, k) - val(g.xColloc(i), yr, g.zStaggr(k))));
            }
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                real zl = g.zStaggr(0) - (g.zStaggr(1) - g.zStaggr(0));
                real zr = g.zStaggr(nz - 1) + (g.zStaggr(nz - 1) - g.zStaggr(nz - 2));
                err = std::max(err, std::fabs(after(i, j, -1) - val(g.xColloc(i), g.yStaggr(j), zl)));
                err = std::max(err, std::fabs(after(i, j, nz) - val(g.xColloc(i), g.yStaggr(j), zr)));
            }
        bad += printResult("Testing periodic data transfer of vfield", err, 1.0e-10);
    }
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("usage: %s kat <ind> <domain> | mgtest <ind> <depth> <cycles> | fieldtest <ind>\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    parser p;
    if (mode == "kat") {
        int ind = argc > 2 ? std::atoi(argv[2]) : 5;
        std::string dom = argc > 3 ? argv[3] : "PPP";
        p.xInd = p.yInd = p.zInd = ind;
        p.vcDepth = ind - 1;
        p.xPer = dom[0] == 'P'; p.yPer = dom[1] == 'P'; p.zPer = dom[2] == 'P';
        grid g(p);
        multigrid_d3 mg(g, p);
        int bad = 0;
        bad += printResult("Testing prolongation operations", mg.testProlong(), 1.0e-10);
        bad += printResult("Testing strided data transfer operations", mg.testTransfer(), 1.0e-10);
        bad += printResult("Testing periodic data transfer operations", mg.testPeriodic(), 1.0e-10);
        return bad ? 1 : 0;
    }
    if (mode == "mgtest") {
        int ind = argc > 2 ? std::atoi(argv[2]) : 5;
        p.xInd = p.yInd = p.zInd = ind;
        p.vcDepth = argc > 3 ? std::atoi(argv[3]) : ind - 1;
        p.vcCount = argc > 4 ? std::atoi(argv[4]) : 25;
        p.testPoisson = true;
        p.printResidual = true;
        p.resType = 2;
        grid g(p);
        multigrid_d3 mg(g, p);
        plainsf Pp(g), rhs(g);
        Pp = 0.0;
        rhs = 1.0;   // mgRHS.F = 1.0 under TEST_POISSON (hydro_d3.cc:332-334)
        mg.mgSolve(Pp, rhs);
        return 0;
    }
    if (mode == "fieldtest") return fieldTests(argc > 2 ? std::atoi(argv[2]) : 5);
    std::printf("unknown mode %s\n", mode.c_str());
    return 2;
}
