// Command-line driver of the C++ drop-in classes (saras_amd/host), in the manner of the reference's
// test programs:
//   saras_mg_driver kat <ind> <domain>      testProlong / testTransfer / testPeriodic, printed like
//                                           src/tests/global/unittest.cc printResult (tol 1e-10)
//   saras_mg_driver mgtest <ind> <depth> <cycles>
//                                           TEST_POISSON r^2/6 case of tests/testPoisson.sh on one GPU
// Exit status is non-zero on a failed check.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "grid.h"
#include "parser.h"
#include "plainsf.h"
#include "poisson.h"

static int printResult(const char *what, real err, real tol) {
    bool ok = err < tol;
    std::printf("%-50s %s (%.3e)\n", what, ok ? "[PASSED]" : "[FAILED]", err);
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("usage: %s kat <ind> <domain> | mgtest <ind> <depth> <cycles>\n", argv[0]);
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
    std::printf("unknown mode %s\n", mode.c_str());
    return 2;
}
