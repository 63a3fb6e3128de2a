// The solver parameters the pressure solve reads, with the field names of lib/io/parser.h:65-102.
// The reference fills them from input/parameters.yaml; here the caller sets them.
#pragma once

#include "sr_array.h"

class parser {
  public:
    int npX = 1, npY = 1;
    int xInd = 5, yInd = 5, zInd = 5;
    int resType = 2, vcDepth = 4, vcCount = 25;
    int gsSmooth = 1, preSmooth = 2, postSmooth = 2;
    int xGrid = 1, yGrid = 1, zGrid = 1;
    bool printResidual = false;
    bool xPer = false, yPer = false, zPer = false;
    real Lx = 1.0, Ly = 1.0, Lz = 1.0;
    real betaX = 1.0, betaY = 1.0, betaZ = 1.0;
    real mgTolerance = 1.0e-6;
    int iScheme = 1;

    // additions of this implementation (no reference counterpart)
    bool testPoisson = false;    // the TEST_POISSON compile flag of the reference, as a run-time switch
    int semantics = 0;           // 0: reference stop rule; 1: projected RHS, relative-residual stop
    int precision = 0;           // 0: FP64; 1: FP32 inner cycles under FP64 refinement
    int padMode = 0;
    int coarseProject = 0;
    real relTolerance = 1.0e-10;
};
