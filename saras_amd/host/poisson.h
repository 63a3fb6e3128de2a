// Drop-in for lib/poisson/poisson.h: the same classes, constructor arguments and public methods,
// with the level hierarchies held on the GPU behind the C ABI of include/saras_mg.h.
#pragma once

#include "grid.h"
#include "parser.h"
#include "plainsf.h"

struct sr_mg;

class poisson {
  protected:
    int vLevel;
    bool zeroBC;

    const grid &mesh;
    const parser &inputParams;

    sr_mg *dev;                 // device hierarchy (replaces pressureData .. smoothedPres)

    // the protected virtuals of poisson.h:103-117, forwarding to the device operators
    virtual void coarsen();
    virtual void prolong();
    virtual void computeResidual();
    virtual void smooth(const int smoothCount);
    virtual real computeError(const int normOrder);
    virtual void solve();
    virtual void imposeBC();
    virtual void updatePads(int which);

    void vCycle();
    void check(int status) const;

  public:
    poisson(const grid &mesh, const parser &solParam);

    void mgSolve(plainsf &inFn, const plainsf &rhs);

    virtual real testProlong();
    virtual real testTransfer();
    virtual real testPeriodic();

    virtual ~poisson();
};

class multigrid_d3 : public poisson {
  public:
    multigrid_d3(const grid &mesh, const parser &solParam);
    ~multigrid_d3() {}
};
