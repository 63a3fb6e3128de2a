// Plain scalar field on the node grid (lib/field/plainsf.h): the padded array F the pressure
// solve reads and writes.
#pragma once

#include "grid.h"
#include "sr_array.h"

class plainsf {
  public:
    explicit plainsf(const grid &gridData)
        : F(gridData.globalSize[0], gridData.globalSize[1], gridData.globalSize[2]) {}
    Array3 F;
    void operator=(real a) { F = a; }
};
