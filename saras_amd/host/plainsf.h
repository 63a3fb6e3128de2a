// Plain scalar field on the node grid (lib/field/plainsf.h): the padded array F the pressure
// solve reads and writes, resident in HBM.  Declared with the other field classes.
#pragma once

#include "field.h"
