// Host-side element matrices (see ke.cpp).
#pragma once

#include "../../include/hexa.h"

namespace hexa {

// 24x24 row-major hex8 brick stiffness for E = 1; mode is HEXA_KE_*.
bool element_stiffness(double dx, double dy, double dz, double nu, int mode, double* Ke);

// 8x8 scalar Laplacian (K8) and mass (M8) of the Helmholtz filter element.
void helmholtz_element(double dx, double dy, double dz, int bug_compat, double* K8, double* M8);

}  // namespace hexa
