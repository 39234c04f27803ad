// The right-hand side kernels and their launches at finite-difference order q = 4.
#include "tgh_launch.hpp"

TGH_INSTANTIATE_ORDER(4)
