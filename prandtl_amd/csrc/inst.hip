// One (dimension, order) pair of the k_visc / k_rhs launch templates and
// every kernel variant they launch, chosen by -DPRANDTL_INST_DIM and
// -DPRANDTL_INST_NP: the Makefile builds one object per pair of
// PRANDTL_ORDERS, in parallel.
#include "dg_launch.h"

namespace prandtl {
PRANDTL_INSTANTIATE(, PRANDTL_INST_DIM, PRANDTL_INST_NP)
} // namespace prandtl
