"""tandem_amd: MI355X-native matrix-free DG operator apply of tandem (elasticity and Poisson SIPG
on curvilinear simplices), exposed through the C ABI in include/tandem_dg.h."""
from .operator import (BC_DIRICHLET, BC_FAULT, BC_NATURAL, BC_NONE, ELASTICITY, POISSON, DGOperator,
                       Elasticity, Mesh, PMultigrid, Poisson, interpolate, interpolation_matrix,
                       reference_nodes)

__all__ = ["BC_DIRICHLET", "BC_FAULT", "BC_NATURAL", "BC_NONE", "ELASTICITY", "POISSON", "DGOperator",
           "Elasticity", "Mesh", "PMultigrid", "Poisson", "interpolate", "interpolation_matrix",
           "reference_nodes"]
