/* PETSc adapter of the MI355X DG operator (C++; available only in builds that find PETSc).
 *
 * PetscDGShellHIP replaces tandem's PetscDGShell (app/common/PetscDGShell.h:12-24,
 * PetscDGShell.cpp:8-55): a MATSHELL of local size block_size x local elements whose MATOP_MULT is
 * tdg_mult on the device arrays of VECHIP vectors, so KSPSolve's vectors stay in HBM.  The operator
 * is borrowed (as tandem borrows its DGOperator) and must outlive the matrix.  Unlike tandem's
 * shell, a failed apply is reported to PETSc (PETSC_ERR_LIB with tdg_last_error()). */
#ifndef TANDEM_DG_PETSC_H
#define TANDEM_DG_PETSC_H

#include "tandem_dg.h"

#include <petscmat.h>

namespace tdg {

class PetscDGShellHIP {
public:
    PetscDGShellHIP(tdg_op* op, MPI_Comm comm);
    ~PetscDGShellHIP();
    PetscDGShellHIP(PetscDGShellHIP const&) = delete;
    PetscDGShellHIP& operator=(PetscDGShellHIP const&) = delete;

    Mat mat() const { return A_; }

private:
    static PetscErrorCode apply(Mat A, Vec x, Vec y);
    Mat A_ = nullptr;
};

} // namespace tdg

#endif /* TANDEM_DG_PETSC_H */
