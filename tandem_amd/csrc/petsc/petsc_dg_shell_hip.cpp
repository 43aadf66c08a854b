// PETSc MATSHELL over the MI355X DG operator (replaces app/common/PetscDGShell.cpp:8-55).  Compiled
// only where PETSc's headers are found (the build passes PETSC_DIR's include directory); elsewhere
// this translation unit is empty.
#if __has_include(<petscmat.h>)

#include "../../../include/tandem_dg_petsc.h"

#include <stdexcept>
#include <string>

namespace tdg {

namespace {
void check_petsc(PetscErrorCode e, char const* what) {
    if (e != PETSC_SUCCESS) throw std::runtime_error(std::string("PETSc error in ") + what);
}
} // namespace

PetscDGShellHIP::PetscDGShellHIP(tdg_op* op, MPI_Comm comm) {
    int64_t bs = 0, n = 0;
    if (tdg_block_size(op, &bs) != TDG_OK || tdg_local_size(op, &n) != TDG_OK)
        throw std::invalid_argument(tdg_last_error());
    check_petsc(MatCreate(comm, &A_), "MatCreate");
    check_petsc(MatSetSizes(A_, PetscInt(n), PetscInt(n), PETSC_DETERMINE, PETSC_DETERMINE), "MatSetSizes");
    check_petsc(MatSetBlockSize(A_, PetscInt(bs)), "MatSetBlockSize");
    check_petsc(MatSetType(A_, MATSHELL), "MatSetType");
    check_petsc(MatShellSetContext(A_, static_cast<void*>(op)), "MatShellSetContext");
    check_petsc(MatShellSetOperation(A_, MATOP_MULT, (void (*)(void))apply), "MatShellSetOperation");
    check_petsc(MatShellSetVecType(A_, VECHIP), "MatShellSetVecType"); // vectors in HBM
    check_petsc(MatSetUp(A_), "MatSetUp");
}

PetscDGShellHIP::~PetscDGShellHIP() { (void)MatDestroy(&A_); }

PetscErrorCode PetscDGShellHIP::apply(Mat A, Vec x, Vec y) {
    tdg_op* op = nullptr;
    PetscFunctionBeginUser;
    PetscCall(MatShellGetContext(A, &op));
    const PetscScalar* xa = nullptr;
    PetscScalar* ya = nullptr;
    PetscCall(VecHIPGetArrayRead(x, &xa));
    PetscCall(VecHIPGetArrayWrite(y, &ya));
    const int rc = tdg_mult(op, reinterpret_cast<double const*>(xa), reinterpret_cast<double*>(ya));
    PetscCall(VecHIPRestoreArrayRead(x, &xa));
    PetscCall(VecHIPRestoreArrayWrite(y, &ya));
    PetscCheck(rc == TDG_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "%s", tdg_last_error());
    PetscFunctionReturn(PETSC_SUCCESS);
}

} // namespace tdg

#endif
