/*
 * tandem_dg.h -- C ABI of the MI355X-native matrix-free DG operator (elasticity and Poisson SIPG on
 * curvilinear simplices).  Plain pointers and sizes only; every function returns 0 on success and a
 * negative code on failure, with the message available from tdg_last_error().  No exceptions
 * cross this boundary.
 *
 * Vectors are element-blocked exactly as tandem lays them out (app/common/PetscVector.cpp:12-40,
 * src/form/DGOperator.h:287-298): element e owns block_size consecutive entries, entry (k, p) of the
 * block (basis function k, component p) at offset k + p * Nbf.
 */
#ifndef TANDEM_DG_H
#define TANDEM_DG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary conditions, same values as tandem's BC enum (src/form/BC.h:6). */
enum { TDG_BC_NONE = 0, TDG_BC_NATURAL = 1, TDG_BC_FAULT = 3, TDG_BC_DIRICHLET = 5 };
/* Local operators: tandem's Elasticity (app/localoperator/Elasticity.h:33) and
 * Poisson (app/localoperator/Poisson.h). */
enum { TDG_ELASTICITY = 0, TDG_POISSON = 1 };
/* Error codes. */
enum { TDG_OK = 0, TDG_ERR_ARG = -1, TDG_ERR_STATE = -2, TDG_ERR_HIP = -3, TDG_ERR_INTERNAL = -4 };

typedef struct tdg_op tdg_op;

/* Mesh input; replaces the LocalSimplexMesh + Curvilinear pair tandem builds its operator from
 * (src/mesh/LocalSimplexMesh.h, src/geometry/Curvilinear.cpp:32-160).  Element vertex lists are
 * sorted ascending on entry, as tandem's Simplex does (src/mesh/Simplex.h:21-27).
 * geometry_nodes (optional) gives, per element, the binom(degree+dim, dim) warp-and-blend geometry
 * nodes x dim coordinates in the frame of the sorted vertices; NULL means the affine map. */
typedef struct {
    int dim;
    int degree;
    int64_t num_vertices;
    double const* vertices;        /* dim per vertex */
    int64_t num_elements;
    int64_t const* elements;       /* dim+1 per element */
    int64_t num_boundary_facets;
    int64_t const* boundary_facets; /* dim per facet */
    int const* boundary_bc;         /* TDG_BC_* per boundary facet */
    double const* geometry_nodes;
} tdg_mesh_desc;

/* Builds topology and FP64 geometry.  Replaces DGOperatorTopo (src/form/DGOperatorTopo.cpp:13-109)
 * and the geometric part of the DGOperator constructor (src/form/DGOperator.h:54-105,
 * src/form/DGCurvilinearCommon.cpp:41-101). */
int tdg_create(tdg_mesh_desc const* mesh, int kind, tdg_op** op);
int tdg_destroy(tdg_op* op);
char const* tdg_last_error(void);

/* Sizes: block_size() and num_local_elements() of AbstractDGOperator
 * (src/form/AbstractDGOperator.h:26-30).  out[0..6] = block size, elements, local size,
 * facets, volume points per element, facet points per facet, dim. */
int tdg_sizes(tdg_op const* op, int64_t* out);

/* Coordinates of the volume quadrature points (elements x Nq x dim) and of the facet quadrature
 * points (facets x nq x dim), where tandem evaluates its coefficient and BC functionals
 * (src/form/DGCurvilinearCommon.h:63-124).  Host buffers. */
int tdg_volume_points(tdg_op const* op, double* coords);
int tdg_facet_points(tdg_op const* op, double* coords);
/* Facet table, facets x 5: element side 0, element side 1 (= side 0 on the boundary), local face
 * side 0, local face side 1, BC.  Mirrors FacetInfo (src/form/FacetInfo.h:18-24). */
int tdg_facets(tdg_op const* op, int64_t* out);

/* Coefficients at the volume quadrature points (host, elements x Nq): lambda and mu for
 * elasticity, K (c1 ignored) for Poisson.  Runs the material part of the precompute
 * (Elasticity.cpp:104-292, Poisson.cpp:133-223) in FP64 on the host. */
int tdg_set_material(tdg_op* op, double const* c0_Q, double const* c1_Q);

/* Allocates the device buffers and copies the FP64 and FP32 operator data to the GPU.  The first
 * tdg_mult after tdg_set_material does this implicitly; call it beforehand to keep allocations out
 * of a timed or captured region. */
int tdg_upload(tdg_op* op);

/* y = A x, device pointers of local_size entries.  Replaces DGOperator<LOp>::apply
 * (src/form/DGOperator.h:239-243, 363-396) as called by PetscDGShell::apply
 * (app/common/PetscDGShell.cpp:48-55).  Asynchronous on the operator's stream. */
int tdg_mult(tdg_op* op, double const* x, double* y);
/* The same operator in FP32 arithmetic (FP32 copies of the FP64 precompute). */
int tdg_mult_f32(tdg_op* op, float const* x, float* y);
/* Y(:, c) = A X(:, c) for ncols columns with leading dimension ld (>= local size). */
int tdg_mult_batched(tdg_op* op, double const* X, double* Y, int ncols, int64_t ld);
int tdg_mult_batched_f32(tdg_op* op, float const* X, float* Y, int ncols, int64_t ld);

/* Stream all device work of this operator is queued on (hipStream_t; NULL = default stream). */
int tdg_set_stream(tdg_op* op, void* stream);

/* Kernel choice: TDG_VARIANT_AUTO (default) runs the matrix-core (MFMA) kernel with its tables
 * staged in LDS where one is compiled for the operator (elasticity, 3D P2-P4, 2D P2/P4/P6) and the
 * thread-per-point kernel otherwise; TDG_VARIANT_POINTS always runs the thread-per-point kernel;
 * TDG_VARIANT_MFMA_TABLES the MFMA kernel reading pre-swizzled operand tables from memory.  All
 * compute the same operator (the reference's element-centric apply); they differ in summation
 * order only. */
enum { TDG_VARIANT_AUTO = 0, TDG_VARIANT_POINTS = 1, TDG_VARIANT_MFMA_TABLES = 2 };
int tdg_set_variant(tdg_op* op, int variant);

/* Right-hand side b (host, local size) from a body force at the volume points (elements x Q x Nq,
 * may be NULL) and Dirichlet data at the facet points (facets x Q x nq, only boundary Dirichlet
 * facets are read; may be NULL).  Replaces DGOperator::rhs (src/form/DGOperator.h:189-237) with
 * Elasticity::rhs_volume / rhs_boundary (Elasticity.cpp:521-696). */
int tdg_rhs(tdg_op const* op, double const* force_Q, double const* dirichlet_q, double* b);

/* Linear solve A x = b on the GPU (b, x device FP64, x overwritten, initial guess 0); replaces
 * PetscLinearSolver::solve = KSPSolve with KSPCG on the shell matrix
 * (app/common/PetscLinearSolver.h:42, PetscLinearSolver.cpp:22-32), without preconditioner.
 * mode 0: FP64 CG.  mode 1: FP32 CG inner solves (to inner_rtol) inside an FP64 iterative
 * refinement loop whose residual b - A x and update x += d are FP64 (restarted each refinement).
 * mode 2: FP32 CG whose residual is replaced by the FP64 true residual, and whose accumulated FP32
 * update is added to the FP64 x, each time the residual drops by reliable_delta (refinement
 * without restarting the Krylov direction).  rtol applies to the FP64 true residual. */
enum { TDG_SOLVE_FP64_CG = 0, TDG_SOLVE_FP32_CG_FP64_IR = 1, TDG_SOLVE_FP32_CG_FP64_RELIABLE = 2 };
typedef struct {
    double rtol;
    int max_iter;
    int mode;
    double inner_rtol;
    int max_outer;
    int check_every;
    double reliable_delta;
} tdg_solve_opts;
typedef struct {
    int iterations;
    int outer_iterations;
    int inner_iterations;
    double rel_residual;
    double seconds;
    int converged;
} tdg_solve_result;
int tdg_solve(tdg_op* op, double const* b, double* x, tdg_solve_opts const* opts, tdg_solve_result* result);

/* Algorithmic bytes (for the given bytes per real) and flops of one apply. */
int tdg_apply_counts(tdg_op const* op, int precision_bytes, double* bytes, double* flops);

/* Named FP64 precompute arrays (host copy), for inspection: absDetJ, JInv, vcoords, volume,
 * JInv0, JInv1, normal, normalLength, fcoords, area, penalty, mat0, mat1, c0Q, c1Q, fc0, fc1.
 * Returns the element count in *n; data copied if out != NULL. */
int tdg_host_array(tdg_op const* op, char const* name, double* out, int64_t* n);

/* Box-mesh builder (input helper, not part of the operator): [0,1]^dim with n[d] cells,
 * scheme 0 = tandem's GenMesh tessellation (src/mesh/GenMesh.cpp:21-143), 1 = 6-tet Kuhn cubes.
 * bc[2d], bc[2d+1]: BC on x_d = 0 and x_d = 1.  Two calls: sizes (arrays NULL), then fill. */
int tdg_box_mesh(int dim, int64_t const* n, int scheme, int const* bc, int64_t* sizes, double* vertices,
                 int64_t* elements, int64_t* boundary_facets, int* boundary_bc);
/* Warp-and-blend reference nodes (src/basis/WarpAndBlend.cpp:45-184), binom(degree+dim,dim) x dim. */
int tdg_reference_nodes(int dim, int degree, double* out);

#ifdef __cplusplus
}
#endif

#endif /* TANDEM_DG_H */
