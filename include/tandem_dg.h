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
/* The same two sizes one at a time: *bs = block_size() (Nbf x components,
 * src/form/AbstractDGOperator.h:26), *n = number_of_local_dofs() = block size x elements
 * (AbstractDGOperator.h:28; the MatShell's local rows and columns, app/common/PetscDGShell.cpp:14-17). */
int tdg_block_size(tdg_op const* op, int64_t* bs);
int tdg_local_size(tdg_op const* op, int64_t* n);

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
/* The FP64 apply y = A x without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation); no counterpart in the
 * reference.  The facet slots of the two-kernel apply live in a fixed 64 MiB device array of the
 * library (each GPU holds it until the process ends): operators whose slots need more
 * (elements x (dim+1) x block size doubles > 2^23, about 20,000 P4 tetrahedra) are rejected with
 * TDG_ERR_ARG.  The operator must be on the device already (tdg_upload, which also sets up the array
 * on the current device, or one tdg_mult), else TDG_ERR_STATE; with TDG_VARIANT_ASSEMBLED it must
 * have been assembled (tdg_assemble or one tdg_mult), and the apply then uses no scratch memory.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU race; calls ordered on one stream do not. */
int tdg_mult_noalloc(tdg_op* op, double const* x, double* y);
/* Y(:, c) = A X(:, c) for ncols columns with leading dimension ld (>= local size). */
int tdg_mult_batched(tdg_op* op, double const* X, double* Y, int ncols, int64_t ld);
int tdg_mult_batched_f32(tdg_op* op, float const* X, float* Y, int ncols, int64_t ld);

/* Stream all device work of this operator is queued on (hipStream_t; NULL = default stream). */
int tdg_set_stream(tdg_op* op, void* stream);

/* Kernel choice: TDG_VARIANT_AUTO (default) runs the two-kernel matrix-core (MFMA) apply where one
 * is compiled for the operator (elasticity, 3D P2-P4, 2D P2-P4 and P6): a facet kernel evaluates each
 * facet once (both sides' traces, the flux and both lifts) into per-element-face slots, an element
 * kernel applies the volume term and adds the slots; the thread-per-point kernel otherwise.
 * TDG_VARIANT_POINTS always runs the thread-per-point kernel.  TDG_VARIANT_MFMA_ELEMENT runs the
 * element-centric MFMA kernel (each element evaluates both sides of its faces, as the reference's
 * element loop does), TDG_VARIANT_MFMA_XCD the same kernel with its workgroups renumbered so that
 * each XCD takes one contiguous range of elements.  TDG_VARIANT_ASSEMBLED multiplies by the
 * assembled block rows (tdg_assemble; assembled by the first apply if needed): a register kernel
 * that reads each block row once per range of columns for block sizes 3, 6, 12 (2D) and 4, 12 (3D),
 * i.e. the P1 / P2 levels of p-multigrid, and a generic kernel for every other order (a cross-check
 * of the matrix-free apply; the autotuner only considers the register kernel).  All compute the
 * same operator; they differ in summation order only. */
enum {
    TDG_VARIANT_AUTO = 0,
    TDG_VARIANT_POINTS = 1,
    TDG_VARIANT_MFMA_XCD = 2,
    TDG_VARIANT_MFMA_ELEMENT = 3,
    TDG_VARIANT_ASSEMBLED = 4,
    TDG_NUM_VARIANTS = 5
};
int tdg_set_variant(tdg_op* op, int variant);

/* Autotuner: times every variant that applies to the operator (the two-kernel MFMA apply, the
 * element-centric MFMA kernel with linear and XCD-grouped workgroup order, the thread-per-point
 * kernel) on ncols columns in FP64 (precision_bytes 8) or FP32 (4), keeps the fastest as the
 * operator's variant and reports it in *chosen and the milliseconds per apply in ms[variant]
 * (TDG_NUM_VARIANTS entries, -1 where a variant does not apply).  No counterpart in the
 * reference (its kernels are generated for one CPU target). */
int tdg_autotune(tdg_op* op, int ncols, int precision_bytes, int* chosen, double* ms);

/* Launch shapes of the two-kernel MFMA apply (TDG_VARIANT_AUTO): TDG_SPLIT_SHAPES compiled shapes
 * per precision (waves per workgroup and the occupancy the register allocation targets for the
 * facet and the element kernel, the element kernel's prefetch depth, the last basis tile of the
 * lifts on the vector ALU or on the matrix cores) times TDG_SPLIT_ORDERS workgroup orders (0: linear,
 * 1: one contiguous range of groups per XCD, 2: the same ranges run by a persistent launch, as many
 * workgroups as are resident, each looping over its blocks).  tdg_autotune_shapes times all of them
 * on ncols columns in FP64 (precision_bytes 8) or FP32 (4), keeps the fastest for that precision,
 * reports it in *shape and *xcd and the milliseconds per apply in ms[TDG_SPLIT_ORDERS shape + xcd]
 * (TDG_SPLIT_ORDERS x TDG_SPLIT_SHAPES entries, -1 if the operator has no two-kernel apply).  tdg_set_shape selects one directly.  tdg_autotune runs the
 * shape search before it compares the variants.  No counterpart in the reference (its kernels are
 * generated for one CPU target). */
enum { TDG_SPLIT_SHAPES = 4, TDG_SPLIT_ORDERS = 3 };
int tdg_autotune_shapes(tdg_op* op, int ncols, int precision_bytes, int* shape, int* xcd, double* ms);
int tdg_set_shape(tdg_op* op, int precision_bytes, int shape, int xcd);

/* Right-hand side b (host, local size) from a body force at the volume points (elements x Q x Nq,
 * may be NULL) and facet data at the facet points (facets x Q x nq, may be NULL): the Dirichlet
 * values on Dirichlet facets and the slip on Fault facets (halved on boundary Fault facets, as
 * bc_boundary does); other facets are not read.  Replaces DGOperator::rhs
 * (src/form/DGOperator.h:189-237) with Elasticity/Poisson rhs_volume, rhs_skeleton and
 * rhs_boundary (Elasticity.cpp:521-696, Poisson.cpp:383-540). */
int tdg_rhs(tdg_op const* op, double const* force_Q, double const* facet_q, double* b);

/* The same right-hand side on the GPU (elasticity, FP64; device pointers) for ncols columns:
 * column c reads force_Q + c ld_force and facet_q + c ld_facet and writes b + c ld_b (b is
 * overwritten).  Asynchronous on the operator's stream.  Replaces DGOperator::rhs as above. */
int tdg_rhs_device(tdg_op* op, double const* force_Q, double const* facet_q, double* b, int ncols,
                   int64_t ld_force, int64_t ld_facet, int64_t ld_b);

/* Facet traction t (device, facets x Q x nq, FP64) of the device vector x with facet data facet_q
 * (device, facets x Q x nq, as for tdg_rhs; may be NULL = zero): on skeleton facets
 * 1/2 (sigma_0 + sigma_1) n - penalty (u_0 - u_1 - f), on boundary facets sigma_0 n - penalty
 * (u_0 - f), with the unit normal of side 0.  Replaces Elasticity::traction_skeleton /
 * traction_boundary (Elasticity.cpp:949-1018; compute_traction*, elasticity.py:242-250). */
int tdg_traction(tdg_op* op, double const* x, double const* facet_q, double* t);

/* Linear solve A x = b on the GPU (b, x device FP64, x overwritten, initial guess 0); replaces
 * PetscLinearSolver::solve = KSPSolve with KSPCG on the shell matrix
 * (app/common/PetscLinearSolver.h:42, PetscLinearSolver.cpp:22-32); precond = TDG_PRECOND_BLOCK_JACOBI uses the
 * element-block Jacobi preconditioner (tandem's PetscDGMatrix + PCBJACOBI with one block per
 * element, PetscLinearSolver.cpp:14-15, PetscDGMatrix.cpp:14-77), built on first use;
 * TDG_PRECOND_BLOCK_JACOBI_F32 is the same preconditioner with FP32-stored inverse blocks applied in
 * FP64 arithmetic by FP64 solves (half the bytes per application; FP32 solves always use them).
 * mode 0: FP64 CG.  mode 1: FP32 CG inner solves (to inner_rtol) inside an FP64 iterative
 * refinement loop whose residual b - A x and update x += d are FP64 (restarted each refinement).
 * mode 2: FP32 CG whose residual is replaced by the FP64 true residual, and whose accumulated FP32
 * update is added to the FP64 x, each time the residual drops by reliable_delta (refinement
 * without restarting the Krylov direction).  rtol applies to the FP64 true residual. */
enum { TDG_SOLVE_FP64_CG = 0, TDG_SOLVE_FP32_CG_FP64_IR = 1, TDG_SOLVE_FP32_CG_FP64_RELIABLE = 2 };
enum { TDG_PRECOND_NONE = 0, TDG_PRECOND_BLOCK_JACOBI = 1, TDG_PRECOND_BLOCK_JACOBI_F32 = 2 };
typedef struct {
    double rtol;
    int max_iter;
    int mode;
    double inner_rtol;
    int max_outer;
    int check_every;
    double reliable_delta;
    int precond;
    int graphs; /* 1: replay the iterations between residual checks as a captured HIP graph (needs a
                 * stream set with tdg_set_stream and an even check_every; otherwise plain launches) */
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

/* ncols right-hand sides solved together (device FP64, column c at B + c ld, ld >= local size):
 * ncols CG recursions in lockstep, each iteration one batched apply over all columns (the
 * discrete Green's function's slip right-hand sides, solved one by one in
 * SeasQDDiscreteGreenOperator.cpp:540-549).  opts->mode TDG_SOLVE_FP64_CG or
 * TDG_SOLVE_FP32_CG_FP64_RELIABLE (per-column FP64 residual replacement); opts->precond as for
 * tdg_solve.  Per column, iterations[c] and rel_residual[c] (host arrays of ncols, may be NULL);
 * result->iterations = lockstep iterations, result->outer_iterations = replacements. */
int tdg_solve_multi(tdg_op* op, double const* B, double* X, int ncols, int64_t ld, tdg_solve_opts const* opts,
                    tdg_solve_result* result, int* iterations, double* rel_residual);

/* Independent meshes held in one operator (created from the disjoint union of their meshes):
 * segment c of the device vectors b, x is [offsets[c], offsets[c+1]) (host array of nseg+1 DoF
 * offsets, whole elements, offsets[0] = 0, offsets[nseg] = local size; no facet may join two
 * segments).  Each segment is its own CG recursion, all advanced in lockstep by one apply of the
 * whole operator per iteration (one launch over the element groups of every mesh).  The reference
 * solves one mesh per KSPSolve (app/common/PetscLinearSolver.h:42); no counterpart for several.
 * Modes, preconditioner and outputs as for tdg_solve_multi, per segment. */
int tdg_solve_segments(tdg_op* op, double const* b, double* x, int nseg, int64_t const* offsets,
                       tdg_solve_opts const* opts, tdg_solve_result* result, int* iterations, double* rel_residual);

/* Element-block Jacobi preconditioner (the inverse of each element's diagonal block of A, probed
 * with the matrix-free apply on an element colouring; assemble_volume/skeleton/boundary restricted
 * to the element, Elasticity.cpp:317-519, Poisson.cpp:225-381).  tdg_block_jacobi_build builds it
 * (also done by the first preconditioned solve) and returns the number of colours used;
 * tdg_block_jacobi_apply computes z = D^-1 r on device FP64 vectors (asynchronous on the operator's
 * stream), tdg_block_jacobi_apply_batched the same on ncols columns at stride ld (each block read
 * once for all columns); tdg_block_jacobi_inverse copies the FP64 inverse block of one element
 * (bs x bs, column-major) to the host. */
int tdg_block_jacobi_build(tdg_op* op, int* colours);
int tdg_block_jacobi_apply(tdg_op* op, double const* r, double* z);
int tdg_block_jacobi_apply_batched(tdg_op* op, double const* R, double* Z, int ncols, int64_t ld);
int tdg_block_jacobi_inverse(tdg_op* op, int64_t element, double* block);

/* Assembled operator: the block matrix tandem's PetscDGMatrix holds (BAIJ, one block row per
 * element with its own block and one block per face neighbour, app/common/PetscDGMatrix.cpp:14-77)
 * as DGOperator::assemble fills it (src/form/DGOperator.h:121-170, assemble_volume / skeleton /
 * boundary, Elasticity.cpp:317-519, Poisson.cpp:225-381).  Here the blocks are probed from the FP64
 * matrix-free apply on a distance-2 element colouring (no two elements within two face-steps share
 * a colour) and kept on the device in FP64 and FP32 until the material or penalty changes:
 * (faces + 2) x bs^2 values per element in each precision.  tdg_assemble builds it (once) and
 * returns the number of colours; tdg_assembled_block copies block `slot` of element `element`'s
 * block row (slot 0: the element's own block, 1 + f: the block of its face-f neighbour, zero
 * without one) to the host, bs x bs row-major.  Fails for meshes where an element meets one
 * neighbour through two faces. */
int tdg_assemble(tdg_op* op, int* colours);
int tdg_assembled_block(tdg_op* op, int64_t element, int slot, double* block);

/* p-multigrid over polynomial degree (tandem: PCMG set up by PetscLinearSolver::setup_mg,
 * app/common/PetscLinearSolver.cpp:53-114, with the levels of MGConfig::levels,
 * app/common/MGConfig.cpp:11-40, and the options of examples/options/mg_cheby.cfg).
 * levels[0] is the fine operator, levels[1..nlevels-1] operators of strictly decreasing degree
 * built by the caller on the same mesh with the same material (tandem forms the coarse matrices
 * as P^T A P; here each coarse level is its own matrix-free operator, and with galerkin_penalty it
 * takes the fine level's facet penalty, which makes it the fine form restricted to the coarse space
 * up to the quadrature of the material).  The coarse operators are switched to the fine operator's
 * stream.  Levels are joined by the nodal interpolation (tdg_interpolation_matrix) and its
 * transpose; each level below the coarsest is smoothed by `smooth` Chebyshev steps on
 * D^-1 A (D: element-block diagonal) over [eig_lo, eig_hi] x lambda_max (lambda_max from eig_iters
 * Lanczos steps); the coarsest level is solved by block-Jacobi CG to coarse_rtol (at most
 * coarse_max_iter iterations; in FP32 with coarse_fp32).  With vcycle_fp32 the V-cycle (smoothing,
 * transfers) runs in FP32 on the FP32 copies of the operators; the outer CG stays FP64.  With
 * coarse_assembled the coarsest operator is switched to TDG_VARIANT_ASSEMBLED at set-up where its
 * block size has the register kernel (P1 and P2 levels): the coarse CG then reads the assembled
 * block rows once per range of columns instead of the per-point geometry once per column.  With
 * coarse_amg as well (elasticity; FP32 or FP64 with coarse_fp32) the coarse CG is preconditioned by a smoothed-
 * aggregation AMG built on those block rows (the reference's coarse preconditioner is PETSc's GAMG,
 * examples/options/mg_cheby.cfg) instead of block Jacobi: rigid-body near-null space, Chebyshev
 * smoothing, Galerkin coarse operators down to a dense solve.
 * tdg_pmg_create keeps the level pointers (the operators must outlive the hierarchy, and a level
 * whose material or penalty changes after tdg_pmg_setup needs a new hierarchy: the set-up's
 * eigenvalue estimates and AMG levels are not rebuilt);
 * tdg_pmg_setup builds the block-Jacobi inverses and the eigenvalue estimates (lambda_max: host
 * array of nlevels-1, may be NULL); tdg_pmg_vcycle applies one V-cycle z = M^-1 r to ncols device
 * FP64 columns at stride ld; tdg_solve_multi_pmg solves ncols right-hand sides together by flexible
 * CG preconditioned with one V-cycle per iteration (FP64; opts->rtol and opts->max_iter are used,
 * the outputs as for tdg_solve_multi; result->outer_iterations = coarse CG iterations summed over
 * the V-cycles). */
typedef struct tdg_pmg tdg_pmg;
typedef struct {
    int smooth;
    double eig_lo, eig_hi;
    int eig_iters;
    double coarse_rtol;
    int coarse_max_iter;
    int galerkin_penalty;
    int coarse_fp32;
    int vcycle_fp32;
    int coarse_assembled;
    int coarse_amg;
} tdg_pmg_opts;
void tdg_pmg_default_opts(tdg_pmg_opts* opts);
int tdg_pmg_create(tdg_op* const* levels, int nlevels, tdg_pmg_opts const* opts, tdg_pmg** out);
int tdg_pmg_destroy(tdg_pmg* pmg);
int tdg_pmg_setup(tdg_pmg* pmg, double* lambda_max);
int tdg_pmg_vcycle(tdg_pmg* pmg, double const* r, double* z, int ncols, int64_t ld);
int tdg_solve_multi_pmg(tdg_pmg* pmg, double const* B, double* X, int ncols, int64_t ld, tdg_solve_opts const* opts,
                        tdg_solve_result* result, int* iterations, double* rel_residual);

/* Adjoint gradient kernel (elasticity): g_lam, g_mu (device, elements x Nq) = d/d(lam_Q, mu_Q) of
 * psi^T A(lam, mu) u for device vectors u and psi, through everything the coefficients enter: the
 * L2 projection to nodal coefficients and the volume weights (Elasticity.cpp:104-175,
 * prepare_volume_post_skeleton :227-253), the facet traces (:189-225) and the penalty
 * (:255-292, differentiated at its extremal nodes).  With A psi = dJ/du it gives
 * dJ/dm = -(g_lam, g_mu) for a misfit J(u(m)) whose right-hand side does not depend on m.
 * Strains are recomputed from u and psi.  Runs on the matrix cores wherever the MFMA apply exists
 * (variants 0 and 2 of tdg_set_variant), thread-per-point with variant 1.  No counterpart in the
 * reference (it has no adjoint); the FP64 form is checked against central differences of the
 * FP64 apply along seeded directions. */
int tdg_material_gradient(tdg_op* op, double const* u, double const* psi, double* g_lam, double* g_mu);
int tdg_material_gradient_f32(tdg_op* op, float const* u, float const* psi, float* g_lam, float* g_mu);

/* Fault adapter and the Green's-function columns (elasticity, FP64, device buffers).
 * tdg_fault_setup prepares the adapter on the Fault facets in facet order (AdapterBase::prepare,
 *   app/localoperator/AdapterBase.cpp:39-100, with Curvilinear::facetBasis,
 *   src/geometry/Curvilinear.cpp:281-316): up and ref_normal are the scenario's vectors (dim
 *   entries); sizes[0] = number of fault facets, sizes[1] = nbf of the fault space (nodal P_N on
 *   the (dim-1)-simplex).  The fault state of one column is nfault x nbf x (dim-1) (per facet
 *   (l, n) column-major), the fault traction nfault x nbf x dim (per facet (k, p) column-major, p = 0
 *   the normal component).
 * tdg_fault_slip: facet data F (facets x dim x nq per column, zero off the fault) of the slips S:
 *   Adapter<Elasticity>::slip (app/localoperator/ElasticityAdapter.cpp:37-49) for every fault facet.
 * tdg_fault_traction: T = fault-basis traction of the solutions X with facet data Fd (may be NULL):
 *   the facet traction (Elasticity::traction_skeleton/boundary, app/localoperator/Elasticity.cpp:
 *   949-1018) on the fault facets followed by Adapter<Elasticity>::traction (ElasticityAdapter.cpp:
 *   21-34); with X the solutions for unit fault-state columns, T holds the corresponding columns of
 *   the discrete Green's function (SeasQDDiscreteGreenOperator.cpp:536-545).
 * With ncols > 1 every stride must hold one column (ld_s >= nfault*nbf*(dim-1), ld_f >=
 * facets*dim*nq, ld_x >= local size, ld_t >= nfault*nbf*dim), else TDG_ERR_ARG; with ncols = 1
 * the strides are ignored. */
int tdg_fault_setup(tdg_op* op, double const* up, double const* ref_normal, int64_t* sizes);
int tdg_fault_slip(tdg_op* op, double const* S, double* F, int ncols, int64_t ld_s, int64_t ld_f);
int tdg_fault_traction(tdg_op* op, double const* X, double const* Fd, double* T, int ncols, int64_t ld_x,
                       int64_t ld_f, int64_t ld_t);

/* Elastic-wave pieces (elasticity, FP64, device vectors of local size).
 * tdg_inverse_mass: y = M^-1 x with the weight-adjusted inverse of the curvilinear element mass
 *   (unit density): Mhat^-1 E diag(-W/|J|) E^T Mhat^-1 per element and component; replaces the
 *   apply_inverse_mass kernel call of Elasticity::wave_rhs (app/localoperator/Elasticity.cpp:873-879,
 *   generated from app/kernels/elasticity.py:236-238).  The sign follows the reference's
 *   negative_rhoInv_W_Jinv_Q, so y = -M^-1 x.
 * tdg_wave_rhs: y = M^-1 (A x - b_facet(facet)) with the facet data of tdg_rhs (may be NULL);
 *   replaces Elasticity::wave_rhs (app/localoperator/Elasticity.cpp:865-880) over all elements. */
int tdg_inverse_mass(tdg_op* op, double const* x, double* y);
int tdg_wave_rhs(tdg_op* op, double const* x, double const* facet, double* y);

/* p-multigrid transfer between nodal spaces of one element (NodalInterpolation::assemble,
 * app/localoperator/NodalInterpolation.cpp:8-38).  tdg_interpolation_matrix writes the host matrix
 * I (n_to x n_from, row-major: I[k n_from + l] = phi^from_l(x^to_k)) for 1 <= from < to <= 8; the
 * reference's per-quantity block-diagonal matrix is I repeated over the quantities.
 * tdg_interpolate applies it on the device to num_elements blocks of num_quantities components:
 * y = I x (prolongation) or, with transpose != 0, y = I^T x (restriction); I is a device copy,
 * stream a hipStream_t (NULL: default stream). */
int tdg_interpolation_matrix(int dim, int to_degree, int from_degree, double* out);
int tdg_interpolate(double const* I, int n_to, int n_from, int num_quantities, int64_t num_elements, int transpose,
                    double const* x, double* y, void* stream);

/* Algorithmic bytes (for the given bytes per real) and flops of one apply. */
int tdg_apply_counts(tdg_op const* op, int precision_bytes, double* bytes, double* flops);

/* Named FP64 precompute arrays (host copy), for inspection: absDetJ, JInv, vcoords, volume,
 * JInv0, JInv1, normal, normalLength, fcoords, area, penalty, mat0, mat1, c0Q, c1Q, fc0, fc1,
 * E_Q (basis at the volume points, k + Nbf q), Dxi_Q (reference gradients at the volume points,
 * k + Nbf (e + D q)), E_q0 ... E_q3 (basis at the facet points of local face f, k + Nbf q),
 * volWeights, facetWeights.
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
