// Launch interface of the HIP kernels of the DG operator.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tdg {

// Device views of one precision of the operator data (see host/operator.h for the layouts).
template <class T> struct ApplyArgs {
    T const* tabV;   // [l][e][q]   volume gradient table
    T const* tabVt;  // [e][q][k]   volume lift table
    T const* tabF;   // [f][l][c][q] facet trace table (c = 0 value, 1 + e gradient)
    T const* tabFt;  // [f][c][q][k] facet lift table
    T const* w;      // facet weights
    int32_t const* perm;  // [f][j]  facet node relabelling (elasticity)
    int32_t const* iperm; // [f][k]
    T const* G;      // [E][Nq][D*D] JInv per volume point
    T const* c0;     // [E][Nq] lamWJ (elasticity) / JWK (poisson)
    T const* c1;     // [E][Nq] muWJ
    T const* fG;     // [F][2][nq][D*D] per side facet geometry
    T const* fc0;    // [F][2][nq] lambda traces / unused
    T const* fc1;    // [F][2][nq] mu traces / unused
    T const* fn;     // [F][nq][D] area-scaled normal of side 0
    T const* fpen;   // [F] penalty
    int4 const* ef;  // [E][D+1] facet, neighbour, flags
    int64_t num_elements;
};

// y_c = A x_c for ncols columns (stride ld).  Returns hipErrorInvalidValue for an unsupported
// (dim, degree) pair.
template <class T>
hipError_t launch_elasticity(int dim, int degree, ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld,
                             hipStream_t stream);
template <class T>
hipError_t launch_poisson(int dim, int degree, ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld,
                          hipStream_t stream);

// Extra device data of the material-gradient (adjoint) kernel, one precision.
template <class T> struct GradArgs {
    T const* tabE;      // [k][q] basis values at the volume points (= the nodal material basis)
    T const* tabEt;     // [q][k] the same table transposed (matrix-core operand order)
    T const* jw;        // [E][Nq] W |det J|
    T const* Minv;      // [E][Nbf][Nbf] inverse material mass matrix (column-major)
    T const* pc;        // [E][2] c0 = min 2 mu_i, c1 = max (D lam_i + 2 mu_i) of the nodal coefficients
    int4 const* pinfo;  // [E] {argmin of 2 mu_i, argmax of D lam_i + 2 mu_i, -, -}
    T const* pcoef;     // [E][D+1] (D+1) c_{N-1} |F| / |K| of each face
};

// g_lam, g_mu (E x Nq) = d/d(lam_Q, mu_Q) of psi^T A u (elasticity).
template <class T>
hipError_t launch_material_gradient(int dim, int degree, ApplyArgs<T> const& a, GradArgs<T> const& g, T const* u,
                                    T const* psi, T* g_lam, T* g_mu, hipStream_t stream);

// Weight-adjusted inverse mass (FP64): reference inverse mass, basis at the volume points and the
// per-point weights -W/|J|.
struct WaveArgs {
    double const* Minv; // [Nbf][Nbf] column-major reference inverse mass
    double const* E;    // [k][q] basis at the volume points
    double const* w;    // [E][Nq] -W / |det J|
    int nbf, nq, nquant;
    int64_t num_elements;
};
hipError_t launch_inverse_mass(WaveArgs const& a, double const* x, double* y, hipStream_t stream);
// y = I x per element and quantity (I: n_to x n_from row-major), or y = I^T x if transpose.
hipError_t launch_interpolate(double const* I, int n_to, int n_from, int nquant, int64_t nelem, int transpose,
                              double const* x, double* y, hipStream_t stream);
hipError_t launch_axpy(double* y, double alpha, double const* x, int64_t n, hipStream_t stream);

// Assembled block rows (one per element): A[slot][row][col][element] with slot 0 the element's own
// block and slot 1 + f the block of its face-f neighbour nb[f][element] (the element itself and a zero
// block where face f has none); block_size x block_size blocks, slots = faces + 1.
template <class T> struct AssembledArgs {
    T const* A;
    int32_t const* nb;
    int64_t num_elements;
    int block_size, slots;
};
// y_c = A x_c for ncols columns at stride ld (the register kernel for the block sizes of
// assembled_register_kernel, the generic kernel otherwise).  With partial: also the per-workgroup
// partial sums of x_c . y_c in partial[c kDotParts + i], i < kDotParts, where the register kernel
// holds one element per lane (block sizes 3, 4, 6 in FP32, 3 in FP64); hipErrorNotSupported (and
// nothing launched) otherwise.
constexpr int kDotParts = 256;
template <class T>
hipError_t launch_assembled(AssembledArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t stream,
                            double* partial = nullptr);
bool assembled_register_kernel(int block_size, int slots);
// Assembly by probing (FP64): X (m columns of E bs) = unit vectors c0 .. c0 + m - 1 on the elements of
// colour c; gather stores the isolated block columns of Y = A X into the assembled layout.
hipError_t launch_assembly_probe(double* X, int32_t const* colour, int c, int c0, int m, int64_t E, int bs,
                                 hipStream_t stream);
hipError_t launch_assembly_gather(double const* Y, int32_t const* colour, int32_t const* nb, double* A, int c, int c0,
                                  int m, int64_t E, int bs, int faces, hipStream_t stream);
hipError_t launch_narrow(double const* a, float* b, int64_t n, hipStream_t stream);

// Volume / facet rule sizes the kernels are compiled for (2N+1 rules of tandem's autorule).
int kernel_rule_sizes(int dim, int degree, int* Nq, int* nq);

// Device views for the right-hand side and facet traction kernels (FP64).
struct RhsArgs {
    double const* tabE;   // [k][q]  basis values at the volume points
    double const* jw;     // [E][Nq] |det J| W
    double const* tabF;   // [l][c][q] face-0 facet table (elasticity)
    double const* w;      // facet weights
    int32_t const* perm;  // [f][l]  facet node relabelling
    int32_t const* iperm; // [f][k]
    double const* fG;     // [F][2][nq][D*D] per side facet geometry
    double const* fc0;    // [F][2][nq] lambda traces
    double const* fc1;    // [F][2][nq] mu traces
    double const* fn;     // [F][nq][D] area-scaled normal of side 0
    double const* fpen;   // [F] penalty
    int4 const* ef;       // [E][D+1] facet, neighbour, flags
    int32_t const* fbc;   // [F] BC | skeleton << 8
    int4 const* fct;      // [F] element 0, element 1, local face 0 | local face 1 << 4
    int64_t num_elements, num_facets;
};

// b(:, c) = rhs of force(:, c) (E x Q x Nq, may be null) and facet data(:, c) (F x Q x nq, may be
// null) for ncols columns.
hipError_t launch_elasticity_rhs(int dim, int degree, RhsArgs const& a, double const* force, double const* fdata,
                                 double* b, int ncols, int64_t ld_force, int64_t ld_fdata, int64_t ld_b,
                                 hipStream_t stream);
// t (F x Q x nq) = facet traction of x with facet data fdata (may be null).
hipError_t launch_elasticity_traction(int dim, int degree, RhsArgs const& a, double const* x, double const* fdata,
                                      double* t, hipStream_t stream);
// Traction on the facets sub[0..nsub) only (rows of t in list order), for ncols columns of x, facet
// data and t at strides ld_x, ld_f, ld_t.
hipError_t launch_elasticity_traction_subset(int dim, int degree, RhsArgs const& a, double const* x,
                                             double const* fdata, double* t, int32_t const* sub, int64_t nsub,
                                             int ncols, int64_t ld_x, int64_t ld_f, int64_t ld_t, hipStream_t s);

// Fault adapter (one precision, FP64): per fault facet m (facet fac[m]) the fault-space basis at the
// facet points e_q[q][l], weights w, normal length nl, fault basis vectors bas[m][q][i][j] (component
// i of vector j) and the inverse weighted facet mass minv[m][k][l].
struct FaultArgs {
    int32_t const* fac;
    int32_t const* fmap; // facet -> fault facet or -1
    double const* e_q;
    double const* w;
    double const* nl;
    double const* bas;
    double const* minv;
    int nbf, nq, dim;
    int64_t nfault, num_facets;
};
// T[c] (nfault x nbf x D, (k, p) column-major per facet) = projection of the facet-point traction
// tq[c] (nfault x D x nq) onto the fault basis; ncols columns at strides ld_q, ld_t.
hipError_t launch_fault_project(FaultArgs const& a, double const* tq, double* T, int ncols, int64_t ld_q,
                                int64_t ld_t, hipStream_t s);
// Facet data F[c] (num_facets x D x nq; zero off the fault) of the fault-basis slip S[c]
// (nfault x nbf x (D-1), (l, n) column-major per facet).
hipError_t launch_fault_slip(FaultArgs const& a, double const* S, double* F, int ncols, int64_t ld_s,
                             int64_t ld_f, hipStream_t s);

} // namespace tdg
