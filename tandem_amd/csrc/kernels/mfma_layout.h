// Tile geometry and device arguments of the MFMA elasticity apply (shared by host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tdg {

// Runtime tile sizes of the MFMA layout for (D, N) with Nbf basis functions, Nq volume and nq facet
// points:
//   KS = ceil(Nbf / 4)   K-steps over the basis,         MT = ceil(Nbf / 16) basis tiles,
//   QB = ceil(Nq / 16)   16-point volume blocks,          FT = ceil(nq / 4)   4-point facet tiles,
//   NCH = D*D + 2        volume channels (JInv, lamWJ, muWJ),
//   NFC = 2 D*D + 4 + D  facet channels (own G', lam, mu, neighbour G', lam, mu, oriented normal).
struct MfmaDims {
    int D, Nbf, Nq, nq, KS, QB, MT, FT, NCH, NFC;
    __host__ __device__ constexpr MfmaDims(int D_, int Nbf_, int Nq_, int nq_)
        : D(D_), Nbf(Nbf_), Nq(Nq_), nq(nq_), KS((Nbf_ + 3) / 4), QB((Nq_ + 15) / 16), MT((Nbf_ + 15) / 16),
          FT((nq_ + 3) / 4), NCH(D_ * D_ + 2), NFC(2 * D_ * D_ + 4 + D_) {}
};

namespace detail {
constexpr int binom(int n, int k) { return k == 0 ? 1 : binom(n - 1, k - 1) * n / k; }
constexpr int vol_points(int D, int N) {
    return D == 2 ? (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : N == 4 ? 19 : N == 5 ? 28 : 37)
                  : (N == 1 ? 6 : N == 2 ? 14 : N == 3 ? 31 : 57);
}
constexpr int fct_points(int D, int N) { return D == 2 ? N + 1 : (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : 19); }
} // namespace detail

template <int D_, int N_> struct MfmaLayout {
    static constexpr int Nbf = detail::binom(N_ + D_, D_);
    static constexpr int Nq = detail::vol_points(D_, N_);
    static constexpr int nq = detail::fct_points(D_, N_);
    static constexpr MfmaDims dims{D_, Nbf, Nq, nq};
    static constexpr int KS = dims.KS, QB = dims.QB, MT = dims.MT, FT = dims.FT, NCH = dims.NCH, NFC = dims.NFC;
};

// Output row of a 16x16x4 MFMA tile held in register i (0..3) of lane group g = lane / 16:
// consecutive rows for the f32 instruction, rows g, g+4, g+8, g+12 for the f64 instruction.
template <class T> __host__ __device__ constexpr int mfma_row(int g, int i) {
    return sizeof(T) == 8 ? g + 4 * i : 4 * g + i;
}
// Inverse: (lane group, register) of tile row rho.
template <class T> __host__ __device__ constexpr int mfma_group(int rho) { return sizeof(T) == 8 ? rho % 4 : rho / 4; }
template <class T> __host__ __device__ constexpr int mfma_reg(int rho) { return sizeof(T) == 8 ? rho / 4 : rho % 4; }

bool mfma_supported(int dim, int degree);

// Arguments of the MFMA kernel: canonical tables (copied to LDS by every block) and the
// element-grouped point data.
template <class T> struct MfmaArgs {
    T const* V;          // [e][q][k]  d phi_k / d xi_e at volume point q
    T const* T0;         // [c][q][k]  face-0 value (c = 0) and d/dxi_(c-1) at facet point q
    int32_t const* perm; // [2][f][k]  facet relabelling perm_f, then its inverse
    T const* w;          // facet weights
    T const* VD;         // [group][q][NCH][16]
    T const* FD;         // [group][f][q][NFC][16]  both sides' G', lam, mu and the oriented normal
                         // (stored per element side: every load is a contiguous 16-element row)
    T const* FP;         // [group][f][16]
    int4 const* FI;      // [group][f][16]
    int64_t num_elements, num_groups;
    int xcd_remap;       // 1: each XCD takes one contiguous range of element groups
};

template <class T>
hipError_t launch_elasticity_mfma(int dim, int degree, MfmaArgs<T> const& a, T const* x, T* y, int ncols,
                                 int64_t ld, hipStream_t stream);

// Arguments of the two-kernel (facet + element) MFMA elasticity apply.  Facets are the ones that
// carry a term (two-sided first, then one-sided), in groups of 8 with the facet innermost:
//   FR    [fgroup][q][NFR][8]   NFR = D*D + D + 4: the tangent columns t_1 .. t_(D-1) of the facet
//                               Jacobian M = J J_f of side 0 (shared by both sides), the third
//                               column c_0 of side 0 and c_1 of side 1, lam_0, mu_0, lam_1, mu_1
//   FPen  [fgroup][8]           penalty
//   FInfo [fgroup][8]           {element 0, element 1, face 0 | face 1 << 4 | kind << 8, 0}
//                               (kind 1 two-sided, 2 one-sided, 0 padding)
// The element kernel reads the volume table V and point data VD of MfmaArgs, and the lifted facet
// contributions L[element][face][Nbf * D].
template <class T> struct SplitArgs {
    T const* V;
    T const* T0;
    int32_t const* perm; // [2][f][k] perm_f, then iperm_f
    T const* w;
    T const* VD;
    T const* FR;
    T const* FPen;
    int4 const* FInfo;
    int64_t num_elements, num_groups, num_facets, num_fgroups, num_two;
    int xcd_remap; // workgroup order: 0 linear; 1 one contiguous range of groups per XCD (one column);
                   // 2 as 1, persistent (the resident workgroups loop over the blocks)
    double* dot;   // one column, optional: per element workgroup the partial sum of x . y
};

// Number of launch shapes of the two-kernel apply (waves per workgroup, occupancy targets, element
// prefetch depth, lift tail on the vector ALU or the matrix cores), per precision.
constexpr int kSplitShapes = 4;
// Workgroup orders (SplitArgs::xcd_remap) searched with each shape.
constexpr int kSplitOrders = 3;

// y = A x for ncols columns (stride ld) through the facet kernel (into the slot buffer L, column
// stride ldL = E (D+1) Nbf D) and the element kernel, with launch shape `shape`.  With a.dot set
// (one column), the element kernel also writes per workgroup the partial sum of x . y into
// a.dot[0 .. split_element_workgroups(shape, num_groups)).
template <class T>
hipError_t launch_elasticity_split(int dim, int degree, int shape, SplitArgs<T> const& a, T const* x, T* y, T* L,
                                   int ncols, int64_t ld, int64_t ldL, hipStream_t stream);
// Workgroups of the element kernel of launch shape `shape` for one column (a multiple of 8).
template <class T> int64_t split_element_workgroups(int shape, int64_t num_groups);

// Zeroes the facet slots listed in `slots` (slot s = entries [s bs, (s + 1) bs) of L).
template <class T> hipError_t launch_zero_slots(T* L, int32_t const* slots, int64_t nslots, int bs, hipStream_t s);

template <class T> struct GradArgs;
// Adjoint material gradient on the matrix cores (same dimensions and degrees as the apply):
// g_lam, g_mu (E x Nq) = d/d(lam_Q, mu_Q) of psi^T A u.
template <class T>
hipError_t launch_material_gradient_mfma(int dim, int degree, MfmaArgs<T> const& a, GradArgs<T> const& g,
                                         T const* u, T const* psi, T* g_lam, T* g_mu, hipStream_t stream);

} // namespace tdg
