// Host side of the MI355X DG operator: FP64 precompute (geometry, material projection, penalty) in
// the layouts the HIP kernels stream, device residency and kernel dispatch.
#pragma once

#include "mesh.h"
#include "refelem.h"
#include "../kernels/kernels.h"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace tdg {

enum OpKind : int { Kind_Elasticity = 0, Kind_Poisson = 1 };

// Face kinds of the element-centric apply (Elasticity.cpp:721-748, Poisson.cpp:573-636).
enum FaceKind : int { Face_Skip = 0, Face_Skeleton = 1, Face_Boundary = 2 };

template <class T> struct DeviceArray {
    T* ptr = nullptr;
    std::size_t n = 0;
    DeviceArray() = default;
    DeviceArray(DeviceArray const&) = delete;
    DeviceArray& operator=(DeviceArray const&) = delete;
    ~DeviceArray() { reset(); }
    void reset();
    void upload(std::vector<T> const& h);
    void alloc(std::size_t count);
};

// Everything one apply reads, in one precision.
template <class T> struct DeviceOperator {
    DeviceArray<T> tabV, tabVt, tabF, tabFt; // reference tables
    DeviceArray<T> W, w;                       // volume / facet weights
    DeviceArray<int32_t> perm, iperm;          // nodal facet permutations (elasticity)
    DeviceArray<T> G, c0, c1;                  // per element & volume point: JInv, lamWJ/muWJ (or JWK)
    DeviceArray<T> fG, fc0, fc1, fn, fpen;     // per facet
    DeviceArray<int32_t> ef;                   // per element face: facet, neighbour, flags
};

template <class T> struct MfmaArgs;
template <class T> struct SplitArgs;

// Operand tables and element-grouped point data of the MFMA elasticity kernel (one precision).
template <class T> struct DeviceMfma {
    DeviceArray<T> V, T0, w;       // canonical reference tables, facet weights
    DeviceArray<T> VD, FD, FP;     // volume point data, two-sided facet records, penalties
    DeviceArray<int32_t> perm2, FI; // perm_f followed by iperm_f; face info
};

// Facet records, penalties and facet info of the two-kernel (facet + element) MFMA apply, and the
// buffer of lifted facet contributions per element face (cap columns; cleared when allocated, so the
// slots of faces without a facet term stay zero).
template <class T> struct DeviceSplit {
    DeviceArray<T> FR, FPen;
    DeviceArray<int32_t> FInfo; // int4 per facet
    DeviceArray<int32_t> idle;  // slots (element * (D+1) + face) of faces without a facet term
    DeviceArray<T> L;
    int64_t nfacets = 0, ntwo = 0;
    int cap = 0;
};

// Extra data of the material-gradient kernel (one precision).
template <class T> struct DeviceGrad {
    DeviceArray<T> tabE, tabEt, jw, Minv, pc, pcoef;
    DeviceArray<int32_t> pinfo; // int4 per element
};

// Extra FP64 data of the right-hand side and traction kernels (elasticity).
struct DeviceRhs {
    DeviceArray<double> tabE, jw;
    DeviceArray<int32_t> fbc, fct; // fct: int4 per facet
};

// Data of the elastic-wave kernels: reference inverse mass, E table, -W/|J| per volume point and
// two work vectors.
struct DeviceWave {
    DeviceArray<double> Minv, E, w, ax, b;
};

// Fault adapter data (FP64) and a work buffer for the facet-point traction of the fault facets.
struct DeviceFault {
    DeviceArray<int32_t> fac, fmap;
    DeviceArray<double> e_q, w, nl, bas, minv, work;
    int64_t nfault = 0;
    int nbf = 0;
};

// Assembled block rows of the operator (kernels/dg_assembled.hip): FP64 blocks probed from the
// matrix-free apply, an FP32 copy, the face neighbours and the number of probe colours.
struct DeviceAssembled {
    DeviceArray<double> A64;
    DeviceArray<float> A32;
    DeviceArray<int32_t> nb;
    int colours = 0;
};

// Variant_Auto uses the two-kernel (facet + element) MFMA apply where one is compiled and the
// thread-per-point kernel otherwise; Variant_Points forces the thread-per-point kernel;
// Variant_MfmaXcd / Variant_MfmaElement run the element-centric MFMA kernel (every element
// evaluates both sides of its faces) with workgroups renumbered per XCD / in linear order;
// Variant_Assembled multiplies by the assembled block rows (assembled on first use).
enum Variant : int {
    Variant_Auto = 0,
    Variant_Points = 1,
    Variant_MfmaXcd = 2,
    Variant_MfmaElement = 3,
    Variant_Assembled = 4
};
constexpr int kNumVariants = 5;

class BlockJacobi;

class Operator {
public:
    Operator(SimplexMesh mesh, int kind, unsigned degree, std::vector<double> const* geometry_nodes);
    ~Operator();

    int kind() const { return kind_; }
    unsigned dim() const { return D_; }
    unsigned degree() const { return N_; }
    std::size_t num_elements() const { return E_; }
    std::size_t num_facets() const { return topo_.facets.size(); }
    unsigned nbf() const { return Nbf_; }
    unsigned nq_vol() const { return Nq_; }
    unsigned nq_fct() const { return nq_; }
    std::size_t block_size() const { return Nbf_ * Q_; }
    std::size_t local_size() const { return block_size() * E_; }
    bool ready() const { return ready_; }

    std::vector<double> const& volume_points() const { return vcoords_; }
    std::vector<double> const& facet_points() const { return fcoords_; }
    Topology const& topology() const { return topo_; }

    // Coefficients at the volume quadrature points: (lambda, mu) or (K, unused).  Projects them
    // onto the nodal P_N material space, builds all per-point data and uploads it.
    void set_material(double const* c0_Q, double const* c1_Q);

    // SIPG penalty per facet; set_penalty replaces it (same facet numbering, after set_material) and
    // drops the device copies (p-multigrid coarse levels take the fine level's penalty).
    std::vector<double> const& penalty() const { return penalty_; }
    void set_penalty(std::vector<double> const& pen);
    unsigned quantities() const { return Q_; }

    // Uploads both precisions to the device (done implicitly by the first apply).
    void upload_device();
    template <class T> DeviceOperator<T> const& device();

    void set_stream(hipStream_t s) { stream_ = s; }
    hipStream_t stream() const { return stream_; }
    void set_variant(int v);
    int variant() const { return variant_; }
    // Times every kernel variant that applies to this operator on ncols columns in the given
    // precision (8 or 4 bytes), keeps the fastest and returns the milliseconds per apply indexed by
    // variant (negative where a variant does not apply).
    std::vector<double> autotune(int ncols, int precision_bytes);
    // Launch shapes of the two-kernel apply: times shape s with workgroup order x (0 linear, 1 one
    // contiguous range per XCD) at index 2 s + x, keeps the fastest for that precision.
    std::vector<double> autotune_shapes(int ncols, int precision_bytes);
    void set_split_shape(int precision_bytes, int shape, int xcd);
    int split_shape(int precision_bytes) const { return split_shape_[precision_bytes == 8 ? 0 : 1]; }
    int split_xcd(int precision_bytes) const { return split_xcd_[precision_bytes == 8 ? 0 : 1]; }
    bool uses_mfma() const;
    bool uses_split() const;
    // Allocates the facet-slot buffers of both precisions for ncols columns (done by the apply when
    // needed; solvers call it before capturing applies into a graph).
    void reserve_columns(int ncols);

    // y = A x on device buffers (x, y of local_size() entries).
    void apply(double const* x, double* y);
    // The same FP64 apply with the facet slots in caller-provided device memory of cap_doubles
    // entries; allocates nothing (the operator must have been uploaded before).  Clears the slots of
    // faces without a facet term, so any previous contents of the memory are harmless.
    void apply_noalloc(double const* x, double* y, double* scratch, std::size_t cap_doubles);
    // Entries of the facet-slot buffer of one column (0 when the apply uses no slots).
    std::size_t slot_entries() const;
    // y = A x and, where the two-kernel apply runs, per element workgroup the partial sums of x . y
    // in partials[0 .. return value) (fixed order, deterministic); returns 0 (and only applies) for
    // the other kernels.  partials needs dot_partials_capacity() entries.
    int64_t apply_dot(double const* x, double* y, double* partials);
    int64_t apply_dot(float const* x, float* y, double* partials);
    int64_t dot_partials_capacity() const;
    void apply(float const* x, float* y);
    // multi-vector apply: columns of length local_size() with leading dimension ld
    void apply_batched(double const* X, double* Y, int ncols, int64_t ld);
    void apply_batched(float const* X, float* Y, int ncols, int64_t ld);
    // The same with the per-column partial sums of x . y (partial[c kDotParts + i], i < kDotParts)
    // from the same kernel, where the assembled register kernel runs; false (nothing done) otherwise.
    bool apply_batched_dot(double const* X, double* Y, int ncols, int64_t ld, double* partial);
    bool apply_batched_dot(float const* X, float* Y, int ncols, int64_t ld, double* partial);

    // Right-hand side, FP64, from the force at the volume points (E x Q x Nq) and facet data at the
    // facet points (F x Q x nq): Dirichlet values on Dirichlet facets, slip on Fault facets (half
    // of it on boundary Fault facets).  Host version and device version (device buffers, ncols
    // columns with leading dimensions ld_*; force or facet data may be null).
    void rhs(double const* force_Q, double const* facet_q, double* b_host) const;
    void rhs_device(double const* force_Q, double const* facet_q, double* b, int ncols, int64_t ld_force,
                    int64_t ld_facet, int64_t ld_b);
    // Facet traction (F x Q x nq) of the device vector x with facet data (may be null), FP64.
    void traction_device(double const* x, double const* facet_q, double* t);

    // Element-block Jacobi preconditioner (built on first use; rebuilt after set_material).
    BlockJacobi& block_jacobi();

    // Assembled form (tandem's PetscDGMatrix filled by DGOperator::assemble): every block row
    // (element) holds its own block and one block per face neighbour, probed from the matrix-free
    // FP64 apply on a distance-2 colouring of the elements; kept on the device in FP64 and FP32 until
    // set_material / set_penalty.  Returns the number of colours (built once).
    int assemble();
    bool assembled() const { return bool(asm_); }
    // Block `slot` (0: own, 1 + f: face-f neighbour) of element e's block row, row-major, to host.
    void assembled_block(int64_t e, int slot, double* out) const;
    // Whether the assembled apply has its register kernel for this block size (P1 / P2 levels).
    bool assembled_fast() const;
    // Host copies of the FP64 block rows (A[slot][row][col][element]) and face neighbours
    // (nb[face][element]); assembles first if needed.
    void assembled_host(std::vector<double>& A, std::vector<int32_t>& nb);
    SimplexMesh const& mesh() const { return mesh_; }

    // g_lam, g_mu (E x Nq, device) = d/d(lam_Q, mu_Q) of psi^T A u for device vectors u, psi
    // (elasticity): the adjoint gradient kernel.
    void material_gradient(double const* u, double const* psi, double* g_lam, double* g_mu);
    void material_gradient(float const* u, float const* psi, float* g_lam, float* g_mu);

    // Elastic-wave operator pieces on device vectors (elasticity, FP64): y = M^-1 x with the
    // weight-adjusted inverse mass, and y = M^-1 (A x - b_facet(f)) (facet data may be null).
    void inverse_mass_device(double const* x, double* y);
    void wave_rhs_device(double const* x, double const* facet_q, double* y);
    // Nodal interpolation matrix from degree `from` to degree `to`, row-major (to x from).
    static std::vector<double> interpolation_matrix(unsigned D, unsigned to, unsigned from);

    // Fault adapter (AdapterBase.cpp:20-100): set up on the Fault facets (facet order) with the up
    // vector and reference normal of the scenario; returns {fault facets, nbf of the fault space}.
    std::array<int64_t, 2> fault_setup(double const* up, double const* ref_normal);
    // Facet data (F x Q x nq per column, zero off the fault) of fault-basis slips S (nfault x nbf x
    // (D-1) per column), and the fault-basis traction T (nfault x nbf x D per column) of solutions X
    // with facet data Fd (may be null): ncols columns at the given strides, device buffers.
    void fault_slip_device(double const* S, double* F, int ncols, int64_t ld_s, int64_t ld_f);
    void fault_traction_device(double const* X, double const* Fd, double* T, int ncols, int64_t ld_x, int64_t ld_f,
                               int64_t ld_t);

    // named host arrays for inspection
    std::vector<double> const* host_array(std::string const& name) const;

    // algorithmic counts of one apply (bytes streamed, flops)
    double apply_bytes(int precision_bytes) const;
    double apply_flops() const;

private:
    void drop_device();
    void build_tables();
    void build_geometry(std::vector<double> const* geometry_nodes);
    template <class T> void upload(DeviceOperator<T>& d);
    template <class T> void upload_mfma(DeviceMfma<T>& d);
    template <class T> void upload_mfma_faces(DeviceMfma<T>& d);
    template <class T> void upload_split(DeviceSplit<T>& d);
    template <class T> SplitArgs<T> split_args(DeviceMfma<T> const& m, DeviceSplit<T> const& s) const;
    template <class T> void reserve_split(DeviceSplit<T>& s, int ncols);
    template <class T> int64_t apply_dot_impl(T const* x, T* y, double* partials);
    void ensure_mfma_faces();
    template <class T> void launch(DeviceOperator<T> const& d, T const* x, T* y, int ncols, int64_t ld);
    template <class T> MfmaArgs<T> mfma_args(DeviceMfma<T> const& d) const;
    template <class T> void launch_mfma(DeviceMfma<T> const& d, T const* x, T* y, int ncols, int64_t ld);
    template <class T> void dispatch(T const* x, T* y, int ncols, int64_t ld);
    template <class T> AssembledArgs<T> assembled_args() const;
    template <class T> bool apply_batched_dot_impl(T const* X, T* Y, int ncols, int64_t ld, double* partial);

    int kind_;
    unsigned D_, N_, Q_, Nbf_, Nq_, nq_;
    std::size_t E_;
    SimplexMesh mesh_;
    Topology topo_;
    RefSimplex ref_;
    QuadRule vol_, fct_;
    bool ready_ = false;
    hipStream_t stream_ = nullptr;

    // reference tables (FP64)
    std::vector<double> E_Q_, Dxi_Q_;                       // solution basis at volume points
    std::vector<std::vector<double>> E_q_, Dxi_q_;          // solution basis at facet points, per face
    std::vector<double> matE_Q_;                            // material basis at volume points
    std::vector<std::vector<double>> matE_q_;               // material basis at facet points, per face
    std::vector<int32_t> perm_, iperm_;                     // facet node permutations (nodal bases)
    std::vector<std::array<double, 9>> faceJinv_;           // Jacobian inverse of facet maps

    // geometry (FP64)
    std::vector<double> absDetJ_, JInv_, vcoords_, volume_;
    std::vector<double> fJInv_[2], normal_, nlen_, fcoords_, area_;
    std::vector<double> fM_[2];                             // facet Jacobians J J_f per side

    // material-dependent (FP64)
    std::vector<double> mat0_, mat1_;                       // projected nodal coefficients
    std::vector<double> c0Q_, c1Q_;                         // lamWJ, muWJ (or JWK)
    std::vector<double> fc0_, fc1_;                         // facet coefficient traces (2 sides)
    std::vector<double> fGp_;                               // facet per-side geometry for kernels
    std::vector<double> penalty_;
    std::vector<int32_t> ef_;

    std::unique_ptr<DeviceOperator<double>> d64_;
    std::unique_ptr<DeviceOperator<float>> d32_;
    std::unique_ptr<DeviceMfma<double>> m64_;
    std::unique_ptr<DeviceMfma<float>> m32_;
    std::unique_ptr<DeviceSplit<double>> s64_;
    std::unique_ptr<DeviceSplit<float>> s32_;
    std::unique_ptr<DeviceRhs> rhs64_;
    std::unique_ptr<BlockJacobi> bj_;
    std::unique_ptr<DeviceAssembled> asm_;
    std::unique_ptr<DeviceGrad<double>> g64_;
    std::unique_ptr<DeviceGrad<float>> g32_;
    std::unique_ptr<DeviceWave> wave_;
    std::unique_ptr<DeviceFault> fault_;
    FaultArgs fault_args() const;
    WaveArgs wave_args();
    template <class T> void upload_grad(DeviceGrad<T>& g);
    template <class T> void grad_dispatch(T const* u, T const* psi, T* gl, T* gm);
    RhsArgs rhs_args();
    int variant_ = Variant_Auto;
    // launch shape of the two-kernel apply, FP64 / FP32 (defaults: the fastest measured on the
    // 24^3-cube P4 workload; the autotuner re-chooses per operator)
    int split_shape_[2] = {0, 1};
    int split_xcd_[2] = {2, 1};
    double time_configs(int ncols, int pb, int count, std::function<bool(int)> const& select, std::vector<double>& ms);
};

} // namespace tdg
