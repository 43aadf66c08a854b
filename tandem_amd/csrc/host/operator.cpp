// Host side of the MI355X DG operator: FP64 precompute and device residency.
//
// The precompute restates tandem's DGOperator constructor (src/form/DGOperator.h:54-105):
// geometry per element and facet (DGCurvilinearCommon.cpp:41-101, Curvilinear.cpp:195-275),
// L2 projection of the coefficients onto the nodal P_N space (Elasticity.cpp:104-164,
// Poisson.cpp:133-160), facet traces of the coefficients (Elasticity.cpp:189-225), the
// weighted volume coefficients (Elasticity.cpp:227-253) and the SIPG penalty
// (Elasticity.cpp:255-292, Poisson.cpp:209-223).  Everything here stays FP64; FP32 device copies
// are rounded from it.
#include "operator.h"

#include <functional>

#include "../kernels/kernels.h"
#include "../kernels/mfma_layout.h"
#include "../solver/block_jacobi.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <stdexcept>
#include <type_traits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace tdg {

#define TDG_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

template <class T> void DeviceArray<T>::reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
}
template <class T> void DeviceArray<T>::alloc(std::size_t count) {
    reset();
    if (count == 0) return;
    TDG_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
    n = count;
}
template <class T> void DeviceArray<T>::upload(std::vector<T> const& h) {
    alloc(h.size());
    if (!h.empty()) TDG_HIP(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
template struct DeviceArray<double>;
template struct DeviceArray<float>;
template struct DeviceArray<int32_t>;

static double trace_constant(unsigned D, unsigned N) { return (N + 1) * (N + D) / static_cast<double>(D); }

static double det(unsigned D, double const* J) {
    if (D == 2) return J[0] * J[3] - J[2] * J[1];
    return J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) + J[6] * (J[1] * J[5] - J[4] * J[2]);
}
static void inverse(unsigned D, double const* J, double* R) {
    std::vector<double> A(J, J + D * D), I(D * D, 0.0);
    for (unsigned i = 0; i < D; ++i) I[i + i * D] = 1.0;
    dense_solve(A, D, I.data(), D);
    std::memcpy(R, I.data(), D * D * sizeof(double));
}

Operator::Operator(SimplexMesh mesh, int kind, unsigned degree, std::vector<double> const* geometry_nodes)
    : kind_(kind), D_(mesh.dim), N_(degree), mesh_(std::move(mesh)), ref_(mesh_.dim) {
    if (D_ != 2 && D_ != 3) throw std::runtime_error("dimension must be 2 or 3");
    if (N_ < 1 || N_ > 6) throw std::runtime_error("polynomial degree must be in [1, 6]");
    if (kind_ != Kind_Elasticity && kind_ != Kind_Poisson) throw std::runtime_error("unknown operator kind");
    E_ = mesh_.num_elements();
    if (E_ == 0) throw std::runtime_error("empty mesh");
    int64_t nv = static_cast<int64_t>(mesh_.vertices.size() / D_);
    for (auto v : mesh_.elements)
        if (v < 0 || v >= nv) throw std::runtime_error("element references a vertex out of range");
    Q_ = kind_ == Kind_Elasticity ? D_ : 1;
    topo_ = build_topology(mesh_);
    build_tables();
    build_geometry(geometry_nodes);
}

Operator::~Operator() = default;

void Operator::build_tables() {
    unsigned order = 2 * N_ + 1; // MinQuadOrder() (app/config.h.in:8-11)
    vol_ = simplex_rule(D_, order);
    fct_ = simplex_rule(D_ - 1, order);
    Nq_ = vol_.size();
    nq_ = fct_.size();
    Basis sol(D_, N_, kind_ == Kind_Elasticity); // Poisson: modal solution space (Poisson.h:130)
    Basis mat(D_, N_, true);
    Nbf_ = sol.size();
    E_Q_ = sol.values(vol_.points);
    Dxi_Q_ = sol.grads(vol_.points);
    matE_Q_ = mat.values(vol_.points);
    E_q_.clear();
    Dxi_q_.clear();
    matE_q_.clear();
    faceJinv_.clear();
    for (unsigned f = 0; f <= D_; ++f) {
        auto pts = ref_.facet_points(f, fct_.points);
        E_q_.push_back(sol.values(pts));
        Dxi_q_.push_back(sol.grads(pts));
        matE_q_.push_back(mat.values(pts));
        faceJinv_.push_back(ref_.face_map_jacobian_inverse(f));
    }
    perm_.assign((D_ + 1) * Nbf_, 0);
    iperm_.assign((D_ + 1) * Nbf_, 0);
    if (sol.nodal()) {
        // The warp-and-blend node set is invariant under the vertex permutations of the simplex, so
        // the basis on face f is the face-0 basis with nodes relabelled: E_q[f](pi_f(j)) = E_q[0](j).
        auto const& nodes = sol.nodes();
        const double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (unsigned f = 0; f <= D_; ++f) {
            auto const& fv = ref_.faces[f];
            for (unsigned j = 0; j < Nbf_; ++j) {
                double y[3] = {0, 0, 0};
                for (unsigned i = 0; i < D_; ++i) {
                    y[i] = rv[fv[0]][i];
                    for (unsigned d = 0; d < D_; ++d) y[i] += (rv[fv[d + 1]][i] - rv[fv[0]][i]) * nodes[j * D_ + d];
                }
                int hit = -1;
                for (unsigned k = 0; k < Nbf_; ++k) {
                    double dist = 0.0;
                    for (unsigned i = 0; i < D_; ++i) dist = std::max(dist, std::fabs(nodes[k * D_ + i] - y[i]));
                    if (dist < 1e-10) hit = static_cast<int>(k);
                }
                if (hit < 0) throw std::runtime_error("node set is not symmetric under facet maps");
                perm_[f * Nbf_ + j] = hit;
                iperm_[f * Nbf_ + hit] = static_cast<int>(j);
            }
            // verify the relabelling reproduces the face tables
            for (unsigned q = 0; q < nq_; ++q)
                for (unsigned j = 0; j < Nbf_; ++j) {
                    double a = E_q_[f][perm_[f * Nbf_ + j] + Nbf_ * q], b = E_q_[0][j + Nbf_ * q];
                    if (std::fabs(a - b) > 1e-10) throw std::runtime_error("facet basis relabelling check failed");
                }
        }
    }
}

void Operator::build_geometry(std::vector<double> const* geometry_nodes) {
    unsigned D = D_, DD = D * D;
    Basis geo(D, N_, true); // geometry degree = solution degree (static.cpp:81)
    unsigned ng = geo.size();
    std::vector<double> X;
    if (geometry_nodes) {
        if (geometry_nodes->size() != E_ * ng * D) throw std::runtime_error("geometry_nodes has the wrong size");
        X = *geometry_nodes;
    } else {
        // affine map of the reference nodes with vertices in ascending order (Curvilinear.cpp:101-106)
        X.resize(E_ * ng * D);
        auto const& rn = geo.nodes();
#pragma omp parallel for schedule(static)
        for (std::size_t e = 0; e < E_; ++e) {
            int64_t const* el = &mesh_.elements[e * (D + 1)];
            for (unsigned p = 0; p < ng; ++p) {
                double s = 0.0;
                for (unsigned d = 0; d < D; ++d) s += rn[p * D + d];
                for (unsigned i = 0; i < D; ++i) {
                    double v = (1.0 - s) * mesh_.vertices[el[0] * D + i];
                    for (unsigned d = 0; d < D; ++d) v = v + rn[p * D + d] * mesh_.vertices[el[d + 1] * D + i];
                    X[(e * ng + p) * D + i] = v;
                }
            }
        }
    }
    auto gE_Q = geo.values(vol_.points);
    auto gD_Q = geo.grads(vol_.points);
    std::vector<std::vector<double>> gE_q, gD_q;
    for (unsigned f = 0; f <= D; ++f) {
        auto pts = ref_.facet_points(f, fct_.points);
        gE_q.push_back(geo.values(pts));
        gD_q.push_back(geo.grads(pts));
    }
    // J(d, j, p) = sum_k X_k[d] dphi_k/dxi_j(p)
    auto jac = [&](std::size_t e, std::vector<double> const& gD, unsigned np, double* J) {
        double const* Xe = &X[e * ng * D];
        for (unsigned p = 0; p < np; ++p)
            for (unsigned j = 0; j < D; ++j)
                for (unsigned d = 0; d < D; ++d) {
                    double s = 0.0;
                    for (unsigned k = 0; k < ng; ++k) s += Xe[k * D + d] * gD[k + ng * (j + D * p)];
                    J[d + D * (j + D * p)] = s;
                }
    };
    auto map = [&](std::size_t e, std::vector<double> const& gE, unsigned np, double* x) {
        double const* Xe = &X[e * ng * D];
        for (unsigned p = 0; p < np; ++p)
            for (unsigned d = 0; d < D; ++d) {
                double s = 0.0;
                for (unsigned k = 0; k < ng; ++k) s += Xe[k * D + d] * gE[k + ng * p];
                x[d + D * p] = s;
            }
    };
    absDetJ_.assign(E_ * Nq_, 0.0);
    JInv_.assign(E_ * Nq_ * DD, 0.0);
    vcoords_.assign(E_ * Nq_ * D, 0.0);
    volume_.assign(E_, 0.0);
    bool bad = false;
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e) {
        std::vector<double> J(Nq_ * DD);
        jac(e, gD_Q, Nq_, J.data());
        double vol = 0.0;
        for (unsigned q = 0; q < Nq_; ++q) {
            double dj = det(D, &J[q * DD]);
            if (!(std::fabs(dj) > 0.0)) bad = true;
            absDetJ_[e * Nq_ + q] = std::fabs(dj);
            inverse(D, &J[q * DD], &JInv_[(e * Nq_ + q) * DD]);
            vol += vol_.weights[q] * std::fabs(dj);
        }
        volume_[e] = vol;
        map(e, gE_Q, Nq_, &vcoords_[e * Nq_ * D]);
    }
    if (bad) throw std::runtime_error("degenerate element (zero Jacobian determinant)");
    std::size_t F = topo_.facets.size();
    for (int s = 0; s < 2; ++s) fJInv_[s].assign(F * nq_ * DD, 0.0);
    for (int s = 0; s < 2; ++s) fM_[s].assign(F * nq_ * DD, 0.0);
    std::vector<std::array<double, 9>> faceJ;
    for (unsigned f = 0; f <= D; ++f) faceJ.push_back(ref_.face_map_jacobian(f));
    // M = J J_f: the Jacobian of the element map composed with the reference map carrying facet 0
    // onto the local facet (columns 0 .. D-2 tangential to the facet, D-1 into the element)
    auto facet_jacobian = [&](double const* J, int lf, double* M) {
        auto const& Jf = faceJ[lf];
        for (unsigned c = 0; c < D; ++c)
            for (unsigned r = 0; r < D; ++r) {
                double v = 0.0;
                for (unsigned k = 0; k < D; ++k) v += J[r + D * k] * Jf[k + D * c];
                M[r + D * c] = v;
            }
    };
    normal_.assign(F * nq_ * D, 0.0);
    nlen_.assign(F * nq_, 0.0);
    fcoords_.assign(F * nq_ * D, 0.0);
    area_.assign(F, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t i = 0; i < F; ++i) {
        Facet const& fc = topo_.facets[i];
        std::vector<double> J(nq_ * DD);
        jac(fc.elem[0], gD_q[fc.local[0]], nq_, J.data());
        auto const& rn = ref_.normals[fc.local[0]];
        double a = 0.0;
        for (unsigned q = 0; q < nq_; ++q) {
            double dj = det(D, &J[q * DD]);
            double* Ji = &fJInv_[0][(i * nq_ + q) * DD];
            inverse(D, &J[q * DD], Ji);
            facet_jacobian(&J[q * DD], fc.local[0], &fM_[0][(i * nq_ + q) * DD]);
            // area-scaled normal n = |det J| J^{-T} N_ref (Curvilinear.cpp:256-267)
            double ss = 0.0;
            for (unsigned r = 0; r < D; ++r) {
                double v = 0.0;
                for (unsigned j = 0; j < D; ++j) v += Ji[j + r * D] * rn[j];
                v *= std::fabs(dj);
                normal_[(i * nq_ + q) * D + r] = v;
                ss += v * v;
            }
            nlen_[i * nq_ + q] = std::sqrt(ss);
            a += fct_.weights[q] * nlen_[i * nq_ + q];
        }
        area_[i] = a;
        map(fc.elem[0], gE_q[fc.local[0]], nq_, &fcoords_[i * nq_ * D]);
        jac(fc.elem[1], gD_q[fc.local[1]], nq_, J.data());
        for (unsigned q = 0; q < nq_; ++q) {
            inverse(D, &J[q * DD], &fJInv_[1][(i * nq_ + q) * DD]);
            facet_jacobian(&J[q * DD], fc.local[1], &fM_[1][(i * nq_ + q) * DD]);
        }
    }
}

void Operator::set_material(double const* c0_Q, double const* c1_Q) {
    unsigned D = D_, DD = D * D, Nm = static_cast<unsigned>(num_basis(D, N_));
    bool el = kind_ == Kind_Elasticity;
    if (!c0_Q || (el && !c1_Q)) throw std::runtime_error("missing coefficient array");
    std::size_t F = topo_.facets.size();
    // L2 projection onto the nodal material space: M c = b, M = matE^T diag(W|J|) matE
    mat0_.assign(E_ * Nm, 0.0);
    mat1_.assign(el ? E_ * Nm : 0, 0.0);
    bool bad = false;
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e) {
        std::vector<double> M(Nm * Nm, 0.0), b(Nm * 2, 0.0);
        for (unsigned q = 0; q < Nq_; ++q) {
            double wj = vol_.weights[q] * absDetJ_[e * Nq_ + q];
            double const* phi = &matE_Q_[Nm * q];
            for (unsigned l = 0; l < Nm; ++l) {
                double pl = phi[l] * wj;
                for (unsigned k = 0; k < Nm; ++k) M[k + l * Nm] += phi[k] * pl;
                b[l] += c0_Q[e * Nq_ + q] * pl;
                if (el) b[l + Nm] += c1_Q[e * Nq_ + q] * pl;
            }
        }
        try {
            dense_solve(M, Nm, b.data(), el ? 2 : 1);
        } catch (...) {
            bad = true;
        }
        std::memcpy(&mat0_[e * Nm], b.data(), Nm * sizeof(double));
        if (el) std::memcpy(&mat1_[e * Nm], b.data() + Nm, Nm * sizeof(double));
    }
    if (bad) throw std::runtime_error("material projection failed (singular mass matrix)");
    // volume coefficients: lamWJ, muWJ (precomputeVolume) or JWK (J_W_K_Q)
    c0Q_.assign(E_ * Nq_, 0.0);
    c1Q_.assign(el ? E_ * Nq_ : 0, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e)
        for (unsigned q = 0; q < Nq_; ++q) {
            double a = 0.0, b = 0.0;
            for (unsigned t = 0; t < Nm; ++t) {
                a += matE_Q_[t + Nm * q] * mat0_[e * Nm + t];
                if (el) b += matE_Q_[t + Nm * q] * mat1_[e * Nm + t];
            }
            double wj = vol_.weights[q] * absDetJ_[e * Nq_ + q];
            c0Q_[e * Nq_ + q] = el ? a * wj : wj * a;
            if (el) c1Q_[e * Nq_ + q] = b * wj;
        }
    // facet coefficient traces (both sides; a boundary facet mirrors side 0) and kernel geometry
    fc0_.assign(F * 2 * nq_, 0.0);
    fc1_.assign(F * 2 * nq_, 0.0);
    fGp_.assign(F * 2 * nq_ * DD, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t i = 0; i < F; ++i) {
        Facet const& fc = topo_.facets[i];
        bool skel = fc.elem[0] != fc.elem[1];
        for (int side = 0; side < 2; ++side) {
            int s = skel ? side : 0;
            int64_t e = fc.elem[s];
            int lf = fc.local[s];
            for (unsigned q = 0; q < nq_; ++q) {
                double a = 0.0, b = 0.0;
                for (unsigned t = 0; t < Nm; ++t) {
                    a += matE_q_[lf][t + Nm * q] * mat0_[e * Nm + t];
                    if (el) b += matE_q_[lf][t + Nm * q] * mat1_[e * Nm + t];
                }
                fc0_[(i * 2 + side) * nq_ + q] = a;
                fc1_[(i * 2 + side) * nq_ + q] = b;
                double const* Ji = &fJInv_[s][(i * nq_ + q) * DD];
                double* out = &fGp_[((i * 2 + side) * nq_ + q) * DD];
                if (el) {
                    // G' = Jf^{-1} JInv: gradients of the face-0-relabelled basis (see build_tables)
                    auto const& Jf = faceJinv_[lf];
                    for (unsigned ep = 0; ep < D; ++ep)
                        for (unsigned j = 0; j < D; ++j) {
                            double v = 0.0;
                            for (unsigned e2 = 0; e2 < D; ++e2) v += Jf[ep + e2 * D] * Ji[e2 + j * D];
                            out[ep + j * D] = v;
                        }
                } else {
                    // K_G_q = K_q JInv (poisson.py:121)
                    for (unsigned k = 0; k < DD; ++k) out[k] = a * Ji[k];
                }
            }
        }
    }
    // SIPG penalty (Elasticity.cpp:279-292 / Poisson.cpp:209-223)
    penalty_.assign(F, 0.0);
    double cN1 = trace_constant(D, N_ - 1);
    auto pside = [&](std::size_t i, int s) {
        int64_t e = topo_.facets[i].elem[s];
        double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
        for (unsigned t = 0; t < Nm; ++t) {
            if (el) {
                lo = std::min(lo, 2.0 * mat1_[e * Nm + t]);
                hi = std::max(hi, D * mat0_[e * Nm + t] + 2.0 * mat1_[e * Nm + t]);
            } else {
                lo = std::min(lo, mat0_[e * Nm + t]);
                hi = std::max(hi, mat0_[e * Nm + t]);
            }
        }
        return (D + 1) * cN1 * (area_[i] / volume_[e]) * (hi * hi / lo);
    };
    for (std::size_t i = 0; i < F; ++i) {
        Facet const& fc = topo_.facets[i];
        penalty_[i] = fc.elem[0] != fc.elem[1] ? (pside(i, 0) + pside(i, 1)) / 4.0 : pside(i, 0);
    }
    // per element face: facet, neighbour, flags = kind | side << 2 | neighbour local face << 4
    unsigned nf = D + 1;
    ef_.assign(E_ * nf * 4, 0);
    for (std::size_t e = 0; e < E_; ++e)
        for (unsigned f = 0; f < nf; ++f) {
            Side const& sd = topo_.sides[e * nf + f];
            bool skel = sd.nb != static_cast<int64_t>(e);
            bool fd = sd.bc == BC_Fault || sd.bc == BC_Dirichlet;
            int kind = (sd.bc == BC_None || (skel && fd)) ? Face_Skeleton : (fd ? Face_Boundary : Face_Skip);
            int32_t* o = &ef_[(e * nf + f) * 4];
            o[0] = static_cast<int32_t>(sd.facet);
            o[1] = static_cast<int32_t>(sd.nb);
            o[2] = kind | (sd.side << 2) | (sd.nb_local << 4);
            o[3] = 0;
        }
    if (E_ >= (std::size_t(1) << 31) || F >= (std::size_t(1) << 31)) throw std::runtime_error("mesh too large for 32-bit indices");
    drop_device();
    ready_ = true;
}

void Operator::drop_device() {
    d64_.reset();
    d32_.reset();
    m64_.reset();
    m32_.reset();
    s64_.reset();
    s32_.reset();
    rhs64_.reset();
    bj_.reset();
    asm_.reset();
    g64_.reset();
    g32_.reset();
}

void Operator::set_penalty(std::vector<double> const& pen) {
    if (!ready_) throw std::logic_error("set_material() must be called before set_penalty()");
    if (pen.size() != penalty_.size()) throw std::invalid_argument("penalty: one value per facet expected");
    penalty_ = pen;
    drop_device();
}

template <class T> static std::vector<T> cast(std::vector<double> const& v);

// Device residency of both precisions (first apply after set_material, or an explicit call).
void Operator::upload_device() {
    if (!ready_) throw std::logic_error("set_material() must be called before the device upload");
    if (!d64_ || !d32_) {
        auto a = std::make_unique<DeviceOperator<double>>();
        auto b = std::make_unique<DeviceOperator<float>>();
        upload(*a);
        upload(*b);
        d64_ = std::move(a);
        d32_ = std::move(b);
    }
    if (uses_mfma() && (!m64_ || !m32_)) {
        auto a = std::make_unique<DeviceMfma<double>>();
        auto b = std::make_unique<DeviceMfma<float>>();
        upload_mfma(*a);
        upload_mfma(*b);
        m64_ = std::move(a);
        m32_ = std::move(b);
    }
    if (uses_split()) {
        if (!s64_ || !s32_) {
            auto a = std::make_unique<DeviceSplit<double>>();
            auto b = std::make_unique<DeviceSplit<float>>();
            upload_split(*a);
            upload_split(*b);
            s64_ = std::move(a);
            s32_ = std::move(b);
        }
    } else if (uses_mfma()) {
        ensure_mfma_faces();
    }
    if (variant_ == Variant_Assembled && !asm_) assemble(); // before any graph capture of the apply
}

// The element-face records of the element-centric MFMA kernel (also read by the matrix-core
// gradient kernel), built on first use.
void Operator::ensure_mfma_faces() {
    if (!m64_ || !m32_) {
        auto a = std::make_unique<DeviceMfma<double>>();
        auto b = std::make_unique<DeviceMfma<float>>();
        upload_mfma(*a);
        upload_mfma(*b);
        m64_ = std::move(a);
        m32_ = std::move(b);
    }
    if (m64_->FD.n == 0) upload_mfma_faces(*m64_);
    if (m32_->FD.n == 0) upload_mfma_faces(*m32_);
}

void Operator::reserve_columns(int ncols) {
    if (ncols < 1) throw std::invalid_argument("reserve_columns: ncols must be positive");
    upload_device();
    if (uses_split()) {
        reserve_split(*s64_, ncols);
        reserve_split(*s32_, ncols);
    }
}

template <class T> void Operator::reserve_split(DeviceSplit<T>& ds, int ncols) {
    if (ds.cap >= ncols) return;
    const std::size_t per = E_ * (D_ + 1) * block_size();
    ds.L.alloc(per * std::size_t(ncols));
    TDG_HIP(hipMemsetAsync(ds.L.ptr, 0, ds.L.n * sizeof(T), stream_));
    TDG_HIP(hipStreamSynchronize(stream_));
    ds.cap = ncols;
}

void Operator::set_variant(int v) {
    if (v < 0 || v >= kNumVariants) throw std::invalid_argument("unknown kernel variant");
    variant_ = v;
}

// Times applies of the operator on ncols zero columns (HIP events on the operator's stream) for each
// configuration in `configs` (set by `select(i)`), keeps the fastest and returns the milliseconds
// of each (negative where select returns false).
double Operator::time_configs(int ncols, int pb, int count, std::function<bool(int)> const& select,
                              std::vector<double>& ms) {
    const std::size_t n = local_size() * std::size_t(ncols);
    void *X = nullptr, *Y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (X) (void)hipFree(X);
        if (Y) (void)hipFree(Y);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    ms.assign(std::size_t(count), -1.0);
    double best = -1.0;
    try {
        TDG_HIP(hipMalloc(&X, n * pb));
        TDG_HIP(hipMalloc(&Y, n * pb));
        TDG_HIP(hipMemsetAsync(X, 0, n * pb, stream_));
        TDG_HIP(hipEventCreate(&e0));
        TDG_HIP(hipEventCreate(&e1));
        const int64_t ld = int64_t(local_size());
        auto run = [&] {
            if (pb == 8) dispatch(static_cast<double const*>(X), static_cast<double*>(Y), ncols, ld);
            else dispatch(static_cast<float const*>(X), static_cast<float*>(Y), ncols, ld);
        };
        // a warm-up pass over every configuration first (uploads, and the GPU's clocks settle under
        // sustained load), then two timed rounds; each configuration keeps its faster round
        for (int round = 0; round < 3; ++round)
            for (int c = 0; c < count; ++c) {
                if (!select(c)) continue;
                upload_device();
                if (round == 0) {
                    for (int r = 0; r < 4; ++r) run();
                    continue;
                }
                constexpr int reps = 5;
                run();
                TDG_HIP(hipEventRecord(e0, stream_));
                for (int r = 0; r < reps; ++r) run();
                TDG_HIP(hipEventRecord(e1, stream_));
                TDG_HIP(hipEventSynchronize(e1));
                float t = 0.0f;
                TDG_HIP(hipEventElapsedTime(&t, e0, e1));
                const double v = double(t) / reps;
                if (ms[std::size_t(c)] < 0.0 || v < ms[std::size_t(c)]) ms[std::size_t(c)] = v;
                if (best < 0.0 || v < best) best = v;
            }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    return best;
}

std::vector<double> Operator::autotune_shapes(int ncols, int pb) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    if (ncols < 1 || (pb != 8 && pb != 4)) throw std::invalid_argument("bad autotune arguments");
    const int p = pb == 8 ? 0 : 1;
    std::vector<double> ms;
    if (!(kind_ == Kind_Elasticity && mfma_supported(D_, N_)))
        return std::vector<double>(kSplitOrders * kSplitShapes, -1.0);
    const int saved_variant = variant_, saved_shape = split_shape_[p], saved_xcd = split_xcd_[p];
    try {
        variant_ = Variant_Auto;
        time_configs(ncols, pb, kSplitOrders * kSplitShapes, [&](int c) {
            split_shape_[p] = c / kSplitOrders;
            split_xcd_[p] = c % kSplitOrders;
            return true;
        }, ms);
    } catch (...) {
        variant_ = saved_variant;
        split_shape_[p] = saved_shape;
        split_xcd_[p] = saved_xcd;
        throw;
    }
    variant_ = saved_variant;
    int best = 0;
    for (int c = 1; c < kSplitOrders * kSplitShapes; ++c)
        if (ms[std::size_t(c)] >= 0.0 && ms[std::size_t(c)] < ms[std::size_t(best)]) best = c;
    split_shape_[p] = best / kSplitOrders;
    split_xcd_[p] = best % kSplitOrders;
    return ms;
}

std::vector<double> Operator::autotune(int ncols, int pb) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    if (ncols < 1 || (pb != 8 && pb != 4)) throw std::invalid_argument("bad autotune arguments");
    const bool mfma = kind_ == Kind_Elasticity && mfma_supported(D_, N_);
    if (mfma) autotune_shapes(ncols, pb); // the two-kernel apply runs its fastest launch shape
    const int saved = variant_;
    std::vector<double> ms;
    try {
        time_configs(ncols, pb, kNumVariants, [&](int v) {
            if (v == Variant_Assembled) {
                if (!assembled_fast()) return false; // the generic kernel is a cross-check only
                variant_ = v;
                return true;
            }
            if (v != Variant_Points && !mfma) return false;
            variant_ = v;
            return true;
        }, ms);
    } catch (...) {
        variant_ = saved;
        throw;
    }
    int best = -1;
    for (int v = 0; v < kNumVariants; ++v)
        if (ms[std::size_t(v)] >= 0.0 && (best < 0 || ms[std::size_t(v)] < ms[std::size_t(best)])) best = v;
    variant_ = best;
    return ms;
}

void Operator::set_split_shape(int pb, int shape, int xcd) {
    if ((pb != 8 && pb != 4) || shape < 0 || shape >= kSplitShapes || xcd < 0 || xcd >= kSplitOrders)
        throw std::invalid_argument("bad launch shape");
    split_shape_[pb == 8 ? 0 : 1] = shape;
    split_xcd_[pb == 8 ? 0 : 1] = xcd;
}

bool Operator::uses_mfma() const {
    return variant_ != Variant_Points && variant_ != Variant_Assembled && kind_ == Kind_Elasticity &&
           mfma_supported(D_, N_);
}

bool Operator::uses_split() const { return variant_ == Variant_Auto && uses_mfma(); }

// Tables and element-grouped point data of the MFMA kernel (kernels/dg_apply_lds.hip): canonical
// reference tables V[e][q][k] and T0[c][q][k] (face 0, c = 0 value, 1 + e gradient), the facet
// relabellings, and per group of 16 elements (element innermost) the volume point data VD, the
// two-sided facet records FD, penalties FP and face info FI.
template <class T> void Operator::upload_mfma(DeviceMfma<T>& dm) {
    const unsigned D = D_, DD = D * D, NF = D + 1;
    const MfmaDims L{int(D), int(Nbf_), int(Nq_), int(nq_)};
    auto T0 = [&](int c, int jb, int q) -> double {
        return c == 0 ? E_q_[0][jb + Nbf_ * q] : Dxi_q_[0][jb + Nbf_ * ((c - 1) + D * q)];
    };
    std::vector<double> V(std::size_t(D) * Nq_ * Nbf_), TT(std::size_t(NF) * nq_ * Nbf_);
    for (unsigned e = 0; e < D; ++e)
        for (unsigned q = 0; q < Nq_; ++q)
            for (unsigned k = 0; k < Nbf_; ++k) V[(e * Nq_ + q) * Nbf_ + k] = Dxi_Q_[k + Nbf_ * (e + D * q)];
    for (unsigned c = 0; c <= D; ++c)
        for (unsigned q = 0; q < nq_; ++q)
            for (unsigned k = 0; k < Nbf_; ++k) TT[(c * nq_ + q) * Nbf_ + k] = T0(int(c), int(k), int(q));
    std::vector<int32_t> P2(perm_);
    P2.insert(P2.end(), iperm_.begin(), iperm_.end());
    // element-grouped volume point data
    const std::size_t ng = (E_ + 15) / 16;
    std::vector<T> VD(ng * Nq_ * L.NCH * 16, T(0));
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e) {
        std::size_t grp = e / 16, j = e % 16;
        for (unsigned q = 0; q < Nq_; ++q) {
            T* o = &VD[((grp * Nq_ + q) * L.NCH) * 16 + j];
            for (unsigned c = 0; c < DD; ++c) o[c * 16] = T(JInv_[(e * Nq_ + q) * DD + c]);
            o[DD * 16] = T(c0Q_[e * Nq_ + q]);
            o[(DD + 1) * 16] = T(c1Q_[e * Nq_ + q]);
        }
    }
    dm.V.upload(cast<T>(V));
    dm.T0.upload(cast<T>(TT));
    dm.perm2.upload(P2);
    dm.w.upload(cast<T>(fct_.weights));
    dm.VD.upload(VD);
}

// Two-sided element-face records of the element-centric MFMA kernel: per group of 16 elements and
// local face, both sides' G' = J_f^-1 J^-1 and coefficient traces and the oriented normal (FD), the
// penalty (FP) and the face info (FI).
template <class T> void Operator::upload_mfma_faces(DeviceMfma<T>& dm) {
    const unsigned D = D_, DD = D * D, NF = D + 1;
    const MfmaDims L{int(D), int(Nbf_), int(Nq_), int(nq_)};
    const std::size_t ng = (E_ + 15) / 16;
    std::vector<T> FP(ng * NF * 16, T(0));
    std::vector<T> FD(ng * NF * nq_ * L.NFC * 16, T(0));
    std::vector<int32_t> FI(ng * NF * 16 * 4, 0);
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e) {
        std::size_t grp = e / 16, j = e % 16;
        for (unsigned f = 0; f < NF; ++f) {
            int32_t const* ef = &ef_[(e * NF + f) * 4];
            int64_t fct = ef[0];
            int side = (ef[2] >> 2) & 1;
            int32_t* fi = &FI[((grp * NF + f) * 16 + j) * 4];
            fi[0] = (ef[2] & 3) | (((ef[2] >> 4) & 3) << 4);
            fi[1] = ef[1];
            FP[(grp * NF + f) * 16 + j] = T(penalty_[fct]);
            for (unsigned q = 0; q < nq_; ++q) {
                T* o = &FD[(((grp * NF + f) * nq_ + q) * L.NFC) * 16 + j];
                double const* g0 = &fGp_[((fct * 2 + side) * nq_ + q) * DD];
                double const* g1 = &fGp_[((fct * 2 + (1 - side)) * nq_ + q) * DD];
                for (unsigned c = 0; c < DD; ++c) {
                    o[c * 16] = T(g0[c]);
                    o[(DD + 2 + c) * 16] = T(g1[c]);
                }
                o[DD * 16] = T(fc0_[(fct * 2 + side) * nq_ + q]);
                o[(DD + 1) * 16] = T(fc1_[(fct * 2 + side) * nq_ + q]);
                o[(2 * DD + 2) * 16] = T(fc0_[(fct * 2 + (1 - side)) * nq_ + q]);
                o[(2 * DD + 3) * 16] = T(fc1_[(fct * 2 + (1 - side)) * nq_ + q]);
                double sg = side ? -1.0 : 1.0;
                for (unsigned d = 0; d < D; ++d) o[(2 * DD + 4 + d) * 16] = T(sg * normal_[(fct * nq_ + q) * D + d]);
            }
        }
    }
    dm.FD.upload(FD);
    dm.FP.upload(FP);
    dm.FI.upload(FI);
}

// Facet records of the two-kernel MFMA apply: the facets that carry a term, two-sided first, then
// one-sided, each in mesh facet order, in groups of 8 (facet innermost).  Per facet point: the
// tangent columns of side 0's facet Jacobian M_0 = J_0 J_f0 (shared by both sides: the facet
// parametrisations of the two sides coincide, vertices being in ascending order), the third
// columns of M_0 and M_1, and both sides' coefficient traces; per facet the penalty and
// {element 0, element 1, face 0 | face 1 << 4 | kind << 8}.  Padding lanes get the identity
// Jacobian so that their (discarded) arithmetic stays finite.
template <class T> void Operator::upload_split(DeviceSplit<T>& ds) {
    const unsigned D = D_, DD = D * D, NF = D + 1;
    const unsigned NFR = DD + D + 4;
    const std::size_t F = topo_.facets.size();
    std::vector<int64_t> order;
    order.reserve(F);
    for (int pass = 0; pass < 2; ++pass)
        for (std::size_t i = 0; i < F; ++i) {
            Facet const& fc = topo_.facets[i];
            int kind = ef_[(std::size_t(fc.elem[0]) * NF + fc.local[0]) * 4 + 2] & 3;
            if ((pass == 0 && kind == Face_Skeleton) || (pass == 1 && kind == Face_Boundary))
                order.push_back(int64_t(i));
        }
    std::size_t ntwo = 0;
    for (int64_t i : order) {
        Facet const& fc = topo_.facets[std::size_t(i)];
        if ((ef_[(std::size_t(fc.elem[0]) * NF + fc.local[0]) * 4 + 2] & 3) == Face_Skeleton) ++ntwo;
    }
    constexpr std::size_t G = 8; // facets per group: the two sides of 8 facets are the 16 MFMA columns
    const std::size_t nact = order.size(), ng = (nact + G - 1) / G;
    std::vector<T> FR(ng * nq_ * NFR * G, T(0)), FPen(ng * G, T(0));
    std::vector<int32_t> FInfo(ng * G * 4, 0);
#pragma omp parallel for schedule(static)
    for (std::size_t p = 0; p < ng * G; ++p) {
        const std::size_t grp = p / G, j = p % G;
        for (unsigned q = 0; q < nq_; ++q) {
            T* o = &FR[((grp * nq_ + q) * NFR) * G + j];
            if (p >= nact) {
                for (unsigned d = 0; d < D; ++d) // identity: t_d = e_d, c_0 = c_1 = e_(D-1)
                    for (unsigned r = 0; r < D; ++r) o[(d * D + r) * G] = T(r == d ? 1.0 : 0.0);
                for (unsigned r = 0; r < D; ++r) o[(DD + r) * G] = T(r == D - 1 ? 1.0 : 0.0);
                for (unsigned c = 0; c < 4; ++c) o[(DD + D + c) * G] = T(1.0);
                continue;
            }
            const std::size_t i = std::size_t(order[p]);
            const bool two = p < ntwo;
            double const* M0 = &fM_[0][(i * nq_ + q) * DD];
            double const* M1 = two ? &fM_[1][(i * nq_ + q) * DD] : M0;
            for (unsigned d = 0; d < D; ++d)
                for (unsigned r = 0; r < D; ++r) o[(d * D + r) * G] = T(M0[r + d * D]);
            for (unsigned r = 0; r < D; ++r) o[(DD + r) * G] = T(M1[r + (D - 1) * D]);
            const int s1 = two ? 1 : 0;
            o[(DD + D) * G] = T(fc0_[(i * 2 + 0) * nq_ + q]);
            o[(DD + D + 1) * G] = T(fc1_[(i * 2 + 0) * nq_ + q]);
            o[(DD + D + 2) * G] = T(fc0_[(i * 2 + s1) * nq_ + q]);
            o[(DD + D + 3) * G] = T(fc1_[(i * 2 + s1) * nq_ + q]);
        }
        if (p < nact) {
            const std::size_t i = std::size_t(order[p]);
            Facet const& fc = topo_.facets[i];
            const bool two = p < ntwo;
            FPen[p] = T(penalty_[i]);
            int32_t* fi = &FInfo[p * 4];
            fi[0] = int32_t(fc.elem[0]);
            fi[1] = int32_t(two ? fc.elem[1] : fc.elem[0]);
            fi[2] = fc.local[0] | ((two ? fc.local[1] : fc.local[0]) << 4) | ((two ? 1 : 2) << 8);
            fi[3] = 0;
        }
    }
    ds.FR.upload(FR);
    ds.FPen.upload(FPen);
    ds.FInfo.upload(FInfo);
    std::vector<int32_t> idle;
    for (std::size_t e = 0; e < E_; ++e)
        for (unsigned f = 0; f < NF; ++f)
            if ((ef_[(e * NF + f) * 4 + 2] & 3) == Face_Skip) idle.push_back(int32_t(e * NF + f));
    ds.idle.upload(idle);
    ds.nfacets = int64_t(nact);
    ds.ntwo = int64_t(ntwo);
    ds.cap = 0;
    reserve_split(ds, 1);
}

template <class T> static std::vector<T> cast(std::vector<double> const& v) {
    return std::vector<T>(v.begin(), v.end());
}

template <class T> void Operator::upload(DeviceOperator<T>& d) {
    unsigned D = D_, Nf = D + 1;
    // volume tables: tabV[l][e][q] = Dxi_Q(l,e,q); tabVt[e][q][k] = Dxi_Q(k,e,q)
    std::vector<double> tV(Nbf_ * D * Nq_), tVt(Nbf_ * D * Nq_);
    for (unsigned l = 0; l < Nbf_; ++l)
        for (unsigned e = 0; e < D; ++e)
            for (unsigned q = 0; q < Nq_; ++q) {
                double v = Dxi_Q_[l + Nbf_ * (e + D * q)];
                tV[(l * D + e) * Nq_ + q] = v;
                tVt[(e * Nq_ + q) * Nbf_ + l] = v;
            }
    // facet tables, channel c = 0: value, c = 1 + e: d/dxi_e.  Elasticity keeps face 0 only.
    unsigned nfaces = kind_ == Kind_Elasticity ? 1 : Nf;
    std::vector<double> tF(nfaces * Nbf_ * (D + 1) * nq_), tFt(tF.size());
    for (unsigned f = 0; f < nfaces; ++f)
        for (unsigned l = 0; l < Nbf_; ++l)
            for (unsigned c = 0; c <= D; ++c)
                for (unsigned q = 0; q < nq_; ++q) {
                    double v = c == 0 ? E_q_[f][l + Nbf_ * q] : Dxi_q_[f][l + Nbf_ * ((c - 1) + D * q)];
                    tF[((f * Nbf_ + l) * (D + 1) + c) * nq_ + q] = v;
                    tFt[((f * (D + 1) + c) * nq_ + q) * Nbf_ + l] = v;
                }
    d.tabV.upload(cast<T>(tV));
    d.tabVt.upload(cast<T>(tVt));
    d.tabF.upload(cast<T>(tF));
    d.tabFt.upload(cast<T>(tFt));
    d.W.upload(cast<T>(vol_.weights));
    d.w.upload(cast<T>(fct_.weights));
    d.perm.upload(perm_);
    d.iperm.upload(iperm_);
    d.G.upload(cast<T>(JInv_));
    d.c0.upload(cast<T>(c0Q_));
    d.c1.upload(cast<T>(c1Q_));
    d.fG.upload(cast<T>(fGp_));
    d.fc0.upload(cast<T>(fc0_));
    d.fc1.upload(cast<T>(fc1_));
    d.fn.upload(cast<T>(normal_));
    d.fpen.upload(cast<T>(penalty_));
    d.ef.upload(ef_);
}

// rhs_volume + rhs_boundary with the interior-penalty lift (Elasticity.cpp:521-696 /
// Poisson.cpp:383-540; rhsVolume, rhs_lift_ip, rhsFacet).  Dirichlet data on boundary facets only.
void Operator::rhs(double const* force_Q, double const* facet_q, double* b) const {
    if (!ready_) throw std::runtime_error("set_material() must be called first");
    unsigned D = D_, DD = D * D;
    std::size_t bs = block_size();
    bool el = kind_ == Kind_Elasticity;
    std::fill(b, b + bs * E_, 0.0);
    if (force_Q) {
#pragma omp parallel for schedule(static)
        for (std::size_t e = 0; e < E_; ++e)
            for (unsigned p = 0; p < Q_; ++p)
                for (unsigned k = 0; k < Nbf_; ++k) {
                    double s = 0.0;
                    for (unsigned q = 0; q < Nq_; ++q)
                        s += absDetJ_[e * Nq_ + q] * vol_.weights[q] * E_Q_[k + Nbf_ * q] * force_Q[(e * Q_ + p) * Nq_ + q];
                    b[e * bs + k + p * Nbf_] += s;
                }
    }
    if (!facet_q) return;
    // rhs_skeleton / rhs_boundary (Elasticity.cpp:545-696, Poisson.cpp:405-540): Dirichlet data on
    // Dirichlet facets, slip on Fault facets (halved on the boundary); skeleton facets add to both
    // sides with c10 = eps/2, c20 = +-penalty and the side-0 normal.
    std::size_t F = topo_.facets.size();
    std::vector<double> fq(Q_ * nq_);
    for (std::size_t i = 0; i < F; ++i) {
        Facet const& fc = topo_.facets[i];
        if (fc.bc != BC_Dirichlet && fc.bc != BC_Fault) continue;
        bool skel = fc.elem[0] != fc.elem[1];
        double scale = (!skel && fc.bc == BC_Fault) ? 0.5 : 1.0;
        for (unsigned j = 0; j < Q_ * nq_; ++j) fq[j] = scale * facet_q[i * Q_ * nq_ + j];
        double pen = penalty_[i];
        for (int side = 0; side < (skel ? 2 : 1); ++side) {
            double const* Ji = &fJInv_[side][i * nq_ * DD];
            auto const& Eq = E_q_[fc.local[side]];
            auto const& Dq = Dxi_q_[fc.local[side]];
            double* B = b + fc.elem[side] * bs;
            double c10 = skel ? -0.5 : -1.0, c20 = side == 0 ? pen : -pen;
            for (unsigned q = 0; q < nq_; ++q) {
                double w = fct_.weights[q];
                double const* n = &normal_[(i * nq_ + q) * D];
                double nl = nlen_[i * nq_ + q];
                double c0s = fc0_[(i * 2 + side) * nq_ + q], c1s = fc1_[(i * 2 + side) * nq_ + q];
                for (unsigned k = 0; k < Nbf_; ++k) {
                    double dx[3];
                    for (unsigned j = 0; j < D; ++j) {
                        double v = 0.0;
                        for (unsigned e = 0; e < D; ++e) v += Ji[q * DD + e + j * D] * Dq[k + Nbf_ * (e + D * q)];
                        dx[j] = v;
                    }
                    if (el) {
                        double fn = 0.0, nd = 0.0, dfd = 0.0;
                        for (unsigned j = 0; j < D; ++j) {
                            fn += fq[j * nq_ + q] * n[j];
                            nd += dx[j] * n[j];
                            dfd += dx[j] * fq[j * nq_ + q];
                        }
                        for (unsigned p = 0; p < D; ++p) {
                            double tt = c0s * dx[p] * fn + c1s * (nd * fq[p * nq_ + q] + dfd * n[p]);
                            B[k + p * Nbf_] += c10 * tt * w + c20 * w * Eq[k + Nbf_ * q] * fq[p * nq_ + q] * nl;
                        }
                    } else {
                        double kn = 0.0;
                        for (unsigned j = 0; j < D; ++j) kn += c0s * dx[j] * n[j];
                        B[k] += c10 * w * kn * fq[q] + c20 * w * Eq[k + Nbf_ * q] * nl * fq[q];
                    }
                }
            }
        }
    }
}

RhsArgs Operator::rhs_args() {
    if (kind_ != Kind_Elasticity) throw std::invalid_argument("the device rhs / traction is built for elasticity");
    upload_device();
    if (!rhs64_) {
        auto r = std::make_unique<DeviceRhs>();
        std::vector<double> tabE(std::size_t(Nbf_) * Nq_), jw(E_ * Nq_);
        for (unsigned k = 0; k < Nbf_; ++k)
            for (unsigned q = 0; q < Nq_; ++q) tabE[k * Nq_ + q] = E_Q_[k + Nbf_ * q];
        for (std::size_t e = 0; e < E_; ++e)
            for (unsigned q = 0; q < Nq_; ++q) jw[e * Nq_ + q] = absDetJ_[e * Nq_ + q] * vol_.weights[q];
        std::size_t F = topo_.facets.size();
        std::vector<int32_t> fbc(F), fct(F * 4, 0);
        for (std::size_t i = 0; i < F; ++i) {
            Facet const& fc = topo_.facets[i];
            bool skel = fc.elem[0] != fc.elem[1];
            fbc[i] = int32_t(fc.bc) | (skel ? 1 << 8 : 0);
            fct[i * 4 + 0] = int32_t(fc.elem[0]);
            fct[i * 4 + 1] = int32_t(fc.elem[1]);
            fct[i * 4 + 2] = int32_t(fc.local[0]) | (int32_t(fc.local[1]) << 4);
        }
        r->tabE.upload(tabE);
        r->jw.upload(jw);
        r->fbc.upload(fbc);
        r->fct.upload(fct);
        rhs64_ = std::move(r);
    }
    DeviceOperator<double> const& d = *d64_;
    RhsArgs a;
    a.tabE = rhs64_->tabE.ptr;
    a.jw = rhs64_->jw.ptr;
    a.tabF = d.tabF.ptr;
    a.w = d.w.ptr;
    a.perm = d.perm.ptr;
    a.iperm = d.iperm.ptr;
    a.fG = d.fG.ptr;
    a.fc0 = d.fc0.ptr;
    a.fc1 = d.fc1.ptr;
    a.fn = d.fn.ptr;
    a.fpen = d.fpen.ptr;
    a.ef = reinterpret_cast<int4 const*>(d.ef.ptr);
    a.fbc = rhs64_->fbc.ptr;
    a.fct = reinterpret_cast<int4 const*>(rhs64_->fct.ptr);
    a.num_elements = int64_t(E_);
    a.num_facets = int64_t(topo_.facets.size());
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    return a;
}

void Operator::rhs_device(double const* force_Q, double const* facet_q, double* b, int ncols, int64_t ld_force,
                          int64_t ld_facet, int64_t ld_b) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    RhsArgs a = rhs_args();
    TDG_HIP(launch_elasticity_rhs(D_, N_, a, force_Q, facet_q, b, ncols, ld_force, ld_facet, ld_b, stream_));
}

// Data of the material-gradient kernel: W|J|, the inverse material mass matrix of every element
// (the projection's M, Elasticity.cpp:104-164), and the penalty's extremal nodes and factors
// (Elasticity.cpp:255-292).
template <class T> void Operator::upload_grad(DeviceGrad<T>& g) {
    const unsigned D = D_, NF = D + 1, Nm = Nbf_;
    std::vector<double> tabE(std::size_t(Nbf_) * Nq_), jw(E_ * Nq_), Minv(E_ * Nm * Nm), pc(E_ * 2), pcoef(E_ * NF);
    std::vector<int32_t> pinfo(E_ * 4, 0);
    for (unsigned k = 0; k < Nbf_; ++k)
        for (unsigned q = 0; q < Nq_; ++q) tabE[k * Nq_ + q] = E_Q_[k + Nbf_ * q];
    const double cN1 = trace_constant(D, N_ - 1);
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < E_; ++e) {
        std::vector<double> M(Nm * Nm, 0.0), I(Nm * Nm, 0.0);
        for (unsigned q = 0; q < Nq_; ++q) {
            double wj = vol_.weights[q] * absDetJ_[e * Nq_ + q];
            jw[e * Nq_ + q] = wj;
            for (unsigned k = 0; k < Nm; ++k)
                for (unsigned l = 0; l < Nm; ++l) M[k + l * Nm] += matE_Q_[k + Nm * q] * wj * matE_Q_[l + Nm * q];
        }
        for (unsigned i = 0; i < Nm; ++i) I[i + i * Nm] = 1.0;
        dense_solve(M, Nm, I.data(), Nm);
        std::copy(I.begin(), I.end(), Minv.begin() + e * Nm * Nm);
        double c0 = std::numeric_limits<double>::max(), c1 = std::numeric_limits<double>::lowest();
        int i0 = 0, i1 = 0;
        for (unsigned i = 0; i < Nm; ++i) {
            double a = 2.0 * mat1_[e * Nm + i], b = D * mat0_[e * Nm + i] + 2.0 * mat1_[e * Nm + i];
            if (a < c0) { c0 = a; i0 = int(i); }
            if (b > c1) { c1 = b; i1 = int(i); }
        }
        pc[e * 2 + 0] = c0;
        pc[e * 2 + 1] = c1;
        pinfo[e * 4 + 0] = i0;
        pinfo[e * 4 + 1] = i1;
        for (unsigned f = 0; f < NF; ++f) {
            std::size_t fct = std::size_t(ef_[(e * NF + f) * 4]);
            pcoef[e * NF + f] = (D + 1) * cN1 * area_[fct] / volume_[e];
        }
    }
    std::vector<double> tabEt(tabE.size());
    for (unsigned k = 0; k < Nbf_; ++k)
        for (unsigned q = 0; q < Nq_; ++q) tabEt[q * Nbf_ + k] = tabE[k * Nq_ + q];
    g.tabE.upload(cast<T>(tabE));
    g.tabEt.upload(cast<T>(tabEt));
    g.jw.upload(cast<T>(jw));
    g.Minv.upload(cast<T>(Minv));
    g.pc.upload(cast<T>(pc));
    g.pcoef.upload(cast<T>(pcoef));
    g.pinfo.upload(pinfo);
}

template <class T> void Operator::grad_dispatch(T const* u, T const* psi, T* gl, T* gm) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    if (kind_ != Kind_Elasticity) throw std::invalid_argument("the material gradient is built for elasticity");
    upload_device();
    auto& slot = [&]() -> auto& {
        if constexpr (std::is_same<T, double>::value) return g64_;
        else return g32_;
    }();
    if (!slot) {
        auto g = std::make_unique<DeviceGrad<T>>();
        upload_grad(*g);
        slot = std::move(g);
    }
    GradArgs<T> g;
    g.tabE = slot->tabE.ptr;
    g.tabEt = slot->tabEt.ptr;
    g.jw = slot->jw.ptr;
    g.Minv = slot->Minv.ptr;
    g.pc = slot->pc.ptr;
    g.pinfo = reinterpret_cast<int4 const*>(slot->pinfo.ptr);
    g.pcoef = slot->pcoef.ptr;
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    if (uses_mfma()) {
        ensure_mfma_faces();
        MfmaArgs<T> b;
        if constexpr (std::is_same<T, double>::value) b = mfma_args(*m64_);
        else b = mfma_args(*m32_);
        TDG_HIP(launch_material_gradient_mfma<T>(D_, N_, b, g, u, psi, gl, gm, stream_));
        return;
    }
    DeviceOperator<T> const& d = device<T>();
    ApplyArgs<T> a;
    a.tabV = d.tabV.ptr;
    a.tabVt = d.tabVt.ptr;
    a.tabF = d.tabF.ptr;
    a.tabFt = d.tabFt.ptr;
    a.w = d.w.ptr;
    a.perm = d.perm.ptr;
    a.iperm = d.iperm.ptr;
    a.G = d.G.ptr;
    a.c0 = d.c0.ptr;
    a.c1 = d.c1.ptr;
    a.fG = d.fG.ptr;
    a.fc0 = d.fc0.ptr;
    a.fc1 = d.fc1.ptr;
    a.fn = d.fn.ptr;
    a.fpen = d.fpen.ptr;
    a.ef = reinterpret_cast<int4 const*>(d.ef.ptr);
    a.num_elements = int64_t(E_);
    TDG_HIP(launch_material_gradient<T>(D_, N_, a, g, u, psi, gl, gm, stream_));
}

// Weight-adjusted inverse mass of the curvilinear element with unit density (Elasticity.cpp:227-253,
// applied as in elasticity.py:236-238): Mhat^-1 E diag(-W/|J|) E^T Mhat^-1, Mhat the nodal mass
// matrix of the reference simplex (exact with the 2N+1 volume rule).
WaveArgs Operator::wave_args() {
    if (kind_ != Kind_Elasticity) throw std::invalid_argument("the wave operator is built for elasticity");
    if (!ready_) throw std::logic_error("set_material() must be called first");
    upload_device();
    if (!wave_) {
        auto wv = std::make_unique<DeviceWave>();
        const unsigned n = Nbf_;
        std::vector<double> M(std::size_t(n) * n, 0.0), Minv(std::size_t(n) * n, 0.0);
        for (unsigned q = 0; q < Nq_; ++q)
            for (unsigned l = 0; l < n; ++l)
                for (unsigned k = 0; k < n; ++k) M[k + l * n] += vol_.weights[q] * E_Q_[k + n * q] * E_Q_[l + n * q];
        for (unsigned k = 0; k < n; ++k) Minv[k + k * n] = 1.0;
        dense_solve(M, int(n), Minv.data(), int(n));
        std::vector<double> tabE(std::size_t(n) * Nq_), w(E_ * Nq_);
        for (unsigned k = 0; k < n; ++k)
            for (unsigned q = 0; q < Nq_; ++q) tabE[k * Nq_ + q] = E_Q_[k + n * q];
        for (std::size_t e = 0; e < E_; ++e)
            for (unsigned q = 0; q < Nq_; ++q) w[e * Nq_ + q] = -vol_.weights[q] / absDetJ_[e * Nq_ + q];
        wv->Minv.upload(Minv);
        wv->E.upload(tabE);
        wv->w.upload(w);
        wv->ax.alloc(local_size());
        wv->b.alloc(local_size());
        wave_ = std::move(wv);
    }
    WaveArgs a;
    a.Minv = wave_->Minv.ptr;
    a.E = wave_->E.ptr;
    a.w = wave_->w.ptr;
    a.nbf = int(Nbf_);
    a.nq = int(Nq_);
    a.nquant = int(Q_);
    a.num_elements = int64_t(E_);
    return a;
}

void Operator::inverse_mass_device(double const* x, double* y) {
    WaveArgs a = wave_args();
    TDG_HIP(launch_inverse_mass(a, x, y, stream_));
}

// apply_<true> of Elasticity.cpp:865-880 followed by the inverse mass: y = M^-1 (A x - b_facet(f)).
void Operator::wave_rhs_device(double const* x, double const* facet_q, double* y) {
    WaveArgs a = wave_args();
    apply(x, wave_->ax.ptr);
    if (facet_q) {
        rhs_device(nullptr, facet_q, wave_->b.ptr, 1, 0, 0, 0);
        TDG_HIP(launch_axpy(wave_->ax.ptr, -1.0, wave_->b.ptr, int64_t(local_size()), stream_));
    }
    TDG_HIP(launch_inverse_mass(a, wave_->ax.ptr, y, stream_));
}

// I[k][l] = phi^from_l(x^to_k) on the nodal bases (NodalInterpolation.cpp:8-38).
std::vector<double> Operator::interpolation_matrix(unsigned D, unsigned to, unsigned from) {
    Basis bt(D, to, true), bf(D, from, true);
    return bf.values(bt.nodes()); // column-major (n_from x n_to) == row-major (n_to x n_from)
}

// Fault adapter: AdapterBase::prepare (app/localoperator/AdapterBase.cpp:39-100) with
// Curvilinear::facetBasis (src/geometry/Curvilinear.cpp:281-316): per fault facet point the normal
// length, the unit normal oriented along the reference normal, the tangents from the up vector, all
// basis vectors negated where the orientation flipped the normal, and the inverse of the facet mass
// weighted by the normal length, on the nodal P_N basis of the (D-1)-simplex.
std::array<int64_t, 2> Operator::fault_setup(double const* up, double const* ref_normal) {
    if (kind_ != Kind_Elasticity) throw std::invalid_argument("the fault adapter is built for elasticity");
    const unsigned D = D_, nq = nq_;
    Basis space(D - 1, N_, true);
    const unsigned nbf = space.size();
    std::vector<double> e_q = space.values(fct_.points); // (l + nbf q)
    std::vector<int32_t> fac, fmap(topo_.facets.size(), -1);
    for (std::size_t f = 0; f < topo_.facets.size(); ++f)
        if (topo_.facets[f].bc == BC_Fault) {
            fmap[f] = int32_t(fac.size());
            fac.push_back(int32_t(f));
        }
    const std::size_t nf = fac.size();
    std::vector<double> nl(nf * nq), bas(nf * nq * D * D, 0.0), minv(nf * nbf * nbf, 0.0);
    const double tol = 10000.0 * std::numeric_limits<double>::epsilon();
    for (std::size_t m = 0; m < nf; ++m) {
        const std::size_t f = std::size_t(fac[m]);
        for (unsigned q = 0; q < nq; ++q) {
            double n[3] = {0, 0, 0}, len = 0.0, dot = 0.0;
            for (unsigned d = 0; d < D; ++d) {
                n[d] = normal_[(f * nq + q) * D + d];
                len += n[d] * n[d];
            }
            len = std::sqrt(len);
            for (unsigned d = 0; d < D; ++d) dot += ref_normal[d] * n[d];
            nl[m * nq + q] = len;
            if (std::fabs(dot) < tol) throw std::invalid_argument("normal and reference normal are almost perpendicular");
            const double flip = dot < 0.0 ? -1.0 : 1.0;
            double nh[3] = {0, 0, 0};
            for (unsigned d = 0; d < D; ++d) nh[d] = flip * n[d] / len; // oriented along the reference normal
            double v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0};
            if (D == 2) {
                const double c = up[0] * nh[1] - up[1] * nh[0];
                if (std::fabs(c) < tol) throw std::invalid_argument("up vector and normal are almost colinear");
                const double sg = c > 0.0 ? 1.0 : -1.0;
                v1[0] = -sg * nh[1];
                v1[1] = sg * nh[0];
            } else {
                double sv[3] = {up[1] * nh[2] - up[2] * nh[1], up[2] * nh[0] - up[0] * nh[2],
                                up[0] * nh[1] - up[1] * nh[0]};
                const double sl = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
                if (sl < tol) throw std::invalid_argument("up vector and normal are almost colinear");
                for (double& v : sv) v /= sl;
                double dv[3] = {sv[1] * nh[2] - sv[2] * nh[1], sv[2] * nh[0] - sv[0] * nh[2],
                                sv[0] * nh[1] - sv[1] * nh[0]};
                const double dl = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                for (unsigned d = 0; d < 3; ++d) {
                    v1[d] = dv[d] / dl;
                    v2[d] = sv[d];
                }
            }
            double* b = &bas[(m * nq + q) * D * D]; // b[i D + j]: component i of vector j
            for (unsigned d = 0; d < D; ++d) {
                b[d * D + 0] = flip * nh[d];
                b[d * D + 1] = flip * v1[d];
                if (D == 3) b[d * D + 2] = flip * v2[d];
            }
        }
        std::vector<double> M(std::size_t(nbf) * nbf, 0.0), I(std::size_t(nbf) * nbf, 0.0);
        for (unsigned q = 0; q < nq; ++q) {
            const double wq = fct_.weights[q] * nl[m * nq + q];
            for (unsigned j = 0; j < nbf; ++j)
                for (unsigned i = 0; i < nbf; ++i) M[i + j * nbf] += wq * e_q[i + nbf * q] * e_q[j + nbf * q];
        }
        for (unsigned i = 0; i < nbf; ++i) I[i + i * nbf] = 1.0;
        dense_solve(M, int(nbf), I.data(), int(nbf));
        std::copy(I.begin(), I.end(), minv.begin() + m * nbf * nbf);
    }
    auto fd = std::make_unique<DeviceFault>();
    fd->fac.upload(fac);
    fd->fmap.upload(fmap);
    fd->e_q.upload(e_q);
    fd->w.upload(fct_.weights);
    fd->nl.upload(nl);
    fd->bas.upload(bas);
    fd->minv.upload(minv);
    fd->nfault = int64_t(nf);
    fd->nbf = int(nbf);
    fault_ = std::move(fd);
    return {int64_t(nf), int64_t(nbf)};
}

FaultArgs Operator::fault_args() const {
    if (!fault_) throw std::logic_error("fault_setup() must be called first");
    FaultArgs a;
    a.fac = fault_->fac.ptr;
    a.fmap = fault_->fmap.ptr;
    a.e_q = fault_->e_q.ptr;
    a.w = fault_->w.ptr;
    a.nl = fault_->nl.ptr;
    a.bas = fault_->bas.ptr;
    a.minv = fault_->minv.ptr;
    a.nbf = fault_->nbf;
    a.nq = int(nq_);
    a.dim = int(D_);
    a.nfault = fault_->nfault;
    a.num_facets = int64_t(topo_.facets.size());
    return a;
}

void Operator::fault_slip_device(double const* S, double* F, int ncols, int64_t ld_s, int64_t ld_f) {
    FaultArgs a = fault_args();
    if (ncols > 1) {
        // columns must not overlap: each stride holds at least one column
        if (ld_s < a.nfault * int64_t(a.nbf) * int64_t(D_ - 1))
            throw std::invalid_argument("ld_s is below one column of fault-basis slips (nfault*nbf*(dim-1))");
        if (ld_f < int64_t(num_facets()) * int64_t(D_) * nq_)
            throw std::invalid_argument("ld_f is below one column of facet data (facets*dim*nq)");
    }
    TDG_HIP(launch_fault_slip(a, S, F, ncols, ld_s, ld_f, stream_));
}

// Traction of the solution columns on the fault facets only, then its projection onto the fault
// basis: the column blocks of the Green's function (SeasQDDiscreteGreenOperator.cpp:536-545).
void Operator::fault_traction_device(double const* X, double const* Fd, double* T, int ncols, int64_t ld_x,
                                     int64_t ld_f, int64_t ld_t) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    FaultArgs a = fault_args();
    if (ncols > 1) {
        if (ld_x < int64_t(local_size())) throw std::invalid_argument("ld_x is below the local size");
        if (Fd && ld_f < int64_t(num_facets()) * int64_t(D_) * nq_)
            throw std::invalid_argument("ld_f is below one column of facet data (facets*dim*nq)");
        if (ld_t < a.nfault * int64_t(a.nbf) * int64_t(D_))
            throw std::invalid_argument("ld_t is below one column of fault-basis traction (nfault*nbf*dim)");
    }
    RhsArgs r = rhs_args();
    const int64_t per = a.nfault * int64_t(D_) * nq_;
    const std::size_t need = std::size_t(per) * std::size_t(ncols);
    if (fault_->work.n < need) fault_->work.alloc(need);
    TDG_HIP(launch_elasticity_traction_subset(D_, N_, r, X, Fd, fault_->work.ptr, a.fac, a.nfault, ncols, ld_x, ld_f,
                                              per, stream_));
    TDG_HIP(launch_fault_project(a, fault_->work.ptr, T, ncols, per, ld_t, stream_));
}

BlockJacobi& Operator::block_jacobi() {
    if (!bj_) bj_ = std::make_unique<BlockJacobi>();
    if (!bj_->built()) bj_->build(*this);
    return *bj_;
}

// Assembled block rows by probing.  Colouring: no two elements within two face-steps of each other
// share a colour, so with x = e_j on every element of colour c, y on an element of that colour is
// column j of its own block, and y on any other element meets at most one element of colour c
// among its face neighbours: column j of the block coupling it to that neighbour.
int Operator::assemble() {
    if (!ready_) throw std::logic_error("set_material() must be called before assembling");
    if (asm_) return asm_->colours;
    const unsigned NF = D_ + 1;
    const int bs = int(block_size());
    const int64_t E = int64_t(E_), n = int64_t(local_size());
    std::vector<int32_t> nb(std::size_t(NF) * E_);
    for (std::size_t e = 0; e < E_; ++e)
        for (unsigned f = 0; f < NF; ++f) {
            const int64_t o = topo_.sides[e * NF + f].nb;
            nb[std::size_t(f) * E_ + e] = int32_t(o);
            for (unsigned g = 0; g < f; ++g)
                if (o != int64_t(e) && nb[std::size_t(g) * E_ + e] == int32_t(o))
                    throw std::runtime_error("assembly: element " + std::to_string(e) +
                                             " meets one neighbour through two faces");
        }
    std::vector<int32_t> col(E_, -1);
    int nc = 0;
    std::vector<char> used;
    for (std::size_t e = 0; e < E_; ++e) {
        used.assign(std::size_t(nc) + 1, 0);
        for (unsigned f = 0; f < NF; ++f) {
            const int64_t o = nb[std::size_t(f) * E_ + e];
            if (o == int64_t(e)) continue;
            if (col[std::size_t(o)] >= 0) used[std::size_t(col[std::size_t(o)])] = 1;
            for (unsigned g = 0; g < NF; ++g) {
                const int64_t p = nb[std::size_t(g) * E_ + std::size_t(o)];
                if (p != int64_t(e) && col[std::size_t(p)] >= 0) used[std::size_t(col[std::size_t(p)])] = 1;
            }
        }
        int c = 0;
        while (used[std::size_t(c)]) ++c;
        col[e] = c;
        nc = std::max(nc, c + 1);
    }
    auto a = std::make_unique<DeviceAssembled>();
    a->colours = nc;
    a->nb.upload(nb);
    DeviceArray<int32_t> dcol;
    dcol.upload(col);
    const std::size_t entries = std::size_t(NF + 1) * std::size_t(bs) * std::size_t(bs) * E_;
    a->A64.alloc(entries);
    TDG_HIP(hipMemsetAsync(a->A64.ptr, 0, entries * sizeof(double), stream_));
    // probe columns per batched apply: at most 1 GiB for each of X and Y
    const int chunk = int(std::max<int64_t>(1, std::min<int64_t>(bs, (int64_t(1) << 30) / (n * 8))));
    DeviceArray<double> X, Y;
    X.alloc(std::size_t(chunk) * std::size_t(n));
    Y.alloc(std::size_t(chunk) * std::size_t(n));
    const int saved = variant_;
    if (variant_ == Variant_Assembled) variant_ = Variant_Auto; // probe the matrix-free apply
    try {
        for (int c = 0; c < nc; ++c)
            for (int c0 = 0; c0 < bs; c0 += chunk) {
                const int m = std::min(chunk, bs - c0);
                TDG_HIP(launch_assembly_probe(X.ptr, dcol.ptr, c, c0, m, E, bs, stream_));
                dispatch(static_cast<double const*>(X.ptr), Y.ptr, m, n);
                TDG_HIP(launch_assembly_gather(Y.ptr, dcol.ptr, a->nb.ptr, a->A64.ptr, c, c0, m, E, bs, int(NF),
                                               stream_));
            }
    } catch (...) {
        variant_ = saved;
        throw;
    }
    variant_ = saved;
    a->A32.alloc(entries);
    TDG_HIP(launch_narrow(a->A64.ptr, a->A32.ptr, int64_t(entries), stream_));
    TDG_HIP(hipStreamSynchronize(stream_));
    asm_ = std::move(a);
    return nc;
}

void Operator::assembled_block(int64_t e, int slot, double* out) const {
    if (!asm_) throw std::logic_error("the operator has not been assembled");
    if (e < 0 || e >= int64_t(E_) || slot < 0 || slot > int(D_ + 1)) throw std::out_of_range("no such block");
    const std::size_t bs = block_size();
    TDG_HIP(hipStreamSynchronize(stream_));
    // entries (slot, i, j) of element e lie E apart
    TDG_HIP(hipMemcpy2D(out, sizeof(double), asm_->A64.ptr + (std::size_t(slot) * bs * bs) * E_ + std::size_t(e),
                        E_ * sizeof(double), sizeof(double), bs * bs, hipMemcpyDeviceToHost));
}

bool Operator::assembled_fast() const { return assembled_register_kernel(int(block_size()), int(D_ + 2)); }

void Operator::assembled_host(std::vector<double>& A, std::vector<int32_t>& nb) {
    assemble();
    TDG_HIP(hipStreamSynchronize(stream_));
    A.resize(asm_->A64.n);
    nb.resize(asm_->nb.n);
    TDG_HIP(hipMemcpy(A.data(), asm_->A64.ptr, A.size() * sizeof(double), hipMemcpyDeviceToHost));
    TDG_HIP(hipMemcpy(nb.data(), asm_->nb.ptr, nb.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
}

template <class T> AssembledArgs<T> Operator::assembled_args() const {
    AssembledArgs<T> a{};
    if constexpr (std::is_same<T, double>::value) a.A = asm_->A64.ptr;
    else a.A = asm_->A32.ptr;
    a.nb = asm_->nb.ptr;
    a.num_elements = int64_t(E_);
    a.block_size = int(block_size());
    a.slots = int(D_ + 2);
    return a;
}

void Operator::traction_device(double const* x, double const* facet_q, double* t) {
    if (!ready_) throw std::logic_error("set_material() must be called first");
    RhsArgs a = rhs_args();
    TDG_HIP(launch_elasticity_traction(D_, N_, a, x, facet_q, t, stream_));
}

std::vector<double> const* Operator::host_array(std::string const& n) const {
    if (n == "absDetJ") return &absDetJ_;
    if (n == "JInv") return &JInv_;
    if (n == "vcoords") return &vcoords_;
    if (n == "volume") return &volume_;
    if (n == "JInv0") return &fJInv_[0];
    if (n == "JInv1") return &fJInv_[1];
    if (n == "normal") return &normal_;
    if (n == "normalLength") return &nlen_;
    if (n == "fcoords") return &fcoords_;
    if (n == "area") return &area_;
    if (n == "penalty") return &penalty_;
    if (n == "mat0") return &mat0_;
    if (n == "mat1") return &mat1_;
    if (n == "c0Q") return &c0Q_;
    if (n == "c1Q") return &c1Q_;
    if (n == "fc0") return &fc0_;
    if (n == "fc1") return &fc1_;
    if (n == "E_Q") return &E_Q_;
    if (n == "Dxi_Q") return &Dxi_Q_;
    if (n.size() == 4 && n.compare(0, 3, "E_q") == 0 && n[3] >= '0' && n[3] < char('0' + E_q_.size()))
        return &E_q_[std::size_t(n[3] - '0')];
    if (n == "volWeights") return &vol_.weights;
    if (n == "facetWeights") return &fct_.weights;
    return nullptr;
}

// Algorithmic bytes of one apply: x and y once, per volume point JInv + 2 coefficients, per facet
// both sides' JInv, normal and unit normal, both sides' 2 coefficients and the penalty.
double Operator::apply_bytes(int pb) const {
    double D = D_, E = double(E_), F = double(topo_.facets.size());
    double nqv = Nq_, nqf = nq_, bs = double(block_size());
    double c = kind_ == Kind_Elasticity ? 2.0 : 1.0;
    return pb * (E * (2.0 * bs + nqv * (D * D + c)) + F * (2.0 * D * D * nqf + 2.0 * D * nqf + 2.0 * c * nqf + 1.0));
}

// Algorithmic flops: volume gradient/divergence GEMMs + per-point transforms, per element face
// own value/gradient traces and the lift, and one flux combination per facet.
double Operator::apply_flops() const {
    double D = D_, E = double(E_), F = double(topo_.facets.size());
    double Nq = Nq_, nq = nq_, Nbf = Nbf_, Q = Q_, nf = D + 1;
    double vol = 4 * D * Nq * Nbf * Q + 4 * D * D * Nq * Q + (3 * D * D + 2 * D) * Nq;
    double face = 2 * nq * Nbf * Q + (2 * D * nq * Nbf * Q + 2 * D * D * nq * Q) +
                  (2 * Nbf * D * nq * Q + 2 * D * D * nq * Q + 2 * Nbf * nq * Q);
    return E * vol + nf * E * face + F * (6 * D * D + 6 * D) * nq;
}

} // namespace tdg

namespace tdg {

template <class T>
void Operator::launch(DeviceOperator<T> const& d, T const* x, T* y, int ncols, int64_t ld) {
    if (!ready_) throw std::logic_error("set_material() must be called before apply");
    ApplyArgs<T> a;
    a.tabV = d.tabV.ptr;
    a.tabVt = d.tabVt.ptr;
    a.tabF = d.tabF.ptr;
    a.tabFt = d.tabFt.ptr;
    a.w = d.w.ptr;
    a.perm = d.perm.ptr;
    a.iperm = d.iperm.ptr;
    a.G = d.G.ptr;
    a.c0 = d.c0.ptr;
    a.c1 = d.c1.ptr;
    a.fG = d.fG.ptr;
    a.fc0 = d.fc0.ptr;
    a.fc1 = d.fc1.ptr;
    a.fn = d.fn.ptr;
    a.fpen = d.fpen.ptr;
    a.ef = reinterpret_cast<int4 const*>(d.ef.ptr);
    a.num_elements = static_cast<int64_t>(E_);
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    hipError_t err = kind_ == Kind_Elasticity ? launch_elasticity<T>(D_, N_, a, x, y, ncols, ld, stream_)
                                              : launch_poisson<T>(D_, N_, a, x, y, ncols, ld, stream_);
    TDG_HIP(err);
}

template <> DeviceOperator<double> const& Operator::device<double>() {
    if (!d64_ || !d32_) upload_device();
    return *d64_;
}
template <> DeviceOperator<float> const& Operator::device<float>() {
    if (!d64_ || !d32_) upload_device();
    return *d32_;
}

template <class T> MfmaArgs<T> Operator::mfma_args(DeviceMfma<T> const& d) const {
    MfmaArgs<T> b;
    b.V = d.V.ptr;
    b.T0 = d.T0.ptr;
    b.perm = d.perm2.ptr;
    b.w = d.w.ptr;
    b.VD = d.VD.ptr;
    b.FD = d.FD.ptr;
    b.FP = d.FP.ptr;
    b.FI = reinterpret_cast<int4 const*>(d.FI.ptr);
    b.num_elements = int64_t(E_);
    b.num_groups = int64_t((E_ + 15) / 16);
    b.xcd_remap = variant_ == Variant_MfmaXcd ? 1 : 0;
    return b;
}

template <class T> SplitArgs<T> Operator::split_args(DeviceMfma<T> const& m, DeviceSplit<T> const& d) const {
    SplitArgs<T> b;
    b.V = m.V.ptr;
    b.T0 = m.T0.ptr;
    b.perm = m.perm2.ptr;
    b.w = m.w.ptr;
    b.VD = m.VD.ptr;
    b.FR = d.FR.ptr;
    b.FPen = d.FPen.ptr;
    b.FInfo = reinterpret_cast<int4 const*>(d.FInfo.ptr);
    b.num_elements = int64_t(E_);
    b.num_groups = int64_t((E_ + 15) / 16);
    b.num_facets = d.nfacets;
    b.num_fgroups = (d.nfacets + 7) / 8;
    b.num_two = d.ntwo;
    b.xcd_remap = 0;
    b.dot = nullptr;
    return b;
}

template <class T>
void Operator::launch_mfma(DeviceMfma<T> const& d, T const* x, T* y, int ncols, int64_t ld) {
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    TDG_HIP(launch_elasticity_mfma<T>(D_, N_, mfma_args(d), x, y, ncols, ld, stream_));
}

template <class T> void Operator::dispatch(T const* x, T* y, int ncols, int64_t ld) {
    if (!ready_) throw std::logic_error("set_material() must be called before apply");
    if (variant_ == Variant_Assembled) {
        if (!asm_) assemble();
        TDG_HIP(launch_assembled<T>(assembled_args<T>(), x, y, ncols, ld, stream_));
        return;
    }
    upload_device();
    if (uses_split()) {
        int kq = 0, kf = 0;
        if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
            throw std::runtime_error("no kernel compiled for this dimension/degree");
        auto& ds = [&]() -> DeviceSplit<T>& {
            if constexpr (std::is_same<T, double>::value) return *s64_;
            else return *s32_;
        }();
        auto const& dm = [&]() -> DeviceMfma<T> const& {
            if constexpr (std::is_same<T, double>::value) return *m64_;
            else return *m32_;
        }();
        reserve_split(ds, ncols);
        const int64_t ldL = int64_t(E_) * (D_ + 1) * int64_t(block_size());
        const int p = sizeof(T) == 8 ? 0 : 1;
        SplitArgs<T> sa = split_args(dm, ds);
        sa.xcd_remap = split_xcd_[p];
        TDG_HIP(launch_elasticity_split<T>(D_, N_, split_shape_[p], sa, x, y, ds.L.ptr, ncols, ncols > 1 ? ld : 0,
                                           ldL, stream_));
    } else if (uses_mfma()) {
        if constexpr (std::is_same<T, double>::value) launch_mfma(*m64_, x, y, ncols, ld);
        else launch_mfma(*m32_, x, y, ncols, ld);
    } else {
        launch(device<T>(), x, y, ncols, ld);
    }
}

void Operator::apply(double const* x, double* y) { dispatch(x, y, 1, 0); }

int64_t Operator::dot_partials_capacity() const {
    if (!(kind_ == Kind_Elasticity && mfma_supported(D_, N_))) return 0;
    const int64_t ng = int64_t((E_ + 15) / 16);
    int64_t m = 0;
    for (int sh = 0; sh < kSplitShapes; ++sh)
        m = std::max({m, split_element_workgroups<double>(sh, ng), split_element_workgroups<float>(sh, ng)});
    return m;
}

template <class T> int64_t Operator::apply_dot_impl(T const* x, T* y, double* partials) {
    if (!ready_) throw std::logic_error("set_material() must be called before apply");
    upload_device();
    if (!uses_split()) {
        dispatch(x, y, 1, 0);
        return 0;
    }
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    auto& ds = [&]() -> DeviceSplit<T>& {
        if constexpr (std::is_same<T, double>::value) return *s64_;
        else return *s32_;
    }();
    auto const& dm = [&]() -> DeviceMfma<T> const& {
        if constexpr (std::is_same<T, double>::value) return *m64_;
        else return *m32_;
    }();
    const int p = sizeof(T) == 8 ? 0 : 1;
    SplitArgs<T> sa = split_args(dm, ds);
    sa.xcd_remap = split_xcd_[p];
    sa.dot = partials;
    const int64_t ldL = int64_t(E_) * (D_ + 1) * int64_t(block_size());
    TDG_HIP(launch_elasticity_split<T>(D_, N_, split_shape_[p], sa, x, y, ds.L.ptr, 1, 0, ldL, stream_));
    return split_element_workgroups<T>(split_shape_[p], int64_t((E_ + 15) / 16));
}

int64_t Operator::apply_dot(double const* x, double* y, double* partials) { return apply_dot_impl(x, y, partials); }
int64_t Operator::apply_dot(float const* x, float* y, double* partials) { return apply_dot_impl(x, y, partials); }

std::size_t Operator::slot_entries() const {
    return uses_split() ? E_ * (D_ + 1) * block_size() : 0;
}

void Operator::apply_noalloc(double const* x, double* y, double* scratch, std::size_t cap) {
    if (!ready_) throw std::logic_error("set_material() must be called before apply");
    if (variant_ == Variant_Assembled) { // no scratch memory; the blocks must already be assembled
        if (!asm_) throw std::logic_error("the allocation-free apply needs the assembled operator: call tdg_assemble "
                                          "(or one tdg_mult) first");
        TDG_HIP(launch_assembled<double>(assembled_args<double>(), x, y, 1, 0, stream_));
        return;
    }
    // everything the selected kernels read must already be on the device: nothing is uploaded here
    const bool have = d64_ && (!uses_mfma() || (m64_ && (uses_split() ? bool(s64_) : m64_->FD.n > 0)));
    if (!have)
        throw std::logic_error("the allocation-free apply needs the operator on the device: call tdg_upload "
                               "(or one tdg_mult) first");
    if (!uses_split()) { // the other kernels use no scratch memory
        dispatch(x, y, 1, 0);
        return;
    }
    const std::size_t need = slot_entries();
    if (need > cap)
        throw std::invalid_argument("the allocation-free apply holds at most " + std::to_string(cap) +
                                    " facet-slot entries; this operator needs " + std::to_string(need) +
                                    " (use tdg_mult)");
    int kq = 0, kf = 0;
    if (kernel_rule_sizes(D_, N_, &kq, &kf) != 0 || kq != int(Nq_) || kf != int(nq_))
        throw std::runtime_error("no kernel compiled for this dimension/degree");
    TDG_HIP(launch_zero_slots<double>(scratch, s64_->idle.ptr, int64_t(s64_->idle.n), int(block_size()), stream_));
    SplitArgs<double> sa = split_args(*m64_, *s64_);
    sa.xcd_remap = split_xcd_[0];
    TDG_HIP(launch_elasticity_split<double>(D_, N_, split_shape_[0], sa, x, y, scratch, 1, 0, int64_t(need), stream_));
}
void Operator::apply(float const* x, float* y) { dispatch(x, y, 1, 0); }
void Operator::apply_batched(double const* X, double* Y, int ncols, int64_t ld) { dispatch(X, Y, ncols, ld); }
void Operator::apply_batched(float const* X, float* Y, int ncols, int64_t ld) { dispatch(X, Y, ncols, ld); }

template <class T> bool Operator::apply_batched_dot_impl(T const* X, T* Y, int ncols, int64_t ld, double* partial) {
    if (!ready_) throw std::logic_error("set_material() must be called before apply");
    if (variant_ != Variant_Assembled || !assembled_fast()) return false;
    if (!asm_) assemble();
    const hipError_t e = launch_assembled<T>(assembled_args<T>(), X, Y, ncols, ld, stream_, partial);
    if (e == hipErrorNotSupported) return false; // no fused kernel for this block size and precision
    TDG_HIP(e);
    return true;
}
bool Operator::apply_batched_dot(double const* X, double* Y, int ncols, int64_t ld, double* partial) {
    return apply_batched_dot_impl(X, Y, ncols, ld, partial);
}
bool Operator::apply_batched_dot(float const* X, float* Y, int ncols, int64_t ld, double* partial) {
    return apply_batched_dot_impl(X, Y, ncols, ld, partial);
}

void Operator::material_gradient(double const* u, double const* psi, double* g_lam, double* g_mu) {
    grad_dispatch(u, psi, g_lam, g_mu);
}
void Operator::material_gradient(float const* u, float const* psi, float* g_lam, float* g_mu) {
    grad_dispatch(u, psi, g_lam, g_mu);
}

} // namespace tdg
