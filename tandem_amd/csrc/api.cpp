// extern "C" boundary of the MI355X DG operator (declared in include/tandem_dg.h).
#include "../../include/tandem_dg.h"

#include "abi.h"
#include "host/mesh.h"
#include "host/operator.h"
#include "host/refelem.h"
#include "kernels/mfma_layout.h"
#include "solver/krylov.h"
#include "solver/block_jacobi.h"
#include "solver/pmg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

using namespace tdg_abi;

extern "C" {

char const* tdg_last_error(void) { return g_error.c_str(); }

int tdg_create(tdg_mesh_desc const* m, int kind, tdg_op** op) {
    return guarded([&] {
        require(m && op, "null argument");
        require(m->dim == 2 || m->dim == 3, "dim must be 2 or 3");
        require(m->num_vertices > 0 && m->vertices, "no vertices");
        require(m->num_elements > 0 && m->elements, "no elements");
        require(m->num_boundary_facets == 0 || (m->boundary_facets && m->boundary_bc), "boundary arrays missing");
        tdg::SimplexMesh mesh;
        mesh.dim = m->dim;
        mesh.vertices.assign(m->vertices, m->vertices + m->num_vertices * m->dim);
        mesh.elements.assign(m->elements, m->elements + m->num_elements * (m->dim + 1));
        for (int64_t e = 0; e < m->num_elements; ++e)
            std::sort(mesh.elements.begin() + e * (m->dim + 1), mesh.elements.begin() + (e + 1) * (m->dim + 1));
        mesh.bnd_facets.assign(m->boundary_facets, m->boundary_facets + m->num_boundary_facets * m->dim);
        mesh.bnd_bc.assign(m->boundary_bc, m->boundary_bc + m->num_boundary_facets);
        for (int b : mesh.bnd_bc)
            require(b == TDG_BC_NONE || b == TDG_BC_NATURAL || b == TDG_BC_FAULT || b == TDG_BC_DIRICHLET,
                    "unknown boundary condition");
        std::vector<double> nodes;
        if (m->geometry_nodes) {
            std::size_t n = m->num_elements * tdg::num_basis(m->dim, m->degree) * m->dim;
            nodes.assign(m->geometry_nodes, m->geometry_nodes + n);
        }
        auto h = std::make_unique<tdg_op>();
        h->impl = std::make_unique<tdg::Operator>(std::move(mesh), kind, m->degree,
                                                  m->geometry_nodes ? &nodes : nullptr);
        *op = h.release();
    });
}

int tdg_destroy(tdg_op* op) {
    delete op;
    return TDG_OK;
}

int tdg_sizes(tdg_op const* op, int64_t* out) {
    return guarded([&] {
        require(op && op->impl && out, "null argument");
        auto const& o = *op->impl;
        out[0] = int64_t(o.block_size());
        out[1] = int64_t(o.num_elements());
        out[2] = int64_t(o.local_size());
        out[3] = int64_t(o.num_facets());
        out[4] = o.nq_vol();
        out[5] = o.nq_fct();
        out[6] = o.dim();
    });
}

int tdg_block_size(tdg_op const* op, int64_t* bs) {
    return guarded([&] {
        require(op && op->impl && bs, "null argument");
        *bs = int64_t(op->impl->block_size());
    });
}

int tdg_local_size(tdg_op const* op, int64_t* n) {
    return guarded([&] {
        require(op && op->impl && n, "null argument");
        *n = int64_t(op->impl->local_size());
    });
}

int tdg_volume_points(tdg_op const* op, double* coords) {
    return guarded([&] {
        require(op && op->impl && coords, "null argument");
        auto const& v = op->impl->volume_points();
        std::memcpy(coords, v.data(), v.size() * sizeof(double));
    });
}

int tdg_facet_points(tdg_op const* op, double* coords) {
    return guarded([&] {
        require(op && op->impl && coords, "null argument");
        auto const& v = op->impl->facet_points();
        std::memcpy(coords, v.data(), v.size() * sizeof(double));
    });
}

int tdg_facets(tdg_op const* op, int64_t* out) {
    return guarded([&] {
        require(op && op->impl && out, "null argument");
        auto const& t = op->impl->topology();
        for (std::size_t i = 0; i < t.facets.size(); ++i) {
            auto const& f = t.facets[i];
            out[5 * i + 0] = f.elem[0];
            out[5 * i + 1] = f.elem[1];
            out[5 * i + 2] = f.local[0];
            out[5 * i + 3] = f.local[1];
            out[5 * i + 4] = f.bc;
        }
    });
}

int tdg_set_material(tdg_op* op, double const* c0_Q, double const* c1_Q) {
    return guarded([&] {
        require(op && op->impl, "null operator");
        op->impl->set_material(c0_Q, c1_Q);
    });
}

int tdg_upload(tdg_op* op) {
    return guarded([&] {
        ready(op).upload_device();
        prepare_noalloc_scratch();
    });
}

int tdg_mult(tdg_op* op, double const* x, double* y) {
    return guarded([&] {
        auto& o = ready(op);
        require(x && y && x != y, "x and y must be distinct device buffers");
        o.apply(x, y);
    });
}

int tdg_mult_f32(tdg_op* op, float const* x, float* y) {
    return guarded([&] {
        auto& o = ready(op);
        require(x && y && x != y, "x and y must be distinct device buffers");
        o.apply(x, y);
    });
}

int tdg_mult_batched(tdg_op* op, double const* X, double* Y, int ncols, int64_t ld) {
    return guarded([&] {
        auto& o = ready(op);
        require(X && Y && X != Y && ncols > 0 && ncols <= 65535 && ld >= int64_t(o.local_size()),
                "bad batched arguments");
        o.apply_batched(X, Y, ncols, ld);
    });
}

int tdg_mult_batched_f32(tdg_op* op, float const* X, float* Y, int ncols, int64_t ld) {
    return guarded([&] {
        auto& o = ready(op);
        require(X && Y && X != Y && ncols > 0 && ncols <= 65535 && ld >= int64_t(o.local_size()),
                "bad batched arguments");
        o.apply_batched(X, Y, ncols, ld);
    });
}

int tdg_set_stream(tdg_op* op, void* stream) {
    return guarded([&] {
        require(op && op->impl, "null operator");
        op->impl->set_stream(static_cast<hipStream_t>(stream));
    });
}

int tdg_set_variant(tdg_op* op, int variant) {
    return guarded([&] {
        require(op && op->impl, "null operator");
        require(variant >= 0 && variant < TDG_NUM_VARIANTS, "unknown kernel variant");
        op->impl->set_variant(variant);
    });
}

int tdg_rhs(tdg_op const* op, double const* force_Q, double const* facet_q, double* b) {
    return guarded([&] {
        require(op && op->impl && b, "null argument");
        op->impl->rhs(force_Q, facet_q, b);
    });
}

int tdg_rhs_device(tdg_op* op, double const* force_Q, double const* facet_q, double* b, int ncols,
                   int64_t ld_force, int64_t ld_facet, int64_t ld_b) {
    return guarded([&] {
        auto& o = ready(op);
        require(b && ncols >= 1, "bad rhs arguments");
        const int64_t E = int64_t(o.num_elements()), F = int64_t(o.num_facets());
        const int64_t Q = int64_t(o.block_size() / o.nbf());
        require(ncols == 1 || (ld_b >= int64_t(o.local_size()) &&
                               (!force_Q || ld_force >= E * Q * int64_t(o.nq_vol())) &&
                               (!facet_q || ld_facet >= F * Q * int64_t(o.nq_fct()))),
                "leading dimensions are smaller than one column");
        o.rhs_device(force_Q, facet_q, b, ncols, ld_force, ld_facet, ld_b);
    });
}

int tdg_traction(tdg_op* op, double const* x, double const* facet_q, double* t) {
    return guarded([&] {
        auto& o = ready(op);
        require(x && t, "null argument");
        o.traction_device(x, facet_q, t);
    });
}

static tdg::SolveOptions solve_options(tdg_solve_opts const* opts) {
    tdg::SolveOptions so;
    so.rtol = opts->rtol;
    so.max_iter = opts->max_iter;
    so.mode = opts->mode;
    so.inner_rtol = opts->inner_rtol;
    so.max_outer = opts->max_outer;
    so.check_every = opts->check_every;
    so.reliable_delta = opts->reliable_delta;
    so.precond = opts->precond;
    so.graphs = opts->graphs;
    require(so.precond == TDG_PRECOND_NONE || so.precond == TDG_PRECOND_BLOCK_JACOBI ||
                so.precond == TDG_PRECOND_BLOCK_JACOBI_F32,
            "unknown preconditioner");
    require(so.mode == TDG_SOLVE_FP64_CG || so.mode == TDG_SOLVE_FP32_CG_FP64_IR ||
                so.mode == TDG_SOLVE_FP32_CG_FP64_RELIABLE,
            "unknown solve mode");
    require(so.rtol > 0 && so.max_iter > 0 && so.check_every > 0 && so.max_outer > 0 &&
                (so.mode != TDG_SOLVE_FP32_CG_FP64_IR || (so.inner_rtol > 0 && so.inner_rtol < 1)) &&
                (so.mode != TDG_SOLVE_FP32_CG_FP64_RELIABLE || (so.reliable_delta > 0 && so.reliable_delta < 1)),
            "bad solver options");
    return so;
}

int tdg_solve(tdg_op* op, double const* b, double* x, tdg_solve_opts const* opts, tdg_solve_result* result) {
    return guarded([&] {
        auto& o = ready(op);
        require(b && x && opts && result && b != x, "bad solve arguments");
        auto r = tdg::solve(o, b, x, solve_options(opts));
        result->iterations = r.iterations;
        result->outer_iterations = r.outer_iterations;
        result->inner_iterations = r.inner_iterations;
        result->rel_residual = r.rel_residual;
        result->seconds = r.seconds;
        result->converged = r.converged;
    });
}

int tdg_solve_multi(tdg_op* op, double const* B, double* X, int ncols, int64_t ld, tdg_solve_opts const* opts,
                    tdg_solve_result* result, int* iterations, double* rel_residual) {
    return guarded([&] {
        auto& o = ready(op);
        require(B && X && opts && result && B != X && ncols >= 1 && ld >= int64_t(o.local_size()),
                "bad multi-column solve arguments");
        auto so = solve_options(opts);
        require(so.mode != TDG_SOLVE_FP32_CG_FP64_IR, "multi-column solves support modes 0 and 2");
        auto r = tdg::solve_multi(o, B, X, ncols, ld, so);
        result->iterations = r.total_iterations;
        result->outer_iterations = r.replacements;
        result->inner_iterations = r.total_iterations;
        double worst = 0.0;
        for (int c = 0; c < ncols; ++c) {
            worst = std::max(worst, r.rel_residual[c]);
            if (iterations) iterations[c] = r.iterations[c];
            if (rel_residual) rel_residual[c] = r.rel_residual[c];
        }
        result->rel_residual = worst;
        result->seconds = r.seconds;
        result->converged = r.converged;
    });
}

struct tdg_pmg {
    std::unique_ptr<tdg::PMG> impl;
};

void tdg_pmg_default_opts(tdg_pmg_opts* opts) {
    if (!opts) return;
    tdg::PMGOptions d;
    opts->smooth = d.smooth;
    opts->eig_lo = d.eig_lo;
    opts->eig_hi = d.eig_hi;
    opts->eig_iters = d.eig_iters;
    opts->coarse_rtol = d.coarse_rtol;
    opts->coarse_max_iter = d.coarse_max_iter;
    opts->galerkin_penalty = d.galerkin_penalty;
    opts->coarse_fp32 = d.coarse_fp32;
    opts->vcycle_fp32 = d.vcycle_fp32;
    opts->coarse_assembled = d.coarse_assembled;
    opts->coarse_amg = d.coarse_amg;
}

int tdg_pmg_create(tdg_op* const* levels, int nlevels, tdg_pmg_opts const* opts, tdg_pmg** out) {
    return guarded([&] {
        require(levels && out && nlevels >= 2, "bad p-multigrid arguments");
        *out = nullptr;
        std::vector<tdg::Operator*> ops;
        for (int l = 0; l < nlevels; ++l) ops.push_back(&ready(levels[l]));
        tdg::PMGOptions o;
        if (opts) {
            o.smooth = opts->smooth;
            o.eig_lo = opts->eig_lo;
            o.eig_hi = opts->eig_hi;
            o.eig_iters = opts->eig_iters;
            o.coarse_rtol = opts->coarse_rtol;
            o.coarse_max_iter = opts->coarse_max_iter;
            o.galerkin_penalty = opts->galerkin_penalty;
            o.coarse_fp32 = opts->coarse_fp32;
            o.vcycle_fp32 = opts->vcycle_fp32;
            o.coarse_assembled = opts->coarse_assembled;
            o.coarse_amg = opts->coarse_amg;
        }
        auto h = std::make_unique<tdg_pmg>();
        h->impl = std::make_unique<tdg::PMG>(ops, o);
        *out = h.release();
    });
}

int tdg_pmg_destroy(tdg_pmg* pmg) {
    return guarded([&] { delete pmg; });
}

int tdg_pmg_setup(tdg_pmg* pmg, double* lambda_max) {
    return guarded([&] {
        require(pmg && pmg->impl, "null p-multigrid hierarchy");
        pmg->impl->setup();
        if (lambda_max)
            for (std::size_t l = 0; l < pmg->impl->lambda_max().size(); ++l) lambda_max[l] = pmg->impl->lambda_max()[l];
    });
}

int tdg_pmg_vcycle(tdg_pmg* pmg, double const* r, double* z, int ncols, int64_t ld) {
    return guarded([&] {
        require(pmg && pmg->impl && r && z && r != z, "bad V-cycle arguments");
        pmg->impl->vcycle(r, z, ncols, ld);
    });
}

int tdg_solve_multi_pmg(tdg_pmg* pmg, double const* B, double* X, int ncols, int64_t ld, tdg_solve_opts const* opts,
                        tdg_solve_result* result, int* iterations, double* rel_residual) {
    return guarded([&] {
        require(pmg && pmg->impl && B && X && opts && result && B != X, "bad p-multigrid solve arguments");
        auto so = solve_options(opts);
        const int64_t before = pmg->impl->coarse_iterations();
        auto r = tdg::solve_multi_pmg(*pmg->impl, B, X, ncols, ld, so);
        result->iterations = r.total_iterations;
        result->outer_iterations = int(pmg->impl->coarse_iterations() - before);
        result->inner_iterations = r.total_iterations;
        double worst = 0.0;
        for (int c = 0; c < ncols; ++c) {
            worst = std::max(worst, r.rel_residual[c]);
            if (iterations) iterations[c] = r.iterations[c];
            if (rel_residual) rel_residual[c] = r.rel_residual[c];
        }
        result->rel_residual = worst;
        result->seconds = r.seconds;
        result->converged = r.converged;
    });
}

int tdg_autotune(tdg_op* op, int ncols, int precision_bytes, int* chosen, double* ms) {
    return guarded([&] {
        auto& o = ready(op);
        require(ncols >= 1 && (precision_bytes == 8 || precision_bytes == 4), "bad autotune arguments");
        auto t = o.autotune(ncols, precision_bytes);
        if (chosen) *chosen = o.variant();
        if (ms)
            for (int v = 0; v < TDG_NUM_VARIANTS; ++v) ms[v] = t[std::size_t(v)];
    });
}

int tdg_autotune_shapes(tdg_op* op, int ncols, int precision_bytes, int* shape, int* xcd, double* ms) {
    return guarded([&] {
        auto& o = ready(op);
        require(ncols >= 1 && (precision_bytes == 8 || precision_bytes == 4), "bad autotune arguments");
        static_assert(TDG_SPLIT_SHAPES == tdg::kSplitShapes, "shape count of the ABI and the library differ");
        static_assert(TDG_SPLIT_ORDERS == tdg::kSplitOrders, "order count of the ABI and the library differ");
        auto t = o.autotune_shapes(ncols, precision_bytes);
        if (shape) *shape = o.split_shape(precision_bytes);
        if (xcd) *xcd = o.split_xcd(precision_bytes);
        if (ms)
            for (int c = 0; c < TDG_SPLIT_ORDERS * TDG_SPLIT_SHAPES; ++c) ms[c] = t[std::size_t(c)];
    });
}

int tdg_set_shape(tdg_op* op, int precision_bytes, int shape, int xcd) {
    return guarded([&] {
        require(op && op->impl, "null operator");
        require((precision_bytes == 8 || precision_bytes == 4) && shape >= 0 && shape < TDG_SPLIT_SHAPES &&
                    xcd >= 0 && xcd < TDG_SPLIT_ORDERS,
                "bad launch shape");
        op->impl->set_split_shape(precision_bytes, shape, xcd);
    });
}

int tdg_solve_segments(tdg_op* op, double const* b, double* x, int nseg, int64_t const* offsets,
                       tdg_solve_opts const* opts, tdg_solve_result* result, int* iterations, double* rel_residual) {
    return guarded([&] {
        auto& o = ready(op);
        require(b && x && opts && result && offsets && b != x && nseg >= 1, "bad segmented solve arguments");
        auto so = solve_options(opts);
        require(so.mode != TDG_SOLVE_FP32_CG_FP64_IR, "segmented solves support modes 0 and 2");
        auto r = tdg::solve_segments(o, b, x, nseg, offsets, so);
        result->iterations = r.total_iterations;
        result->outer_iterations = r.replacements;
        result->inner_iterations = r.total_iterations;
        double worst = 0.0;
        for (int c = 0; c < nseg; ++c) {
            worst = std::max(worst, r.rel_residual[c]);
            if (iterations) iterations[c] = r.iterations[c];
            if (rel_residual) rel_residual[c] = r.rel_residual[c];
        }
        result->rel_residual = worst;
        result->seconds = r.seconds;
        result->converged = r.converged;
    });
}

int tdg_block_jacobi_build(tdg_op* op, int* colours) {
    return guarded([&] {
        auto& bj = ready(op).block_jacobi();
        if (colours) *colours = bj.colours();
    });
}

int tdg_block_jacobi_apply(tdg_op* op, double const* r, double* z) {
    return guarded([&] {
        auto& o = ready(op);
        require(r && z && r != z, "bad block Jacobi arguments");
        o.block_jacobi().apply(r, z, o.stream());
    });
}

int tdg_block_jacobi_apply_batched(tdg_op* op, double const* R, double* Z, int ncols, int64_t ld) {
    return guarded([&] {
        auto& o = ready(op);
        require(R && Z && R != Z && ncols >= 1 && ld >= int64_t(o.local_size()), "bad block Jacobi arguments");
        o.block_jacobi().apply(R, Z, o.stream(), ncols, ld);
    });
}

int tdg_block_jacobi_inverse(tdg_op* op, int64_t element, double* block) {
    return guarded([&] {
        auto& o = ready(op);
        require(block, "null argument");
        o.block_jacobi().inverse_block(element, block);
    });
}

int tdg_assemble(tdg_op* op, int* colours) {
    return guarded([&] {
        const int c = ready(op).assemble();
        if (colours) *colours = c;
    });
}

int tdg_assembled_block(tdg_op* op, int64_t element, int slot, double* block) {
    return guarded([&] {
        auto& o = ready(op);
        require(block, "null argument");
        require(element >= 0 && element < int64_t(o.num_elements()) && slot >= 0 && slot <= int(o.dim()) + 1,
                "no such block");
        o.assemble();
        o.assembled_block(element, slot, block);
    });
}

int tdg_material_gradient(tdg_op* op, double const* u, double const* psi, double* g_lam, double* g_mu) {
    return guarded([&] {
        require(u && psi && g_lam && g_mu, "null argument");
        ready(op).material_gradient(u, psi, g_lam, g_mu);
    });
}

int tdg_material_gradient_f32(tdg_op* op, float const* u, float const* psi, float* g_lam, float* g_mu) {
    return guarded([&] {
        require(u && psi && g_lam && g_mu, "null argument");
        ready(op).material_gradient(u, psi, g_lam, g_mu);
    });
}

int tdg_fault_setup(tdg_op* op, double const* up, double const* ref_normal, int64_t* sizes) {
    return guarded([&] {
        require(op && op->impl && up && ref_normal && sizes, "null argument");
        auto r = op->impl->fault_setup(up, ref_normal);
        sizes[0] = r[0];
        sizes[1] = r[1];
    });
}

int tdg_fault_slip(tdg_op* op, double const* S, double* F, int ncols, int64_t ld_s, int64_t ld_f) {
    return guarded([&] {
        auto& o = ready(op);
        require(S && F && ncols >= 1, "bad fault slip arguments");
        o.fault_slip_device(S, F, ncols, ld_s, ld_f);
    });
}

int tdg_fault_traction(tdg_op* op, double const* X, double const* Fd, double* T, int ncols, int64_t ld_x,
                       int64_t ld_f, int64_t ld_t) {
    return guarded([&] {
        auto& o = ready(op);
        require(X && T && ncols >= 1, "bad fault traction arguments");
        o.fault_traction_device(X, Fd, T, ncols, ld_x, ld_f, ld_t);
    });
}

int tdg_inverse_mass(tdg_op* op, double const* x, double* y) {
    return guarded([&] {
        require(x && y, "null argument");
        ready(op).inverse_mass_device(x, y);
    });
}

int tdg_wave_rhs(tdg_op* op, double const* x, double const* facet, double* y) {
    return guarded([&] {
        require(x && y, "null argument");
        ready(op).wave_rhs_device(x, facet, y);
    });
}

int tdg_interpolation_matrix(int dim, int to_degree, int from_degree, double* out) {
    return guarded([&] {
        require(out && (dim == 2 || dim == 3), "bad arguments");
        require(from_degree >= 1 && to_degree > from_degree && to_degree <= 8, "need 1 <= from < to <= 8");
        auto I = tdg::Operator::interpolation_matrix(unsigned(dim), unsigned(to_degree), unsigned(from_degree));
        std::memcpy(out, I.data(), I.size() * sizeof(double));
    });
}

int tdg_interpolate(double const* I, int n_to, int n_from, int num_quantities, int64_t num_elements, int transpose,
                    double const* x, double* y, void* stream) {
    return guarded([&] {
        require(I && x && y, "null argument");
        require(n_to > 0 && n_from > 0 && num_quantities > 0 && num_elements >= 0, "bad sizes");
        if (num_elements == 0) return;
        hipError_t e = tdg::launch_interpolate(I, n_to, n_from, num_quantities, num_elements, transpose ? 1 : 0, x,
                                               y, static_cast<hipStream_t>(stream));
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
    });
}

int tdg_apply_counts(tdg_op const* op, int pb, double* bytes, double* flops) {
    return guarded([&] {
        require(op && op->impl && bytes && flops, "null argument");
        *bytes = op->impl->apply_bytes(pb);
        *flops = op->impl->apply_flops();
    });
}

int tdg_host_array(tdg_op const* op, char const* name, double* out, int64_t* n) {
    return guarded([&] {
        require(op && op->impl && name && n, "null argument");
        auto v = op->impl->host_array(name);
        require(v != nullptr, "unknown array name");
        *n = int64_t(v->size());
        if (out) std::memcpy(out, v->data(), v->size() * sizeof(double));
    });
}

int tdg_box_mesh(int dim, int64_t const* n, int scheme, int const* bc, int64_t* sizes, double* vertices,
                 int64_t* elements, int64_t* boundary_facets, int* boundary_bc) {
    return guarded([&] {
        require(n && bc && sizes, "null argument");
        require(dim == 2 || dim == 3, "dim must be 2 or 3");
        std::array<uint64_t, 3> nn{1, 1, 1};
        for (int d = 0; d < dim; ++d) {
            require(n[d] > 0, "cell counts must be positive");
            nn[d] = uint64_t(n[d]);
        }
        std::array<int, 6> b{};
        for (int i = 0; i < 2 * dim; ++i) b[i] = bc[i];
        auto m = tdg::box_mesh(dim, nn, scheme, b);
        sizes[0] = int64_t(m.vertices.size() / dim);
        sizes[1] = int64_t(m.num_elements());
        sizes[2] = int64_t(m.bnd_bc.size());
        if (vertices) std::memcpy(vertices, m.vertices.data(), m.vertices.size() * sizeof(double));
        if (elements) std::memcpy(elements, m.elements.data(), m.elements.size() * sizeof(int64_t));
        if (boundary_facets) std::memcpy(boundary_facets, m.bnd_facets.data(), m.bnd_facets.size() * sizeof(int64_t));
        if (boundary_bc) std::memcpy(boundary_bc, m.bnd_bc.data(), m.bnd_bc.size() * sizeof(int));
    });
}

int tdg_reference_nodes(int dim, int degree, double* out) {
    return guarded([&] {
        require(out && (dim == 2 || dim == 3) && degree >= 1, "bad arguments");
        auto v = tdg::warp_blend_nodes(dim, degree);
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

} // extern "C"
