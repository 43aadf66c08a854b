"""Loader of the native library libtandem_dg.so (the C ABI declared in include/tandem_dg.h).

The library is built in-tree by tandem_amd/build.py; there is no fallback implementation: if the
library is missing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtandem_dg.so")

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class TdgError(RuntimeError):
    pass


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("max_iter", C.c_int), ("mode", C.c_int), ("inner_rtol", C.c_double),
                ("max_outer", C.c_int), ("check_every", C.c_int), ("reliable_delta", C.c_double),
                ("precond", C.c_int), ("graphs", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("inner_iterations", C.c_int),
                ("rel_residual", C.c_double), ("seconds", C.c_double), ("converged", C.c_int)]


class PMGOpts(C.Structure):
    _fields_ = [("smooth", C.c_int), ("eig_lo", C.c_double), ("eig_hi", C.c_double), ("eig_iters", C.c_int),
                ("coarse_rtol", C.c_double), ("coarse_max_iter", C.c_int), ("galerkin_penalty", C.c_int),
                ("coarse_fp32", C.c_int), ("vcycle_fp32", C.c_int), ("coarse_assembled", C.c_int),
                ("coarse_amg", C.c_int)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("degree", C.c_int),
        ("num_vertices", C.c_int64),
        ("vertices", C.c_void_p),
        ("num_elements", C.c_int64),
        ("elements", C.c_void_p),
        ("num_boundary_facets", C.c_int64),
        ("boundary_facets", C.c_void_p),
        ("boundary_bc", C.c_void_p),
        ("geometry_nodes", C.c_void_p),
    ]


EXPORTS = [
    "tdg_create", "tdg_destroy", "tdg_last_error", "tdg_sizes", "tdg_block_size", "tdg_local_size",
    "tdg_volume_points", "tdg_facet_points",
    "tdg_facets", "tdg_set_material", "tdg_upload", "tdg_mult", "tdg_mult_f32", "tdg_mult_noalloc", "tdg_mult_batched", "tdg_mult_batched_f32",
    "tdg_set_stream", "tdg_set_variant", "tdg_solve", "tdg_solve_multi", "tdg_solve_segments", "tdg_autotune", "tdg_autotune_shapes", "tdg_set_shape", "tdg_rhs", "tdg_rhs_device", "tdg_traction", "tdg_block_jacobi_build", "tdg_block_jacobi_apply", "tdg_block_jacobi_apply_batched", "tdg_block_jacobi_inverse", "tdg_assemble", "tdg_assembled_block", "tdg_material_gradient",
    "tdg_material_gradient_f32", "tdg_fault_setup", "tdg_fault_slip", "tdg_fault_traction", "tdg_inverse_mass", "tdg_wave_rhs", "tdg_interpolation_matrix", "tdg_interpolate",
    "tdg_apply_counts", "tdg_host_array", "tdg_box_mesh", "tdg_reference_nodes",
    "tdg_pmg_default_opts", "tdg_pmg_create", "tdg_pmg_destroy", "tdg_pmg_setup", "tdg_pmg_vcycle",
    "tdg_solve_multi_pmg",
]


def lib():
    """The loaded library (raises TdgError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TdgError("native library %s is missing: run `python -m tandem_amd.build` first" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.tdg_last_error.restype = C.c_char_p
    L.tdg_create.argtypes = [C.POINTER(MeshDesc), C.c_int, C.POINTER(vp)]
    L.tdg_destroy.argtypes = [vp]
    L.tdg_sizes.argtypes = [vp, _ip64]
    L.tdg_block_size.argtypes = [vp, _ip64]
    L.tdg_local_size.argtypes = [vp, _ip64]
    L.tdg_volume_points.argtypes = [vp, _dp]
    L.tdg_facet_points.argtypes = [vp, _dp]
    L.tdg_facets.argtypes = [vp, _ip64]
    L.tdg_set_material.argtypes = [vp, vp, vp]
    L.tdg_upload.argtypes = [vp]
    L.tdg_mult.argtypes = [vp, vp, vp]
    L.tdg_mult_f32.argtypes = [vp, vp, vp]
    L.tdg_mult_noalloc.argtypes = [vp, vp, vp]
    L.tdg_mult_batched.argtypes = [vp, vp, vp, C.c_int, C.c_int64]
    L.tdg_mult_batched_f32.argtypes = [vp, vp, vp, C.c_int, C.c_int64]
    L.tdg_set_stream.argtypes = [vp, vp]
    L.tdg_set_variant.argtypes = [vp, C.c_int]
    L.tdg_rhs.argtypes = [vp, vp, vp, _dp]
    L.tdg_rhs_device.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int64, C.c_int64, C.c_int64]
    L.tdg_traction.argtypes = [vp, vp, vp, vp]
    L.tdg_block_jacobi_build.argtypes = [vp, C.POINTER(C.c_int)]
    L.tdg_block_jacobi_apply.argtypes = [vp, vp, vp]
    L.tdg_block_jacobi_apply_batched.argtypes = [vp, vp, vp, C.c_int, C.c_int64]
    L.tdg_block_jacobi_inverse.argtypes = [vp, C.c_int64, _dp]
    L.tdg_assemble.argtypes = [vp, C.POINTER(C.c_int)]
    L.tdg_assembled_block.argtypes = [vp, C.c_int64, C.c_int, _dp]
    L.tdg_material_gradient.argtypes = [vp, vp, vp, vp, vp]
    L.tdg_material_gradient_f32.argtypes = [vp, vp, vp, vp, vp]
    L.tdg_fault_setup.argtypes = [vp, _dp, _dp, _ip64]
    L.tdg_fault_slip.argtypes = [vp, vp, vp, C.c_int, C.c_int64, C.c_int64]
    L.tdg_fault_traction.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int64, C.c_int64, C.c_int64]
    L.tdg_inverse_mass.argtypes = [vp, vp, vp]
    L.tdg_wave_rhs.argtypes = [vp, vp, vp, vp]
    L.tdg_interpolation_matrix.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
    L.tdg_interpolate.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, vp, vp, vp]
    L.tdg_solve.argtypes = [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveResult)]
    L.tdg_pmg_default_opts.argtypes = [C.POINTER(PMGOpts)]
    L.tdg_pmg_default_opts.restype = None
    L.tdg_pmg_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(PMGOpts), C.POINTER(vp)]
    L.tdg_pmg_destroy.argtypes = [vp]
    L.tdg_pmg_setup.argtypes = [vp, _dp]
    L.tdg_pmg_vcycle.argtypes = [vp, vp, vp, C.c_int, C.c_int64]
    L.tdg_solve_multi_pmg.argtypes = [vp, vp, vp, C.c_int, C.c_int64, C.POINTER(SolveOpts), C.POINTER(SolveResult),
                                      _ip32, _dp]
    L.tdg_solve_multi.argtypes = [vp, vp, vp, C.c_int, C.c_int64, C.POINTER(SolveOpts), C.POINTER(SolveResult),
                                  _ip32, _dp]
    L.tdg_solve_segments.argtypes = [vp, vp, vp, C.c_int, _ip64, C.POINTER(SolveOpts), C.POINTER(SolveResult),
                                     _ip32, _dp]
    L.tdg_autotune.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp]
    L.tdg_autotune_shapes.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
    L.tdg_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.tdg_apply_counts.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tdg_host_array.argtypes = [vp, C.c_char_p, vp, C.POINTER(C.c_int64)]
    L.tdg_box_mesh.argtypes = [C.c_int, _ip64, C.c_int, _ip32, _ip64, vp, vp, vp, vp]
    L.tdg_reference_nodes.argtypes = [C.c_int, C.c_int, _dp]
    for name in EXPORTS:
        if getattr(L, name, None) is None:
            raise TdgError("symbol %s missing from %s" % (name, LIB_PATH))
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise TdgError("tandem_dg error %d: %s" % (rc, lib().tdg_last_error().decode()))
    return rc


def ptr(a):
    """ctypes pointer of a numpy array (None -> NULL)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)
