"""ctypes declarations of include/aphi_gpu.h and the loader of libaphi_amd.so.

The library is built in-tree (hybrid-aphi-solver_amd/lib/libaphi_amd.so). There is
no fallback: if it is missing, loading fails loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libaphi_amd.so")

APHI_OK = 0
APHI_INVALID_ARGUMENT = 1
APHI_LOGIC_ERROR = 2
APHI_RUNTIME_ERROR = 3
APHI_DEVICE_ERROR = 4
APHI_OUT_OF_MEMORY = 5
APHI_PIVOT_FAILED = 6

APHI_NATURAL, APHI_ROW_SCALED, APHI_SCALED_PHI = 0, 1, 2
APHI_ORDER_NATURAL, APHI_ORDER_RCM, APHI_ORDER_AMD = 0, 1, 2
APHI_FP64, APHI_FP32 = 0, 1

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
DP = C.POINTER(C.c_double)


class CubeSpec(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("side", C.c_double),
        ("bar_i0", C.c_int), ("bar_i1", C.c_int), ("bar_j0", C.c_int), ("bar_j1", C.c_int),
        ("sigma", C.c_double), ("eps_r", C.c_double), ("mu_r", C.c_double),
        ("air_sigma", C.c_double), ("air_eps_r", C.c_double), ("air_mu_r", C.c_double),
        ("v_top_re", C.c_double), ("v_top_im", C.c_double),
        ("v_bottom_re", C.c_double), ("v_bottom_im", C.c_double),
        ("ordering", C.c_int), ("relax_small", C.c_int), ("relax_fill", C.c_double),
        ("top_current", C.c_int), ("i_top_re", C.c_double), ("i_top_im", C.c_double),
    ]


class ProblemCounts(C.Structure):
    _fields_ = [(k, C.c_int64) for k in (
        "nodes", "edges", "faces", "tets", "num_a", "num_phi", "num_ports", "num_total", "nnz",
        "predicted_nnz", "num_roots", "tree_edges", "dirichlet_edges", "free_edges", "num_super",
        "num_levels", "factor_entries", "cb_pool_entries", "contributions")] + [
        ("flops", C.c_double), ("analyze_ms", C.c_double), ("build_ms", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_tets", C.c_int32), ("num_bodies", C.c_int32), ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("coords", DP), ("tet_nodes", I32P), ("tet_body", I32P),
        ("body_sigma", DP), ("body_eps_r", DP), ("body_mu_r", DP),
        ("target_ptr", I64P), ("entry", C.POINTER(C.c_uint32)), ("kind", C.POINTER(C.c_uint8)),
        ("num_values", C.c_int32), ("values", DP),
        ("row_ptr", I32P), ("col_index", I32P),
        ("factor_perm", I32P), ("col_ref", I32P),
        ("num_super", C.c_int32), ("num_levels", C.c_int32),
        ("super_first", I32P), ("super_parent", I32P), ("super_row_ptr", I64P), ("super_rows", I32P),
        ("super_rel", I32P), ("super_level", I32P), ("panel_offset", I64P), ("cb_offset", I64P),
        ("cb_pool_entries", C.c_int64), ("slot_to_panel", I64P),
        ("flops", C.c_double), ("predicted_nnz", C.c_int64),
        ("tet_dofs", I32P), ("tet_edge_signs", C.POINTER(C.c_int8)), ("tet_phi_coeff", DP), ("tet_phi_fixed", DP),
    ]


class GaugeFreeDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_tets", C.c_int32), ("num_bodies", C.c_int32), ("unknowns", C.c_int32),
        ("nnz", C.c_int64),
        ("coords", DP), ("tet_nodes", I32P), ("tet_body", I32P), ("body_sigma", DP), ("body_eps_r", DP),
        ("tet_on", I32P), ("column", I32P), ("prescribed", DP),
        ("num_ports", C.c_int32), ("port_column", I32P), ("port_value", DP),
        ("target_ptr", I64P), ("entry", C.POINTER(C.c_uint32)), ("kind", C.POINTER(C.c_uint8)),
        ("num_values", C.c_int32), ("values", DP),
    ]


class FieldsDesc(C.Structure):
    _fields_ = [("num_tets", C.c_int32), ("num_p2", C.c_int32), ("tet_p2", I32P), ("node_ptr", I64P),
                ("node_entry", I32P)]


class Fields(C.Structure):
    _fields_ = [("a_tet", DP), ("b_tet", DP), ("e_tet", DP), ("j_tet", DP), ("sigma_tet", DP),
                ("a_node", DP), ("b_node", DP), ("h_node", DP), ("e_node", DP), ("j_node", DP),
                ("vertex_weight", DP), ("on_material_interface", C.POINTER(C.c_int8)),
                ("num_interface_nodes", C.c_int32), ("num_orphan_vertices", C.c_int32)]


class BoundDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_tets", C.c_int32), ("num_bodies", C.c_int32), ("num_total", C.c_int32),
        ("num_ports", C.c_int32),
        ("coords", DP), ("tet_nodes", I32P), ("tet_edge_signs", C.POINTER(C.c_int8)),
        ("body_sigma", DP), ("body_eps_r", DP), ("body_mu_r", DP), ("body_ptr", I32P), ("body_tets", I32P),
        ("edge_index", I32P), ("edge_coeff", DP), ("phi_index", I32P), ("phi_coeff", DP), ("phi_fixed", DP),
        ("port_index", I32P), ("port_current", I32P), ("port_amplitude", DP),
        ("row_ptr", I32P), ("col_index", I32P),
        ("scatter_live", I32P), ("scatter_pos", I32P), ("scatter_offset", I64P), ("scatter_slot", I32P),
    ]


class FactorStats(C.Structure):
    _fields_ = [("ok", C.c_int), ("columns_done", C.c_int), ("smallest_pivot", C.c_double),
                ("largest_pivot", C.c_double), ("largest_multiplier", C.c_double), ("nnz", C.c_int64),
                ("milliseconds", C.c_double), ("flops", C.c_double), ("update_launches", C.c_int64),
                ("update_ms", C.c_double), ("update_flops", C.c_double)]


class SolveOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("equilibration_iterations", C.c_int), ("pivot_floor", C.c_double),
                ("max_refinements", C.c_int), ("tolerance", C.c_double)]


class SolveReport(C.Structure):
    _fields_ = [("ok", C.c_int), ("refinement_steps", C.c_int), ("precision_used", C.c_int),
                ("backward_error", C.c_double), ("scaled_backward_error", C.c_double),
                ("residual", C.c_double), ("equilibration_min", C.c_double), ("equilibration_max", C.c_double),
                ("smallest_pivot", C.c_double), ("largest_pivot", C.c_double),
                ("largest_multiplier", C.c_double), ("factor_nnz", C.c_int64), ("columns_done", C.c_int),
                ("factorize_ms", C.c_double), ("solve_ms", C.c_double), ("total_ms", C.c_double)]


# name -> (restype, argtypes); mirrors include/aphi_gpu.h
SIGNATURES = {
    "aphi_last_error": (C.c_char_p, []),
    "aphi_version": (C.c_char_p, []),
    "aphi_problem_create_cube": (C.c_int, [C.POINTER(CubeSpec), C.POINTER(C.c_void_p)]),
    "aphi_problem_create_csr": (C.c_int, [C.c_int32, I32P, I32P, C.c_int, C.c_int64, C.c_int, C.c_double,
                                          C.POINTER(C.c_void_p)]),
    "aphi_problem_create_bound": (C.c_int, [C.POINTER(BoundDesc), C.c_int, C.c_int, C.c_double,
                                            C.POINTER(C.c_void_p)]),
    "aphi_problem_destroy": (None, [C.c_void_p]),
    "aphi_problem_counts_get": (C.c_int, [C.c_void_p, C.POINTER(ProblemCounts)]),
    "aphi_problem_get_int": (C.c_int, [C.c_void_p, C.c_char_p, I64P, C.c_int64, I64P]),
    "aphi_problem_get_double": (C.c_int, [C.c_void_p, C.c_char_p, DP, C.c_int64, I64P]),
    "aphi_frequency_sweep": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, DP]),
    "aphi_write_potential": (C.c_int, [C.c_void_p, C.c_char_p, DP, C.c_double, C.c_int, C.c_int, I64P]),
    "aphi_problem_describe": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc)]),
    "aphi_problem_create_gauge_free": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "aphi_problem_describe_gauge_free": (C.c_int, [C.c_void_p, C.POINTER(GaugeFreeDesc)]),
    "aphi_tet_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p]),
    "aphi_problem_describe_fields": (C.c_int, [C.c_void_p, C.POINTER(FieldsDesc)]),
    "aphi_write_vtk_fields": (C.c_int, [C.c_void_p, C.c_char_p, DP, C.c_double, C.c_int, C.POINTER(Fields), I64P]),
    "aphi_fields_attach": (C.c_int, [C.c_void_p, C.POINTER(FieldsDesc)]),
    "aphi_compute_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.POINTER(Fields)]),
    "aphi_mesh_release_workspace": (C.c_int, [C.c_void_p]),
    "aphi_philox4x32_10": (C.c_int, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "aphi_random_normal_host": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, DP]),
    "aphi_random_uniform_host": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, DP]),
    "aphi_problem_create_cubes": (C.c_int, [C.POINTER(CubeSpec), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "aphi_random_normal": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_void_p]),
    "aphi_gauge_free_attach": (C.c_int, [C.c_void_p, C.POINTER(GaugeFreeDesc)]),
    "aphi_gauge_free_refill": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    "aphi_gauge_free_potential": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, DP]),
    "aphi_context_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "aphi_context_destroy": (None, [C.c_void_p]),
    "aphi_context_synchronize": (C.c_int, [C.c_void_p]),
    "aphi_mesh_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "aphi_mesh_destroy": (None, [C.c_void_p]),
    "aphi_mesh_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "aphi_mesh_set_solve_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "aphi_mesh_set_update_tiles": (C.c_int, [C.c_void_p, C.c_int]),
    "aphi_mesh_device_bytes": (C.c_int64, [C.c_void_p]),
    "aphi_mesh_tune_update_tiles": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int8), DP, DP]),
    "aphi_mesh_get_update_tile_table": (C.c_int, [C.c_void_p, C.POINTER(C.c_int8)]),
    "aphi_mesh_set_update_tile_table": (C.c_int, [C.c_void_p, C.POINTER(C.c_int8)]),
    "aphi_refill": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "aphi_system_download": (C.c_int, [C.c_void_p, DP, DP]),
    "aphi_system_upload": (C.c_int, [C.c_void_p, DP, DP]),
    "aphi_rhs_device": (C.c_void_p, [C.c_void_p]),
    "aphi_rhs_copy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "aphi_equilibrate": (C.c_int, [C.c_void_p, C.c_int, DP]),
    "aphi_factorize": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.POINTER(FactorStats)]),
    "aphi_factor_diagonal": (C.c_int, [C.c_void_p, DP]),
    "aphi_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    "aphi_solve_symmetric": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                       C.POINTER(SolveOptions), C.POINTER(SolveReport), DP]),
    "aphi_backward_error": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, DP, DP]),
    "aphi_terminal_current": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_int, DP,
                                        C.c_double, C.c_double, DP, DP, DP]),
    "aphi_batch_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "aphi_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), I64P, I64P]),
    "aphi_batch_refill": (C.c_int, [C.c_void_p, DP, C.c_int]),
    "aphi_batch_solve_symmetric": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                             C.POINTER(SolveOptions), C.POINTER(SolveReport), DP]),
    "aphi_batch_backward_error": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                            DP, DP]),
    "aphi_batch_terminal_current": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, DP, C.c_int, DP, DP, DP,
                                              DP, DP, DP]),
    "aphi_terminal_current_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_int,
                                                DP, C.c_double, C.c_double, DP, DP, DP]),
}

_lib = None


def load():
    """Loads libaphi_amd.so (once) and declares every exported function."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libaphi_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    # PyTorch (the device allocator) ships its own HIP runtime: load it first, so
    # the library binds to that same runtime instead of starting a second one
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
