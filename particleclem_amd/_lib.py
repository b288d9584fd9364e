"""ctypes binding of the C ABI in include/pc_hydro.h (libpc_hydro.so).

The library is the product: if it is missing this module raises; there is
no fallback path.
"""
import ctypes as C
import os

from .ctypes_types import Array4, Box, LevelDesc, Params

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpc_hydro.so")

# every symbol include/pc_hydro.h declares
EXPORTS = [
    "pc_params_default", "pc_ctx_create", "pc_ctx_destroy",
    "pc_ctx_num_species", "pc_ctx_nvar", "pc_ctx_qvar", "pc_last_error",
    "pc_ctoprim_v1", "pc_srctoprim_v1", "pc_umdrv_v1",
    "pc_hydro_source_box_v1", "pc_hydro_source_box_noalloc_v1",
    "pc_mol_src_box_v1", "pc_level_create_v1",
    "pc_level_destroy_v1",
    "pc_level_box_offset_v1", "pc_level_box_size_v1", "pc_level_state_ptr_v1",
    "pc_level_state_size_v1", "pc_level_write_box_v1", "pc_level_read_box_v1",
    "pc_level_set_dt_v1", "pc_level_get_dt_v1", "pc_level_set_time_v1",
    "pc_level_get_time_v1", "pc_level_init_dt_v1",
    "pc_level_step_v1", "pc_level_fill_ghosts_v1", "pc_level_estdt_v1",
    "pc_level_write_plotfile_v1", "pc_write_plotfile_v1", "pc_fluxreg_create_v1",
    "pc_fluxreg_destroy_v1", "pc_fluxreg_nfaces_v1", "pc_fluxreg_reset_v1",
    "pc_fluxreg_crse_add_v1", "pc_fluxreg_fine_add_v1", "pc_fluxreg_reflux_v1",
    "pc_average_down_v1",
    "pc_level_enable_timing_v1", "pc_level_kernel_times_v1",
    "pc_level_status_v1", "pc_level_prim_to_cons_v1",
    "pc_level_autotune_v1", "pc_level_plan_v1", "pc_level_set_plan_v1",
    "pc_level_set_graph_v1", "pc_level_set_branches_v1", "pc_level_graph_info_v1",
    "pc_level_set_precision_v1",
    "pc_batch_create_v1", "pc_batch_destroy_v1", "pc_batch_slot_bytes_v1",
    "pc_batch_memory_v1", "pc_batch_free_ranges_v1", "pc_batch_slot_extent_v1",
    "pc_batch_add_v1", "pc_batch_step_v1", "pc_batch_set_branches_v1",
    "pc_batch_read_v1", "pc_batch_status_v1", "pc_batch_remove_v1",
    "pc_batch_snapshot_v1", "pc_batch_snapshot_status_v1",
]

PC_OK = 0
KERNELS = ["ghost", "ctoprim", "divu", "trace_x", "trace_y", "trace_z",
           "flux1_x", "flux1_y", "flux1_z", "trans_xy", "trans_xz",
           "trans_yx", "trans_yz", "trans_zx", "trans_zy", "final_x",
           "final_y", "final_z", "update", "dt", "mol_x", "mol_y", "mol_z",
           "mol_update1", "mol_update2"]
# stage plans (pc_level_plan_v1 stage ids)
STAGES = ["ctoprim", "divu", "trace", "flux1_x", "flux1_y", "flux1_z",
          "trans_xy", "trans_xz", "trans_yx", "trans_yz", "trans_zx",
          "trans_zy", "final_x", "final_y", "final_z", "update", "ghost",
          "mol_prim", "mol_x", "mol_y", "mol_z", "mol_update"]
ERRORS = {-1: "PC_ERR_ARG", -2: "PC_ERR_GROW", -3: "PC_ERR_PPM_TYPE",
          -4: "PC_ERR_PLM_IORDER", -5: "PC_ERR_HIP", -6: "PC_ERR_NAN",
          -7: "PC_ERR_UNSUPPORTED", -8: "PC_ERR_NOMEM", -9: "PC_ERR_DT",
          -10: "PC_ERR_MOL_IORDER", -11: "PC_ERR_IO"}

_lib = None


class HydroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, "
            "gfx950).  There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    i3 = P(C.c_int)
    d3 = P(C.c_double)
    L.pc_params_default.argtypes = [P(Params)]
    L.pc_params_default.restype = None
    L.pc_ctx_create.argtypes = [C.c_int, P(vp)]
    L.pc_ctx_destroy.argtypes = [vp]
    L.pc_ctx_num_species.argtypes = [vp]
    L.pc_ctx_nvar.argtypes = [vp]
    L.pc_ctx_qvar.argtypes = [vp]
    L.pc_last_error.argtypes = [vp]
    L.pc_last_error.restype = C.c_char_p
    L.pc_ctoprim_v1.argtypes = [vp, Box, Array4, Array4, Array4, vp]
    L.pc_srctoprim_v1.argtypes = [vp, Box, Array4, Array4, Array4, Array4, vp]
    L.pc_umdrv_v1.argtypes = [vp, C.c_double, Box, i3, i3, i3, i3, Array4,
                              Array4, Array4, Array4, Array4, d3, C.c_double,
                              C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_double, P(Array4), P(Array4), Array4, d3, vp]
    L.pc_hydro_source_box_v1.argtypes = [vp, Box, i3, i3, i3, i3, Array4,
                                         Array4, Array4, d3, C.c_double,
                                         P(Params), P(Array4), P(Array4),
                                         Array4, vp]
    L.pc_hydro_source_box_noalloc_v1.argtypes = L.pc_hydro_source_box_v1.argtypes
    L.pc_mol_src_box_v1.argtypes = [vp, Box, Array4, Array4, P(Params),
                                    P(Array4), P(Array4), Array4, vp]
    L.pc_level_write_plotfile_v1.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_double,
                                             C.c_int, d3, vp]
    L.pc_fluxreg_create_v1.argtypes = [vp, P(Box), C.c_int, P(Box), C.c_int, i3, i3, i3,
                                       C.c_int, P(vp)]
    L.pc_fluxreg_destroy_v1.argtypes = [vp]
    L.pc_fluxreg_nfaces_v1.argtypes = [vp]
    L.pc_fluxreg_reset_v1.argtypes = [vp, vp]
    L.pc_fluxreg_crse_add_v1.argtypes = [vp, C.c_int, Box, P(Array4), C.c_double, vp]
    L.pc_fluxreg_fine_add_v1.argtypes = [vp, C.c_int, Box, P(Array4), C.c_double, vp]
    L.pc_fluxreg_reflux_v1.argtypes = [vp, C.c_int, Box, Array4, Array4, C.c_double, vp]
    L.pc_average_down_v1.argtypes = [vp, Box, Array4, Array4, C.c_int, vp]
    L.pc_write_plotfile_v1.argtypes = [C.c_char_p, C.c_int, P(C.c_char_p), C.c_int,
                                       P(Box), P(P(C.c_double)), i3, i3, d3, d3,
                                       C.c_double, C.c_int]
    L.pc_level_autotune_v1.argtypes = [vp, C.c_int, vp]
    L.pc_level_plan_v1.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_double)]
    L.pc_level_set_graph_v1.argtypes = [vp, C.c_int]
    L.pc_level_set_branches_v1.argtypes = [vp, C.c_int]
    L.pc_level_graph_info_v1.argtypes = [vp, P(C.c_int), P(C.c_int)]
    L.pc_level_set_precision_v1.argtypes = [vp, C.c_int]
    ll = C.c_longlong
    L.pc_batch_create_v1.argtypes = [vp, P(Params), C.c_int, ll, vp, C.c_int, P(vp)]
    L.pc_batch_destroy_v1.argtypes = [vp]
    L.pc_batch_slot_bytes_v1.argtypes = [vp, Box]
    L.pc_batch_slot_bytes_v1.restype = ll
    L.pc_batch_memory_v1.argtypes = [vp, P(ll), P(ll), P(ll)]
    L.pc_batch_free_ranges_v1.argtypes = [vp, P(ll), P(ll), C.c_int, P(C.c_int)]
    L.pc_batch_slot_extent_v1.argtypes = [vp, C.c_int, P(ll), P(ll)]
    L.pc_batch_add_v1.argtypes = [vp, Box, i3, i3, d3, Array4, C.c_int, vp, P(C.c_int)]
    L.pc_batch_step_v1.argtypes = [vp, P(C.c_int), C.c_int, vp]
    L.pc_batch_set_branches_v1.argtypes = [vp, C.c_int]
    L.pc_batch_read_v1.argtypes = [vp, C.c_int, vp]
    L.pc_batch_status_v1.argtypes = [vp, C.c_int, P(C.c_double), P(C.c_int), P(C.c_int)]
    L.pc_batch_remove_v1.argtypes = [vp, C.c_int, vp]
    L.pc_batch_snapshot_v1.argtypes = [vp, C.c_int, vp]
    L.pc_batch_snapshot_status_v1.argtypes = [vp, C.c_int, P(C.c_double), P(C.c_int),
                                              P(C.c_int)]
    L.pc_level_set_plan_v1.argtypes = [vp] + [C.c_int] * 5
    L.pc_level_create_v1.argtypes = [vp, P(LevelDesc), P(Params), P(vp)]
    L.pc_level_destroy_v1.argtypes = [vp]
    L.pc_level_box_offset_v1.argtypes = [vp, C.c_int]
    L.pc_level_box_offset_v1.restype = C.c_longlong
    L.pc_level_box_size_v1.argtypes = [vp, C.c_int]
    L.pc_level_box_size_v1.restype = C.c_longlong
    L.pc_level_state_ptr_v1.argtypes = [vp]
    L.pc_level_state_ptr_v1.restype = vp
    L.pc_level_state_size_v1.argtypes = [vp]
    L.pc_level_state_size_v1.restype = C.c_longlong
    L.pc_level_write_box_v1.argtypes = [vp, C.c_int, vp]
    L.pc_level_read_box_v1.argtypes = [vp, C.c_int, vp]
    L.pc_level_set_dt_v1.argtypes = [vp, C.c_double]
    L.pc_level_get_dt_v1.argtypes = [vp, d3]
    L.pc_level_set_time_v1.argtypes = [vp, C.c_double]
    L.pc_level_get_time_v1.argtypes = [vp, d3]
    L.pc_level_init_dt_v1.argtypes = [vp, vp]
    L.pc_level_step_v1.argtypes = [vp, C.c_int, vp]
    L.pc_level_fill_ghosts_v1.argtypes = [vp, vp]
    L.pc_level_estdt_v1.argtypes = [vp, d3, vp]
    L.pc_level_enable_timing_v1.argtypes = [vp, C.c_int]
    L.pc_level_kernel_times_v1.argtypes = [vp, d3, C.c_int]
    L.pc_level_status_v1.argtypes = [vp, P(C.c_int), P(C.c_int)]
    L.pc_level_prim_to_cons_v1.argtypes = [vp, C.c_int, Array4, C.c_int, vp]
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != PC_OK:
        msg = lib().pc_last_error(ctx).decode() if ctx else ""
        raise HydroError(rc, msg)
    return rc
