"""TEST INFRASTRUCTURE ONLY: ctypes loader for the CPU oracle (oracle/hexed_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this; the product never does.
"""
import ctypes
import os

from hexed_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_hexed.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH}")
    lib = ctypes.CDLL(LIB_PATH)
    M = ctypes.POINTER(abi.hx_host_mesh)
    PDE = ctypes.POINTER(abi.hx_pde)
    KO = ctypes.POINTER(abi.hx_kernel_options)
    D = ctypes.c_double
    I = ctypes.c_int
    sigs = {
        "oracle_last_error": (ctypes.c_char_p, []),
        "oracle_set_threads": (None, [I]),
        "oracle_get_threads": (I, []),
        "oracle_write_face": (I, [M, I, PDE]),
        "oracle_neighbor": (I, [M, I, PDE, I]),
        "oracle_local": (I, [M, I, PDE, D, I, I, I]),
        "oracle_neighbor_reconcile": (I, [M, I, PDE]),
        "oracle_reconcile_ldg_flux": (I, [M, I, PDE, D, I, I, I]),
        "oracle_max_dt": (I, [M, I, PDE, I, I, D, D, ctypes.POINTER(D)]),
        "oracle_prolong_refined": (I, [M, PDE, I, I]),
        "oracle_restrict_refined": (I, [M, PDE, I, I]),
        "oracle_apply_state_bcs": (I, [M]),
        "oracle_apply_flux_bcs": (I, [M]),
        "oracle_characteristics": (I, [I] + [ctypes.POINTER(ctypes.c_double)] * 5),
        "oracle_compute_euler": (I, [M, KO]),
        "oracle_compute_navier_stokes": (I, [M, KO, abi.hx_transport, abi.hx_transport]),
        "oracle_max_dt_euler": (I, [M, KO, D, D, I, ctypes.POINTER(D)]),
        "oracle_max_dt_navier_stokes": (I, [M, KO, D, D, I, abi.hx_transport, abi.hx_transport, ctypes.POINTER(D)]),
        "oracle_update": (I, [M, ctypes.POINTER(abi.hx_update_options), ctypes.POINTER(abi.hx_update_status)]),
        "oracle_compute_residual": (I, [M, ctypes.POINTER(abi.hx_update_options)]),
        "oracle_apply_advection_bcs": (I, [M]),
        "oracle_apply_avc_diff_bcs": (I, [M]),
        "oracle_apply_avc_diff_flux_bcs": (I, [M]),
        "oracle_compute_advection": (I, [M, KO, D]),
        "oracle_compute_smooth_av": (I, [M, KO, D, D]),
        "oracle_update_art_visc_smoothness": (I, [M, ctypes.POINTER(abi.hx_av_options), ctypes.POINTER(D)]),
        "oracle_set_vertices": (I, [M, I, ctypes.POINTER(I), ctypes.POINTER(I), I, ctypes.POINTER(I)]),
        "oracle_calc_jacobian": (I, [M, ctypes.POINTER(abi.hx_vertex_data)]),
        "oracle_is_admissible": (I, [M, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "oracle_fix_coefficient": (I, [M, ctypes.POINTER(I), ctypes.POINTER(D)]),
        "oracle_stab_art_visc": (I, [M, D, ctypes.POINTER(D)]),
        "oracle_set_art_visc_admis": (I, [M, D, ctypes.POINTER(D), ctypes.POINTER(D)]),
        "oracle_compute_fix_therm_admis": (I, [M, KO]),
        "oracle_max_dt_fix_therm_admis": (I, [M, KO, D, D, I, ctypes.POINTER(D)]),
        "oracle_fix_admissibility": (I, [M, ctypes.POINTER(abi.hx_update_options), D, ctypes.POINTER(I)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise RuntimeError(load().oracle_last_error().decode())
