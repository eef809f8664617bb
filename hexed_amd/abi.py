"""ctypes mirror of include/hexed_amd.h and loaders for the built shared libraries."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libhexed_amd.so")

HX_OK, HX_ERR_INVALID, HX_ERR_HIP, HX_ERR_NOT_FINITE, HX_ERR_UNSUPPORTED, HX_ERR_NO_DEVICE = range(6)
HX_CARTESIAN, HX_DEFORMED = 0, 1
HX_PDE_EULER, HX_PDE_NAVIER_STOKES, HX_PDE_ADVECTION, HX_PDE_SMOOTH_AV, HX_PDE_FIX_ADMIS = range(5)
HX_FP64, HX_FP32_MFMA = 0, 1
(HX_BC_FREESTREAM, HX_BC_COPY, HX_BC_NONPENETRATION, HX_BC_OUTFLOW, HX_BC_RIEMANN_INVARIANTS, HX_BC_PRESSURE_OUTFLOW,
 HX_BC_NO_SLIP) = range(7)
HX_THERMAL_HEAT_FLUX, HX_THERMAL_ENERGY, HX_THERMAL_EQUILIBRIUM = range(3)
(HX_FIELD_STATE, HX_FIELD_RCACHE, HX_FIELD_TSS, HX_FIELD_AV, HX_FIELD_FORCING, HX_FIELD_ADVECTION,
 HX_FIELD_VERTEX_TSS, HX_FIELD_METRICS, HX_FIELD_FACES, HX_FIELD_NORMALS) = range(10)

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)


class hx_transport(ctypes.Structure):
    _fields_ = [("const_val", ctypes.c_double), ("ref_val", ctypes.c_double), ("ref_temp", ctypes.c_double),
                ("temp_offset", ctypes.c_double), ("is_viscous", ctypes.c_int)]


def inviscid():
    """Transport_model::inviscid() (Transport_model.hpp:159)."""
    return hx_transport(0., 0., 1., 1., 0)


def constant(value):
    """Transport_model::constant (Transport_model.hpp:161)."""
    return hx_transport(value, 0., 1., 1., 1)


def sutherland(ref_value, ref_temp, temp_offset):
    """Transport_model::sutherland (Transport_model.hpp:166-169)."""
    return hx_transport(0., ref_value, ref_temp, temp_offset, 1)


class hx_pde(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("visc", hx_transport), ("cond", hx_transport),
                ("advect_length", ctypes.c_double), ("diff_time", ctypes.c_double), ("cheby_step", ctypes.c_double)]


def euler_pde():
    return hx_pde(HX_PDE_EULER, inviscid(), inviscid(), 0., 0., 0.)


def ns_pde(visc, cond):
    return hx_pde(HX_PDE_NAVIER_STOKES, visc, cond, 0., 0., 0.)


class hx_bc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("state", ctypes.c_double * 5), ("pressure", ctypes.c_double),
                ("thermal", ctypes.c_int), ("thermal_value", ctypes.c_double), ("emissivity", ctypes.c_double),
                ("heat_transfer_coef", ctypes.c_double), ("temperature", ctypes.c_double),
                ("coercion", ctypes.c_double)]


class hx_host_mesh(ctypes.Structure):
    _fields_ = [
        ("n_dim", ctypes.c_int), ("row_size", ctypes.c_int), ("n_elem", ctypes.c_int * 2),
        ("state", c_double_p * 2), ("rcache", c_double_p * 2), ("tss", c_double_p * 2), ("av", c_double_p * 2),
        ("forcing", c_double_p * 2), ("advection", c_double_p * 2), ("nom_size", c_double_p * 2),
        ("vertex_tss", c_double_p * 2), ("face_slot", c_int_p * 2), ("metrics", c_double_p), ("face_nslot", c_int_p),
        ("n_slot", ctypes.c_int), ("faces", c_double_p), ("n_nslot", ctypes.c_int), ("normals", c_double_p),
        ("n_con", ctypes.c_int * 2), ("con_slot", c_int_p * 2), ("con_dir", c_int_p * 2), ("con_nslot", c_int_p),
        ("n_ref", ctypes.c_int), ("ref_slot", c_int_p), ("ref_stretch", c_int_p),
        ("n_bc_con", ctypes.c_int), ("bc_con", c_int_p), ("bc_serial", c_int_p),
        ("n_bc", ctypes.c_int), ("bcs", ctypes.POINTER(hx_bc)), ("bc_cache", c_double_p),
    ]


class hx_kernel_options(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("i_stage", ctypes.c_int), ("compute_residual", ctypes.c_int),
                ("use_filter", ctypes.c_int)]


class hx_update_options(ctypes.Structure):
    _fields_ = [("max_safety", ctypes.c_double), ("max_time_step", ctypes.c_double), ("n_cheby_flow", ctypes.c_int),
                ("flow_iters", ctypes.c_int), ("local_time", ctypes.c_int), ("use_filter", ctypes.c_int),
                ("use_ldg", ctypes.c_int), ("visc", hx_transport), ("cond", hx_transport),
                ("fix_admis", ctypes.c_int), ("fix_admis_max_safety", ctypes.c_double)]


def update_options(max_safety=0.7, max_time_step=1e300, n_cheby_flow=1, flow_iters=1, local_time=False, use_filter=False,
                   visc=None, cond=None, use_ldg=None, fix_admis=False, fix_admis_max_safety=0.7):
    """Defaults of Solver (Solver.cpp:143-174); use_ldg follows Solver::use_ldg (Solver.cpp:117-120) unless given."""
    visc = visc if visc is not None else inviscid()
    cond = cond if cond is not None else inviscid()
    if use_ldg is None:
        use_ldg = bool(visc.is_viscous or cond.is_viscous)
    return hx_update_options(max_safety, max_time_step, n_cheby_flow, flow_iters, int(local_time), int(use_filter),
                             int(use_ldg), visc, cond, int(fix_admis), fix_admis_max_safety)


class hx_launch_shape(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("workgroups_per_cu", ctypes.c_int), ("ms", ctypes.c_double)]


class hx_av_options(ctypes.Structure):
    _fields_ = [("advect_length", ctypes.c_double), ("advect_iters", ctypes.c_int), ("advect_max_safety", ctypes.c_double),
                ("diff_ratio", ctypes.c_double), ("diff_max_safety", ctypes.c_double), ("n_cheby", ctypes.c_int),
                ("diff_iters", ctypes.c_int), ("visc_mult", ctypes.c_double), ("unscaled_max", ctypes.c_double),
                ("freestream_mass", ctypes.c_double), ("freestream_energy", ctypes.c_double), ("av_rs", ctypes.c_int)]


def av_options(advect_length, row_size, freestream_mass, freestream_energy, advect_iters=1, advect_max_safety=.7,
               diff_ratio=.3, diff_max_safety=.7, n_cheby=1, diff_iters=1, visc_mult=1e2, unscaled_max=5., av_rs=None):
    """Settings of Solver::update_art_visc_smoothness with the reference's defaults (Solver.cpp:144,154-163)."""
    return hx_av_options(advect_length, advect_iters, advect_max_safety, diff_ratio, diff_max_safety, n_cheby, diff_iters,
                         visc_mult, unscaled_max, freestream_mass, freestream_energy,
                         row_size if av_rs is None else av_rs)


def advection_pde(advect_length):
    p = hx_pde()
    p.kind = HX_PDE_ADVECTION
    p.advect_length = advect_length
    return p


def smooth_av_pde(diff_time, cheby_step):
    p = hx_pde()
    p.kind = HX_PDE_SMOOTH_AV
    p.diff_time = diff_time
    p.cheby_step = cheby_step
    return p


class hx_vertex_data(ctypes.Structure):
    _fields_ = [("n_vertex", ctypes.c_int), ("vertex_id", ctypes.POINTER(ctypes.c_int) * 2),
                ("position", ctypes.POINTER(ctypes.c_double)), ("node_adj", ctypes.POINTER(ctypes.c_double)),
                ("n_hanging", ctypes.c_int), ("hanging", ctypes.POINTER(ctypes.c_int))]


class hx_update_status(ctypes.Structure):
    _fields_ = [("time_step", ctypes.c_double), ("flow_time", ctypes.c_double), ("iteration", ctypes.c_int),
                ("fix_admis_iters", ctypes.c_int)]


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library (built by __graft_entry__.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"hexed_amd: {path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    sigs = {
        "hx_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "hx_ctx_destroy": (ctypes.c_int, [P]),
        "hx_last_error": (ctypes.c_char_p, [P]),
        "hx_ctx_synchronize": (ctypes.c_int, [P]),
        "hx_ctx_stream": (P, [P]),
        "hx_kernels_built": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "hx_ctx_set_timing": (ctypes.c_int, [P, ctypes.c_int]),
        "hx_ctx_timing_reset": (ctypes.c_int, [P]),
        "hx_ctx_timing": (ctypes.c_int, [P, ctypes.c_char_p, c_double_p, ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_longlong)]),
        "hx_ctx_timing_list": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_size_t]),
        "hx_stream_copy": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.c_int, c_double_p]),
        "hx_basis_table": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, c_double_p, ctypes.c_int]),
        "hx_basis_max_cfl": (ctypes.c_double, [ctypes.c_int]),
        "hx_basis_step_ratio": (ctypes.c_double, [ctypes.c_int]),
        "hx_mesh_create": (ctypes.c_int, [P, ctypes.POINTER(hx_host_mesh), ctypes.POINTER(P)]),
        "hx_mesh_destroy": (ctypes.c_int, [P]),
        "hx_mesh_validate": (ctypes.c_int, [ctypes.POINTER(hx_host_mesh), ctypes.c_char_p, ctypes.c_size_t]),
        "hx_mesh_upload": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, c_double_p, ctypes.c_size_t]),
        "hx_mesh_download": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, c_double_p, ctypes.c_size_t]),
        "hx_mesh_set_precision": (ctypes.c_int, [P, ctypes.c_int]),
        "hx_mesh_device_ptr": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P),
                                              ctypes.POINTER(ctypes.c_size_t)]),
        "hx_mesh_mark_changed": (ctypes.c_int, [P]),
        "hx_mesh_set_face_fold": (ctypes.c_int, [P, ctypes.c_int]),
        "hx_write_face": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde)]),
        "hx_neighbor": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde), ctypes.c_int]),
        "hx_local": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde), ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]),
        "hx_neighbor_reconcile": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde)]),
        "hx_reconcile_ldg_flux": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde), ctypes.c_double,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "hx_max_dt": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(hx_pde), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, c_double_p]),
        "hx_prolong_refined": (ctypes.c_int, [P, ctypes.POINTER(hx_pde), ctypes.c_int, ctypes.c_int]),
        "hx_restrict_refined": (ctypes.c_int, [P, ctypes.POINTER(hx_pde), ctypes.c_int, ctypes.c_int]),
        "hx_apply_state_bcs": (ctypes.c_int, [P]),
        "hx_apply_flux_bcs": (ctypes.c_int, [P]),
        "hx_compute_euler": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options)]),
        "hx_compute_navier_stokes": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), hx_transport, hx_transport]),
        "hx_max_dt_euler": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, c_double_p]),
        "hx_max_dt_navier_stokes": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, hx_transport, hx_transport,
                                                   c_double_p]),
        "hx_compute_write_face": (ctypes.c_int, [P]),
        "hx_compute_prolong": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "hx_compute_restrict": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "hx_update": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.POINTER(hx_update_status)]),
        "hx_update_async": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.c_int,
                                           ctypes.POINTER(hx_update_status)]),
        "hx_update_status_sync": (ctypes.c_int, [P, ctypes.POINTER(hx_update_status)]),
        "hx_update_no_alloc": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.POINTER(hx_update_status)]),
        "hx_group_create": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(P)]),
        "hx_group_destroy": (ctypes.c_int, [P]),
        "hx_group_fused": (ctypes.c_int, [P]),
        "hx_group_set_shape": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "hx_group_update": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.POINTER(hx_update_status)]),
        "hx_group_autotune": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(hx_launch_shape), ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int)]),
        "hx_mesh_set_cases": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "hx_update_cases": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.POINTER(hx_update_status)]),
        "hx_compute_residual": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options)]),
        "hx_apply_advection_bcs": (ctypes.c_int, [P]),
        "hx_apply_avc_diff_bcs": (ctypes.c_int, [P]),
        "hx_apply_avc_diff_flux_bcs": (ctypes.c_int, [P]),
        "hx_compute_advection": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double]),
        "hx_compute_smooth_av": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double, ctypes.c_double]),
        "hx_compute_write_face_advection": (ctypes.c_int, [P, ctypes.c_double]),
        "hx_compute_write_face_smooth_av": (ctypes.c_int, [P]),
        "hx_compute_prolong_advection": (ctypes.c_int, [P, ctypes.c_double]),
        "hx_max_dt_advection": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_double, c_double_p]),
        "hx_max_dt_smooth_av": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double, ctypes.c_double,
                                               ctypes.c_int, c_double_p]),
        "hx_update_art_visc_smoothness": (ctypes.c_int, [P, ctypes.POINTER(hx_av_options), c_double_p]),
        "hx_calc_jacobian": (ctypes.c_int, [P, ctypes.POINTER(hx_vertex_data)]),
        "hx_stab_art_visc": (ctypes.c_int, [P, ctypes.c_double, c_double_p]),
        "hx_set_art_visc_admis": (ctypes.c_int, [P, ctypes.POINTER(hx_vertex_data), ctypes.c_double, c_double_p]),
        "hx_is_admissible": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int)]),
        "hx_fix_admissibility": (ctypes.c_int, [P, ctypes.POINTER(hx_update_options), ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_int)]),
        "hx_compute_fix_therm_admis": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options)]),
        "hx_max_dt_fix_therm_admis": (ctypes.c_int, [P, ctypes.POINTER(hx_kernel_options), ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_int, c_double_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# names exported by include/hexed_amd.h, checked by the CPU test suite
EXPORTED_SYMBOLS = [
    "hx_ctx_create", "hx_ctx_destroy", "hx_last_error", "hx_ctx_synchronize", "hx_ctx_stream", "hx_kernels_built",
    "hx_ctx_set_timing", "hx_ctx_timing_reset", "hx_ctx_timing", "hx_ctx_timing_list", "hx_stream_copy",
    "hx_basis_table", "hx_basis_max_cfl", "hx_basis_step_ratio", "hx_mesh_create", "hx_mesh_destroy", "hx_mesh_validate",
    "hx_mesh_upload", "hx_mesh_download", "hx_mesh_set_precision", "hx_mesh_device_ptr", "hx_mesh_mark_changed", "hx_mesh_set_face_fold",
    "hx_write_face",
    "hx_neighbor", "hx_local", "hx_neighbor_reconcile", "hx_reconcile_ldg_flux", "hx_max_dt", "hx_prolong_refined",
    "hx_restrict_refined", "hx_apply_state_bcs", "hx_apply_flux_bcs", "hx_compute_euler", "hx_compute_navier_stokes",
    "hx_max_dt_euler", "hx_max_dt_navier_stokes", "hx_compute_write_face", "hx_compute_prolong", "hx_compute_restrict",
    "hx_update", "hx_update_async", "hx_update_status_sync", "hx_update_no_alloc", "hx_group_create", "hx_group_destroy",
    "hx_group_fused", "hx_group_set_shape", "hx_group_update", "hx_group_autotune", "hx_compute_residual", "hx_apply_advection_bcs", "hx_apply_avc_diff_bcs", "hx_apply_avc_diff_flux_bcs",
    "hx_compute_advection", "hx_compute_smooth_av", "hx_compute_write_face_advection", "hx_compute_write_face_smooth_av",
    "hx_compute_prolong_advection", "hx_max_dt_advection", "hx_max_dt_smooth_av", "hx_update_art_visc_smoothness",
    "hx_mesh_set_cases", "hx_update_cases", "hx_is_admissible", "hx_fix_admissibility", "hx_compute_fix_therm_admis",
    "hx_max_dt_fix_therm_admis", "hx_calc_jacobian", "hx_stab_art_visc", "hx_set_art_visc_admis",
]
