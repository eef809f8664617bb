"""Handle lifetimes and the fused time step's edge cases, through the raw C ABI.

- A context destroyed while one of its meshes is alive detaches the mesh: the mesh's later calls fail with
  HX_ERR_INVALID and hx_mesh_destroy releases the handle without touching the freed context.
- hx_mesh_mark_changed retires the Max_dt that the previous hx_update fused into its last stage, for callers that
  write device memory through hx_mesh_device_ptr.
- A non-finite point never wins the fused Max_dt minimum, as with std::min(dt, NaN) in the reference's Max_dt
  (Spatial.hpp:820-827): the fused and the separately launched Max_dt give the same time step as the oracle.
"""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, init_faces, o_call, say, smooth_state

pytestmark = pytest.mark.gpu

HIP_MEMCPY_HOST_TO_DEVICE = 1


def test_context_destroyed_before_mesh():
    say("[gpu] lifetime: hx_ctx_destroy before hx_mesh_destroy (raw C ABI)")
    lib = abi.load_library()
    m = mesh.box(3, 4, 2, bc=(abi.HX_BC_FREESTREAM, freestream(3)))
    smooth_state(m)
    s = m.struct()
    ctx, msh = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.hx_ctx_create(0, ctypes.byref(ctx)) == abi.HX_OK
    assert lib.hx_mesh_create(ctx, ctypes.byref(s), ctypes.byref(msh)) == abi.HX_OK
    assert lib.hx_compute_write_face(msh) == abi.HX_OK
    assert lib.hx_ctx_destroy(ctx) == abi.HX_OK
    # the detached mesh refuses work and downloads instead of reaching freed memory
    out = np.zeros_like(m.state[0])
    assert lib.hx_mesh_download(msh, abi.HX_FIELD_STATE, 0, abi.dptr(out), out.size) == abi.HX_ERR_INVALID
    assert lib.hx_compute_write_face(msh) == abi.HX_ERR_INVALID
    st = abi.hx_update_status()
    assert lib.hx_update(msh, ctypes.byref(abi.update_options()), ctypes.byref(st)) == abi.HX_ERR_INVALID
    assert lib.hx_mesh_mark_changed(msh) == abi.HX_ERR_INVALID
    assert lib.hx_mesh_destroy(msh) == abi.HX_OK
    # a new context works afterwards
    ctx2, msh2 = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.hx_ctx_create(0, ctypes.byref(ctx2)) == abi.HX_OK
    assert lib.hx_mesh_create(ctx2, ctypes.byref(s), ctypes.byref(msh2)) == abi.HX_OK
    assert lib.hx_mesh_download(msh2, abi.HX_FIELD_STATE, 0, abi.dptr(out), out.size) == abi.HX_OK
    np.testing.assert_array_equal(out, m.state[0])
    assert lib.hx_mesh_destroy(msh2) == abi.HX_OK
    assert lib.hx_ctx_destroy(ctx2) == abi.HX_OK


def test_python_handles_close_in_either_order():
    say("[gpu] lifetime: Python context closed before its mesh")
    m = mesh.box(2, 4, 3, bc=(abi.HX_BC_FREESTREAM, freestream(2)))
    smooth_state(m)
    ctx = Context(0)
    dev = Device_mesh(ctx, m.copy())
    ctx.close()
    dev.close()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def test_mark_changed_after_device_pointer_write(ctx):
    say("[gpu] fused max_dt: write through hx_mesh_device_ptr, then hx_mesh_mark_changed")
    m = mesh.box(3, 4, 3, bc=(abi.HX_BC_FREESTREAM, freestream(3)))
    smooth_state(m)
    init_faces(m)
    dev = Device_mesh(ctx, m.copy())
    lib = ctx.lib
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    ctx.check(lib.hx_mesh_device_ptr(dev.ptr, abi.HX_FIELD_STATE, 0, ctypes.byref(p), ctypes.byref(n)))
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    s = m.struct()
    o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
    dev.update(opts, st_d)
    assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
    # the caller speeds the flow up through its own pointer (stale for the fused minimum of the update above)
    u = dev.download("state", 0, out=np.zeros_like(m.state[0]))
    u[:, 0] *= 3.0
    u[:, -1] += 4.0 * u[:, 0] ** 2 / u[:, 3]
    m.state[0][:] = u
    ctx.synchronize()
    assert _hip().hipMemcpy(p, u.ctypes.data, u.nbytes, HIP_MEMCPY_HOST_TO_DEVICE) == 0
    ctx.check(lib.hx_mesh_mark_changed(dev.ptr), "hx_mesh_mark_changed")
    o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
    dev.update(opts, st_d)
    assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step, \
        f"dt {st_d.time_step!r} vs oracle {st_o.time_step!r}"
    dev.close()


def _nan_mesh():
    """1D Euler, row size 4: 64 elements per workgroup of the Euler Local, so element 0 is lane 0 of the first
    wavefront of workgroup 0. A NaN there, a fast flow in elements 20..39 of the same workgroup (the smallest time
    step of the mesh) and a slow flow elsewhere."""
    m = mesh.box(1, 4, 256, bc=(abi.HX_BC_FREESTREAM, freestream(1, mach=0.1)))

    def f(pos):
        x = pos[0]
        fast = (x > 20 / 256) & (x < 40 / 256)
        vel = np.where(fast, 0.9, 0.1) * (1 + 0.01 * np.sin(6 * x))
        rho = 1.0 + 0.0 * x
        p = 1 / 1.4 + 0.0 * x
        return [rho * vel, rho, p / 0.4 + 0.5 * rho * vel * vel]
    mesh.set_state(m, f)
    m.state[0][0, 1, 0] = np.nan
    init_faces(m)
    return m


def test_nan_point_never_wins_fused_minimum(ctx):
    say("[gpu] fused max_dt: a NaN point in lane 0 of a workgroup")
    m = _nan_mesh()
    fused, plain = Device_mesh(ctx, m.copy()), Device_mesh(ctx, m.copy())
    opts = abi.update_options(max_safety=0.5)
    st_o, st_f, st_p = abi.hx_update_status(), abi.hx_update_status(), abi.hx_update_status()
    s = m.struct()
    for step in range(3):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        fused.update(opts, st_f)
        ctx.check(ctx.lib.hx_mesh_mark_changed(plain.ptr))  # Max_dt launched anew every step
        plain.update(opts, st_p)
        assert np.isfinite(st_o.time_step)
        assert st_f.time_step == st_p.time_step, f"step {step}: fused {st_f.time_step!r} vs {st_p.time_step!r}"
        assert abs(st_f.time_step - st_o.time_step) <= 1e-13 * st_o.time_step, \
            f"step {step}: dt {st_f.time_step!r} vs oracle {st_o.time_step!r}"
    fused.close()
    plain.close()
