"""Python handle on the C ABI: one HIP context plus one device mesh built from a `mesh.HostMesh`.

This is the host-side mirror used by the tests, the smoke check and the benchmark; it calls only the C ABI of
libhexed_amd.so (include/hexed_amd.h) and fails loudly when the library or a GPU is missing.
"""
import ctypes
import weakref

import numpy as np

from . import abi

FIELDS = {
    "state": abi.HX_FIELD_STATE, "rcache": abi.HX_FIELD_RCACHE, "tss": abi.HX_FIELD_TSS, "av": abi.HX_FIELD_AV,
    "vertex_tss": abi.HX_FIELD_VERTEX_TSS, "metrics": abi.HX_FIELD_METRICS, "faces": abi.HX_FIELD_FACES,
    "normals": abi.HX_FIELD_NORMALS, "forcing": abi.HX_FIELD_FORCING, "advection": abi.HX_FIELD_ADVECTION,
}


class HxError(RuntimeError):
    pass


class Context:
    def __init__(self, device=0):
        self.lib = abi.load_library()
        self.ptr = ctypes.c_void_p()
        self._meshes = weakref.WeakSet()  # live meshes; a context outlives every mesh made on it
        self._groups = weakref.WeakSet()  # live groups, closed before the meshes they hold
        rc = self.lib.hx_ctx_create(device, ctypes.byref(self.ptr))
        if rc != abi.HX_OK:
            raise HxError(f"hx_ctx_create failed with status {rc}")

    def check(self, rc, what=""):
        if rc != abi.HX_OK:
            raise HxError(f"{what}: status {rc}: {self.lib.hx_last_error(self.ptr).decode()}")

    def synchronize(self):
        self.check(self.lib.hx_ctx_synchronize(self.ptr), "synchronize")

    def set_timing(self, enable=True):
        self.check(self.lib.hx_ctx_set_timing(self.ptr, int(enable)), "set_timing")

    def timing_reset(self):
        self.check(self.lib.hx_ctx_timing_reset(self.ptr), "timing_reset")

    def timings(self):
        """{category: (total_ms, launches, work_units)} for every category recorded since the last reset."""
        buf = ctypes.create_string_buffer(1 << 16)
        self.check(self.lib.hx_ctx_timing_list(self.ptr, buf, len(buf)), "timing_list")
        out = {}
        for name in buf.value.decode().split("\n"):
            if not name:
                continue
            ms, n, u = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
            self.check(self.lib.hx_ctx_timing(self.ptr, name.encode(), ctypes.byref(ms), ctypes.byref(n),
                                              ctypes.byref(u)), "timing")
            out[name] = (ms.value, n.value, u.value)
        return out

    def stream_copy(self, nbytes, reps=10):
        """Achieved GB/s of a plain copy kernel (one read + one write per byte)."""
        out = ctypes.c_double()
        self.check(self.lib.hx_stream_copy(self.ptr, nbytes, reps, ctypes.byref(out)), "stream_copy")
        return out.value

    @property
    def stream(self):
        return self.lib.hx_ctx_stream(self.ptr)

    def close(self):
        if self.ptr:
            for g in list(self._groups):
                g.close()
            for m in list(self._meshes):
                m.close()
            self.lib.hx_ctx_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Device_mesh:
    """A mesh resident in device memory; `host` is the HostMesh it was created from (and downloads go into)."""

    def __init__(self, ctx, host):
        self.ctx = ctx
        self.lib = ctx.lib
        self.host = host
        self.ptr = ctypes.c_void_p()
        self._struct = host.struct()
        ctx.check(self.lib.hx_mesh_create(ctx.ptr, ctypes.byref(self._struct), ctypes.byref(self.ptr)), "hx_mesh_create")
        ctx._meshes.add(self)

    def close(self):
        # valid in either order: a mesh whose context was destroyed first holds no device memory any more, and
        # hx_mesh_destroy then releases only its handle
        if self.ptr:
            # refused while the mesh belongs to a live group (close the group first): the handle is kept
            self.ctx.check(self.lib.hx_mesh_destroy(self.ptr), "hx_mesh_destroy")
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision):
        self.ctx.check(self.lib.hx_mesh_set_precision(self.ptr, precision), "set_precision")

    def set_face_fold(self, mode):
        """Whether update folds the face fluxes into the Local: 1 always, 0 never (Neighbor then Local, as the
        reference sequences them), -1 where 2 n_dim <= row_size (the default)."""
        self.ctx.check(self.lib.hx_mesh_set_face_fold(self.ptr, int(mode)), "set_face_fold")

    def _field_array(self, name, cls):
        h = self.host
        if name in ("faces", "normals", "metrics"):
            return getattr(h, name)
        arr = getattr(h, name)[cls]
        return arr

    def download(self, name, cls=0, out=None):
        arr = self._field_array(name, cls) if out is None else out
        if arr is None:
            n = ctypes.c_size_t()
            p = ctypes.c_void_p()
            self.ctx.check(self.lib.hx_mesh_device_ptr(self.ptr, FIELDS[name], cls, ctypes.byref(p), ctypes.byref(n)))
            arr = np.zeros(n.value)
        self.ctx.check(self.lib.hx_mesh_download(self.ptr, FIELDS[name], cls, abi.dptr(arr), arr.size), f"download {name}")
        return arr

    def upload(self, name, cls=0, arr=None):
        arr = self._field_array(name, cls) if arr is None else np.ascontiguousarray(arr, dtype=np.float64)
        self.ctx.check(self.lib.hx_mesh_upload(self.ptr, FIELDS[name], cls, abi.dptr(arr), arr.size), f"upload {name}")

    # kernels -------------------------------------------------------------------------------------------------
    def write_face(self, cls, pde=None):
        self.ctx.check(self.lib.hx_write_face(self.ptr, cls, ctypes.byref(pde or abi.euler_pde())), "write_face")

    def neighbor(self, cls, pde=None, i_stage=0):
        self.ctx.check(self.lib.hx_neighbor(self.ptr, cls, ctypes.byref(pde or abi.euler_pde()), i_stage), "neighbor")

    def local(self, cls, dt, i_stage=0, compute_residual=False, use_filter=False, pde=None):
        self.ctx.check(self.lib.hx_local(self.ptr, cls, ctypes.byref(pde or abi.euler_pde()), dt, i_stage,
                                         int(compute_residual), int(use_filter)), "local")

    def neighbor_reconcile(self, cls, pde):
        self.ctx.check(self.lib.hx_neighbor_reconcile(self.ptr, cls, ctypes.byref(pde)), "neighbor_reconcile")

    def reconcile_ldg_flux(self, cls, pde, dt, compute_residual=False, use_filter=False):
        self.ctx.check(self.lib.hx_reconcile_ldg_flux(self.ptr, cls, ctypes.byref(pde), dt, 0, int(compute_residual),
                                                      int(use_filter)), "reconcile_ldg_flux")

    def max_dt(self, cls, pde=None, is_local=False, safety_conv=0.7, safety_diff=0.7):
        out = ctypes.c_double()
        self.ctx.check(self.lib.hx_max_dt(self.ptr, cls, ctypes.byref(pde or abi.euler_pde()), int(is_local), 0,
                                          safety_conv, safety_diff, ctypes.byref(out)), "max_dt")
        return out.value

    def prolong(self, scale=False, offset=False, pde=None):
        self.ctx.check(self.lib.hx_prolong_refined(self.ptr, ctypes.byref(pde or abi.euler_pde()), int(scale),
                                                   int(offset)), "prolong")

    def restrict(self, scale=True, offset=False, pde=None):
        self.ctx.check(self.lib.hx_restrict_refined(self.ptr, ctypes.byref(pde or abi.euler_pde()), int(scale),
                                                    int(offset)), "restrict")

    def apply_state_bcs(self):
        self.ctx.check(self.lib.hx_apply_state_bcs(self.ptr), "state_bcs")

    def apply_flux_bcs(self):
        self.ctx.check(self.lib.hx_apply_flux_bcs(self.ptr), "flux_bcs")

    def compute_euler(self, dt, i_stage, compute_residual=False, use_filter=False):
        o = abi.hx_kernel_options(dt, i_stage, int(compute_residual), int(use_filter))
        self.ctx.check(self.lib.hx_compute_euler(self.ptr, ctypes.byref(o)), "compute_euler")

    def compute_navier_stokes(self, dt, visc, cond, compute_residual=False, use_filter=False):
        o = abi.hx_kernel_options(dt, 0, int(compute_residual), int(use_filter))
        self.ctx.check(self.lib.hx_compute_navier_stokes(self.ptr, ctypes.byref(o), visc, cond), "compute_navier_stokes")

    def compute_write_face(self):
        self.ctx.check(self.lib.hx_compute_write_face(self.ptr), "compute_write_face")

    # artificial-viscosity generator --------------------------------------------------------------------------
    def update_art_visc_smoothness(self, av_opts):
        r = ctypes.c_double()
        self.ctx.check(self.lib.hx_update_art_visc_smoothness(self.ptr, ctypes.byref(av_opts), ctypes.byref(r)),
                       "update_art_visc_smoothness")
        return r.value

    def compute_advection(self, dt, i_stage, advect_length):
        o = abi.hx_kernel_options(dt, i_stage, 0, 0)
        self.ctx.check(self.lib.hx_compute_advection(self.ptr, ctypes.byref(o), advect_length), "compute_advection")

    def compute_smooth_av(self, dt, diff_time, cheby_step):
        o = abi.hx_kernel_options(dt, 0, 0, 0)
        self.ctx.check(self.lib.hx_compute_smooth_av(self.ptr, ctypes.byref(o), diff_time, cheby_step),
                       "compute_smooth_av")

    def apply_advection_bcs(self):
        self.ctx.check(self.lib.hx_apply_advection_bcs(self.ptr), "advection bcs")

    def apply_avc_diff_bcs(self):
        self.ctx.check(self.lib.hx_apply_avc_diff_bcs(self.ptr), "avc diff bcs")

    # mesh metrics (Solver::calc_jacobian, Solver.cpp:273-381) ---------------------------------------------------
    def calc_jacobian(self, vertex_data=None):
        """Metrics, normals and vertex time step scale on the device from the host mesh's vertex data."""
        vd = vertex_data if vertex_data is not None else self.host.vertex_data()
        self.ctx.check(self.lib.hx_calc_jacobian(self.ptr, ctypes.byref(vd)), "calc_jacobian")

    def stab_art_visc(self, char_speed):
        """Stab_art_visc's element values (Cartesian elements first)."""
        out = np.zeros(sum(self.host.n_elem))
        self.ctx.check(self.lib.hx_stab_art_visc(self.ptr, char_speed, out.ctypes.data_as(abi.c_double_p)),
                       "stab_art_visc")
        return out

    def set_art_visc_admis(self, char_speed, vertex_data=None):
        """Solver::set_art_visc_admis into the Laplacian artificial viscosity; returns the element values."""
        vd = vertex_data if vertex_data is not None else self.host.vertex_data()
        out = np.zeros(sum(self.host.n_elem))
        self.ctx.check(self.lib.hx_set_art_visc_admis(self.ptr, ctypes.byref(vd), char_speed,
                                                      out.ctypes.data_as(abi.c_double_p)), "set_art_visc_admis")
        return out

    # state files (Solver::write_state / read_state, Solver.cpp:232-271) -----------------------------------------
    def _field_present(self, name, cls):
        n, p = ctypes.c_size_t(), ctypes.c_void_p()
        self.ctx.check(self.lib.hx_mesh_device_ptr(self.ptr, FIELDS[name], cls, ctypes.byref(p), ctypes.byref(n)),
                       "device_ptr")
        return bool(p.value)

    def write_state(self, name):
        """`name`.state.h5 from the device fields, in the reference's element storage layout."""
        from hexed_amd import state_file
        h = self.host
        per_elem = dict(state=h.nv, tss=1, av=2, forcing=4, advection=h.row_size, rcache=h.nrc)

        def get(field, c):
            return self.download(field, c, out=np.zeros((h.n_elem[c], per_elem[field], h.nq)))
        state_file.write(name, state_file.elements_of(h, get))

    def read_state(self, name):
        """The fields of `name`.state.h5 onto the device. An optional field (artificial viscosity, forcing, advection)
        that the device does not hold and the file holds as zero stays absent."""
        from hexed_amd import state_file

        def store(field, c, arr):
            if field in ("av", "forcing", "advection") and not arr.any() and not self._field_present(field, c):
                return
            self.upload(field, c, arr)
        state_file.set_elements(self.host, state_file.read(name), store)

    # thermodynamic admissibility ------------------------------------------------------------------------------
    def is_admissible(self):
        out = ctypes.c_int()
        self.ctx.check(self.lib.hx_is_admissible(self.ptr, ctypes.byref(out)), "is_admissible")
        return bool(out.value)

    def fix_admissibility(self, opts, stability_ratio=0.7):
        """Solver::fix_admissibility with the flow settings `opts`; returns its iteration count."""
        out = ctypes.c_int()
        self.ctx.check(self.lib.hx_fix_admissibility(self.ptr, ctypes.byref(opts), stability_ratio, ctypes.byref(out)),
                       "fix_admissibility")
        return out.value

    def compute_fix_therm_admis(self, dt):
        o = abi.hx_kernel_options(dt, 0, 0, 0)
        self.ctx.check(self.lib.hx_compute_fix_therm_admis(self.ptr, ctypes.byref(o)), "compute_fix_therm_admis")

    def max_dt_fix_therm_admis(self, msc, msd, local_time):
        out = ctypes.c_double()
        o = abi.hx_kernel_options(0., 0, 0, 0)
        self.ctx.check(self.lib.hx_max_dt_fix_therm_admis(self.ptr, ctypes.byref(o), msc, msd, int(local_time),
                                                          ctypes.byref(out)), "max_dt_fix_therm_admis")
        return out.value

    def update(self, opts, status=None):
        status = status if status is not None else abi.hx_update_status()
        self.ctx.check(self.lib.hx_update(self.ptr, ctypes.byref(opts), ctypes.byref(status)), "update")
        return status

    def update_async(self, opts, n_updates=1, start=None):
        """n_updates Solver::update calls enqueued with the time step kept on the device (the steady-state update
        replayed from a captured graph); `start`, an hx_update_status, sets the device status first."""
        self.ctx.check(self.lib.hx_update_async(self.ptr, ctypes.byref(opts), int(n_updates),
                                                ctypes.byref(start) if start is not None else None), "update_async")

    def status_sync(self, status=None):
        """Waits for the device and returns the status that update_async accumulated."""
        status = status if status is not None else abi.hx_update_status()
        self.ctx.check(self.lib.hx_update_status_sync(self.ptr, ctypes.byref(status)), "update_status_sync")
        return status

    # batch of independent cases --------------------------------------------------------------------------------
    def set_cases(self, n_cases, case_id):
        """case_id: per class, the case of every element (int32 arrays; None for an empty class)."""
        self._case_id = [None if c is None else np.ascontiguousarray(c, dtype=np.int32) for c in case_id]
        ptrs = [abi.iptr(c) if c is not None else None for c in self._case_id]
        self.ctx.check(self.lib.hx_mesh_set_cases(self.ptr, n_cases, ptrs[0], ptrs[1]), "set_cases")
        self.n_cases = n_cases

    def update_cases(self, opts, status=None):
        """One Solver::update of every case; `status` is a ctypes array of n_cases hx_update_status."""
        status = status if status is not None else (abi.hx_update_status * self.n_cases)()
        self.ctx.check(self.lib.hx_update_cases(self.ptr, ctypes.byref(opts), status), "update_cases")
        return status

    def compute_residual(self, opts):
        self.ctx.check(self.lib.hx_compute_residual(self.ptr, ctypes.byref(opts)), "compute_residual")


class Device_group:
    """Meshes stepped together (hx_group): each kernel of an Euler stage launched once for all of them, both element
    classes and every row size (3D, row sizes 4 to 6). Every mesh needs its cases registered (Device_mesh.set_cases);
    the group must be closed before its meshes."""

    def __init__(self, ctx, meshes):
        self.ctx = ctx
        self.lib = ctx.lib
        self.meshes = list(meshes)
        self.n_cases = sum(getattr(m, "n_cases", 0) for m in self.meshes)
        ptrs = (ctypes.c_void_p * len(self.meshes))(*[m.ptr.value for m in self.meshes])
        self.ptr = ctypes.c_void_p()
        ctx.check(self.lib.hx_group_create(ctx.ptr, ptrs, len(self.meshes), ctypes.byref(self.ptr)), "hx_group_create")
        ctx._groups.add(self)

    @property
    def fused(self):
        return bool(self.lib.hx_group_fused(self.ptr))

    def set_shape(self, order, workgroups_per_cu=0):
        self.ctx.check(self.lib.hx_group_set_shape(self.ptr, order, workgroups_per_cu), "hx_group_set_shape")

    def update(self, opts, status=None):
        """One Solver::update of every case; `status` is a ctypes array of n_cases hx_update_status (meshes in order)."""
        status = status if status is not None else (abi.hx_update_status * self.n_cases)()
        self.ctx.check(self.lib.hx_group_update(self.ptr, ctypes.byref(opts), status), "hx_group_update")
        return status

    def autotune(self, use_filter=False, reps=5):
        """Times the fused Local under every candidate launch shape and keeps the fastest; returns
        [(order, workgroups_per_cu, ms)] of all candidates."""
        table = (abi.hx_launch_shape * 16)()
        n = ctypes.c_int()
        self.ctx.check(self.lib.hx_group_autotune(self.ptr, int(use_filter), reps, table, 16, ctypes.byref(n)),
                       "hx_group_autotune")
        return [(table[k].order, table[k].workgroups_per_cu, table[k].ms) for k in range(n.value)]

    def close(self):
        if self.ptr:
            self.lib.hx_group_destroy(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
