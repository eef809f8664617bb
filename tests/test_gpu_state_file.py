"""State files on the GPU: a file written from device fields equals the oracle's, a file read onto the device restores
every field, and a device restart (write_state, read_state into a new device mesh, faces written, updates) continues
as the uninterrupted run. Bar: relative L2 <= 1e-12."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, mesh, state_file
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, init_faces, o_call, rel_l2, say, smooth_state
from test_gpu_parity import compare_mesh
from test_state_file import HAVE_HDF5, mixed_mesh

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not HAVE_HDF5, reason="the HDF5 C library is not in this image")]

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_read_write_fields(ctx, tmp_path):
    say("[gpu] state file fields")
    host = mixed_mesh()
    name = str(tmp_path / "fields")
    state_file.write_mesh(name, host)
    blank = host.copy()
    for c in (0, 1):
        for f in ("state", "tss", "av", "forcing", "advection", "rcache"):
            getattr(blank, f)[c][:] = 0.
    dev = Device_mesh(ctx, blank)
    dev.read_state(name)
    compare_mesh("read_state", host, dev, fields=("state", "tss", "av", "forcing", "advection", "rcache"), tol=0.)
    dev.write_state(name + "_dev")
    np.testing.assert_array_equal(state_file.read(name + "_dev"), state_file.read(name))


@pytest.mark.parametrize("ns", [False, True])
def test_restart(ctx, tmp_path, ns):
    say(f"[gpu] state file restart ns={ns}")
    nd, rs = 3, 4

    def fresh():
        m = mesh.box(nd, rs, 3, deformed=True, jitter=0.1, seed=4, bc=(abi.HX_BC_FREESTREAM, freestream(nd)),
                     with_av=True)
        smooth_state(m, seed=4)
        init_faces(m)
        return m
    opts = abi.update_options(max_safety=0.5, visc=abi.constant(2e-3) if ns else None,
                              cond=abi.constant(3e-3) if ns else None)
    # the oracle, uninterrupted
    ref = fresh()
    s = ref.struct()
    o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(abi.hx_update_status()))
    after_one = ref.copy()
    for _ in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(abi.hx_update_status()))
    # the device: one update, write, read into a new device mesh, faces as Solver::initialize writes them, two more
    dev = Device_mesh(ctx, fresh())
    dev.update(opts)
    name = str(tmp_path / "restart")
    dev.write_state(name)
    file_data = state_file.read(name)
    want = state_file.elements_of(after_one, state_file.host_fields(after_one))
    assert rel_l2(file_data[:, :after_one.nv], want[:, :after_one.nv]) <= TOL
    dev2 = Device_mesh(ctx, fresh())
    dev2.read_state(name)
    dev2.compute_write_face()
    dev2.ctx.check(dev2.lib.hx_compute_prolong(dev2.ptr, 0, 0), "prolong")
    for _ in range(2):
        dev2.update(opts)
    compare_mesh("restart", ref, dev2, fields=("state",), tol=TOL)
