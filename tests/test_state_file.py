"""State files in the reference's format (Solver::write_state / read_state, Solver.cpp:232-271): the element storage
layout of Element.cpp:114-194, the HDF5 dataset the reference reads and writes, and a restart that continues exactly
as an uninterrupted run (oracle)."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, mesh, state_file

from helpers import freestream, init_faces, o_call, smooth_state

try:
    state_file._hdf5()
    HAVE_HDF5 = True
except RuntimeError:
    HAVE_HDF5 = False
needs_hdf5 = pytest.mark.skipif(not HAVE_HDF5, reason="the HDF5 C library is not in this image")


def mixed_mesh(nd=2, rs=4):
    """Cartesian and deformed elements in one mesh, with every optional field present."""
    b = mesh.Builder(nd, rs, root_size=0.5)
    bc = b.add_bc(abi.HX_BC_FREESTREAM, freestream(nd))
    for y, deformed in ((0, False), (2, True)):  # a Cartesian pair and, apart from it, a deformed pair
        e0 = b.add_element(0, deformed, [0, y])
        e1 = b.add_element(0, deformed, [1, y])
        if deformed:
            b.connect_deformed(e0, e1, (0, 0), (True, False))
        else:
            b.connect_cartesian(e0, e1, 0)
        for e, d, s in [(e0, 0, 0), (e1, 0, 1), (e0, 1, 0), (e0, 1, 1), (e1, 1, 0), (e1, 1, 1)]:
            b.connect_boundary(e, d, s, bc)
    m = b.build(with_av=True).enable_av_generator()
    smooth_state(m, seed=1)
    rng = np.random.default_rng(2)
    for c in (0, 1):
        for name in ("tss", "av", "forcing", "advection", "rcache"):
            arr = getattr(m, name)[c]
            arr[:] = rng.uniform(0.1, 1.0, size=arr.shape)
    return m


def test_n_var_numeric():
    # Storage_params.cpp:32-35 with two stages and 4 forcing variables
    assert state_file.n_var_numeric(3, 6) == 5 + 3 + 4 + 6 + 6
    assert state_file.n_var_numeric(2, 2) == 4 + 3 + 4 + 2 + 4
    assert state_file.n_var_numeric(3, 8) == 5 + 3 + 4 + 8 + 8


def test_element_layout():
    """Each field at the offset of its Element accessor (Element.cpp:119-188): state(), time_step_scale() at n_dof,
    bulk_av_coef() and laplacian_av_coef() next, art_visc_forcing(), advection_state(), residual_cache()."""
    m = mixed_mesh()
    nd, rs, nv, nq = m.n_dim, m.row_size, m.nv, m.nq
    data = state_file.elements_of(m, state_file.host_fields(m))
    assert data.shape == (m.n_elem[0] + m.n_elem[1], state_file.n_var_numeric(nd, rs), nq)
    for c, E in ((0, 0), (1, m.n_elem[0])):
        np.testing.assert_array_equal(data[E, :nv], m.state[c][0])
        np.testing.assert_array_equal(data[E, nv], m.tss[c][0])
        np.testing.assert_array_equal(data[E, nv + 1], m.av[c][0, 0])
        np.testing.assert_array_equal(data[E, nv + 2], m.av[c][0, 1])
        np.testing.assert_array_equal(data[E, nv + 3:nv + 7], m.forcing[c][0])
        np.testing.assert_array_equal(data[E, nv + 7:nv + 7 + rs], m.advection[c][0])
        np.testing.assert_array_equal(data[E, nv + 7 + rs:], m.rcache[c][0])


@needs_hdf5
def test_file_round_trip(tmp_path):
    m = mixed_mesh()
    name = str(tmp_path / "case")
    state_file.write_mesh(name, m)
    # the dataset the reference writes: "state", doubles, [n_elem][n_var_numeric][n_qpoint]
    h5 = state_file._hdf5()
    f = h5.H5Fopen((name + ".state.h5").encode(), 0, 0)
    assert f >= 0
    d = h5.H5Dopen2(f, b"state", 0)
    assert d >= 0
    t = h5.H5Dget_type(d)
    assert h5.H5Tget_class(t) == state_file.H5T_FLOAT and h5.H5Tget_size(t) == 8
    sp = h5.H5Dget_space(d)
    dims = (ctypes.c_ulonglong * 3)()
    assert h5.H5Sget_simple_extent_ndims(sp) == 3
    h5.H5Sget_simple_extent_dims(sp, dims, None)
    assert tuple(dims) == (m.n_elem[0] + m.n_elem[1], state_file.n_var_numeric(2, 4), m.nq)
    for closer, h in ((h5.H5Sclose, sp), (h5.H5Tclose, t), (h5.H5Dclose, d), (h5.H5Fclose, f)):
        closer(h)
    # read back into zeroed fields
    z = m.copy()
    for c in (0, 1):
        for name_ in ("state", "tss", "av", "forcing", "advection", "rcache"):
            getattr(z, name_)[c][:] = 0.
    state_file.read_mesh(name, z)
    for c in (0, 1):
        for name_ in ("state", "tss", "av", "forcing", "advection", "rcache"):
            np.testing.assert_array_equal(getattr(z, name_)[c], getattr(m, name_)[c])
    # a mesh of another size refuses the file
    other = mesh.box(2, 4, 3, with_av=True)
    with pytest.raises(ValueError, match="state file holds"):
        state_file.read_mesh(name, other)


@needs_hdf5
@pytest.mark.parametrize("ns", [False, True])
def test_restart_continues_exactly(tmp_path, ns):
    """Three updates in one go equal one update, write_state, read_state into a freshly built mesh, faces written
    as Solver::initialize writes them, and two more updates (oracle)."""
    nd, rs = 3, 4
    def fresh():
        m = mesh.box(nd, rs, 3, deformed=True, jitter=0.1, seed=4, bc=(abi.HX_BC_FREESTREAM, freestream(nd)),
                     with_av=True)
        smooth_state(m, seed=4)
        init_faces(m)
        return m
    opts = abi.update_options(max_safety=0.5, visc=abi.constant(2e-3) if ns else None,
                              cond=abi.constant(3e-3) if ns else None)
    a = fresh()
    sa, st = a.struct(), abi.hx_update_status()
    for _ in range(3):
        o_call("oracle_update", ctypes.byref(sa), ctypes.byref(opts), ctypes.byref(st))
    b = fresh()
    sb = b.struct()
    o_call("oracle_update", ctypes.byref(sb), ctypes.byref(opts), ctypes.byref(abi.hx_update_status()))
    name = str(tmp_path / "restart")
    state_file.write_mesh(name, b)
    r = fresh()
    state_file.read_mesh(name, r)
    init_faces(r)
    sr = r.struct()
    for _ in range(2):
        o_call("oracle_update", ctypes.byref(sr), ctypes.byref(opts), ctypes.byref(abi.hx_update_status()))
    np.testing.assert_array_equal(r.state[1], a.state[1])
    np.testing.assert_array_equal(r.rcache[1], a.rcache[1])
