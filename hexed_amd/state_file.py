"""State files in the reference's format (Solver::write_state / read_state, Solver.cpp:232-271).

A file `<name>.state.h5` holds one HDF5 dataset "state" of doubles, shape [n_elem][n_var_numeric][n_qpoint]: every
element's whole storage in the order of Accessible_mesh::elements() (the Cartesian elements, then the deformed ones,
Accessible_mesh.cpp:133). n_var_numeric = nv + 3 + n_forcing + row_size + max(nv, row_size) (Storage_params.cpp:32-35,
two stages, n_forcing 4), and each element's variables are laid out as Element.cpp:114-194 keeps them:

    [0, nv)                      conserved state
    nv                           time step scale
    nv + 1, nv + 2               bulk and Laplacian artificial viscosity
    [nv + 3, nv + 7)             artificial-viscosity forcing
    [nv + 7, nv + 7 + rs)        advection state
    [nv + 7 + rs, n_var_numeric) residual cache (stage-0 rate)

The files are read and written through the HDF5 C library of the image (libhdf5, found next to the Python prefix or
through HEXED_AMD_HDF5); `elements_of` / `set_elements` convert between that layout and a HostMesh or Device_mesh.
"""
import ctypes
import os
import sys

import numpy as np

N_FORCING = 4


def n_var_numeric(nd, rs):
    nv = nd + 2
    return nv + 3 + N_FORCING + rs + max(nv, rs)


def _layout(nd, rs):
    nv = nd + 2
    return dict(state=(0, nv), tss=(nv, nv + 1), av=(nv + 1, nv + 3), forcing=(nv + 3, nv + 7),
                advection=(nv + 7, nv + 7 + rs), rcache=(nv + 7 + rs, nv + 7 + rs + max(nv, rs)))


def elements_of(m, fields):
    """[n_elem][n_var_numeric][nq] storage of every element of `m`, Cartesian first. `fields(name, cls)` returns the
    [n][k][nq] array of a field (None when the mesh has none: those variables are zero)."""
    nd, rs, nq = m.n_dim, m.row_size, m.nq
    out = np.zeros((m.n_elem[0] + m.n_elem[1], n_var_numeric(nd, rs), nq))
    lo = 0
    for c in (0, 1):
        n = m.n_elem[c]
        if n == 0:
            continue
        for name, (a, b) in _layout(nd, rs).items():
            arr = fields(name, c)
            if arr is not None:
                out[lo:lo + n, a:b] = np.asarray(arr).reshape(n, b - a, nq)
        lo += n
    return out


def set_elements(m, data, store):
    """Inverse of elements_of: store(name, cls, [n][k][nq] array) for every field of every class."""
    nd, rs = m.n_dim, m.row_size
    data = np.asarray(data)
    expect = (m.n_elem[0] + m.n_elem[1], n_var_numeric(nd, rs), m.nq)
    if data.shape != expect:
        # the reference's read_state checks the storage parameters the same way (Solver.cpp:224-228)
        raise ValueError(f"state file holds {data.shape}, the mesh needs {expect}")
    lo = 0
    for c in (0, 1):
        n = m.n_elem[c]
        if n == 0:
            continue
        for name, (a, b) in _layout(nd, rs).items():
            store(name, c, np.ascontiguousarray(data[lo:lo + n, a:b]))
        lo += n


def host_fields(m):
    def get(name, c):
        if name == "tss":
            return m.tss[c][:, None, :]
        arr = getattr(m, name)[c]
        return arr
    return get


# ---- the HDF5 C library ------------------------------------------------------------------------------------------

_h5 = None


def _hdf5():
    global _h5
    if _h5 is not None:
        return _h5
    candidates = [os.environ.get("HEXED_AMD_HDF5", ""), os.path.join(sys.prefix, "lib", "libhdf5.so"),
                  "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    lib = None
    for c in candidates:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError:
            continue
    if lib is None:
        raise RuntimeError("state files need the HDF5 C library (libhdf5.so); set HEXED_AMD_HDF5 to its path")
    hid, herr, hsize = ctypes.c_int64, ctypes.c_int, ctypes.c_ulonglong
    sig = {
        "H5open": (herr, []),
        "H5Fcreate": (hid, [ctypes.c_char_p, ctypes.c_uint, hid, hid]),
        "H5Fopen": (hid, [ctypes.c_char_p, ctypes.c_uint, hid]),
        "H5Fclose": (herr, [hid]),
        "H5Screate_simple": (hid, [ctypes.c_int, ctypes.POINTER(hsize), ctypes.POINTER(hsize)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid, ctypes.POINTER(hsize), ctypes.POINTER(hsize)]),
        "H5Sclose": (herr, [hid]),
        "H5Dcreate2": (hid, [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]),
        "H5Dopen2": (hid, [hid, ctypes.c_char_p, hid]),
        "H5Dget_space": (hid, [hid]),
        "H5Dget_type": (hid, [hid]),
        "H5Dwrite": (herr, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
        "H5Dread": (herr, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
        "H5Dclose": (herr, [hid]),
        "H5Tget_class": (ctypes.c_int, [hid]),
        "H5Tget_size": (ctypes.c_size_t, [hid]),
        "H5Tclose": (herr, [hid]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    if lib.H5open() < 0:
        raise RuntimeError("H5open failed")
    lib.native_double = hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    _h5 = lib
    return lib


H5F_ACC_RDONLY, H5F_ACC_TRUNC, H5P_DEFAULT, H5S_ALL, H5T_FLOAT = 0, 2, 0, 0, 1


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def write(name, data):
    """Solver::write_state (Solver.cpp:253-271): `name`.state.h5 with dataset "state" = data (doubles)."""
    h5 = _hdf5()
    data = np.ascontiguousarray(data, dtype=np.float64)
    if data.ndim != 3:
        raise ValueError("state data is [n_elem][n_var_numeric][n_qpoint]")
    f = _ok(h5.H5Fcreate((name + ".state.h5").encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        dims = (ctypes.c_ulonglong * 3)(*data.shape)
        sp = _ok(h5.H5Screate_simple(3, dims, None), "H5Screate_simple")
        try:
            d = _ok(h5.H5Dcreate2(f, b"state", h5.native_double, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2")
            try:
                _ok(h5.H5Dwrite(d, h5.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data), "H5Dwrite")
            finally:
                h5.H5Dclose(d)
        finally:
            h5.H5Sclose(sp)
    finally:
        h5.H5Fclose(f)


def read(name):
    """Solver::read_state (Solver.cpp:232-251): the "state" dataset of `name`.state.h5, converted to doubles."""
    h5 = _hdf5()
    f = _ok(h5.H5Fopen((name + ".state.h5").encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    try:
        d = _ok(h5.H5Dopen2(f, b"state", H5P_DEFAULT), "H5Dopen2 state")
        try:
            sp = _ok(h5.H5Dget_space(d), "H5Dget_space")
            try:
                if h5.H5Sget_simple_extent_ndims(sp) != 3:
                    raise RuntimeError("the state dataset is not 3-dimensional")
                dims = (ctypes.c_ulonglong * 3)()
                h5.H5Sget_simple_extent_dims(sp, dims, None)
            finally:
                h5.H5Sclose(sp)
            out = np.empty(tuple(int(x) for x in dims))
            _ok(h5.H5Dread(d, h5.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread")
        finally:
            h5.H5Dclose(d)
    finally:
        h5.H5Fclose(f)
    return out


def write_mesh(name, m):
    """write_state of a HostMesh."""
    write(name, elements_of(m, host_fields(m)))


def read_mesh(name, m):
    """read_state into a HostMesh (fields it lacks, such as forcing on a flow-only mesh, are skipped)."""
    def store(field, c, arr):
        if field == "tss":
            m.tss[c][:] = arr[:, 0]
            return
        tgt = getattr(m, field)[c]
        if tgt is not None:
            tgt[:] = arr
    set_elements(m, read(name), store)
