"""The C ABI library on a host without a GPU: it loads, exports every function include/hexed_amd.h declares, serves
the basis tables, and rejects malformed meshes before any device work (no kernels are launched here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from hexed_amd import abi, basis_gen, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^[a-z_ ]+\**\s*\**(hx_[a-z0-9_]+)\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(abi.EXPORTED_SYMBOLS) == names


def test_basis_tables_served_by_library_match_json():
    lib = abi.load_library()
    tables = basis_gen.load_tables()
    for rs in range(2, 9):
        for name in ("node", "weight", "diff_mat", "boundary", "orthogonal", "filter", "prolong", "restrict", "eigenvals"):
            n = lib.hx_basis_table(rs, name.encode(), None, 0)
            out = np.zeros(n)
            assert lib.hx_basis_table(rs, name.encode(), abi.dptr(out), n) == n
            np.testing.assert_array_equal(out, np.ravel(tables[rs][name]))
        assert lib.hx_basis_max_cfl(rs) == -2 * 0.9 / tables[rs]["eigenvals"][0]
        assert lib.hx_basis_step_ratio(rs) == 0.5 / 0.9
    assert lib.hx_basis_table(9, b"node", None, 0) == -1
    assert lib.hx_basis_table(6, b"nonsense", None, 0) == -1


def test_kernels_built_for_every_factory_combination():
    """kernel_factory instantiates n_dim 1..3 x row_size 2..8 (kernel_factory.hpp:64-119)."""
    lib = abi.load_library()
    for nd in (1, 2, 3):
        for rs in range(2, 9):
            assert lib.hx_kernels_built(nd, rs, abi.HX_PDE_EULER) == 1
            assert lib.hx_kernels_built(nd, rs, abi.HX_PDE_NAVIER_STOKES) == 1
    assert lib.hx_kernels_built(4, 6, abi.HX_PDE_EULER) == 0
    assert lib.hx_kernels_built(3, 9, abi.HX_PDE_EULER) == 0
    assert lib.hx_kernels_built(3, 1, abi.HX_PDE_EULER) == 0


def validate(m_struct):
    lib = abi.load_library()
    buf = ctypes.create_string_buffer(256)
    rc = lib.hx_mesh_validate(ctypes.byref(m_struct), buf, len(buf))
    return rc, buf.value.decode()


def test_mesh_validation():
    m = mesh.box(3, 4, 2, deformed=True, jitter=0.1)
    s = m.struct()
    assert validate(s) == (abi.HX_OK, "")
    s.row_size = 9
    assert validate(s) == (abi.HX_ERR_INVALID, "demand for invalid kernel")
    s.row_size = 4
    s.n_dim = 0
    assert validate(s)[1] == "demand for invalid kernel"
    s.n_dim = 3
    m.face_slot[1][0, 0] = m.faces.shape[0]
    assert validate(m.struct())[1] == "invalid mesh: face slot out of range"
    m.face_slot[1][0, 0] = 0
    m.con_dir[1][0] = 3  # i_dim 3 does not exist in 3D
    assert validate(m.struct())[1] == "invalid mesh: connection direction out of range"
    m.con_dir[1][0] = 0
    m.bc_serial[0] = 5
    assert validate(m.struct())[1] == "invalid mesh: boundary condition serial out of range"
    m.bc_serial[0] = 0
    m.face_nslot[0, 0] = -2
    assert validate(m.struct())[1] == "invalid mesh: face normal slot out of range"
    m.face_nslot[0, 0] = -1
    assert validate(m.struct()) == (abi.HX_OK, "")


def test_no_device_reports_status_not_crash():
    """Without a GPU the context constructor returns HX_ERR_NO_DEVICE; the product never falls back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = abi.load_library()
    p = ctypes.c_void_p()
    assert lib.hx_ctx_create(0, ctypes.byref(p)) == abi.HX_ERR_NO_DEVICE
    assert not p.value
