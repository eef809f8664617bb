"""Hanging faces: the HIP path against the oracle on seeded 2:1-balanced tree meshes, through
Prolong_refined / Restrict_refined (Spatial.hpp:138-286) alone and inside whole time steps.

Bar: relative L2 <= 1e-12 per conserved variable (floating point; summation order differs between host and gfx950)."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, configs
from hexed_amd.device import Context, Device_mesh

from helpers import faces_rel_l2, init_faces, o_call, per_var_rel_l2, rel_l2, say, update_folds

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (n_dim, row_size, base, levels, refinement probability, seed)
TREES = [
    (3, 4, 3, 2, 0.2, 1),
    (3, 6, 2, 2, 0.3, 2),
    (2, 5, 4, 3, 0.2, 3),
    (3, 2, 3, 2, 0.3, 4),
    (2, 8, 3, 2, 0.3, 5),
]


def ids(t):
    return f"nd{t[0]}-rs{t[1]}-b{t[2]}-l{t[3]}"


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def tree(t):
    nd, rs, base, levels, frac, seed = t
    m = configs.euler_tree(base=base, rs=rs, levels=levels, frac=frac, seed=seed, nd=nd)
    assert m.ref_slot.shape[0] > 0
    init_faces(m)
    return m


def compare(tag, host, dev, tol=TOL, ldg=True):
    """ldg False: the faces without their LDG blocks (after a folded Navier-Stokes update)."""
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert faces_rel_l2(got, host.faces, host, ldg) <= tol, f"{tag}: faces"
    for name in ("state", "rcache"):
        ref = getattr(host, name)[0]
        errs = per_var_rel_l2(dev.download(name, 0, out=np.zeros_like(ref)), ref)
        assert max(errs) <= tol, f"{tag}: {name} {errs}"


@pytest.mark.parametrize("t", TREES, ids=ids)
@pytest.mark.parametrize("scale,offset", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_prolong_restrict_kernels(ctx, t, scale, offset):
    """Both kernels on random face data, with and without the stretch flags, against the oracle."""
    say(f"[gpu] prolong/restrict {ids(t)} scale {scale} offset {offset}")
    host = tree(t)
    rng = np.random.default_rng(t[-1])
    host.faces[:] = rng.standard_normal(host.faces.shape)
    if t[0] == 3:
        host.ref_stretch[:] = rng.integers(0, 2, size=host.ref_stretch.shape)
    elif t[0] == 2:
        host.ref_stretch[:, 0] = rng.integers(0, 2, size=host.ref_stretch.shape[0])
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    nv = t[0] + 2
    o_call("oracle_prolong_refined", ctypes.byref(s), None, scale, offset)
    dev.prolong(bool(scale), bool(offset))
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL
    o_call("oracle_restrict_refined", ctypes.byref(s), None, scale, offset)
    dev.restrict(bool(scale), bool(offset))
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL
    assert nv == host.nv
    dev.close()


@pytest.mark.parametrize("t", TREES, ids=ids)
def test_update_euler_tree(ctx, t):
    say(f"[gpu] tree update euler {ids(t)}")
    host = tree(t)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare(f"euler tree update {it}", host, dev)
    dev.close()


@pytest.mark.parametrize("t", TREES[:3], ids=ids)
def test_update_navier_stokes_tree(ctx, t):
    """The LDG stage prolongs and restricts the LDG blocks too (kernels_diffusive.cpp:8-29)."""
    say(f"[gpu] tree update navier-stokes {ids(t)}")
    nd, rs, base, levels, frac, seed = t
    host = configs.euler_tree(base=base, rs=rs, levels=levels, frac=frac, seed=seed, nd=nd)
    host.av = [np.full((host.n_elem[0], 2, host.nq), 1e-3), None]
    init_faces(host)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    visc, cond = abi.constant(2e-3), abi.constant(2.8e-3)
    opts = abi.update_options(max_safety=0.3, visc=visc, cond=cond)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare(f"ns tree update {it}", host, dev, ldg=not update_folds(host))
    dev.close()


@pytest.mark.parametrize("fold", [0, 1])
@pytest.mark.parametrize("t", TREES[:2], ids=ids)
def test_update_euler_tree_fold_modes(ctx, t, fold):
    """hx_update on 3D tree meshes with the face fluxes folded (where the Neighbor on the mortar connections and the
    restriction run as one launch per stage, k_mortar_restrict) and not, each forced at both row sizes, against the
    oracle after every step."""
    say(f"[gpu] tree update euler fold {fold} {ids(t)}")
    host = tree(t)
    dev = Device_mesh(ctx, host.copy())
    dev.set_face_fold(fold)
    s = host.struct()
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare(f"euler tree update fold {fold} {it}", host, dev)
    dev.close()


def test_freestream_tree(ctx):
    fs = configs.freestream_state(3, 1.0, [0.8, 0.1, -0.2], 1.0 / 1.4)
    host = configs.tree_mesh(base=3, rs=6, levels=2, frac=0.3, seed=7, bc=(abi.HX_BC_FREESTREAM, fs))
    for v in range(5):
        host.state[0][:, v, :] = fs[v]
    init_faces(host)
    dev = Device_mesh(ctx, host.copy())
    for precision in (abi.HX_FP64, abi.HX_FP32_MFMA):
        dev.set_precision(precision)
        st = abi.hx_update_status()
        for _ in range(2):
            dev.update(abi.update_options(max_safety=0.5), st)
        got = dev.download("state", 0, out=np.zeros_like(host.state[0]))
        tol = 1e-13 if precision == abi.HX_FP64 else 1e-6
        for v in range(5):
            assert np.abs(got[:, v] - fs[v]).max() <= tol * 2.2, f"precision {precision} variable {v}"
    dev.close()
