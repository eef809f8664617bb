"""FP64 form of the HIP path against the CPU oracle, kernel by kernel and stage by stage.

Inputs are identical seeded states on identical meshes (built once on the host, copied to both sides).
Bar: relative L2 <= 1e-12 per conserved variable after every stage (floating point; not bit-exact because
summation order and FMA contraction differ between the host and gfx950 code).
"""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import (faces_rel_l2, freestream, init_faces, o_call, per_var_rel_l2, rel_l2, say, smooth_state,
                     update_folds)

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (n_dim, row_size, deformed, elements per axis)
CONFIGS = [
    (3, 6, False, 3),
    (3, 6, True, 3),
    (2, 6, True, 4),
    (2, 3, False, 5),
    (3, 4, True, 3),
    (3, 8, True, 2),
    (3, 7, False, 2),
    (1, 5, True, 6),
    (3, 2, False, 3),
]


def ids(c):
    return f"nd{c[0]}-rs{c[1]}-{'def' if c[2] else 'car'}-n{c[3]}"


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def build(cfg, bc_kind=abi.HX_BC_FREESTREAM, with_av=False, seed=0):
    nd, rs, deformed, n = cfg
    m = mesh.box(nd, rs, n, deformed=deformed, jitter=0.1 if deformed else 0.0, seed=seed,
                 bc=(bc_kind, freestream(nd)), with_av=with_av)
    smooth_state(m, seed=seed)
    if with_av:
        rng = np.random.default_rng(seed + 7)
        for c in (0, 1):
            if m.n_elem[c]:
                m.av[c][:] = rng.uniform(0.0, 2e-3, size=m.av[c].shape)
    init_faces(m)
    return m


def compare_mesh(tag, host, dev, fields=("state", "rcache", "faces"), tol=TOL, ldg=True):
    """ldg False: the faces without their LDG blocks (after a folded Navier-Stokes update)."""
    worst = 0.0
    for name in fields:
        if name == "faces":
            got = dev.download("faces", 0, out=np.zeros_like(host.faces))
            err = faces_rel_l2(got, host.faces, host, ldg)
            assert err <= tol, f"{tag}: faces rel L2 {err:.3e}"
            worst = max(worst, err)
            continue
        for c in (0, 1):
            if host.n_elem[c] == 0:
                continue
            ref = getattr(host, name)[c]
            got = dev.download(name, c, out=np.zeros_like(ref))
            if ref.ndim == 3:
                errs = per_var_rel_l2(got, ref)
            else:
                errs = [rel_l2(got, ref)]
            assert max(errs) <= tol, f"{tag}: {name}[{c}] rel L2 {errs}"
            worst = max(worst, max(errs))
    return worst


@pytest.mark.parametrize("cfg", CONFIGS, ids=ids)
def test_neighbor_local_stages(ctx, cfg):
    """Neighbor, Local stage 0 and stage 1 (compute_euler sequence), compared after every call."""
    say(f"[gpu] neighbor/local stages {ids(cfg)}")
    host = build(cfg)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    dt = 1e-3
    for stage in (0, 1):
        o_call("oracle_apply_state_bcs", ctypes.byref(s))
        dev.apply_state_bcs()
        for c in (0, 1):
            o_call("oracle_neighbor", ctypes.byref(s), c, None, stage)
            dev.neighbor(c, i_stage=stage)
        compare_mesh(f"neighbor stage {stage}", host, dev, fields=("faces",))
        for c in (0, 1):
            o_call("oracle_local", ctypes.byref(s), c, None, dt, stage, 0, 0)
            dev.local(c, dt, i_stage=stage)
        compare_mesh(f"local stage {stage}", host, dev)


@pytest.mark.parametrize("cfg", CONFIGS, ids=ids)
def test_max_dt(ctx, cfg):
    say(f"[gpu] max_dt {ids(cfg)}")
    host = build(cfg)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    for c in (0, 1):
        if host.n_elem[c] == 0:
            continue
        for is_local in (0, 1):
            ref = ctypes.c_double()
            o_call("oracle_max_dt", ctypes.byref(s), c, None, is_local, 0, 0.7, 0.7, ctypes.byref(ref))
            got = dev.max_dt(c, is_local=bool(is_local))
            assert abs(got - ref.value) <= 1e-13 * abs(ref.value), (c, is_local, got, ref.value)
            tss = dev.download("tss", c, out=np.zeros_like(host.tss[c]))
            assert rel_l2(tss, host.tss[c]) <= 1e-13


@pytest.mark.parametrize("cfg", CONFIGS, ids=ids)
def test_update_euler(ctx, cfg):
    """Three Solver::update calls (global time step, 2 stages each) with freestream boundaries."""
    say(f"[gpu] update euler {ids(cfg)}")
    host = build(cfg)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(3):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare_mesh(f"update {it}", host, dev)
    assert st_d.iteration == st_o.iteration == 3


@pytest.mark.parametrize("cfg", [(3, 6, False, 3), (3, 6, True, 3), (2, 5, True, 4), (3, 4, False, 3)], ids=ids)
def test_update_filter_local_time(ctx, cfg):
    """Modal filter and local time stepping on (Spatial.hpp:472-481, Max_dt is_local)."""
    say(f"[gpu] update filter/local-time {ids(cfg)}")
    host = build(cfg, seed=3)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    opts = abi.update_options(max_safety=0.4, use_filter=True, local_time=True)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        compare_mesh(f"update {it}", host, dev, fields=("state", "rcache", "faces", "tss"))


NS_CONFIGS = [(3, 6, False, 3), (3, 6, True, 3), (2, 6, True, 4), (3, 4, True, 3), (3, 8, True, 2), (2, 3, False, 4)]


def sutherland_air():
    # Case.hil:83-90 (dimensional air), scaled so the nondimensional test state sees Re ~ 1e2
    return abi.sutherland(1.716e-5 * 2e3, 273., 111.), abi.sutherland(0.0241 * 2e3, 273., 194.)


@pytest.mark.parametrize("cfg", [(3, 6, True, 3), (3, 6, False, 3), (2, 5, True, 4), (3, 7, True, 2)], ids=ids)
def test_update_navier_stokes_filter(ctx, cfg):
    """Modal filter on the LDG stage 0 residual (Spatial.hpp:472-481): the convective and viscous rates are each
    filtered along every axis before the update."""
    say(f"[gpu] update navier-stokes filter {ids(cfg)}")
    host = build(cfg, with_av=True, seed=6)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    visc, cond = sutherland_air()
    opts = abi.update_options(max_safety=0.3, visc=visc, cond=cond, use_filter=True)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        compare_mesh(f"update {it}", host, dev, ldg=not update_folds(host))


@pytest.mark.parametrize("cfg", NS_CONFIGS, ids=ids)
@pytest.mark.parametrize("bc_kind", [abi.HX_BC_FREESTREAM, abi.HX_BC_NONPENETRATION, abi.HX_BC_OUTFLOW])
def test_update_navier_stokes(ctx, cfg, bc_kind):
    """LDG stage 0 (compute_navier_stokes) + Euler stage 1 with artificial viscosity fields present."""
    say(f"[gpu] update navier-stokes {ids(cfg)} bc {bc_kind}")
    host = build(cfg, bc_kind=bc_kind, with_av=True, seed=5)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    visc, cond = sutherland_air()
    opts = abi.update_options(max_safety=0.3, visc=visc, cond=cond)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare_mesh(f"ns update {it}", host, dev, ldg=not update_folds(host))


@pytest.mark.parametrize("cfg", [(3, 6, True, 3), (3, 4, True, 3), (2, 3, False, 5), (3, 5, False, 3)], ids=ids)
@pytest.mark.parametrize("fold", [0, 1])
@pytest.mark.parametrize("ns", [False, True], ids=["euler", "ns"])
def test_update_face_fold_modes(ctx, cfg, fold, ns):
    """hx_update with the face fluxes folded into the Local (1) and with the Neighbor-then-Local sequence (0), each
    forced on row sizes where the default picks the other, against the oracle after every step."""
    say(f"[gpu] update face fold {fold} {'ns' if ns else 'euler'} {ids(cfg)}")
    host = build(cfg, with_av=ns, seed=9)
    dev = Device_mesh(ctx, host.copy())
    dev.set_face_fold(fold)
    s = host.struct()
    if ns:
        visc, cond = sutherland_air()
        opts = abi.update_options(max_safety=0.3, visc=visc, cond=cond)
    else:
        opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare_mesh(f"fold {fold} update {it}", host, dev, ldg=not (ns and fold))


@pytest.mark.parametrize("cfg", [(3, 6, True, 3), (2, 4, False, 4)], ids=ids)
def test_residual(ctx, cfg):
    """Solver::compute_residual: residual cache only, state untouched."""
    say(f"[gpu] compute_residual {ids(cfg)}")
    host = build(cfg, with_av=True, seed=11)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    visc, cond = sutherland_air()
    for opts in (abi.update_options(), abi.update_options(visc=visc, cond=cond)):
        o_call("oracle_compute_residual", ctypes.byref(s), ctypes.byref(opts))
        dev.compute_residual(opts)
        compare_mesh("residual", host, dev)


# ---- every connection orientation and the reference's rotated test meshes ----------------------------------------

import itertools  # noqa: E402

import meshes  # noqa: E402

ORIENTATIONS = [(nd, dims, signs)
                for nd in (2, 3)
                for dims in itertools.product(range(nd), repeat=2)
                for signs in itertools.product((0, 1), repeat=2)
                if not (dims[0] == dims[1] and signs[0] == signs[1])]


def run_both(ctx, host, opts, steps=2):
    """The oracle's updates against hx_update's, with the face fluxes folded as the row size decides and with the
    reference's Neighbor-then-Local sequence (where the whole face buffer, LDG blocks included, must match)."""
    start = host.copy()
    for fold in (-1, 0):
        h = start.copy()
        dev = Device_mesh(ctx, h.copy())
        dev.set_face_fold(fold)
        s = h.struct()
        st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
        ldg = not (opts.use_ldg and update_folds(h, fold))
        for it in range(steps):
            o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
            dev.update(opts, st_d)
            compare_mesh(f"fold {fold} update {it}", h, dev, ldg=ldg)
        dev.close()


@pytest.mark.parametrize("nd,dims,signs", ORIENTATIONS, ids=lambda v: str(v))
def test_orientation_parity(ctx, nd, dims, signs):
    """Two deformed elements joined in every orientation (face permutation + normal signs on the device)."""
    say(f"[gpu] orientation nd{nd} dims {dims} signs {signs}")
    b, host = meshes.pair(nd, 4, dims, signs, bc_kind=abi.HX_BC_COPY)
    smooth_state(host, seed=13)
    host.av = [np.zeros((n, 2, host.nq)) for n in host.n_elem]
    init_faces(host)
    run_both(ctx, host, abi.update_options(max_safety=0.05, visc=abi.constant(1e-2), cond=abi.constant(1.4e-2),
                                           use_ldg=True), steps=1)
    run_both(ctx, host, abi.update_options(max_safety=0.05), steps=1)


@pytest.mark.parametrize("name", ["all_cartesian", "all_deformed_rot0", "all_deformed_rot1"])
def test_reference_test_meshes(ctx, name):
    say(f"[gpu] reference test mesh {name}")
    if name == "all_cartesian":
        host = meshes.all_cartesian(abi.HX_BC_NONPENETRATION, with_av=True)
    else:
        host = meshes.all_deformed(int(name[-1]), abi.HX_BC_NONPENETRATION, with_av=True)
    smooth_state(host, seed=17)
    init_faces(host)
    run_both(ctx, host, abi.update_options(max_safety=0.1, visc=abi.constant(1e-2), cond=abi.constant(1.4e-2),
                                           use_ldg=True))
