"""The artificial-viscosity generator on the GPU against the oracle: every kernel of pde::Advection and
pde::Smooth_art_visc (Write_face, Neighbor, Local, Neighbor_reconcile, Reconcile_ldg_flux, Max_dt), their boundary
hooks, and Solver::update_art_visc_smoothness as a whole, including the reference's own convergence scenario
(test_Solver.cpp:829-878).

Bar: relative L2 <= 1e-12 per field (floating point; summation order differs between host and gfx950); the
9000-iteration scenario is compared to 1e-9 after its scaling checks."""
import ctypes
import math

import numpy as np
import pytest

from hexed_amd import abi, configs, mesh
from hexed_amd.device import Context, Device_mesh

import oracle
from helpers import freestream, init_faces, o_call, rel_l2, say
from test_av import channel, reference_opts, tanh_state, ROW_SIZE

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (n_dim, row_size, kind of mesh)
MESHES = [(3, 4, "cartesian"), (3, 6, "deformed"), (2, 5, "deformed"), (3, 4, "tree"), (2, 3, "cartesian")]


def ids(c):
    return f"nd{c[0]}-rs{c[1]}-{c[2]}"


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def build(cfg, seed=0):
    nd, rs, kind = cfg
    if kind == "tree":
        m = configs.tree_mesh(base=2, rs=rs, levels=2, frac=0.3, seed=seed, nd=nd)
    else:
        m = mesh.box(nd, rs, 3, deformed=kind == "deformed", jitter=0.1 if kind == "deformed" else 0.0, seed=seed,
                     bc=(abi.HX_BC_FREESTREAM, freestream(nd)))
    m.enable_av_generator()
    rng = np.random.default_rng(seed + 1)
    for c in (0, 1):
        n = m.n_elem[c]
        if not n:
            continue
        m.state[c][:, :nd] = rng.uniform(-1, 1, size=(n, nd, m.nq))
        m.state[c][:, nd] = rng.uniform(0.8, 1.2, size=(n, m.nq))
        m.state[c][:, nd + 1] = rng.uniform(2.0, 3.0, size=(n, m.nq))
        m.advection[c][:] = rng.uniform(0., 2., size=m.advection[c].shape)
        m.forcing[c][:] = rng.uniform(0., 1., size=m.forcing[c].shape)
        m.tss[c][:] = rng.uniform(0.5, 1.5, size=m.tss[c].shape)
        m.rcache[c][:] = rng.uniform(-1, 1, size=m.rcache[c].shape)
    m.faces[:] = rng.uniform(-1, 1, size=m.faces.shape)
    return m


def compare(tag, host, dev, fields):
    for name in fields:
        if name == "faces":
            got = dev.download("faces", 0, out=np.zeros_like(host.faces))
            assert rel_l2(got, host.faces) <= TOL, f"{tag}: faces {rel_l2(got, host.faces):.3e}"
            continue
        for c in (0, 1):
            if host.n_elem[c] == 0:
                continue
            ref = getattr(host, name)[c]
            got = dev.download(name, c, out=np.zeros_like(ref))
            err = rel_l2(got, ref)
            assert err <= TOL, f"{tag}: {name}[{c}] {err:.3e}"


def both(host, dev, oracle_name, oracle_args, dev_name, dev_args):
    s = host.struct()
    o_call(oracle_name, ctypes.byref(s), *oracle_args)
    dev.ctx.check(getattr(dev.lib, dev_name)(dev.ptr, *dev_args), dev_name)


@pytest.mark.parametrize("cfg", MESHES, ids=ids)
def test_advection_kernels(ctx, cfg):
    say(f"[gpu] advection kernels {ids(cfg)}")
    host = build(cfg)
    dev = Device_mesh(ctx, host.copy())
    L = 0.3
    pde = abi.advection_pde(L)
    for c in (0, 1):
        both(host, dev, "oracle_write_face", (c, ctypes.byref(pde)), "hx_write_face", (c, ctypes.byref(pde)))
    compare("write_face", host, dev, ["faces"])
    both(host, dev, "oracle_prolong_refined", (ctypes.byref(pde), 0, 0), "hx_prolong_refined", (ctypes.byref(pde), 0, 0))
    compare("prolong", host, dev, ["faces"])
    both(host, dev, "oracle_apply_advection_bcs", (), "hx_apply_advection_bcs", ())
    compare("advection bcs", host, dev, ["faces"])
    for c in (0, 1):
        both(host, dev, "oracle_neighbor", (c, ctypes.byref(pde), 0), "hx_neighbor", (c, ctypes.byref(pde), 0))
    compare("neighbor", host, dev, ["faces"])
    both(host, dev, "oracle_restrict_refined", (ctypes.byref(pde), 1, 0), "hx_restrict_refined", (ctypes.byref(pde), 1, 0))
    for stage in (0, 1):
        for c in (0, 1):
            both(host, dev, "oracle_local", (c, ctypes.byref(pde), 0.7, stage, 0, 0),
                 "hx_local", (c, ctypes.byref(pde), 0.7, stage, 0, 0))
        compare(f"local stage {stage}", host, dev, ["advection", "rcache", "faces", "state"])
    out_o, out_d = ctypes.c_double(), ctypes.c_double()
    for c in (0, 1):
        both(host, dev, "oracle_max_dt", (c, ctypes.byref(pde), 1, 0, 0.7, 1., ctypes.byref(out_o)),
             "hx_max_dt", (c, ctypes.byref(pde), 1, 0, 0.7, 1., ctypes.byref(out_d)))
    compare("max_dt", host, dev, ["tss"])
    dev.close()


@pytest.mark.parametrize("cfg", MESHES, ids=ids)
def test_smooth_av_kernels(ctx, cfg):
    say(f"[gpu] smooth-av kernels {ids(cfg)}")
    host = build(cfg, seed=4)
    dev = Device_mesh(ctx, host.copy())
    pde = abi.smooth_av_pde(0.02, 1.3)
    for c in (0, 1):
        both(host, dev, "oracle_write_face", (c, ctypes.byref(pde)), "hx_write_face", (c, ctypes.byref(pde)))
    both(host, dev, "oracle_apply_avc_diff_bcs", (), "hx_apply_avc_diff_bcs", ())
    compare("write_face + bcs", host, dev, ["faces"])
    for c in (0, 1):
        both(host, dev, "oracle_neighbor", (c, ctypes.byref(pde), 0), "hx_neighbor", (c, ctypes.byref(pde), 0))
    compare("neighbor", host, dev, ["faces"])
    for c in (0, 1):
        both(host, dev, "oracle_local", (c, ctypes.byref(pde), 0.4, 0, 0, 0), "hx_local", (c, ctypes.byref(pde), 0.4, 0, 0, 0))
    compare("local", host, dev, ["forcing", "faces"])
    both(host, dev, "oracle_apply_avc_diff_flux_bcs", (), "hx_apply_avc_diff_flux_bcs", ())
    for c in (0, 1):
        both(host, dev, "oracle_neighbor_reconcile", (c, ctypes.byref(pde)), "hx_neighbor_reconcile", (c, ctypes.byref(pde)))
    compare("neighbor reconcile", host, dev, ["faces"])
    for c in (0, 1):
        both(host, dev, "oracle_reconcile_ldg_flux", (c, ctypes.byref(pde), 0.4, 0, 0, 0),
             "hx_reconcile_ldg_flux", (c, ctypes.byref(pde), 0.4, 0, 0, 0))
    compare("reconcile", host, dev, ["forcing", "faces"])
    out_o, out_d = ctypes.c_double(), ctypes.c_double()
    for c in (0, 1):
        both(host, dev, "oracle_max_dt", (c, ctypes.byref(pde), 1, 0, 1., 0.7, ctypes.byref(out_o)),
             "hx_max_dt", (c, ctypes.byref(pde), 1, 0, 1., 0.7, ctypes.byref(out_d)))
    compare("max_dt", host, dev, ["tss"])
    dev.close()


SMOOTHNESS_CASES = [((3, 4, "deformed"), 4), ((3, 6, "tree"), 4), ((2, 6, "cartesian"), 4),
                    # no advection (the projection then runs as its own pass) and a single advection iteration
                    ((3, 4, "deformed"), 0), ((3, 4, "deformed"), 1)]


@pytest.mark.parametrize("cfg,advect_iters", SMOOTHNESS_CASES, ids=lambda x: ids(x) if isinstance(x, tuple) else f"adv{x}")
def test_update_art_visc_smoothness(ctx, cfg, advect_iters):
    say(f"[gpu] update_art_visc_smoothness {ids(cfg)}, {advect_iters} advection iterations")
    nd, rs, kind = cfg
    host = build(cfg, seed=9)
    host.state = [configs_state(host, c) for c in (0, 1)]
    init_faces(host)
    dev = Device_mesh(ctx, host.copy())
    opts = abi.av_options(0.2, rs, 1.0, 2.5, advect_iters=advect_iters, diff_iters=3, n_cheby=2)
    r_o = ctypes.c_double()
    s = host.struct()
    oracle.check(oracle.load().oracle_update_art_visc_smoothness(ctypes.byref(s), ctypes.byref(opts), ctypes.byref(r_o)))
    r_d = dev.update_art_visc_smoothness(opts)
    assert r_d == pytest.approx(r_o.value, rel=1e-12)
    compare("update_art_visc_smoothness", host, dev, ["av", "forcing", "advection", "state", "tss", "faces"])
    dev.close()


def configs_state(m, c):
    """A smooth flow with a steep shear layer, so that the indicator has something to find."""
    st = m.state[c]
    if m.n_elem[c] == 0:
        return st
    pos = m.qpos[c]
    nd = m.n_dim
    x = pos[:, 0]
    st = st.copy()
    st[:, 0] = 1. - .1 * np.tanh((x - .5) / .05)
    for d in range(1, nd):
        st[:, d] = 0.05 * np.sin(3 * pos[:, d])
    st[:, nd] = 1.
    st[:, nd + 1] = 2.5
    return st


def test_reference_convergence_scenario(ctx):
    """test_Solver.cpp:829-878 on the GPU: the scaling checks, and the coefficient against the oracle."""
    say("[gpu] reference artificial-viscosity convergence scenario")
    host = channel()
    tanh_state(host, 0.04)
    dev = Device_mesh(ctx, host.copy())
    lib = oracle.load()
    r = ctypes.c_double()
    oracle.check(lib.oracle_update_art_visc_smoothness(ctypes.byref(host.struct()), ctypes.byref(reference_opts(0.01)),
                                                       ctypes.byref(r)))
    dev.update_art_visc_smoothness(reference_opts(0.01))
    got = dev.download("av", 0, out=np.zeros_like(host.av[0]))
    assert rel_l2(got, host.av[0]) <= 1e-9
    init_max = got[:, 0].max()
    dev.update_art_visc_smoothness(reference_opts(0.02))
    got = dev.download("av", 0, out=np.zeros_like(host.av[0]))
    assert math.log2(got[:, 0].max() / init_max) > ROW_SIZE - 1.
    dev.close()
