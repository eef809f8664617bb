"""hx_group (meshes stepped together, each kernel of an Euler stage launched once for all meshes, both element
classes and row sizes 4 to 6) against stepping every mesh alone with hx_update_cases: the same time step per case and
bit-identical states after every step, in both arithmetic forms, with Cartesian and deformed cases, a tree mesh with
hanging faces in the group, the filter and Chebyshev stages, every launch shape of the fused Local and the autotuner's
choice; the fused launches must be the ones that ran (one per kernel per stage); the per-mesh fallback for
Navier-Stokes; the lifetime and registration errors."""
import numpy as np
import pytest

from hexed_amd import abi, configs
from hexed_amd.device import Context, Device_group, Device_mesh, HxError

from helpers import init_faces, say

pytestmark = pytest.mark.gpu

STEPS = 3


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def batch_meshes(deformed=False, with_tree=False):
    """Cases of row sizes 4, 5 and 6 with uneven sizes and Mach numbers, one batch mesh per row size (a tree mesh with
    hanging faces in place of the row-size-4 batch when asked)."""
    cases = configs.batch_cases(n_cases=9, seed=3, min_elem=8, max_elem=216)
    groups = configs.euler_batch(cases)
    out = [b for rs, (b, ks) in sorted(groups.items())]
    if deformed:
        out.append(configs.euler_box(3, 5, mach=0.5, seed=11, deformed=True, jitter=0.1))
    if with_tree:
        out[0] = configs.euler_tree(base=2, rs=4, levels=2, frac=0.3, seed=1)
    return out


def devices(ctx, hosts, precision):
    out = []
    for h in hosts:
        init_faces(h)
        d = Device_mesh(ctx, h.copy())
        d.set_precision(precision)
        ids = getattr(h, "case_id", None)
        if ids is None or not any(c is not None and len(c) for c in ids):
            ids = [np.zeros(h.n_elem[c], np.int32) if h.n_elem[c] else None for c in (0, 1)]
        d.set_cases(int(max(int(np.max(c)) for c in ids if c is not None and len(c))) + 1, ids)
        out.append(d)
    return out


def compare(tag, hosts, singles, grouped):
    for h, a, b in zip(hosts, singles, grouped):
        for c in (0, 1):
            if h.n_elem[c]:
                sa = a.download("state", c, out=np.zeros_like(h.state[c]))
                sb = b.download("state", c, out=np.zeros_like(h.state[c]))
                assert np.array_equal(sa, sb), f"{tag}: row size {h.row_size} class {c}"


def run(ctx, hosts, opts, precision, shape=None, tune=False):
    singles = devices(ctx, hosts, precision)
    members = devices(ctx, hosts, precision)
    g = Device_group(ctx, members)
    assert g.fused
    if shape is not None:
        g.set_shape(*shape)
    if tune:
        table = g.autotune(use_filter=bool(opts.use_filter), reps=2)
        say("  autotune: " + ", ".join(f"({o},{w}) {ms:.3f} ms" for o, w, ms in table))
        # the tuner's residual-mode launches leave the state unchanged; rewrite the singles' faces the same way
        for d in members:
            d.compute_write_face()
        for d in singles:
            d.compute_write_face()
    st_s = [(abi.hx_update_status * d.n_cases)() for d in singles]
    st_g = (abi.hx_update_status * g.n_cases)()
    ctx.set_timing(True)
    ctx.timing_reset()
    for step in range(STEPS):
        for d, s in zip(singles, st_s):
            d.update_cases(opts, s)
        g.update(opts, st_g)
        off = 0
        for s in st_s:
            for k in range(len(s)):
                assert st_g[off + k].time_step == s[k].time_step, (step, off + k)
                assert st_g[off + k].flow_time == s[k].flow_time
            off += len(s)
        compare(f"step {step}", hosts, singles, members)
    t = ctx.timings()
    ctx.set_timing(False)
    n_stage = 2 * STEPS * opts.n_cheby_flow * opts.flow_iters
    assert t["fused/local/euler"][1] == n_stage and t["fused/neighbor/euler"][1] == n_stage
    g.close()
    for d in singles + members:
        d.close()


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
def test_group_matches_cases(ctx, precision):
    say(f"[gpu] group of batch meshes (row sizes 4-6) vs hx_update_cases, precision {precision}")
    run(ctx, batch_meshes(), abi.update_options(max_safety=0.6), precision)


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
def test_group_deformed_tree_filter(ctx, precision):
    say(f"[gpu] group with deformed cases and a tree mesh, filter and 2 Chebyshev stages, precision {precision}")
    opts = abi.update_options(max_safety=0.5, n_cheby_flow=2, use_filter=True)
    run(ctx, batch_meshes(deformed=True, with_tree=True), opts, precision)


@pytest.mark.parametrize("shape", [(2, 0), (0, 2), (0, 8), (1, 4), (1, 16)], ids=lambda s: f"order{s[0]}-cu{s[1]}")
def test_group_launch_shapes(ctx, shape):
    say(f"[gpu] group launch shape {shape}")
    run(ctx, batch_meshes(deformed=True), abi.update_options(max_safety=0.6), abi.HX_FP64, shape=shape)


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
def test_group_autotuned(ctx, precision):
    say(f"[gpu] group after autotuning, precision {precision}")
    run(ctx, batch_meshes(), abi.update_options(max_safety=0.6), precision, tune=True)


def test_group_navier_stokes_steps_meshes_alone(ctx):
    say("[gpu] group with Navier-Stokes options (per-mesh fallback)")
    visc, cond = configs.sutherland_air()
    opts = abi.update_options(max_safety=0.5, visc=visc, cond=cond, use_ldg=True)
    hosts = [configs.ns_deformed(2, rs=5, seed=1), configs.ns_deformed(3, rs=4, seed=2)]
    singles, members = devices(ctx, hosts, abi.HX_FP64), devices(ctx, hosts, abi.HX_FP64)
    g = Device_group(ctx, members)
    st_s = [(abi.hx_update_status * d.n_cases)() for d in singles]
    st_g = (abi.hx_update_status * g.n_cases)()
    for _ in range(2):
        for d, s in zip(singles, st_s):
            d.update_cases(opts, s)
        g.update(opts, st_g)
    compare("navier-stokes", hosts, singles, members)
    g.close()
    for d in singles + members:
        d.close()


def test_group_errors(ctx):
    say("[gpu] group errors")
    m = configs.euler_box(2, 4, seed=1)
    init_faces(m)
    d = Device_mesh(ctx, m)
    with pytest.raises(HxError, match="cases registered"):
        Device_group(ctx, [d])
    d.set_cases(1, [np.zeros(m.n_elem[0], np.int32), None])
    with pytest.raises(HxError, match="twice"):
        Device_group(ctx, [d, d])
    g = Device_group(ctx, [d])
    with pytest.raises(HxError, match="destroy the group first"):
        d.close()
    with pytest.raises(HxError, match="launch shape"):
        g.set_shape(0, 0)
    g.close()
    d.close()
    m2 = configs.euler_box(2, 3, seed=1)  # row size 3 is outside the fused set: the meshes step one at a time
    init_faces(m2)
    d2 = Device_mesh(ctx, m2)
    d2.set_cases(1, [np.zeros(m2.n_elem[0], np.int32), None])
    g2 = Device_group(ctx, [d2])
    assert not g2.fused
    g2.update(abi.update_options())
    g2.close()
    d2.close()
