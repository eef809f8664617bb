"""hx_update_async (device-resident time step, the steady-state update replayed from a captured HIP graph) against
hx_update (time step read back and passed by value) on twin meshes: after every replay the states are bit-identical
and the status equal, with the host's own downloads and synchronizations between replays; several updates per call;
global and local time stepping, Chebyshev stages, the filter and a maximum time step; Euler and Navier-Stokes on box
and tree meshes; both arithmetic forms; a recapture when the options change; mixing the two entry points; and the
rejected uses."""
import numpy as np
import pytest

from hexed_amd import abi, configs
from hexed_amd.device import Context, Device_mesh, HxError

from helpers import init_faces, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def twins(ctx, m, precision=abi.HX_FP64):
    init_faces(m)
    a, b = Device_mesh(ctx, m.copy()), Device_mesh(ctx, m.copy())
    a.set_precision(precision)
    b.set_precision(precision)
    return m, a, b


def assert_same(tag, m, a, b):
    for c in (0, 1):
        if m.n_elem[c]:
            sa = a.download("state", c, out=np.zeros_like(m.state[c]))
            sb = b.download("state", c, out=np.zeros_like(m.state[c]))
            assert np.array_equal(sa, sb), f"{tag}: class {c} max |diff| {np.max(np.abs(sa - sb)):.3e}"


def assert_status(tag, sa, sb):
    assert (sa.time_step, sa.flow_time, sa.iteration) == (sb.time_step, sb.flow_time, sb.iteration), \
        f"{tag}: {(sa.time_step, sa.flow_time, sa.iteration)} vs {(sb.time_step, sb.flow_time, sb.iteration)}"


def ns_opts(**kw):
    visc, cond = configs.sutherland_air()
    return abi.update_options(visc=visc, cond=cond, use_ldg=True, **kw)


CASES = {
    "euler-box": (lambda: configs.euler_box(3, 4, mach=0.6, seed=1), lambda: abi.update_options(max_safety=0.6)),
    "euler-deformed-cheby-filter": (lambda: configs.euler_box(3, 5, mach=0.4, seed=2, deformed=True, jitter=0.1),
                                    lambda: abi.update_options(max_safety=0.5, n_cheby_flow=2, use_filter=True)),
    "euler-local-time": (lambda: configs.euler_box(2, 6, mach=0.7, seed=3),
                         lambda: abi.update_options(max_safety=0.5, local_time=True)),
    "euler-max-time-step": (lambda: configs.euler_box(2, 4, mach=0.5, seed=4),
                            lambda: abi.update_options(max_safety=0.6, max_time_step=1e-3, flow_iters=2)),
    "ns-deformed": (lambda: configs.ns_deformed(2, rs=6, seed=1), lambda: ns_opts(max_safety=0.5)),
    "euler-tree": (lambda: configs.euler_tree(base=2, rs=4, levels=2, frac=0.3, seed=1),
                   lambda: abi.update_options(max_safety=0.5)),
}


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replays_match_update(ctx, name, precision):
    say(f"[gpu] graph replays vs hx_update: {name} precision {precision}")
    make_mesh, make_opts = CASES[name]
    m, a, b = twins(ctx, make_mesh(), precision)
    opts = make_opts()
    sa = abi.hx_update_status()
    for step in range(5):
        a.update(opts, sa)
        b.update_async(opts, 1)
        assert_status(f"{name} step {step}", sa, b.status_sync())
        assert_same(f"{name} step {step}", m, a, b)
    # several updates per call
    for _ in range(4):
        a.update(opts, sa)
    b.update_async(opts, 4)
    assert_status(f"{name} 4 more", sa, b.status_sync())
    assert_same(f"{name} 4 more", m, a, b)
    a.close()
    b.close()


def test_graph_recapture_and_mixing(ctx):
    say("[gpu] graph recapture on new options, mixed with hx_update")
    m, a, b = twins(ctx, configs.euler_box(3, 6, mach=0.8, seed=5))
    o1, o2 = abi.update_options(max_safety=0.6), abi.update_options(max_safety=0.4)
    sa, sb = abi.hx_update_status(), abi.hx_update_status()
    for opts in (o1, o2, o1):
        for _ in range(3):
            a.update(opts, sa)
        b.update_async(opts, 3)
        assert_status("recapture", sa, b.status_sync(sb))
        assert_same("recapture", m, a, b)
    # hx_update on the mesh an asynchronous sequence left (it takes over the fused Max_dt), then asynchronous updates
    # that start from the host status
    a.update(o1, sa)
    b.update(o1, sb)
    assert_status("mixed sync", sa, sb)
    assert_same("mixed sync", m, a, b)
    for _ in range(2):
        a.update(o1, sa)
    b.update_async(o1, 2, start=sb)
    assert_status("mixed async", sa, b.status_sync())
    assert_same("mixed async", m, a, b)
    a.close()
    b.close()


def test_graph_after_face_fold_change(ctx):
    """Changing the face-fold mode between asynchronous sequences recaptures the graph: after each change the replays
    stay bit-identical to hx_update in the new mode (folded and unfolded results differ in rounding)."""
    say("[gpu] graph after a face-fold mode change")
    m, a, b = twins(ctx, configs.euler_box(3, 6, mach=0.8, seed=8))
    opts = abi.update_options(max_safety=0.6)
    sa = abi.hx_update_status()
    for mode in (-1, 0, 1, -1):
        a.set_face_fold(mode)
        b.set_face_fold(mode)
        for _ in range(3):
            a.update(opts, sa)
        b.update_async(opts, 3)
        assert_status(f"fold mode {mode}", sa, b.status_sync())
        assert_same(f"fold mode {mode}", m, a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("local_time", [False, True], ids=["global", "local-time"])
def test_graph_after_state_upload(ctx, local_time):
    """A state uploaded between asynchronous sequences, its faces left as they were: the face slots then no longer hold
    the state's extrapolation, and the next update must read them as the reference does (the graph, which forms each
    element's own face states from its state, waits until an update has rewritten them)."""
    say(f"[gpu] graph after a state upload, local time {local_time}")
    m, a, b = twins(ctx, configs.euler_box(3, 6, mach=0.8, seed=6))
    opts = abi.update_options(max_safety=0.6, local_time=local_time)
    sa = abi.hx_update_status()
    for _ in range(3):
        a.update(opts, sa)
    b.update_async(opts, 3)
    assert_status("before upload", sa, b.status_sync())
    assert_same("before upload", m, a, b)
    for c in (0, 1):
        if m.n_elem[c]:
            s = a.download("state", c, out=np.zeros_like(m.state[c]))
            s[:, 0] *= 1.5
            a.upload("state", c, s)
            b.upload("state", c, s)
    for _ in range(3):
        a.update(opts, sa)
    b.update_async(opts, 3)
    assert_status("after upload", sa, b.status_sync())
    assert_same("after upload", m, a, b)
    a.close()
    b.close()


def test_graph_rejected_uses(ctx):
    say("[gpu] graph entry point errors")
    m = configs.euler_box(2, 4, seed=1)
    init_faces(m)
    d = Device_mesh(ctx, m.copy())
    with pytest.raises(HxError, match="no asynchronous update"):
        d.status_sync()
    with pytest.raises(HxError, match="admissibility"):
        d.update_async(abi.update_options(fix_admis=True), 1)
    d.set_cases(1, [np.zeros(m.n_elem[0], np.int32), None])
    with pytest.raises(HxError, match="batch"):
        d.update_async(abi.update_options(), 1)
    d.close()
