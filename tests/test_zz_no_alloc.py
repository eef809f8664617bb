"""hx_update_no_alloc (hx_update with its scratch in a file-scope device array, for callers that cannot allocate)
against hx_update, through the C ABI: states bit-identical and the same status, step after step, on two meshes in a
row, the second one smaller (the array keeps the first mesh's contents); a Navier-Stokes mesh; the array shared by two
meshes stepped alternately (each call must then run its own Max_dt instead of the other mesh's fused minima); and the
rejected admissibility fix."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, configs
from hexed_amd.device import Context, Device_mesh

from helpers import init_faces, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def no_alloc_update(dev, opts, st):
    rc = dev.lib.hx_update_no_alloc(dev.ptr, ctypes.byref(opts), ctypes.byref(st))
    dev.ctx.check(rc, "update_no_alloc")


def twins(ctx, m):
    init_faces(m)
    return m, Device_mesh(ctx, m.copy()), Device_mesh(ctx, m.copy())


def assert_same(tag, m, a, b):
    for c in (0, 1):
        if m.n_elem[c]:
            sa = a.download("state", c, out=np.zeros_like(m.state[c]))
            sb = b.download("state", c, out=np.zeros_like(m.state[c]))
            assert np.array_equal(sa, sb), f"{tag}: class {c} max |diff| {np.max(np.abs(sa - sb)):.3e}"


def run_twins(tag, m, a, b, opts, steps):
    sa, sb = abi.hx_update_status(), abi.hx_update_status()
    for k in range(steps):
        a.update(opts, sa)
        no_alloc_update(b, opts, sb)
        assert (sa.time_step, sa.flow_time, sa.iteration) == (sb.time_step, sb.flow_time, sb.iteration), (tag, k)
        assert_same(f"{tag} step {k}", m, a, b)


def test_no_alloc_matches_update_twice_smaller_second(ctx):
    say("[gpu] hx_update_no_alloc vs hx_update: 3D row size 6 box, then a smaller row size 4 box")
    opts = abi.update_options(max_safety=0.6)
    for tag, build in (("large", lambda: configs.euler_box(6, 6, mach=0.7, seed=1, deformed=True, jitter=0.1)),
                       ("small", lambda: configs.euler_box(2, 4, mach=0.4, seed=2))):
        m, a, b = twins(ctx, build())
        run_twins(tag, m, a, b, opts, 4)
        a.close()
        b.close()


def test_no_alloc_navier_stokes_and_chebyshev(ctx):
    say("[gpu] hx_update_no_alloc vs hx_update: Navier-Stokes, 2 Chebyshev stages")
    visc, cond = configs.sutherland_air()
    m, a, b = twins(ctx, configs.ns_deformed(3, rs=5, seed=4))
    run_twins("ns", m, a, b, abi.update_options(max_safety=0.5, n_cheby_flow=2, visc=visc, cond=cond, use_ldg=True), 3)
    a.close()
    b.close()


def test_no_alloc_two_meshes_share_the_array(ctx):
    say("[gpu] hx_update_no_alloc on two meshes stepped alternately")
    opts = abi.update_options(max_safety=0.6)
    m1, a1, b1 = twins(ctx, configs.euler_box(3, 5, mach=0.5, seed=5))
    m2, a2, b2 = twins(ctx, configs.euler_box(4, 5, mach=0.8, seed=6))
    s = [abi.hx_update_status() for _ in range(4)]
    for k in range(4):
        for m, a, b, sa, sb in ((m1, a1, b1, s[0], s[1]), (m2, a2, b2, s[2], s[3])):
            a.update(opts, sa)
            no_alloc_update(b, opts, sb)
            assert (sa.time_step, sa.flow_time) == (sb.time_step, sb.flow_time), k
            assert_same(f"alternate step {k}", m, a, b)
    for d in (a1, b1, a2, b2):
        d.close()


def test_no_alloc_rejects_the_fix(ctx):
    say("[gpu] hx_update_no_alloc rejects fix_admis")
    m = configs.euler_box(2, 4, seed=1)
    init_faces(m)
    d = Device_mesh(ctx, m)
    st = abi.hx_update_status()
    rc = d.lib.hx_update_no_alloc(d.ptr, ctypes.byref(abi.update_options(fix_admis=True)), ctypes.byref(st))
    assert rc == abi.HX_ERR_UNSUPPORTED
    assert b"allocates" in d.lib.hx_last_error(ctx.ptr)
    d.close()
