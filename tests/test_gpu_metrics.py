"""Solver::calc_jacobian on the GPU against the oracle: deformed metrics (with face node adjustments), element face
normals, boundary and connection normals, and the vertex time step scale shared by minimum with hanging vertices
matched; then an update on the device-computed metrics. Bar: relative L2 <= 1e-12."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, configs, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, init_faces, o_call, rel_l2, say, smooth_state
from test_gpu_parity import compare_mesh
from test_state_file import mixed_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def compare_geometry(tag, host, dev):
    for c in (0, 1):
        if host.n_elem[c]:
            got = dev.download("vertex_tss", c, out=np.zeros_like(host.vertex_tss[c]))
            assert rel_l2(got, host.vertex_tss[c]) <= TOL, f"{tag}: vertex_tss[{c}]"
    if host.n_elem[1]:
        got = dev.download("metrics", 1, out=np.zeros_like(host.metrics))
        assert rel_l2(got, host.metrics) <= TOL, f"{tag}: metrics {rel_l2(got, host.metrics):.3e}"
    if host.normals.size:
        got = dev.download("normals", 0, out=np.zeros_like(host.normals))
        assert rel_l2(got, host.normals) <= TOL, f"{tag}: normals {rel_l2(got, host.normals):.3e}"
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL, f"{tag}: faces {rel_l2(got, host.faces):.3e}"


CASES = [(1, 5, 4), (2, 2, 4), (2, 6, 3), (3, 3, 3), (3, 6, 2), (3, 8, 2), (2, 8, 3), (3, 4, 3)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "nd%d_rs%d_n%d" % c)
@pytest.mark.parametrize("warp", [False, True])
def test_calc_jacobian_box(ctx, case, warp):
    nd, rs, n = case
    say(f"[gpu] calc_jacobian box {case} warp={warp}")
    host = mesh.box(nd, rs, n, deformed=True, jitter=0.15, seed=nd * 10 + rs,
                    bc=(abi.HX_BC_FREESTREAM, freestream(nd)), with_av=True)
    if warp:
        rng = np.random.default_rng(rs)
        host.node_adj = rng.uniform(-0.03, 0.03, size=(host.n_elem[1], 2 * nd, host.nfq))
    for arr in (host.metrics, host.normals, host.vertex_tss[1]):
        arr[:] = 0.
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    vd = host.vertex_data()
    o_call("oracle_calc_jacobian", ctypes.byref(s), ctypes.byref(vd))
    dev.calc_jacobian()
    compare_geometry("box", host, dev)


@pytest.mark.parametrize("nd,seed", [(2, 3), (3, 1)])
def test_calc_jacobian_tree(ctx, nd, seed):
    say(f"[gpu] calc_jacobian tree nd={nd}")
    host = configs.tree_mesh(base=2, rs=4, levels=2, frac=0.4, seed=seed, nd=nd)
    host.vertex_tss[0][:] = 0.
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    vd = host.vertex_data()
    o_call("oracle_calc_jacobian", ctypes.byref(s), ctypes.byref(vd))
    dev.calc_jacobian()
    compare_geometry("tree", host, dev)


def test_calc_jacobian_mixed_then_update(ctx):
    """Cartesian and deformed elements in one mesh; an update on the device-computed geometry equals the oracle's
    update on the oracle-computed geometry."""
    say("[gpu] calc_jacobian mixed mesh + update")
    host = mixed_mesh(nd=2, rs=5)
    host.metrics[:] = 0.
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    vd = host.vertex_data()
    o_call("oracle_calc_jacobian", ctypes.byref(s), ctypes.byref(vd))
    dev.calc_jacobian()
    compare_geometry("mixed", host, dev)
    # faces as Solver::initialize writes them, then two updates
    init_faces(host)
    dev.compute_write_face()
    dev.ctx.check(dev.lib.hx_compute_prolong(dev.ptr, 0, 0), "prolong")
    for c in (0, 1):
        host.tss[c][:] = 1.
        dev.upload("tss", c, host.tss[c])
        host.av[c][:] = 0.
        dev.upload("av", c, host.av[c])
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare_mesh(f"update {it}", host, dev, fields=("state",), tol=TOL)
