"""hx_update computes the next step's Max_dt inside the last stage's Local and uses it at the next hx_update only
if nothing changed the mesh in between and the time-step settings are the same. These tests change things between
updates -- an uploaded state, the safety factor, a kernel called on its own, the PDE -- and compare every step with
the oracle, which always runs Max_dt (Solver::update, Solver.cpp:834-886). Time steps must match to 1e-13 and
states to the parity bar."""
import ctypes

import pytest

from hexed_amd import abi
from hexed_amd.device import Context, Device_mesh

from helpers import o_call, say
from test_gpu_parity import build, compare_mesh

pytestmark = pytest.mark.gpu

CFG = (3, 6, True, 3)


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def step(host, dev, opts, st_o, st_d, tag):
    s = host.struct()
    o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
    dev.update(opts, st_d)
    assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step, \
        f"{tag}: dt {st_d.time_step!r} vs oracle {st_o.time_step!r}"
    compare_mesh(tag, host, dev, fields=("state",))


def test_state_upload_between_updates(ctx):
    say("[gpu] fused max_dt: state upload between updates")
    host = build(CFG)
    dev = Device_mesh(ctx, host.copy())
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    step(host, dev, opts, st_o, st_d, "step 1")
    step(host, dev, opts, st_o, st_d, "step 2 (fused dt)")
    # a faster flow: the time step must shrink accordingly
    for c in (0, 1):
        if host.n_elem[c]:
            host.state[c][:, 0] *= 3.0
            host.state[c][:, -1] += 4.0 * host.state[c][:, 0] ** 2 / host.state[c][:, host.n_dim]
            dev.upload("state", c, host.state[c])
    dt_before = st_o.time_step
    step(host, dev, opts, st_o, st_d, "step 3 after upload")
    assert st_o.time_step < 0.9 * dt_before


def test_settings_change_between_updates(ctx):
    say("[gpu] fused max_dt: safety factor and Chebyshev stages change")
    host = build(CFG)
    dev = Device_mesh(ctx, host.copy())
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    step(host, dev, abi.update_options(max_safety=0.5), st_o, st_d, "safety 0.5")
    step(host, dev, abi.update_options(max_safety=0.3), st_o, st_d, "safety 0.3")
    step(host, dev, abi.update_options(max_safety=0.3, n_cheby_flow=3), st_o, st_d, "3 chebyshev stages")
    step(host, dev, abi.update_options(max_safety=0.3, n_cheby_flow=3, flow_iters=2), st_o, st_d, "2 flow iterations")
    step(host, dev, abi.update_options(max_safety=0.3, local_time=True), st_o, st_d, "local time")
    step(host, dev, abi.update_options(max_safety=0.3), st_o, st_d, "global again")


def test_kernel_call_between_updates(ctx):
    say("[gpu] fused max_dt: a kernel sequence called between updates")
    host = build(CFG)
    dev = Device_mesh(ctx, host.copy())
    opts = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    step(host, dev, opts, st_o, st_d, "step 1")
    # one more stage-0 Euler pass with a large dt changes the state outside hx_update
    s = host.struct()
    ko = abi.hx_kernel_options(40 * st_o.time_step, 0, 0, 0)
    o_call("oracle_apply_state_bcs", ctypes.byref(s))
    o_call("oracle_compute_euler", ctypes.byref(s), ctypes.byref(ko))
    dev.apply_state_bcs()
    dev.compute_euler(40 * st_o.time_step, 0)
    compare_mesh("after compute_euler", host, dev, fields=("state",))
    step(host, dev, opts, st_o, st_d, "step 2 after compute_euler")


def test_euler_then_navier_stokes(ctx):
    say("[gpu] fused max_dt: switching between Euler and Navier-Stokes settings")
    host = build(CFG, with_av=True)
    dev = Device_mesh(ctx, host.copy())
    ns = abi.update_options(max_safety=0.5, visc=abi.constant(1e-3), cond=abi.constant(1.4e-3), use_ldg=True)
    eu = abi.update_options(max_safety=0.5)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for k, opts in enumerate([eu, ns, ns, eu, eu, ns]):
        step(host, dev, opts, st_o, st_d, f"step {k} {'ns' if opts.use_ldg else 'euler'}")


def test_batch_state_upload_between_updates(ctx):
    """hx_update_cases takes each case's time step from its previous stage 1 only while nothing changed the mesh: after
    an upload, the per-case steps are those of a fresh batch mesh holding the uploaded state (Max_dt run anew)."""
    import numpy as np
    from hexed_amd import configs, mesh
    from helpers import init_faces
    say("[gpu] fused max_dt: batch, state upload between updates")
    cases = [configs.euler_box(2, 4, mach=0.3, seed=1), configs.euler_box(3, 4, mach=0.7, seed=2),
             configs.euler_box(2, 4, mach=0.5, seed=3, deformed=True, jitter=0.1)]
    for m in cases:
        init_faces(m)
    b = mesh.concat(cases)
    opts = abi.update_options(max_safety=0.5)

    def batch_dev(host):
        d = Device_mesh(ctx, host.copy())
        d.set_cases(len(cases), [c.astype(np.int32) for c in host.case_id])
        return d
    dev = batch_dev(b)
    st = (abi.hx_update_status * len(cases))()
    dev.update_cases(opts, st)
    dev.update_cases(opts, st)  # (per-case steps from the fused values)
    # a faster flow in every case
    after = b.copy()
    for c in (0, 1):
        if after.n_elem[c]:
            u = dev.download("state", c, out=np.zeros_like(after.state[c]))
            u[:, 0] *= 3.0
            u[:, -1] += 4.0 * u[:, 0] ** 2 / u[:, after.n_dim]
            after.state[c][:] = u
            dev.upload("state", c, u)
    dev.update_cases(opts, st)
    fresh = batch_dev(after)
    st2 = (abi.hx_update_status * len(cases))()
    fresh.update_cases(opts, st2)
    for k in range(len(cases)):
        assert st[k].time_step == st2[k].time_step, f"case {k}: dt {st[k].time_step!r} vs {st2[k].time_step!r}"
    dev.close()
    fresh.close()
