"""Riemann_invariants, Pressure_outflow and No_slip (with each thermal condition) on the GPU against the oracle:
the state and flux conditions alone (the flux condition reading the state cache its state condition wrote), the
no-slip advection hook of the artificial-viscosity generator, and whole Navier-Stokes updates with each condition
on every boundary face. Bar: relative L2 <= 1e-12."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, init_faces, o_call, rel_l2, say, smooth_state
from test_gpu_parity import compare_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12

BCS = {
    "riemann": (abi.HX_BC_RIEMANN_INVARIANTS, {}),
    "pressure_outflow": (abi.HX_BC_PRESSURE_OUTFLOW, {"pressure": 0.68}),
    "no_slip_adiabatic": (abi.HX_BC_NO_SLIP, {"thermal": abi.HX_THERMAL_HEAT_FLUX}),
    "no_slip_heat_flux": (abi.HX_BC_NO_SLIP, {"thermal": abi.HX_THERMAL_HEAT_FLUX, "thermal_value": 0.05}),
    "no_slip_isothermal": (abi.HX_BC_NO_SLIP, {"thermal": abi.HX_THERMAL_ENERGY, "thermal_value": 1.9}),
    "no_slip_equilibrium": (abi.HX_BC_NO_SLIP, {"thermal": abi.HX_THERMAL_EQUILIBRIUM, "emissivity": 0.5,
                                                  "heat_transfer_coef": 2.0, "temperature": 0.002, "coercion": 1.5}),
}


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def build(name, nd=3, rs=4, n=3, with_av=False):
    kind, params = BCS[name]
    fs = freestream(nd, mach=0.4)
    m = mesh.box(nd, rs, n, deformed=True, jitter=0.1, seed=5, bc=(kind, fs, params) if params else (kind, fs),
                 with_av=with_av)
    smooth_state(m, mach=0.4, seed=5)
    init_faces(m)
    return m


@pytest.mark.parametrize("name", sorted(BCS))
def test_state_and_flux_conditions(ctx, name):
    say(f"[gpu] boundary conditions {name}")
    host = build(name)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    o_call("oracle_apply_state_bcs", ctypes.byref(s))
    dev.apply_state_bcs()
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL, f"state conditions: {rel_l2(got, host.faces):.3e}"
    # viscous fluxes in the LDG blocks of every face, then the flux conditions
    nv, nfq = host.nv, host.nfq
    rng = np.random.default_rng(7)
    host.faces[:, nv * nfq:2 * nv * nfq] = rng.uniform(-1, 1, size=(host.faces.shape[0], nv * nfq))
    dev.upload("faces", 0, host.faces)
    o_call("oracle_apply_flux_bcs", ctypes.byref(s))
    dev.apply_flux_bcs()
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL, f"flux conditions: {rel_l2(got, host.faces):.3e}"


def test_no_slip_advection_hook(ctx):
    say("[gpu] no-slip advection boundary condition")
    host = build("no_slip_adiabatic", with_av=True).enable_av_generator()
    rng = np.random.default_rng(3)
    host.faces[:] = rng.uniform(-1, 1, size=host.faces.shape)
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    o_call("oracle_apply_advection_bcs", ctypes.byref(s))
    dev.apply_advection_bcs()
    got = dev.download("faces", 0, out=np.zeros_like(host.faces))
    assert rel_l2(got, host.faces) <= TOL


@pytest.mark.parametrize("name", ["riemann", "pressure_outflow", "no_slip_adiabatic", "no_slip_isothermal",
                                  "no_slip_equilibrium"])
def test_update_navier_stokes(ctx, name):
    say(f"[gpu] navier-stokes update with {name}")
    host = build(name, with_av=True)
    for c in (0, 1):
        if host.n_elem[c]:
            host.av[c][:] = 1e-3
    dev = Device_mesh(ctx, host.copy())
    s = host.struct()
    opts = abi.update_options(max_safety=0.5, visc=abi.constant(2e-3), cond=abi.constant(3e-3), use_ldg=True)
    st_o, st_d = abi.hx_update_status(), abi.hx_update_status()
    for it in range(3):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st_o))
        dev.update(opts, st_d)
        assert abs(st_d.time_step - st_o.time_step) <= 1e-13 * st_o.time_step
        compare_mesh(f"{name} update {it}", host, dev, fields=("state",))


@pytest.mark.parametrize("name", ["riemann", "pressure_outflow", "no_slip_isothermal", "no_slip_equilibrium"])
def test_fp32_form_with_walls(ctx, name):
    """The FP32-MFMA element kernels next to these conditions (which run in FP64) stay within the FP32 form's smooth-flow
    bound of the FP64 form: relative L2 <= 1e-7 per variable after 3 Navier-Stokes updates (measured up to 1.4e-8)."""
    say(f"[gpu] fp32-mfma vs fp64 with {name}")
    host = build(name, with_av=True)
    for c in (0, 1):
        if host.n_elem[c]:
            host.av[c][:] = 1e-3
    opts = abi.update_options(max_safety=0.5, visc=abi.constant(2e-3), cond=abi.constant(3e-3), use_ldg=True)
    out = {}
    for precision in (abi.HX_FP64, abi.HX_FP32_MFMA):
        dev = Device_mesh(ctx, host.copy())
        dev.set_precision(precision)
        for _ in range(3):
            dev.update(opts)
        out[precision] = dev.download("state", 1, out=np.zeros_like(host.state[1]))
        dev.close()
    errs = [rel_l2(out[abi.HX_FP32_MFMA][:, v], out[abi.HX_FP64][:, v]) for v in range(host.nv)]
    say("    rel L2 per variable " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) <= 1e-7
