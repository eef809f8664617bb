"""The oracle's Riemann_invariants, Pressure_outflow and No_slip boundary conditions (Boundary_condition.cpp:78-448)
and the characteristic decomposition they rest on (pde.hpp:182-260), pinned by the reference's own tests
(test_Characteristics.cpp, test_Boundary_condition.cpp: "Riemann_invariants", "No_slip") restated on a one-element
mesh, and by closed forms for Pressure_outflow, which the reference does not test. The GPU conditions are compared
with the oracle in test_gpu_bcs.py."""
import ctypes
import math

import numpy as np
import pytest

from hexed_amd import abi, mesh

import oracle
from helpers import o_call

R_AIR = 287.05287
SIGMA = 2 * math.pi ** 5 * 1.380649e-23 ** 4 / (15 * 299792458. ** 2 * 6.62607015e-34 ** 3)


def P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def characteristics(ref, normal, state):
    nd = len(normal)
    vals, parts = np.zeros(3), np.zeros((nd + 2, 3))
    rc = oracle.load().oracle_characteristics(nd, P(np.asarray(ref, float)), P(np.asarray(normal, float)),
                                              P(np.asarray(state, float)), P(vals), P(parts))
    assert rc == 0
    return vals, parts


def euler_flux(u, n):
    nd = len(n)
    m, rho, e = np.asarray(u[:nd]), u[nd], u[nd + 1]
    p = .4 * (e - .5 * m @ m / rho)
    vn = m @ n / rho
    return np.concatenate([m * vn + p * np.asarray(n), [rho * vn, (e + p) * vn]])


def test_characteristics():
    """test_Characteristics.cpp:4-43: the parts sum to the state and each is an eigenvector of the flux Jacobian."""
    mass, veloc, pres = 1.225, np.array([10., 4., 12.]), 101325.
    ref = np.concatenate([mass * veloc, [mass, pres / .4 + .5 * mass * veloc @ veloc]])
    normal = np.array([1., 1., 1.])
    state1 = np.array([1., 2., 3., 1.3, 2e5])
    vals, parts = characteristics(ref, normal, state1)
    assert np.linalg.norm((parts.sum(axis=1) - state1) / state1) < 1e-12
    diff = 1e-6
    for i in range(3):
        vec = parts[:, i]
        perturb = (euler_flux(ref + diff * vec, normal) - euler_flux(ref, normal)) / np.linalg.norm(normal)
        assert np.linalg.norm(perturb / (diff * vals[i] * vec) - 1) < 1e-5


def one_element(nd, kind, fs=None, **params):
    fs = list(fs) if fs is not None else [0.] * (nd + 2)
    return mesh.box(nd, 8, 1, deformed=True, bc=(kind, fs, params) if params else (kind, fs))


def boundary_face(m, i_dim, sign):
    """(state block, LDG block, ghost state, ghost LDG, normal) arrays of the boundary face (i_dim, sign) of the
    one-element box; views into the host mesh."""
    nv, nfq = m.nv, m.nfq
    k = 2 * i_dim + sign
    con = m.bc_con[k]
    s_in, s_gh = m.con_slot[1][con]
    fin, fgh = m.faces[s_in], m.faces[s_gh]
    return (fin[:nv * nfq].reshape(nv, nfq), fin[nv * nfq:2 * nv * nfq].reshape(nv, nfq),
            fgh[:nv * nfq].reshape(nv, nfq), fgh[nv * nfq:2 * nv * nfq].reshape(nv, nfq),
            m.normals[m.con_nslot[con][0]])


def apply(m, which):
    s = m.struct()
    o_call("oracle_apply_state_bcs" if which == "state" else "oracle_apply_flux_bcs", ctypes.byref(s))


def test_riemann_invariants_supersonic():
    """test_Boundary_condition.cpp:76-118: supersonic outflow keeps the inside state and takes no viscous flux,
    supersonic inflow takes the freestream and keeps the viscous flux."""
    fs = [10., 30., -20., 1.3, 4e5]
    m = one_element(3, abi.HX_BC_RIEMANN_INVARIANTS, fs)
    st, ldg, gst, gldg, nrml = boundary_face(m, 1, 0)
    inside = np.array([1 / 1.2, -600 / 1.2, 1 / 1.2, 1.2, 101325 / .4 + .5 * 1.2 * 360002])
    for q in range(m.nfq):
        s = inside.copy()
        if q:
            s[1] *= -1
        st[:, q] = s
        nrml[:, q] = [0., 1., 0.]
    apply(m, "state")
    for q in range(m.nfq):
        np.testing.assert_allclose(gst[:, q], fs if q else inside, rtol=1e-12)
    ldg[:] = 1.
    apply(m, "flux")
    for q in range(m.nfq):
        np.testing.assert_allclose(gldg[:, q], 1. if q else 0., atol=1e-9)


def test_riemann_invariants_matching_freestream():
    """A subsonic boundary state equal to the freestream is its own ghost (every characteristic agrees)."""
    fs = np.array([0.3, 0.1, 1.0, 1 / .4 + .5 * 0.1])
    m = one_element(2, abi.HX_BC_RIEMANN_INVARIANTS, fs)
    st, _, gst, _, nrml = boundary_face(m, 0, 1)
    st[:] = fs[:, None]
    nrml[:] = np.array([[0.8], [0.6]])
    apply(m, "state")
    np.testing.assert_allclose(gst, np.repeat(fs[:, None], m.nfq, axis=1), rtol=1e-12)


def no_slip_face(thermal, **params):
    """test_Boundary_condition.cpp:194-203: 2D deformed element, face (0, 0), normal 0.7 (1, 1)/sqrt 2"""
    m = one_element(2, abi.HX_BC_NO_SLIP, thermal=thermal, **params)
    faces = boundary_face(m, 0, 0)
    faces[4][:] = .7 / math.sqrt(2.)
    return m, faces


STATE = np.array([1., 1., 1.2, 1e5 / 0.4 + 0.5 * 1.2 * 2.])
FLUX = np.array([10., -20., 1.3, 10.])


def test_no_slip_isothermal():
    m, (st, ldg, gst, gldg, _) = no_slip_face(abi.HX_THERMAL_ENERGY, thermal_value=1e6)
    st[:] = STATE[:, None]
    apply(m, "state")
    np.testing.assert_allclose(gst[:2], -1., rtol=1e-12)
    np.testing.assert_allclose(gst[2], 1.2, rtol=1e-12)
    np.testing.assert_allclose(np.sqrt(gst[3] * st[3]), 1e6 * 1.2, rtol=1e-12)
    ldg[:] = FLUX[:, None]
    apply(m, "flux")
    np.testing.assert_allclose(gldg, (FLUX * [1, 1, -1, 1])[:, None] * np.ones(m.nfq), rtol=1e-12)


def test_no_slip_specified_heat_flux():
    m, (st, ldg, gst, gldg, _) = no_slip_face(abi.HX_THERMAL_HEAT_FLUX, thermal_value=3.)
    st[:] = STATE[:, None]
    apply(m, "state")
    np.testing.assert_allclose(gst[:2], -1., rtol=1e-12)
    np.testing.assert_allclose(gst[2:], STATE[2:, None] * np.ones(m.nfq), rtol=1e-12)
    ldg[:] = FLUX[:, None]
    apply(m, "flux")
    np.testing.assert_allclose(gldg[:3], (FLUX[:3] * [1, 1, -1])[:, None] * np.ones(m.nfq), rtol=1e-12)
    np.testing.assert_allclose((gldg[3] + ldg[3]) / 2, -3. * .7, rtol=1e-12)


def test_no_slip_emissivity():
    """The flux condition reads the state cached by the state condition (the inside state is zeroed between)."""
    m, (st, ldg, gst, gldg, _) = no_slip_face(abi.HX_THERMAL_EQUILIBRIUM, emissivity=.8)
    state = STATE.copy()
    state[3] = 1e5 / .4
    st[:] = state[:, None]
    ldg[:] = FLUX[:, None]
    apply(m, "state")
    st[:] = 0.
    apply(m, "flux")
    temp = 1e5 / 1.2 / R_AIR
    np.testing.assert_allclose((gldg[3] + ldg[3]) / 2, -.8 * SIGMA * temp ** 4 * .7, rtol=1e-12)


@pytest.mark.parametrize("sign", [0, 1])
def test_pressure_outflow(sign):
    """Subsonic outflow (and any inflow): the ghost pressure is the specified one, momentum and density kept;
    supersonic outflow: the inside state. The viscous flux is reversed."""
    p_spec = 9e4
    m = one_element(3, abi.HX_BC_PRESSURE_OUTFLOW, pressure=p_spec)
    st, ldg, gst, gldg, nrml = boundary_face(m, 2, sign)
    outward = 1 if sign else -1
    rho, p = 1.2, 1e5
    c = math.sqrt(1.4 * p / rho)
    speeds = [0.5 * c * outward, 2.0 * c * outward, -0.5 * c * outward]  # subsonic out, supersonic out, inflow
    for q in range(m.nfq):
        vz = speeds[q % 3]
        u = np.array([5., -3., vz]) * [1, 1, 1]
        st[:, q] = np.concatenate([rho * u, [rho, p / .4 + .5 * rho * u @ u]])
        nrml[:, q] = [0., 0., 2.]
    ldg[:] = np.arange(m.nv)[:, None] + 1.
    apply(m, "state")
    for q in range(m.nfq):
        ke = .5 * st[:3, q] @ st[:3, q] / rho
        np.testing.assert_allclose(gst[:4, q], st[:4, q], rtol=1e-14)
        if q % 3 == 1:
            np.testing.assert_allclose(gst[4, q], st[4, q], rtol=1e-14)
        else:
            np.testing.assert_allclose(.4 * (gst[4, q] - ke), p_spec, rtol=1e-12)
    apply(m, "flux")
    np.testing.assert_allclose(gldg, -ldg, rtol=0)
