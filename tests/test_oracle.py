"""Pins the CPU oracle (oracle/hexed_oracle.cpp) against what the reference's own tests and the mathematics fix:
closed-form time steps (test_Max_dt.cpp), analytic residuals on the reference's test meshes (test_Solver.cpp
test_marching / test_visc), discrete conservation (test_conservation), free-stream preservation, the exact
translated isentropic vortex with its convergence order, the tensor-product modal filter, and the geometric
consistency of face permutations with the vertex correspondence of every connection orientation
(test_Face_permutation.cpp)."""
import ctypes
import itertools
import math

import numpy as np
import pytest

from hexed_amd import abi, configs, mesh

import meshes
from helpers import GAMMA, init_faces, o_call

BASIS = mesh.basis


def oracle_residual(m, opts):
    s = m.struct()
    o_call("oracle_compute_residual", ctypes.byref(s), ctypes.byref(opts))


def set_local_tss(m, opts):
    s = m.struct()
    kop = abi.hx_kernel_options(0., 0, 0, 0)
    out = ctypes.c_double()
    if opts.use_ldg:
        o_call("oracle_max_dt_navier_stokes", ctypes.byref(s), ctypes.byref(kop), 0.7, 0.7, 1, opts.visc, opts.cond,
               ctypes.byref(out))
    else:
        o_call("oracle_max_dt_euler", ctypes.byref(s), ctypes.byref(kop), 0.7, 0.7, 1, ctypes.byref(out))


def physical_residual(m, c):
    """Physical_residual (Qpoint_func.cpp:37-45): residual cache divided by the time step scale."""
    return m.rcache[c][:, :m.nv] / m.tss[c][:, None, :]


def quadrature_volume_weights(m, c):
    """w_q * h^nd * det_q for integral_field (Solver.cpp:1113-1131)."""
    nd, rs = m.n_dim, m.row_size
    w = BASIS(rs)["weight"]
    wq = w
    for _ in range(nd - 1):
        wq = np.multiply.outer(wq, w).reshape(-1)
    vol = m.nom_size[c] ** nd
    det = m.metrics[:, nd * nd] if c == 1 else np.ones((m.n_elem[0], m.nq))
    return wq[None, :] * vol[:, None] * det


# ---- time step (test_Max_dt.cpp) ------------------------------------------------------------------------------

def single_elements(nd, rs, n, deformed, h, vert_pos=None):
    """n unconnected elements (no faces are touched by Max_dt)."""
    nvert = 2 ** nd
    m = mesh.HostMesh(nd, rs, 0 if deformed else n, n if deformed else 0, 1, 1, 0, 0)
    c = 1 if deformed else 0
    m.nom_size[c][:] = h
    if deformed:
        if vert_pos is None:
            corner = np.stack(np.unravel_index(np.arange(nvert), (2,) * nd), axis=1).astype(float)
            vert_pos = np.repeat(corner[None] * h, n, axis=0)
        metrics, fn, vtss, qpos = mesh.deformed_metrics(nd, rs, vert_pos, m.nom_size[1])
        m.metrics[:] = metrics
        m.vertex_tss[1][:] = vtss
    else:
        m.vertex_tss[0][:] = h / nd  # Element.cpp:17
    return m


def test_max_dt_cartesian_1d():
    """test_Max_dt.cpp:13-38: the global step is set by the fastest sound speed."""
    rs = 2
    m = single_elements(1, rs, 2, False, 1.027)
    mass, c0, c1 = 0.9, 400., 200.
    e0 = mass * c0 ** 2 / (GAMMA * (GAMMA - 1))
    e1 = mass * c1 ** 2 / (GAMMA * (GAMMA - 1))
    m.state[0][0] = [[0, 0], [mass, mass], [e1, e0]]
    m.state[0][1] = [[0, 0], [mass, mass], [e0, e1]]
    m.tss[0][:] = 0.6
    m.tss[0][:, 1] = 0.17
    s = m.struct()
    out = ctypes.c_double()
    o_call("oracle_max_dt", ctypes.byref(s), 0, None, 0, 0, 0.7, 0.7, ctypes.byref(out))
    max_cfl = -2 * 0.9 / BASIS(rs)["eigenvals"][0]
    assert out.value == pytest.approx(0.7 * max_cfl / (400 / 1.027), rel=1e-14)
    assert np.all(m.tss[0] == 1.0)


def test_max_dt_cartesian_2d_local():
    """test_Max_dt.cpp:39-57: local time step scale within 1% of max_cfl/360/2."""
    rs = 2
    m = single_elements(2, rs, 1, False, 1.0)
    m.state[0][0] = [[2.25] * 4, [24.5] * 4, [1.225] * 4, [101235 / 0.4 + 0.5 * 1.225 * 500] * 4]
    s = m.struct()
    out = ctypes.c_double()
    o_call("oracle_max_dt", ctypes.byref(s), 0, None, 1, 1, 1., 1., ctypes.byref(out))
    max_cfl = -2 * 0.9 / BASIS(rs)["eigenvals"][0]
    assert out.value == 1.0
    np.testing.assert_allclose(m.tss[0][0], max_cfl / 360. / 2., rtol=0.01)


def test_max_dt_deformed_3d():
    """test_Max_dt.cpp:60-104: sound speed, Jacobian and vertex time-step scale all enter the step."""
    rs, h = 4, 0.3
    m = single_elements(3, rs, 3, True, h)
    for e in range(3):
        m.state[1][e, 0:3] = 0
        m.state[1][e, 3] = 1.4
        m.state[1][e, 4] = (1e6 if e == 1 else 1e4) / 0.4
    s = m.struct()
    out = ctypes.c_double()
    o_call("oracle_max_dt", ctypes.byref(s), 1, None, 1, 1, 1., 1., ctypes.byref(out))
    max_cfl = -2 * 0.9 / BASIS(rs)["eigenvals"][0]
    for e in range(3):
        np.testing.assert_allclose(m.tss[1][e], max_cfl / ((1e3 if e == 1 else 1e2) / h) / 3., rtol=1e-12)
    # squash element 1 by half in x and y
    corner = np.stack(np.unravel_index(np.arange(8), (2,) * 3), axis=1).astype(float) * h
    vp = np.repeat(corner[None], 3, axis=0)
    vp[1, :, 0:2] *= 0.5
    metrics, fn, vtss, qpos = mesh.deformed_metrics(3, rs, vp, m.nom_size[1])
    m.metrics[:] = metrics
    m.vertex_tss[1][:] = vtss
    s = m.struct()
    o_call("oracle_max_dt", ctypes.byref(s), 1, None, 0, 0, 1., 1., ctypes.byref(out))
    assert out.value == pytest.approx(max_cfl / (1e3 * (1. + 2 * 2.) / 3. / h) / 3., rel=1e-12)


# ---- analytic residuals on the reference's test meshes (test_Solver.cpp) ---------------------------------------

VELOCS = [0.3, -0.7, 0.8]
WAVE = [-0.1, 0.3, 0.2]


def nonuniform_mass(nd):
    """Nonuniform_mass and its exact time derivative Nonuniform_residual (test_Solver.cpp:103-170)."""
    def state(pos):
        sp = sum(WAVE[d] * pos[d] for d in range(nd))
        vsq = sum(v * v for v in VELOCS[:nd])
        mass = 1. + 0.1 * np.sin(sp)
        return [mass * VELOCS[d] for d in range(nd)] + [mass, 1e5 / 0.4 + mass * 0.5 * vsq]

    def residual(pos):
        sp = sum(WAVE[d] * pos[d] for d in range(nd))
        vsq = sum(v * v for v in VELOCS[:nd])
        scaled = sum(VELOCS[d] * WAVE[d] for d in range(nd))
        dm = -0.1 * scaled * np.cos(sp)
        return [dm * VELOCS[d] for d in range(nd)] + [dm, dm * 0.5 * vsq]
    return state, residual


TEST_MESHES = {
    "all_cartesian": lambda bc, av=False: meshes.all_cartesian(bc, with_av=av),
    "all_deformed_rot0": lambda bc, av=False: meshes.all_deformed(0, bc, with_av=av),
    "all_deformed_rot1": lambda bc, av=False: meshes.all_deformed(1, bc, with_av=av),
}


@pytest.mark.parametrize("name", sorted(TEST_MESHES))
def test_marching_nonuniform_mass(name):
    """test_marching (test_Solver.cpp:555-586): the residual reproduces the analytic time derivative to
    1e-3|state| (the reference's margin), and in fact to 1e-6 relative to the state scale."""
    m = TEST_MESHES[name](abi.HX_BC_COPY)
    state, residual = nonuniform_mass(m.n_dim)
    mesh.set_state(m, state)
    init_faces(m)
    opts = abi.update_options(local_time=True)
    set_local_tss(m, opts)
    oracle_residual(m, opts)
    for c in (0, 1):
        if not m.n_elem[c]:
            continue
        got = physical_residual(m, c)
        pos = m.qpos[c]
        want = np.stack(residual([pos[:, d] for d in range(m.n_dim)]), axis=1)
        st = m.state[c]
        assert np.all(np.abs(got - want) <= 1e-3 * np.abs(st))
        assert np.abs(got - want).max() <= 1e-6 * np.abs(st).max()


@pytest.mark.parametrize("name", sorted(TEST_MESHES))
def test_viscous_sinusoid(name):
    """test_visc (test_Solver.cpp:588-615): divergence-free sinusoidal velocity with constant viscosity 3; the
    momentum rate is -n_dim*3*momentum/1.2 (margin 1.2e-3 for row size 8)."""
    m = TEST_MESHES[name](abi.HX_BC_COPY, True)
    nd = m.n_dim

    def state(pos):
        s = sum(pos)
        cp = np.cos(s)
        mom = [1.2 * cp * (1 - nd if d == 0 else 1) for d in range(nd)]
        return mom + [1.2 + 0 * s, 1e5 / 0.4 + .5 * 1.2 * (nd - 1 + (1 - nd) * (1 - nd)) * cp * cp]
    mesh.set_state(m, state)
    init_faces(m)
    opts = abi.update_options(local_time=True, visc=abi.constant(3.), use_ldg=True)
    set_local_tss(m, opts)
    oracle_residual(m, opts)
    for c in (0, 1):
        if not m.n_elem[c]:
            continue
        got = physical_residual(m, c)
        for v in range(nd):
            want = -nd * 3. * m.state[c][:, v] / 1.2
            assert np.abs(got[:, v] - want).max() <= 1.2e-3


@pytest.mark.parametrize("name", sorted(TEST_MESHES))
@pytest.mark.parametrize("visc", [False, True])
def test_conservation(name, visc):
    """test_conservation (test_Solver.cpp:617-639): a random state with nonpenetration walls and constant artificial
    viscosity 30 has zero integrated mass and energy residual (to round-off, not the reference's 1e-5)."""
    m = TEST_MESHES[name](abi.HX_BC_NONPENETRATION, True)
    nd = m.n_dim
    rng = np.random.default_rng(406)
    for c in (0, 1):
        if not m.n_elem[c]:
            continue
        shape = (m.n_elem[c], m.nq)
        pert = lambda: 1. + 0.1 * rng.integers(0, 1000, size=shape) / 1000
        mass = 1.05
        for d in range(nd):
            m.state[c][:, d] = mass * VELOCS[d] * pert()
        m.state[c][:, nd] = mass * pert()
        m.state[c][:, nd + 1] = 1e5 / 0.4 * pert()
        m.av[c][:, 0] = 30.
    init_faces(m)
    opts = abi.update_options(local_time=True, use_ldg=visc)
    set_local_tss(m, opts)
    oracle_residual(m, opts)
    for v in (nd, nd + 1):
        total_res, total_state = 0., 0.
        for c in (0, 1):
            if not m.n_elem[c]:
                continue
            wv = quadrature_volume_weights(m, c)
            total_res += np.sum(wv * physical_residual(m, c)[:, v])
            total_state += np.sum(wv * m.state[c][:, v])
        assert abs(total_res) <= 1e-11 * abs(total_state), (v, total_res, total_state)


@pytest.mark.parametrize("nd,rs,deformed", [(2, 6, False), (3, 4, False), (2, 6, True), (3, 5, True), (3, 6, True)])
@pytest.mark.parametrize("visc", [False, True])
def test_free_stream_preservation(nd, rs, deformed, visc):
    """A uniform state stays uniform through full updates on jittered deformed meshes (discrete metric identities,
    flux consistency, LDG gradient of a constant)."""
    fs = configs.freestream_state(nd, 1.1, [0.4, -0.2, 0.3][:nd], 0.9)
    m = mesh.box(nd, rs, 3, deformed=deformed, jitter=0.15, seed=3, bc=(abi.HX_BC_FREESTREAM, fs), with_av=True)
    mesh.set_state(m, lambda pos: [np.full_like(pos[0], v) for v in fs])
    c = 1 if deformed else 0
    m.av[c][:, 0] = 0.02
    init_faces(m)
    opts = abi.update_options(max_safety=0.5, visc=abi.constant(1e-3) if visc else None,
                              cond=abi.constant(1e-3) if visc else None, use_ldg=visc)
    s = m.struct()
    st = abi.hx_update_status()
    for _ in range(3):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st))
    dev = np.abs(m.state[c] - np.array(fs)[None, :, None]).max()
    assert dev <= 1e-12 * max(abs(x) for x in fs)


# ---- 2D isentropic vortex: exact translation and convergence ----------------------------------------------

def vortex_error(n, rs, steps, safety, max_time_step=1e300):
    m, exact = configs.isentropic_vortex(n=n, rs=rs)
    init_faces(m)
    opts = abi.update_options(max_safety=safety, max_time_step=max_time_step)
    s = m.struct()
    st = abi.hx_update_status()
    for _ in range(steps):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st))
    pos = m.qpos[0]
    want = np.stack(exact([pos[:, 0], pos[:, 1]], st.flow_time), axis=1)
    return np.abs(m.state[0] - want).max(), st.flow_time


def test_isentropic_vortex_translation():
    """Config 1: 16x16 quads, row size 6, 100 steps at max_safety 0.5 track the translated exact vortex."""
    err, t = vortex_error(16, 6, 100, 0.5)
    assert t > 0.25
    assert err < 1e-4, err


def test_isentropic_vortex_convergence_order():
    """Halving the element size at row size 6 (same small time step, same flow time) cuts the error by more than
    2^4 -- high-order convergence that a dropped or mis-signed term in any kernel would destroy."""
    err_c, t_c = vortex_error(8, 6, 40, 0.25, max_time_step=5e-4)
    err_f, t_f = vortex_error(16, 6, 40, 0.25, max_time_step=5e-4)
    assert t_c == pytest.approx(t_f, rel=1e-12)
    assert err_c / err_f > 16, (err_c, err_f)


# ---- modal filter: tensor-product application computed independently with numpy -----------------------------------

@pytest.mark.parametrize("nd,rs", [(2, 5), (3, 4)])
def test_filter_is_tensor_product(nd, rs):
    """Local with use_filter applies the 1D filter along every axis to the residual (Spatial.hpp:472-481)."""
    m0 = mesh.box(nd, rs, 2, deformed=True, jitter=0.1, seed=2,
                  bc=(abi.HX_BC_FREESTREAM, configs.freestream_state(nd, 1., [0.3] + [0.] * (nd - 1), 1.)))
    rng = np.random.default_rng(5)
    mesh.set_state(m0, lambda pos: configs.freestream_state(nd, 1. + 0.1 * np.sin(3 * pos[0]), [0.3, 0.1, -0.2][:nd],
                                                            1. + 0.05 * np.cos(2 * pos[-1])))
    init_faces(m0)
    m1 = m0.copy()
    res = []
    for m, uf in ((m0, 0), (m1, 1)):
        s = m.struct()
        opts = abi.update_options(use_filter=bool(uf))
        o_call("oracle_compute_residual", ctypes.byref(s), ctypes.byref(opts))
        res.append(m.rcache[1][:, :m.nv].copy())
    # the filter acts on the rate before the point-wise scale dt*tss/(h*det) (Spatial.hpp:472-489)
    F = BASIS(rs)["filter"]
    n = m0.n_elem[1]
    mult = (m0.tss[1] / m0.nom_size[1][:, None] / m0.metrics[:, nd * nd])[:, None, :]
    r = (res[0] / mult).reshape((n, m0.nv) + (rs,) * nd)
    for d in range(nd):
        r = np.moveaxis(np.moveaxis(r, 2 + d, -1) @ F.T, -1, 2 + d)
    np.testing.assert_allclose(res[1], r.reshape(n, m0.nv, m0.nq) * mult, atol=1e-12 * np.abs(res[0]).max())
    assert rng is not None


# ---- face permutation vs vertex correspondence (test_Face_permutation.cpp) -----------------------------------------

def face_positions(m, b, elem_idx, i_dim, sign):
    """Physical positions [nd][nfq] of the face quadrature points of a deformed element (boundary extrapolation)."""
    nd, rs = m.n_dim, m.row_size
    pos = m.qpos[1][elem_idx].reshape((nd,) + (rs,) * nd)
    B = BASIS(rs)["boundary"][sign]
    face = np.tensordot(pos, B, axes=([1 + i_dim], [0]))
    return face.reshape(nd, -1)


ORIENTATIONS = [(nd, dims, signs)
                for nd in (2, 3)
                for dims in itertools.product(range(nd), repeat=2)
                for signs in itertools.product((0, 1), repeat=2)
                if not (dims[0] == dims[1] and signs[0] == signs[1])]


@pytest.mark.parametrize("nd,dims,signs", ORIENTATIONS)
def test_face_permutation_matches_geometry(nd, dims, signs):
    """For every connection orientation, face 1's quadrature points reordered by match_faces coincide with face 0's."""
    rs = 4
    b, m = meshes.pair(nd, rs, dims, signs)
    p0 = face_positions(m, b, 0, dims[0], signs[0])
    p1 = face_positions(m, b, 1, dims[1], signs[1])
    perm = mesh.face_permutation_index(nd, rs, mesh.pack_dir(dims[0], dims[1], signs[0], signs[1]))
    np.testing.assert_allclose(p1[:, perm], p0, atol=1e-12)


def test_invalid_kernel_errors():
    """kernel_factory rejects n_dim outside 1..3 and row_size outside 2..8 (kernel_factory.hpp:114-116)."""
    import oracle
    lib = oracle.load()
    m = mesh.box(2, 3, 2)
    s = m.struct()
    s.row_size = 9
    assert lib.oracle_write_face(ctypes.byref(s), 0, None) != 0
    assert lib.oracle_last_error().decode() == "demand for invalid kernel"
    s.row_size = 3
    s.n_dim = 4
    assert lib.oracle_local(ctypes.byref(s), 0, None, 1e-3, 0, 0, 0) != 0
    s.n_dim = 2
    # HEXED_ASSERTs of Local (Spatial.hpp:322-323)
    assert lib.oracle_local(ctypes.byref(s), 0, None, 1e-3, 1, 1, 0) != 0
    assert lib.oracle_last_error().decode() == "residual calculation is a single-stage operation"
    pde = abi.ns_pde(abi.constant(1.), abi.constant(1.))
    assert lib.oracle_local(ctypes.byref(s), 0, ctypes.byref(pde), 1e-3, 1, 0, 0) != 0
    assert lib.oracle_last_error().decode() == "two-stage stabilization is not applicable to diffusion equations"
    assert math.isfinite(1.0)
