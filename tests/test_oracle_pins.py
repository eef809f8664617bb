"""Pins of the CPU oracle's remaining functions, each against what the mathematics fixes, written independently of the
oracle's own formulation (physical variables, textbook operators), and each shown to reject a plausible mistake:

- the local Lax-Friedrichs flux of Neighbor (Spatial.hpp:601-705) with the reference's total-energy sound speed
  estimate (pde.hpp:161-166), on hand-built connections, Cartesian and deformed with flipped normals;
- Max_dt's characteristic speed for a moving state (Spatial.hpp:790-818);
- Prolong_refined / Restrict_refined on the exponential of test_Prolong_refined.cpp / test_Restrict_refined.cpp,
  with and without stretching, and the analytic residual of test_Solver.cpp's Nonuniform_mass on a tree mesh with
  hanging faces (with the fine-face halves swapped as the planted mistake);
- the Navier-Stokes diffusive flux (pde.hpp:130-158, Transport_model.hpp:35-38) against a symbolic manufactured
  solution: non-solenoidal velocity, nonuniform temperature, Sutherland viscosity and conductivity, bulk and
  Laplacian artificial viscosity, on the reference's All_cartesian and rotated All_deformed meshes at row size 8;
- the artificial-viscosity generator's parts: pde::Advection on an exactly representable advected profile,
  pde::Smooth_art_visc on a cosine mode's diffusion, and the Legendre projection of update_art_visc_smoothness
  (Solver.cpp:532-546) on polynomials with a known top coefficient.
"""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, configs, mesh

import meshes
from helpers import GAMMA, init_faces, o_call

BASIS = mesh.basis
R_AIR = 287.05287  # constants::specific_gas_air (constants.hpp:44)


# ---- local Lax-Friedrichs flux ----------------------------------------------------------------------------------

def conserved(rho, vel, p):
    vel = np.asarray(vel, dtype=float)
    return np.concatenate([rho * vel, [rho, p / (GAMMA - 1) + 0.5 * rho * vel @ vel]])


def physical_flux(rho, vel, p, n):
    """Euler flux through a surface with (area-weighted) normal n, from primitive variables."""
    vel, n = np.asarray(vel, dtype=float), np.asarray(n, dtype=float)
    un = vel @ n
    energy = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
    return np.concatenate([rho * vel * un + p * n, [rho * un], [(energy + p) * un]])


def char_speed_estimate(rho, vel, p):
    """The reference's characteristic speed: a sound speed from the total energy, sqrt(gamma (gamma - 1) E / rho),
    which is not below the true sound speed, plus the flow speed (pde.hpp:161-166)."""
    vel = np.asarray(vel, dtype=float)
    energy = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
    return np.sqrt(GAMMA * (GAMMA - 1) * energy / rho) + np.sqrt(vel @ vel)


def llf(left, right, n):
    """F* = (F(u0) + F(u1) + max(lambda0, lambda1) |n| (u0 - u1)) / 2 with left/right = (rho, vel, p)."""
    lam = max(char_speed_estimate(*left), char_speed_estimate(*right))
    return 0.5 * (physical_flux(*left, n) + physical_flux(*right, n)
                  + lam * np.linalg.norm(n) * (conserved(*left) - conserved(*right)))


def one_connection(nd, deformed, dims=None, signs=(1, 0), normal=None, rs=2):
    """A mesh of one connection between two face slots (no elements): Neighbor reads and writes only these."""
    m = mesh.HostMesh(nd, rs, 0, 0, 2, 2 if deformed else 0, 0 if deformed else 1, 1 if deformed else 0)
    c = 1 if deformed else 0
    dims = dims or (0, 0)
    m.con_slot[c][0] = [0, 1]
    m.con_dir[c][0] = mesh.pack_dir(dims[0], dims[1], signs[0], signs[1])
    if deformed:
        m.con_nslot[0] = [0, 1]
        m.normals[0] = np.asarray(normal, dtype=float)[:, None]
        m.normals[1] = np.asarray(normal, dtype=float)[:, None]
    return m, c


def set_face(m, slot, state, block=0):
    nv, nfq = m.nv, m.nfq
    m.faces[slot, block * nv * nfq:(block + 1) * nv * nfq] = np.repeat(np.asarray(state, float), nfq)


def get_face(m, slot, block=0):
    nv, nfq = m.nv, m.nfq
    return m.faces[slot, block * nv * nfq:(block + 1) * nv * nfq].reshape(nv, nfq)


def run_neighbor(m, c, pde=None):
    s = m.struct()
    o_call("oracle_neighbor", ctypes.byref(s), c, None if pde is None else ctypes.byref(pde), 0)


def test_llf_stationary_contact():
    """A stationary contact (equal pressure, no velocity, unequal density): the exact flux is pure pressure, so all
    of the mass flux is the dissipation term, scaled by the faster side's speed sqrt(gamma p / rho1)."""
    left, right = (1.0, [0., 0., 0.], 1.0), (0.5, [0., 0., 0.], 1.0)
    for d in range(3):
        m, c = one_connection(3, False, dims=(d, d))
        set_face(m, 0, conserved(*left))
        set_face(m, 1, conserved(*right))
        run_neighbor(m, c)
        n = np.eye(3)[d]
        want = np.concatenate([1.0 * n, [0.5 * np.sqrt(GAMMA * 1.0 / 0.5) * (1.0 - 0.5)], [0.]])
        for slot in (0, 1):
            np.testing.assert_allclose(get_face(m, slot), want[:, None].repeat(m.nfq, 1), rtol=0, atol=1e-15)


MOVING = [((1.0, [2.0, 0.5, -0.3], 1.0), (0.8, [1.5, -0.3, 0.2], 0.9)),   # left faster
          ((0.6, [0.2, 0.1, 0.0], 0.5), (1.3, [-2.5, 0.4, 0.7], 1.6))]   # right faster


@pytest.mark.parametrize("left,right", MOVING)
def test_llf_moving_states_cartesian(left, right):
    """Moving states, where the total-energy sound speed exceeds the textbook sqrt(gamma p / rho) by >30%: the flux on
    every Cartesian direction matches F*; the textbook speed, the minimum of the two speeds or a dropped
    dissipation term each miss by far more than the round-off bound."""
    for d in range(3):
        m, c = one_connection(3, False, dims=(d, d))
        set_face(m, 0, conserved(*left))
        set_face(m, 1, conserved(*right))
        run_neighbor(m, c)
        n = np.eye(3)[d]
        want = llf(left, right, n)
        got = get_face(m, 0)[:, 0]
        np.testing.assert_allclose(get_face(m, 1)[:, 0], got, rtol=0, atol=0)
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-14 * scale
        # the alternatives a plausible mistake would compute
        textbook = 0.5 * (physical_flux(*left, n) + physical_flux(*right, n) + max(
            np.sqrt(GAMMA * left[2] / left[0]) + np.linalg.norm(left[1]),
            np.sqrt(GAMMA * right[2] / right[0]) + np.linalg.norm(right[1])) * (conserved(*left) - conserved(*right)))
        slow = 0.5 * (physical_flux(*left, n) + physical_flux(*right, n) + min(
            char_speed_estimate(*left), char_speed_estimate(*right)) * (conserved(*left) - conserved(*right)))
        central = 0.5 * (physical_flux(*left, n) + physical_flux(*right, n))
        for wrong in (textbook, slow, central):
            assert np.abs(got - wrong).max() > 1e-3 * scale


@pytest.mark.parametrize("left,right", MOVING)
@pytest.mark.parametrize("signs", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_llf_deformed_area_weighted_normal(left, right, signs):
    """Deformed connection with a non-unit area-weighted normal: the flux uses n = sign0 * normal, with sign_i = -1
    where face_sign[i] == i (Kernel_connection.hpp:17), its dissipation is scaled by |n|, and side i stores
    sign_i * F* (Spatial.hpp:660-676)."""
    normal = np.array([0.3, 1.7, -0.4])
    m, c = one_connection(3, True, dims=(1, 2), signs=signs, normal=normal)
    set_face(m, 0, conserved(*left))
    set_face(m, 1, conserved(*right))
    run_neighbor(m, c)
    sign = [1 - 2 * int(signs[i] == i) for i in (0, 1)]
    want = llf(left, right, sign[0] * normal)
    for side in (0, 1):
        got = get_face(m, side)
        assert np.abs(got - sign[side] * want[:, None]).max() <= 1e-14 * np.abs(want).max()
    # |n| dropped from the dissipation, or the unsigned normal, would miss
    unit = 0.5 * (physical_flux(*left, sign[0] * normal) + physical_flux(*right, sign[0] * normal)
                  + max(char_speed_estimate(*left), char_speed_estimate(*right)) * (conserved(*left) - conserved(*right)))
    assert np.abs(get_face(m, 0)[:, 0] - unit).max() > 1e-3 * np.abs(want).max()
    if sign[0] < 0:
        assert np.abs(get_face(m, 0)[:, 0] - sign[0] * llf(left, right, normal)).max() > 1e-3 * np.abs(want).max()


def test_llf_consistency_and_ldg_average():
    """Equal states give the physical flux; the Navier-Stokes Neighbor also stores the LDG average (u0 + u1) / 2 in
    both LDG blocks (Spatial.hpp:644-651)."""
    st = (1.2, [0.7, -0.4, 0.25], 0.8)
    m, c = one_connection(3, False, dims=(2, 2))
    set_face(m, 0, conserved(*st))
    set_face(m, 1, conserved(*st))
    run_neighbor(m, c)
    want = physical_flux(*st, np.eye(3)[2])
    assert np.abs(get_face(m, 0)[:, 0] - want).max() <= 1e-15 * np.abs(want).max()
    left, right = MOVING[0]
    m, c = one_connection(3, False, dims=(1, 1))
    set_face(m, 0, conserved(*left))
    set_face(m, 1, conserved(*right))
    run_neighbor(m, c, abi.ns_pde(abi.constant(1e-3), abi.constant(1e-3)))
    avg = 0.5 * (conserved(*left) + conserved(*right))
    for side in (0, 1):
        np.testing.assert_allclose(get_face(m, side, block=1)[:, 0], avg, rtol=1e-15)
        np.testing.assert_allclose(get_face(m, side)[:, 0], llf(left, right, np.eye(3)[1]), rtol=1e-14)


def test_max_dt_moving_state_total_energy_sound_speed():
    """Max_dt of a fast uniform flow on one Cartesian element: dt = safety * max_cfl * spacing / lambda, with lambda
    the reference's estimate (20% above the textbook sqrt(gamma p / rho) + |u| here); spacing = h / n_dim
    (Element.cpp:17), max_cfl = -2 * 0.9 / min_eig_convection (Basis.cpp:6-9)."""
    nd, rs, h = 2, 4, 0.37
    m = mesh.HostMesh(nd, rs, 1, 0, 1, 0, 0, 0)
    m.nom_size[0][:] = h
    m.vertex_tss[0][:] = h / nd
    st = (0.9, [2.7, -0.8], 0.6)
    m.state[0][0] = conserved(*st)[:, None]
    s = m.struct()
    out = ctypes.c_double()
    o_call("oracle_max_dt", ctypes.byref(s), 0, None, 0, 0, 0.6, 0.6, ctypes.byref(out))
    max_cfl = -2 * 0.9 / BASIS(rs)["eigenvals"][0]
    lam = char_speed_estimate(*st)
    assert out.value == pytest.approx(0.6 * max_cfl * (h / nd) / lam, rel=1e-14)
    textbook = np.sqrt(GAMMA * st[2] / st[0]) + np.linalg.norm(st[1])
    assert abs(out.value - 0.6 * max_cfl * (h / nd) / textbook) > 0.1 * out.value


# ---- hanging faces -------------------------------------------------------------------------------------------

def exp_face(rs, nv, x0, x1):
    """exp(x0 + 0.5 x1) + i_var at face point (x0, x1), dimension 0 slowest (test_Prolong_refined.cpp:17-21)."""
    return np.stack([(np.exp(x0[:, None] + 0.5 * x1[None, :]) + v).ravel() for v in range(nv)])


def refined_face_mesh(rs, stretch):
    m = mesh.HostMesh(3, rs, 0, 0, 5, 0, 0, 0, n_ref=1)
    m.ref_slot[0] = [0, 1, 2, 3, 4]
    m.ref_stretch[0] = stretch
    return m


CASES = [((0, 0), [(0, 0), (0, 1), (1, 0), (1, 1)]), ((1, 0), [(None, 0), (None, 1)]), ((0, 1), [(0, None), (1, None)])]


@pytest.mark.parametrize("stretch,halves", CASES, ids=["none", "stretch0", "stretch1"])
def test_prolong_exponential(stretch, halves):
    """test_Prolong_refined.cpp: fine face (i, j) of the four (fewer where stretched) holds the coarse face's
    exp(x0 + 0.5 x1) at its own points (x + half) / 2 along each unstretched dimension; margin 1e-4 in the reference,
    1e-8 here at row size 8."""
    rs = 8
    node = BASIS(rs)["node"]
    m = refined_face_mesh(rs, stretch)
    m.faces[0, :5 * rs * rs] = exp_face(rs, 5, node, node).ravel()
    s = m.struct()
    o_call("oracle_prolong_refined", ctypes.byref(s), None, 0, 0)
    for f, (h0, h1) in enumerate(halves):
        x0 = node if h0 is None else (node + h0) / 2
        x1 = node if h1 is None else (node + h1) / 2
        want = exp_face(rs, 5, x0, x1)
        got = m.faces[1 + f, :5 * rs * rs].reshape(5, -1)
        assert np.abs(got - want).max() <= 1e-8, (f, np.abs(got - want).max())
    # the halves exchanged would be wrong by the function's variation over half a face
    if len(halves) > 1:
        x0 = node if halves[1][0] is None else (node + halves[1][0]) / 2
        x1 = node if halves[1][1] is None else (node + halves[1][1]) / 2
        assert np.abs(m.faces[1, :5 * rs * rs].reshape(5, -1) - exp_face(rs, 5, x0, x1)).max() > 0.1


@pytest.mark.parametrize("stretch,halves", CASES, ids=["none", "stretch0", "stretch1"])
def test_restrict_exponential(stretch, halves):
    """test_Restrict_refined.cpp: fine faces sampling exp(x0 + 0.5 x1) on their halves restrict to the coarse face's
    values, halved along a stretched dimension (Restrict_refined's default scale divides by 2)."""
    rs = 8
    node = BASIS(rs)["node"]
    m = refined_face_mesh(rs, stretch)
    for f, (h0, h1) in enumerate(halves):
        x0 = node if h0 is None else (node + h0) / 2
        x1 = node if h1 is None else (node + h1) / 2
        m.faces[1 + f, :5 * rs * rs] = exp_face(rs, 5, x0, x1).ravel()
    s = m.struct()
    o_call("oracle_restrict_refined", ctypes.byref(s), None, 1, 0)
    factor = 0.5 if any(stretch) else 1.0
    got = m.faces[0, :5 * rs * rs].reshape(5, -1)
    assert np.abs(got - factor * exp_face(rs, 5, node, node)).max() <= 1e-8


def test_restrict_offset_and_unscaled():
    """The offset form acts on the LDG block at (n_dim + 2) nfq, and scale = false keeps stretched sums unhalved."""
    rs = 6
    node = BASIS(rs)["node"]
    m = refined_face_mesh(rs, (1, 0))
    off = 5 * rs * rs
    for f in range(2):
        m.faces[1 + f, off:2 * off] = exp_face(rs, 5, node, (node + f) / 2).ravel()
    s = m.struct()
    o_call("oracle_restrict_refined", ctypes.byref(s), None, 0, 1)
    assert np.abs(m.faces[0, off:2 * off].reshape(5, -1) - exp_face(rs, 5, node, node)).max() <= 1e-6
    assert np.all(m.faces[0, :off] == 0)


def nonuniform_mass_state(nd, wave, pressure=1e5):
    """Nonuniform_mass (test_Solver.cpp:103-170) with wave numbers `wave`, and its exact time derivative. The
    reference's pressure of 1e5 leaves the energy residual (a variation of 1e-5 of the energy) at round-off level 1e-3;
    a pressure of order one keeps every variable's residual well above round-off."""
    veloc = [0.3, -0.7, 0.8][:nd]

    def state(pos):
        sp = sum(wave[d] * pos[d] for d in range(nd))
        mass = 1. + 0.1 * np.sin(sp)
        return [mass * v for v in veloc] + [mass, pressure / 0.4 + mass * 0.5 * sum(v * v for v in veloc)]

    def residual(pos):
        sp = sum(wave[d] * pos[d] for d in range(nd))
        dm = -0.1 * sum(v * w for v, w in zip(veloc, wave)) * np.cos(sp)
        return [dm * v for v in veloc] + [dm, dm * 0.5 * sum(v * v for v in veloc)]
    return state, residual


def tree_residual_error(swap_halves):
    m = configs.tree_mesh(base=2, rs=8, levels=2, frac=0.4, seed=3, bc=(abi.HX_BC_COPY, [0.] * 5))
    assert m.ref_slot.shape[0] > 0
    if swap_halves:
        m.ref_slot[:, [1, 2]] = m.ref_slot[:, [2, 1]]
    state, residual = nonuniform_mass_state(3, [-0.2, 0.6, 0.4], pressure=1.0)
    mesh.set_state(m, state)
    init_faces(m)
    s = m.struct()
    kop = abi.hx_kernel_options(0., 0, 0, 0)
    out = ctypes.c_double()
    o_call("oracle_max_dt_euler", ctypes.byref(s), ctypes.byref(kop), 0.7, 0.7, 1, ctypes.byref(out))
    o_call("oracle_compute_residual", ctypes.byref(s), ctypes.byref(abi.update_options(local_time=True)))
    got = m.rcache[0][:, :5] / m.tss[0][:, None, :]
    pos = m.qpos[0]
    want = np.stack(residual([pos[:, d] for d in range(3)]), axis=1)
    return [np.abs(got[:, v] - want[:, v]).max() / np.abs(want[:, v]).max() for v in range(5)]


def test_nonuniform_mass_on_tree_with_hanging_faces():
    """test_marching's analytic residual (test_Solver.cpp:555-586) through hanging faces, per variable relative to the
    residual itself; the same mesh with each refined face's first two fine faces exchanged (the fine-face -> half
    mapping wrong) misses by orders of magnitude."""
    err = tree_residual_error(False)
    assert max(err) <= 1e-8, err
    assert min(tree_residual_error(True)) > 1.0


# ---- Navier-Stokes diffusive flux: manufactured solution -------------------------------------------------------

SUTH_VISC = (1.716e-5, 273., 111.)   # Case.hil:83-90
SUTH_COND = (0.0241, 273., 194.)
BULK_AV, LAP_AV = 3e-5, 2e-5


def manufactured(nd, variant=""):
    """Symbolic state of a compressible flow (non-solenoidal velocity, nonuniform temperature) and the time
    derivative that the viscous terms alone give it: -div F_v, with F_v of textbook form -- Newtonian stress with
    Stokes' hypothesis plus a bulk artificial viscosity, Fourier conduction with Sutherland coefficients of the
    temperature T = p / (rho R), and a Laplacian artificial viscosity on every conserved variable. `variant` drops or
    alters one term, for the sensitivity checks."""
    import sympy as sp
    X = sp.symbols("x0:%d" % nd)
    rho = 1.2 + 0.1 * sp.sin(0.4 * X[0] + 0.3 * X[1] - (0.2 * X[2] if nd > 2 else 0))
    vel = [40 + 8 * sp.sin(0.5 * X[0] + (0.2 * X[2] if nd > 2 else 0.1 * X[1])),
           -15 + 6 * sp.cos(0.4 * X[1] - 0.3 * X[0])]
    if nd > 2:
        vel.append(10 + 5 * sp.sin(0.3 * X[2] + 0.25 * X[1]))
    p = 1e5 * (1 + 0.02 * sp.cos(0.3 * X[0] + 0.2 * X[1] + (0.35 * X[2] if nd > 2 else 0)))
    r_gas = R_AIR * (1.4 if variant == "gas constant" else 1.0)
    T = p / (rho * r_gas)

    def suth(c):
        if variant == "constant transport":
            return c[0]
        return c[0] * (T / c[1]) ** sp.Rational(3, 2) * (c[1] + c[2]) / (T + c[2])
    mu, k = suth(SUTH_VISC), suth(SUTH_COND)
    grad = [[sp.diff(v, x) for x in X] for v in vel]
    div = sum(grad[i][i] for i in range(nd))
    lam = (0 if variant == "bulk" else BULK_AV * rho) - (0 if variant == "dilatation" else sp.Rational(2, 3) * mu)
    stress = [[mu * (grad[i][j] + grad[j][i]) + (lam * div if i == j else 0) for j in range(nd)] for i in range(nd)]
    energy = p / (GAMMA - 1) + rho * sum(v * v for v in vel) / 2
    cons = [rho * v for v in vel] + [rho, energy]
    lap = 0 if variant == "laplacian" else LAP_AV
    heat = [0 if variant == "conduction" else -k * sp.diff(T, x) for x in X]
    flux = []
    for i in range(nd):
        flux.append([-lap * sp.diff(cons[i], x) - stress[i][j] for j, x in enumerate(X)])
    flux.append([-lap * sp.diff(rho, x) for x in X])
    flux.append([-lap * sp.diff(energy, x) - sum(stress[i][j] * vel[i] for i in range(nd)) + heat[j]
                 for j, x in enumerate(X)])
    rate = [-sum(sp.diff(f[j], X[j]) for j in range(nd)) for f in flux]
    return sp.lambdify(X, cons, "numpy"), sp.lambdify(X, rate, "numpy")


def viscous_rate(m):
    """The oracle's physical residual of Navier-Stokes minus that of Euler on the same state: the viscous terms."""
    rates = []
    for use_ldg in (True, False):
        w = m.copy()
        init_faces(w)
        opts = abi.update_options(local_time=True, visc=abi.sutherland(*SUTH_VISC), cond=abi.sutherland(*SUTH_COND),
                                  use_ldg=use_ldg)
        s = w.struct()
        kop = abi.hx_kernel_options(0., 0, 0, 0)
        out = ctypes.c_double()
        o_call("oracle_max_dt_euler", ctypes.byref(s), ctypes.byref(kop), 0.7, 0.7, 1, ctypes.byref(out))
        o_call("oracle_compute_residual", ctypes.byref(s), ctypes.byref(opts))
        c = 1 if w.n_elem[1] else 0
        rates.append(w.rcache[c][:, :w.nv] / w.tss[c][:, None, :])
    return rates[0] - rates[1]


NS_MESHES = {"all_cartesian": lambda: meshes.all_cartesian(abi.HX_BC_COPY, with_av=True),
             "all_deformed_rot0": lambda: meshes.all_deformed(0, abi.HX_BC_COPY, with_av=True),
             "all_deformed_rot1": lambda: meshes.all_deformed(1, abi.HX_BC_COPY, with_av=True)}


def ns_errors(m, variant=""):
    """Per variable: max |oracle - expected| over the largest expected value of the unaltered flux."""
    c = 1 if m.n_elem[1] else 0
    pos = [m.qpos[c][:, d] for d in range(m.n_dim)]

    def rate(v):
        return np.stack([np.broadcast_to(r, pos[0].shape) for r in manufactured(m.n_dim, v)[1](*pos)], axis=1)
    want, scale = rate(variant), np.abs(rate("")).max(axis=(0, 2))
    return [np.abs(m.visc_rate[:, v] - want[:, v]).max() / scale[v] for v in range(m.nv)]


@pytest.fixture(scope="module", params=sorted(NS_MESHES))
def ns_mesh(request):
    m = NS_MESHES[request.param]()
    c = 1 if m.n_elem[1] else 0
    cons, _ = manufactured(m.n_dim)
    pos = [m.qpos[c][:, d] for d in range(m.n_dim)]
    vals = cons(*pos)
    for v in range(m.nv):
        m.state[c][:, v] = vals[v]
    m.av[c][:, 0] = BULK_AV
    m.av[c][:, 1] = LAP_AV
    m.visc_rate = viscous_rate(m)
    return m


NS_TOL = 5e-6  # measured <= 8.3e-7 (row size 8, element size 1)


def test_ns_viscous_residual_manufactured(ns_mesh):
    """Every conserved variable's viscous rate matches -div F_v at row size 8 (element size 1, smooth fields)."""
    err = ns_errors(ns_mesh)
    assert max(err) <= NS_TOL, err


@pytest.mark.parametrize("variant", ["dilatation", "bulk", "laplacian", "conduction", "gas constant",
                                     "constant transport"])
def test_ns_manufactured_rejects_mistakes(ns_mesh, variant):
    """The same comparison against a flux with one term dropped or altered fails in at least one variable: the bar
    above separates the reference's flux from each of these."""
    assert max(ns_errors(ns_mesh, variant)) > 10 * NS_TOL


# ---- artificial-viscosity generator ---------------------------------------------------------------------------

def av_line(rs, n, h=1.0):
    m = mesh.box(1, rs, n, lo=0.0, hi=n * h, bc=(abi.HX_BC_FREESTREAM, [1.0, 1.0, 2.5]), with_av=True)
    return m.enable_av_generator()


def test_advection_linear_profile_stage():
    """pde::Advection (pde.hpp:265-349), one stage-0 Local: the i-th advected scalar moves with velocity xi_i v,
    xi_i = 2 node_i - 1, under the source 2/L with the implicit sink of write_update (pde.hpp:287-295). For a profile
    linear in x the DG derivative is exact, so each interior point must hold
    (phi + dt tss (2/L - xi_i v b_i)) / (1 + tss 2/L)."""
    rs, n, L, dt, v = 5, 6, 0.8, 0.7, 0.65
    m = av_line(rs, n)
    node = BASIS(rs)["node"]
    xi = 2 * node - 1
    a = np.linspace(0.5, 1.5, rs)
    b = np.linspace(-0.4, 0.3, rs)
    x = m.qpos[0][:, 0]
    m.state[0][:, 0] = v
    m.state[0][:, 1] = 1.0
    m.state[0][:, 2] = 2.5
    for i in range(rs):
        m.advection[0][:, i] = a[i] + b[i] * x
    m.tss[0][:] = 0.3 + 0.1 * np.sin(x)
    before = m.advection[0].copy()
    pde = abi.advection_pde(L)
    s = m.struct()
    for c in (0, 1):
        o_call("oracle_write_face", ctypes.byref(s), c, ctypes.byref(pde))
    o_call("oracle_apply_advection_bcs", ctypes.byref(s))
    o_call("oracle_neighbor", ctypes.byref(s), 0, ctypes.byref(pde), 0)
    o_call("oracle_local", ctypes.byref(s), 0, ctypes.byref(pde), dt, 0, 0, 0)
    tss = m.tss[0][:, None, :]
    want = (before + dt * tss * (2 / L - xi[None, :, None] * v * b[None, :, None])) / (1 + tss * 2 / L)
    interior = slice(1, n - 1)
    err = np.abs(m.advection[0][interior] - want[interior]).max()
    assert err <= 1e-13, err
    # the nodes on [0, 1] instead of [-1, 1] would be a different equation
    wrong = (before + dt * tss * (2 / L - node[None, :, None] * v * b[None, :, None])) / (1 + tss * 2 / L)
    assert np.abs(m.advection[0][interior] - wrong[interior]).max() > 1e-3


def test_smooth_art_visc_cosine_decay():
    """pde::Smooth_art_visc (pde.hpp:355-431) through compute_smooth_av's LDG sequence: with a diffusion time so long
    that source and sink vanish, each of the three smoothed forcing components takes one explicit diffusion step,
    f += dt tss f'', and a cosine mode cos(k x) is multiplied by 1 - dt tss k^2 away from the boundaries."""
    rs, n, k, dt = 8, 10, 0.9, 0.5
    m = av_line(rs, n)
    x = m.qpos[0][:, 0]
    m.state[0][:, 0] = 1.0
    m.state[0][:, 1] = 1.0
    m.state[0][:, 2] = 2.5
    amps = [0.0, 1.0, 0.4, -2.0]
    for i in range(4):
        m.forcing[0][:, i] = amps[i] * np.cos(k * x)
    m.tss[0][:] = 0.2
    before = m.forcing[0].copy()
    s = m.struct()
    sav = abi.hx_pde()
    sav.kind = abi.HX_PDE_SMOOTH_AV
    sav.diff_time, sav.cheby_step = 1e15, 1.0
    for c in (0, 1):
        o_call("oracle_write_face", ctypes.byref(s), c, ctypes.byref(sav))
    o_call("oracle_apply_avc_diff_bcs", ctypes.byref(s))
    kop = abi.hx_kernel_options(dt, 0, 0, 0)
    o_call("oracle_compute_smooth_av", ctypes.byref(s), ctypes.byref(kop), 1e15, 1.0)
    interior = slice(2, n - 2)
    factor = 1 - dt * 0.2 * k * k
    for i in (1, 2, 3):
        want = factor * before[interior, i]
        err = np.abs(m.forcing[0][interior, i] - want).max() / abs(amps[i])
        assert err <= 2e-7, (i, err)  # discretization: (k h)^rs / rs! at row size 8
        # the wrong sign of the Laplacian grows the mode instead
        assert np.abs(m.forcing[0][interior, i] - (2 - factor) * before[interior, i]).max() > 1e-2 * abs(amps[i])
    np.testing.assert_array_equal(m.forcing[0][:, 0], before[:, 0])


@pytest.mark.parametrize("rs,av_rs", [(6, 6), (6, 4), (8, 8)])
def test_forcing_projection_top_legendre_coefficient(rs, av_rs):
    """update_art_visc_smoothness's projection (Solver.cpp:532-546): with advected scalars g(node_i) from a polynomial
    g of degree < row size whose orthonormal-Legendre coefficient of degree av_rs - 1 on [0, 1] is c, Gauss quadrature
    gives proj = c exactly and the first forcing component becomes c^2 * 2 E / rho (no advection or diffusion
    iterations)."""
    from numpy.polynomial import legendre as npleg
    m = av_line(rs, 3)
    node = BASIS(rs)["node"]
    rng = np.random.default_rng(rs * 10 + av_rs)
    nq = m.nq
    c_top = rng.uniform(0.5, 1.5, size=(3, nq))
    for e in range(3):
        for q in range(nq):
            coef = rng.uniform(-1, 1, size=rs)  # Legendre series on [0, 1] (argument 2 s - 1), degree rs - 1
            coef[av_rs - 1] = c_top[e, q] * np.sqrt(2 * (av_rs - 1) + 1)  # orthonormal: sqrt(2n + 1) P_n(2 s - 1)
            m.advection[0][e, :, q] = npleg.legval(2 * node - 1, coef)
    rho = 1.0 + 0.2 * rng.random((3, nq))
    energy = 2.0 + rng.random((3, nq))
    m.state[0][:, 0] = 0.3
    m.state[0][:, 1] = rho
    m.state[0][:, 2] = energy
    init_faces(m)
    opts = abi.av_options(0.1, rs, 1.0, 2.5, advect_iters=0, diff_iters=0, av_rs=av_rs)
    r = ctypes.c_double()
    o_call("oracle_update_art_visc_smoothness", ctypes.byref(m.struct()), ctypes.byref(opts), ctypes.byref(r))
    want = c_top ** 2 * 2 * energy / rho
    np.testing.assert_allclose(m.forcing[0][:, 0], want, rtol=1e-12)


def test_art_visc_saturation_and_residual():
    """update_art_visc_smoothness's clean-up (Solver.cpp:552-575): with no advection or diffusion iterations the last
    forcing component keeps its input f3, and the bulk coefficient becomes us_max F / (us_max + F) with
    F = visc_mult * L * f3 and us_max = L * unscaled_max * sqrt(2 E_inf / rho_inf); the residual is the L2 norm (Gauss
    weights times the element volume h^n_dim) of the coefficient's change. The unsaturated F, or a residual without
    the volume, would land elsewhere."""
    rs, n, L, visc_mult, unscaled_max, mass, energy = 5, 4, 0.2, 3.0, 0.7, 1.0, 2.5
    m = av_line(rs, n, h=0.5)
    rng = np.random.default_rng(17)
    m.state[0][:, 0] = 0.3
    m.state[0][:, 1] = 1.0 + 0.2 * rng.random((n, m.nq))
    m.state[0][:, 2] = 2.0 + rng.random((n, m.nq))
    f3 = rng.uniform(0.0, 2.0, size=(n, m.nq))
    m.forcing[0][:, 3] = f3
    old = rng.uniform(0.0, 0.5, size=(n, m.nq))
    m.av[0][:, 0] = old
    init_faces(m)
    opts = abi.av_options(L, rs, mass, energy, advect_iters=0, diff_iters=0, visc_mult=visc_mult,
                          unscaled_max=unscaled_max)
    r = ctypes.c_double()
    o_call("oracle_update_art_visc_smoothness", ctypes.byref(m.struct()), ctypes.byref(opts), ctypes.byref(r))
    us_max = L * unscaled_max * np.sqrt(2 * energy / mass)
    big_f = visc_mult * L * f3
    want = us_max * big_f / (us_max + big_f)
    np.testing.assert_allclose(m.av[0][:, 0], want, rtol=1e-13)
    assert np.abs(m.av[0][:, 0] - big_f).max() > 1e-2
    w = BASIS(rs)["weight"]
    resid = np.sqrt(np.sum((old - want) ** 2 * w[None, :] * 0.5))
    assert r.value == pytest.approx(resid, rel=1e-12)
    assert abs(r.value - np.sqrt(np.sum((old - want) ** 2 * w[None, :]))) > 1e-3 * resid
