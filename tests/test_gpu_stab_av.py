"""Stab_art_visc (stabilizing_art_visc.cpp:8-61) and Solver::set_art_visc_admis (Solver.cpp:636-659) on the HIP path
against the oracle (itself pinned by closed forms in test_stab_av.py): element values and the Laplacian artificial
viscosity at every point, relative to the largest value, <= 1e-12; on box meshes (1D to 3D, Cartesian and deformed,
row sizes 3 to 8) whose elements carry a top-mode specific-volume perturbation of seeded amplitude, so that the
element values spread over the whole sine ramp (0, on the ramp, 1), and on 2D and 3D tree meshes with hanging
vertices."""
import ctypes

import numpy as np
import pytest

from hexed_amd import abi, configs, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, o_call, say, set_vertices

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def perturb(m, seed):
    """Per element, 1/rho = 1 + b P_{rs-1}(x_d) / 2 along a seeded dimension d, plus a small smooth part; b is chosen
    so that the indicator's log10, about log10(b^2 / (4 n_dim)), is uniform within 1 of the ramp centre
    -4.25 log10(rs - 1): a quarter of the elements below the ramp, half on it, a quarter above."""
    nd, rs = m.n_dim, m.row_size
    node = mesh.basis(rs)["node"]
    idx = np.stack(np.unravel_index(np.arange(rs ** nd), (rs,) * nd))
    x = node[idx]
    top = np.polynomial.legendre.legval(2 * x - 1, np.eye(rs)[rs - 1]) * np.sqrt(2 * rs - 1)
    rng = np.random.default_rng(seed)
    for c in (0, 1):
        n = m.n_elem[c]
        if not n:
            continue
        target = -4.25 * np.log10(rs - 1) + rng.uniform(-1, 1, size=n)
        amp = 2 * np.sqrt(nd * 10.0 ** target)
        dim = rng.integers(0, nd, size=n)
        inv_rho = 1.0 + amp[:, None] * top[dim] * 0.5 + 0.01 * np.sin(3 * x[0])[None, :]
        m.state[c][:, nd] = 1.0 / inv_rho
        for v in range(nd):
            m.state[c][:, v] = 0.1 * m.state[c][:, nd]
        m.state[c][:, nd + 1] = 2.5
        m.av[c][:, 0] = 0.25
        m.av[c][:, 1] = 0.0


def compare(ctx, m, tag, char_speed=3.0):
    set_vertices(m)
    dev = Device_mesh(ctx, m.copy())
    s = m.struct()
    u_o = np.zeros(sum(m.n_elem))
    o_call("oracle_stab_art_visc", ctypes.byref(s), char_speed, u_o.ctypes.data_as(DP))
    u_d = dev.stab_art_visc(char_speed)
    scale = np.max(np.abs(u_o))
    assert scale > 0
    assert np.max(np.abs(u_d - u_o)) <= TOL * scale, f"{tag}: Stab_art_visc {np.max(np.abs(u_d - u_o)) / scale:.3e}"
    lo, hi = (rs_ramp(m.row_size, u_o, char_speed, m))
    say(f"  {tag}: {len(u_o)} elements, {lo} at 0, {hi} at the maximum, {len(u_o) - lo - hi} on the ramp")
    o_call("oracle_set_art_visc_admis", ctypes.byref(s), char_speed, u_o.ctypes.data_as(DP), None)
    u_d = dev.set_art_visc_admis(char_speed)
    assert np.max(np.abs(u_d - u_o)) <= TOL * scale, tag
    for c in (0, 1):
        if m.n_elem[c]:
            got = dev.download("av", c, out=np.zeros_like(m.av[c]))
            err = np.max(np.abs(got[:, 1] - m.av[c][:, 1]))
            assert err <= TOL * scale, f"{tag}: class {c} Laplacian AV {err / scale:.3e}"
            assert np.array_equal(got[:, 0], m.av[c][:, 0]), f"{tag}: bulk AV changed"
    dev.close()


def rs_ramp(rs, u, char_speed, m):
    h = np.concatenate([m.nom_size[c] for c in (0, 1) if m.n_elem[c]])
    top = (rs - 1) * char_speed * h
    return int(np.sum(u == 0)), int(np.sum(np.isclose(u, top, rtol=1e-14)))


BOXES = [(1, 5, False, 8), (2, 4, False, 6), (2, 6, True, 5), (3, 3, True, 4), (3, 6, False, 4), (3, 6, True, 4),
         (3, 8, True, 3)]


@pytest.mark.parametrize("cfg", BOXES, ids=lambda c: f"nd{c[0]}-rs{c[1]}-{'def' if c[2] else 'car'}-n{c[3]}")
def test_stab_art_visc_boxes(ctx, cfg):
    nd, rs, deformed, n = cfg
    say(f"[gpu] stabilizing AV vs oracle: box nd {nd} rs {rs} deformed {deformed}")
    m = mesh.box(nd, rs, n, deformed=deformed, jitter=0.1 if deformed else 0.0, seed=2,
                 bc=(abi.HX_BC_FREESTREAM, freestream(nd)), with_av=True)
    perturb(m, seed=nd * 10 + rs)
    compare(ctx, m, f"box nd{nd} rs{rs}")


@pytest.mark.parametrize("nd,rs", [(2, 5), (3, 4)])
def test_stab_art_visc_tree(ctx, nd, rs):
    say(f"[gpu] stabilizing AV vs oracle: tree nd {nd} rs {rs}")
    m = configs.tree_mesh(base=3, rs=rs, levels=2, frac=0.3, seed=4, nd=nd)
    assert len(m.hanging) > 0
    for c in (0, 1):
        if m.n_elem[c]:
            m.av[c] = np.zeros((m.n_elem[c], 2, m.nq))
    perturb(m, seed=7)
    compare(ctx, m, f"tree nd{nd} rs{rs}")


def set_topology(m, kind, ids0, n0, perm):
    """Vertex ids of `m` rewritten in place: the original ids, the same sharing under renumbered ids, or every element
    with vertices of its own."""
    nvert = 2 ** m.n_dim
    off = 0
    for c in (0, 1):
        if kind == "original":
            m.vertex_ids[c][...] = ids0[c]
        elif kind == "renumbered":
            m.vertex_ids[c][...] = perm[ids0[c]]
        else:
            m.vertex_ids[c][...] = off + np.arange(m.n_elem[c] * nvert, dtype=np.int32).reshape(-1, nvert)
        off += m.n_elem[c] * nvert
    m.n_vertex = off if kind == "unshared" else n0


@pytest.mark.parametrize("shape", ["box3", "tree2"])
def test_set_art_visc_admis_vertex_cache(ctx, shape):
    """The mesh keeps the vertex topology of its last call on the device (hx_calc_jacobian's included): repeated calls,
    the vertex ids renumbered in place (the same host arrays, the same sharing), every element given vertices of its
    own (no sharing, more vertices) and the original ids again each match the oracle given the same topology."""
    say(f"[gpu] set_art_visc_admis vertex cache: {shape}")
    if shape == "box3":
        m = mesh.box(3, 4, 3, deformed=True, jitter=0.1, seed=3, bc=(abi.HX_BC_FREESTREAM, freestream(3)),
                     with_av=True)
    else:
        m = configs.tree_mesh(base=3, rs=5, levels=2, frac=0.3, seed=4, nd=2)
        assert len(m.hanging) > 0
        for c in (0, 1):
            if m.n_elem[c]:
                m.av[c] = np.zeros((m.n_elem[c], 2, m.nq))
    perturb(m, seed=11)
    dev = Device_mesh(ctx, m.copy())
    h = dev.host
    # hx_calc_jacobian fills the same cache; the first set_art_visc_admis below then finds the topology there
    if shape == "box3":
        dev.calc_jacobian()
    ids0 = [m.vertex_ids[c].copy() for c in (0, 1)]
    n0 = int(m.n_vertex)
    perm = np.random.default_rng(5).permutation(n0).astype(np.int32)
    char_speed = 3.0
    lap = {}
    for kind in ("original", "original", "renumbered", "unshared", "original"):
        for x in (m, h):
            set_topology(x, kind, ids0, n0, perm)
        set_vertices(m)
        s = m.struct()
        u_o = np.zeros(sum(m.n_elem))
        o_call("oracle_set_art_visc_admis", ctypes.byref(s), char_speed, u_o.ctypes.data_as(DP), None)
        u_d = dev.set_art_visc_admis(char_speed)
        scale = np.max(np.abs(u_o))
        assert scale > 0
        assert np.max(np.abs(u_d - u_o)) <= TOL * scale, f"{shape} {kind}: element values"
        got = [dev.download("av", c, out=np.zeros_like(m.av[c]))[:, 1] if m.n_elem[c] else None for c in (0, 1)]
        for c in (0, 1):
            if m.n_elem[c]:
                err = np.max(np.abs(got[c] - m.av[c][:, 1]))
                assert err <= TOL * scale, f"{shape} {kind}: class {c} Laplacian AV {err / scale:.3e}"
        lap[kind] = np.concatenate([g.ravel() for g in got if g is not None])
    # the topologies lead to different viscosities, so a stale device topology would have failed above
    assert not np.array_equal(lap["original"], lap["unshared"])
    assert np.array_equal(lap["original"], lap["renumbered"])
    dev.close()
