"""Tree meshes with hanging faces: the builder's wiring of Refined_connection (connection.hpp:168-286)
and the oracle on it. The GPU side is compared with the oracle in test_gpu_tree.py."""
import ctypes
import collections

import numpy as np
import pytest

from hexed_amd import abi, configs

from helpers import init_faces, o_call


def test_leaves_are_balanced_and_tile_the_domain():
    base, levels = 4, 3
    leaves = configs.tree_leaves(base, levels, 0.15, seed=2)
    # volumes add up to the domain
    vol = sum(8.0 ** -lvl for lvl, _ in leaves)
    assert abs(vol - base ** 3) < 1e-12
    # 2:1 across faces: the covering leaf of every same-level face neighbour is at most one level coarser
    for lvl, p in leaves:
        for d in range(3):
            for s in (-1, 1):
                q = list(p)
                q[d] += s
                if not 0 <= q[d] < base * 2 ** lvl:
                    continue
                for up in range(0, lvl + 1):
                    if (lvl - up, tuple(x >> up for x in q)) in leaves:
                        assert up <= 1
                        break
    assert max(lvl for lvl, _ in leaves) == levels


def test_default_config_counts():
    leaves = configs.tree_leaves()
    c = collections.Counter(lvl for lvl, _ in leaves)
    assert dict(c) == {0: 2835, 1: 9355, 2: 5663, 3: 1608}


def test_tree_mesh_wiring():
    m = configs.tree_mesh(base=3, rs=4, levels=2, frac=0.2, seed=1)
    n = m.n_elem[0]
    nd = m.n_dim
    # every slot is referenced exactly once: element faces + mortar faces, plus the boundary ghosts
    used = collections.Counter(m.face_slot[0].ravel().tolist())
    used.update(m.ref_slot[:, 1:][m.ref_slot[:, 1:] >= 0].ravel().tolist())
    for k in m.bc_con:
        used.update([int(m.con_slot[1][k, 1])])
    assert sorted(used) == list(range(m.faces.shape[0]))
    assert set(used.values()) == {1}
    # each coarse slot is a face of its coarse element; each fine connection pairs a mortar with a fine face
    coarse = set(m.ref_slot[:, 0].tolist())
    assert coarse <= set(m.face_slot[0].ravel().tolist())
    mortars = set(m.ref_slot[:, 1:].ravel().tolist()) - {-1}
    assert len(mortars) == 4 * m.ref_slot.shape[0]
    fine_cons = [tuple(c) for c in m.con_slot[0] if c[0] in mortars or c[1] in mortars]
    assert len(fine_cons) == len(mortars)
    assert n == sum(m.tree_counts.values())
    assert m.ref_slot.shape[0] > 0 and nd == 3


@pytest.mark.parametrize("rs", [4, 6])
def test_oracle_freestream_preserved_on_tree(rs):
    """A uniform flow stays uniform through hanging faces: prolongation interpolates constants exactly and
    restriction of the four mortar fluxes reproduces them (R0 P0 + R1 P1 = I)."""
    fs = configs.freestream_state(3, 1.0, [0.8, 0.1, -0.2], 1.0 / 1.4)
    m = configs.tree_mesh(base=2, rs=rs, levels=2, frac=0.4, seed=3, bc=(abi.HX_BC_FREESTREAM, fs))
    assert m.ref_slot.shape[0] > 0
    for v in range(5):
        m.state[0][:, v, :] = fs[v]
    init_faces(m)
    s = m.struct()
    opts = abi.update_options(max_safety=0.5)
    st = abi.hx_update_status()
    for _ in range(2):
        o_call("oracle_update", ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st))
    for v in range(5):
        assert np.abs(m.state[0][:, v, :] - fs[v]).max() <= 1e-13 * max(abs(x) for x in fs)
