"""The reference's test meshes (test_Solver.cpp:416-500), rebuilt through hexed_amd.mesh.Builder."""
from hexed_amd import abi, mesh


def all_cartesian(bc_kind, bc_state=None, rs=8, with_av=False):
    """2x2x2 Cartesian elements with one boundary condition on every outer face (test_Solver.cpp:428-456)."""
    b = mesh.Builder(3, rs, 1.0)
    bc = b.add_bc(bc_kind, bc_state or [0.0] * 5)
    sn = []
    strides = [4, 2, 1]
    for i_elem in range(8):
        inds = [(i_elem // strides[d]) % 2 for d in range(3)]
        sn.append(b.add_element(0, False, inds))
        for d in range(3):
            if inds[d]:
                b.connect_cartesian(sn[i_elem - strides[d]], sn[i_elem], d)
            b.connect_boundary(sn[i_elem], d, inds[d], bc)
    return b.build(with_av=with_av)


def all_deformed(rot_dir, bc_kind, bc_state=None, rs=8, with_av=False):
    """3x3 deformed quads whose centre element is rotated; `rot_dir` picks the rotation (test_Solver.cpp:458-500)."""
    b = mesh.Builder(2, rs, 1.0)
    bc = b.add_bc(bc_kind, bc_state or [0.0] * 4)
    sn = []
    strides = [3, 1]
    for i_elem in range(9):
        inds = [(i_elem // strides[d]) % 3 for d in range(2)]
        sn.append(b.add_element(0, True, inds))
        for d in range(2):
            if inds[d] > 0:
                elems = [i_elem - strides[d], i_elem]
                i_dims, signs = [d, d], [1, 0]
                for s in range(2):
                    if elems[s] == 4:
                        i_dims[s] = 1 - i_dims[s]
                        if d == rot_dir:
                            signs[s] = 1 - signs[s]
                b.connect_deformed(sn[elems[0]], sn[elems[1]], i_dims, signs)
            if inds[d] != 1:
                b.connect_boundary(sn[i_elem], d, int(inds[d] > 0), bc)
    return b.build(with_av=with_av)


def pair(nd, rs, dims, signs, bc_kind=abi.HX_BC_COPY):
    """Two deformed elements joined by a connection of arbitrary direction; all other faces are boundaries."""
    b = mesh.Builder(nd, rs, 1.0)
    bc = b.add_bc(bc_kind, [0.0] * (nd + 2))
    e0 = b.add_element(0, True, [0] * nd)
    pos1 = [0] * nd
    pos1[dims[0]] = 1 if signs[0] else -1
    e1 = b.add_element(0, True, pos1)
    b.connect_deformed(e0, e1, dims, signs)
    for e, (dd, ss) in ((e0, (dims[0], signs[0])), (e1, (dims[1], signs[1]))):
        for d in range(nd):
            for s in (0, 1):
                if (d, s) != (dd, int(ss)):
                    b.connect_boundary(e, d, s, bc)
    m = b.build()
    return b, m
