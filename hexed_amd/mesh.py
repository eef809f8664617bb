"""Host-side mesh description (the `hx_host_mesh` of include/hexed_amd.h) and builders for the box meshes the
benchmarks and tests use.

The reference builds meshes through `Mesh::add_element` / `connect_*` (Mesh.hpp:24-230) and derives geometry in
`Solver::calc_jacobian` (Solver.cpp:273-381) and `Deformed_element::set_jacobian` (Deformed_element.cpp:60-136).
Mesh construction is setup-time work outside the accelerated path; this module produces the same kernel-visible
data directly:
  - element state / time step scale / nominal size / vertex time-step scale / face slots;
  - connection slots and packed `Connection_direction`s, Cartesian connections oriented {{d, d}, {1, 0}}
    (connection.hpp:29-36), deformed ones as given, boundary connections {{d, d}, {s, !s}} (connection.hpp:377);
  - for deformed elements, the reference-level normals and Jacobian determinant at every quadrature point and the
    area-weighted face normals, shared across each connection by the sign-corrected average of
    Solver.cpp:327-352, and the vertex time-step scale of Deformed_element.cpp:116-135 reduced by minimum over
    the elements sharing each vertex (Solver.cpp:380, Vertex::vector_min).
"""
import ctypes

import numpy as np

from . import abi
from .basis_gen import load_tables

_TABLES = None


def basis(rs):
    global _TABLES
    if _TABLES is None:
        _TABLES = load_tables()
    return _TABLES[rs]


def pack_dir(i_dim0, i_dim1, sign0, sign1):
    return int(i_dim0) | (int(i_dim1) << 2) | (int(sign0) << 4) | (int(sign1) << 5)


def unpack_dir(p):
    return (p & 3, (p >> 2) & 3), (bool((p >> 4) & 1), bool((p >> 5) & 1))


def flip_normal(face_sign, i_side):
    return face_sign[i_side] == bool(i_side)


def face_permutation_index(nd, rs, packed):
    """Index map `perm` with match_faces(face1)[i] = face1[perm[i]] (Spatial.hpp:85-129): transpose, then flip."""
    i_dim, face_sign = unpack_dir(packed)
    nfq = rs ** (nd - 1)
    idx = np.arange(nfq)
    transpose = (i_dim[0] == 0 and i_dim[1] == 2) or (i_dim[0] == 2 and i_dim[1] == 0)
    flip_t = (i_dim[0] != i_dim[1]) and (flip_normal(face_sign, 0) == flip_normal(face_sign, 1))
    if nd == 3:
        m = idx.reshape(rs, rs)  # m[c, r] = index of column-major (r, c)
        if transpose:
            m = m.T
        if flip_t:
            colwise = (i_dim[0] > 3 - i_dim[0] - i_dim[1]) != transpose
            m = m[:, ::-1] if colwise else m[::-1, :]
        return np.ascontiguousarray(m).reshape(-1)
    if nd == 2 and flip_t:
        return idx[::-1].copy()
    return idx


class HostMesh:
    """numpy arrays laid out as `hx_host_mesh`; `struct()` returns the ctypes view (arrays must stay alive)."""

    def __init__(self, nd, rs, n_car, n_def, n_slot, n_nslot, n_con_car, n_con_def, n_ref=0, with_av=False):
        self.n_dim, self.row_size = nd, rs
        self.nq, self.nfq, self.nv, self.nvert = rs ** nd, rs ** (nd - 1), nd + 2, 2 ** nd
        self.face_sz = max(2 * self.nv, nd + rs) * self.nfq
        self.n_elem = [n_car, n_def]
        self.state = [np.zeros((n, self.nv, self.nq)) for n in self.n_elem]
        self.nrc = max(self.nv, rs)  # residual-cache variables (Storage_params.cpp:32-35)
        self.rcache = [np.zeros((n, self.nrc, self.nq)) for n in self.n_elem]
        self.tss = [np.ones((n, self.nq)) for n in self.n_elem]
        self.av = [np.zeros((n, 2, self.nq)) for n in self.n_elem] if with_av else [None, None]
        self.forcing = [None, None]    # [n][4][nq] artificial-viscosity forcing (Element.cpp:134)
        self.advection = [None, None]  # [n][rs][nq] advected scalars of the smoothness indicator (Element.cpp:139)
        self.nom_size = [np.zeros(n) for n in self.n_elem]
        self.vertex_tss = [np.zeros((n, self.nvert)) for n in self.n_elem]
        self.face_slot = [np.zeros((n, 2 * nd), np.int32) for n in self.n_elem]
        self.metrics = np.zeros((n_def, nd * nd + 1, self.nq))
        self.face_nslot = np.full((n_def, 2 * nd), -1, np.int32)
        self.faces = np.zeros((n_slot, self.face_sz))
        self.normals = np.zeros((n_nslot, nd, self.nfq))
        self.con_slot = [np.zeros((n_con_car, 2), np.int32), np.zeros((n_con_def, 2), np.int32)]
        self.con_dir = [np.zeros(n_con_car, np.int32), np.zeros(n_con_def, np.int32)]
        self.con_nslot = np.zeros((n_con_def, 2), np.int32)
        self.ref_slot = np.full((n_ref, 5), -1, np.int32)
        self.ref_stretch = np.zeros((n_ref, 2), np.int32)
        self.bc_con = np.zeros(0, np.int32)
        self.bc_serial = np.zeros(0, np.int32)
        # (kind, freestream state[, {pressure, thermal, thermal_value, emissivity, heat_transfer_coef, temperature,
        # coercion}]) per boundary condition
        self.bcs = []
        self.bc_cache = None  # [n_bc_con][nv][nfq], allocated by struct()
        # geometry kept for initial conditions: qpoint positions per class [n][nd][nq]
        self.qpos = [None, None]
        # vertex connectivity (not part of hx_host_mesh; the oracle's fix_admissibility needs it): the shared vertex of
        # every element vertex per class [n][nvert], and per hanging face [12]: i_dim, is_positive, stretch[2], then
        # (class, element) of the fine elements in order (Hanging_vertex_matcher.hpp:25)
        self.n_vertex = 0
        self.vertex_ids = [np.zeros((n, self.nvert), np.int32) for n in self.n_elem]
        self.hanging = np.zeros((0, 12), np.int32)
        # vertex positions of the deformed elements [n_def][nvert][nd] (Vertex::pos) and their face node adjustments
        # [n_def][2 nd][nfq] (Deformed_element::node_adjustments; zero: unwarped faces)
        self.def_vertex_pos = np.zeros((n_def, self.nvert, nd))
        self.node_adj = None

    def vertex_data(self):
        """hx_vertex_data over this mesh's arrays (the arrays must outlive the struct)."""
        self._vd_arrays = [np.ascontiguousarray(self.vertex_ids[c], dtype=np.int32) for c in range(2)]
        self._vd_hang = np.ascontiguousarray(self.hanging, dtype=np.int32)
        self._vd_pos = np.ascontiguousarray(self.def_vertex_pos, dtype=np.float64)
        self._vd_adj = None if self.node_adj is None else np.ascontiguousarray(self.node_adj, dtype=np.float64)
        v = abi.hx_vertex_data()
        v.n_vertex = int(self.n_vertex)
        for c in range(2):
            v.vertex_id[c] = abi.iptr(self._vd_arrays[c])
        v.position = abi.dptr(self._vd_pos)
        v.node_adj = abi.dptr(self._vd_adj) if self._vd_adj is not None else None
        v.n_hanging = int(self._vd_hang.shape[0])
        v.hanging = abi.iptr(self._vd_hang) if self._vd_hang.shape[0] else None
        return v

    def enable_av_generator(self):
        """Allocate the arrays Solver::update_art_visc_smoothness works on (zeros, as Element.cpp:134-142 does)."""
        for c in range(2):
            n = self.n_elem[c]
            if self.av[c] is None:
                self.av[c] = np.zeros((n, 2, self.nq))
            self.forcing[c] = np.zeros((n, 4, self.nq))
            self.advection[c] = np.zeros((n, self.row_size, self.nq))
        return self

    @property
    def n_dof(self):
        return sum(self.n_elem) * self.nq

    def copy(self):
        out = HostMesh.__new__(HostMesh)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(out, k, v.copy())
            elif isinstance(v, list):
                setattr(out, k, [x.copy() if isinstance(x, np.ndarray) else x for x in v])
            else:
                setattr(out, k, v)
        return out

    def struct(self):
        m = abi.hx_host_mesh()
        m.n_dim, m.row_size = self.n_dim, self.row_size
        for c in range(2):
            m.n_elem[c] = self.n_elem[c]
            m.state[c] = abi.dptr(self.state[c])
            m.rcache[c] = abi.dptr(self.rcache[c])
            m.tss[c] = abi.dptr(self.tss[c])
            m.av[c] = abi.dptr(self.av[c])
            m.forcing[c] = abi.dptr(self.forcing[c])
            m.advection[c] = abi.dptr(self.advection[c])
            m.nom_size[c] = abi.dptr(self.nom_size[c])
            m.vertex_tss[c] = abi.dptr(self.vertex_tss[c])
            m.face_slot[c] = abi.iptr(self.face_slot[c])
            m.n_con[c] = len(self.con_dir[c])
            m.con_slot[c] = abi.iptr(self.con_slot[c])
            m.con_dir[c] = abi.iptr(self.con_dir[c])
        m.metrics = abi.dptr(self.metrics)
        m.face_nslot = abi.iptr(self.face_nslot)
        m.n_slot = self.faces.shape[0]
        m.faces = abi.dptr(self.faces)
        m.n_nslot = self.normals.shape[0]
        m.normals = abi.dptr(self.normals)
        m.con_nslot = abi.iptr(self.con_nslot)
        m.n_ref = self.ref_slot.shape[0]
        m.ref_slot = abi.iptr(self.ref_slot)
        m.ref_stretch = abi.iptr(self.ref_stretch)
        m.n_bc_con = len(self.bc_con)
        m.bc_con = abi.iptr(self.bc_con)
        m.bc_serial = abi.iptr(self.bc_serial)
        self._bc_arr = (abi.hx_bc * max(1, len(self.bcs)))()
        for i, bc in enumerate(self.bcs):
            kind, st = bc[0], bc[1]
            extra = bc[2] if len(bc) > 2 else {}
            b = self._bc_arr[i]
            b.kind = kind
            for k, v in enumerate(st):
                b.state[k] = v
            b.coercion = 2.0  # No_slip's default heat_flux_coercion (Boundary_condition.hpp:198)
            for name, value in extra.items():
                setattr(b, name, value)
        m.n_bc = len(self.bcs)
        m.bcs = ctypes.cast(self._bc_arr, ctypes.POINTER(abi.hx_bc))
        n_bc_con = len(self.bc_con)
        if self.bc_cache is None or self.bc_cache.shape[0] != n_bc_con:
            self.bc_cache = np.zeros((n_bc_con, self.nv, self.nfq))
        m.bc_cache = abi.dptr(self.bc_cache)
        return m


# ---- geometry helpers -----------------------------------------------------------------------------------------

def node_coords(nd, rs):
    """[nq][nd] reference coordinates of the quadrature points (dimension 0 slowest)."""
    nodes = basis(rs)["node"]
    q = np.arange(rs ** nd)
    return np.stack([nodes[(q // rs ** (nd - 1 - d)) % rs] for d in range(nd)], axis=1)


def vertex_weights(nd, rs):
    """[nq][nvert] n-linear interpolation weights; vertex index bit (nd-1-d) is the coordinate along d."""
    x = node_coords(nd, rs)
    w = np.ones((rs ** nd, 2 ** nd))
    for v in range(2 ** nd):
        for d in range(nd):
            bit = (v >> (nd - 1 - d)) & 1
            w[:, v] *= x[:, d] if bit else 1 - x[:, d]
    return w


def _along(mat, arr, axis):
    """Apply `mat` (m x rs) along `axis` of `arr`."""
    moved = np.moveaxis(arr, axis, -1)
    return np.moveaxis(moved @ mat.T, -1, axis)


def _cofactor(J, r, c):
    """Signed cofactor (r, c) of the stacked nd x nd matrices J[..., i, j] = J[i][j] given as J[i][j] arrays."""
    nd = len(J)
    if nd == 1:
        return np.ones_like(J[0][0])
    if nd == 2:
        return (1 - 2 * ((r + c) % 2)) * J[1 - r][1 - c]
    r1, r2, c1, c2 = (r + 1) % 3, (r + 2) % 3, (c + 1) % 3, (c + 2) % 3
    return J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1]


def deformed_metrics(nd, rs, vert_pos, nom_size, qpos=None):
    """Deformed_element::set_jacobian restated (Deformed_element.cpp:60-136), vectorised over elements. The quadrature
    point positions are the n-linear interpolation of the vertices (unwarped faces) unless `qpos` gives them.

    vert_pos: [n][nvert][nd]. Returns (metrics [n][nd*nd+1][nq], face_normals [n][2nd][nd][nfq],
    vertex_tss [n][nvert], qpos [n][nd][nq]).
    """
    b = basis(rs)
    n = vert_pos.shape[0]
    nq = rs ** nd
    if qpos is None:
        W = vertex_weights(nd, rs)                                # [nq][nvert]
        qpos = np.einsum("qv,nvd->ndq", W, vert_pos)               # [n][nd][nq]
    shape = (n,) + (rs,) * nd
    J = [[None] * nd for _ in range(nd)]
    for i in range(nd):
        p = qpos[:, i].reshape(shape)
        for j in range(nd):
            J[i][j] = _along(b["diff_mat"], p, 1 + j) / nom_size.reshape((n,) + (1,) * nd)
    det = sum(J[0][c] * _cofactor(J, 0, c) for c in range(nd))
    metrics = np.zeros((n, nd * nd + 1, nq))
    for i in range(nd):
        for j in range(nd):
            metrics[:, i * nd + j] = _cofactor(J, j, i).reshape(n, nq)
    metrics[:, nd * nd] = det.reshape(n, nq)
    nfq = rs ** (nd - 1)
    fn = np.zeros((n, 2 * nd, nd, nfq))
    for i in range(nd):
        for s in range(2):
            FJ = [[_along(b["boundary"][s:s + 1], J[a][c], 1 + i) for c in range(nd)] for a in range(nd)]
            for j in range(nd):
                fn[:, 2 * i + s, j] = _cofactor(FJ, j, i).reshape(n, nfq)
    # vertex time step scale: extrapolate normals and determinant to the vertices
    def to_vertices(arr):
        a = arr.reshape(shape)
        for d in range(nd):
            a = _along(b["boundary"], a, 1 + d)
        return a.reshape(n, 2 ** nd)
    vdet = to_vertices(metrics[:, nd * nd])
    norm_sum = np.zeros((n, 2 ** nd))
    for i in range(nd):
        sq = np.zeros((n, 2 ** nd))
        for j in range(nd):
            sq += to_vertices(metrics[:, i * nd + j]) ** 2
        norm_sum += np.sqrt(sq)
    vtss = nom_size[:, None] * vdet / norm_sum
    return metrics, fn, vtss, qpos


# ---- box builder ----------------------------------------------------------------------------------------------

def box(nd, rs, n, lo=0.0, hi=1.0, deformed=False, jitter=0.0, seed=0, bc=(abi.HX_BC_FREESTREAM, None), with_av=False):
    """A box of n^nd equal elements on [lo, hi]^nd with one flow boundary condition on every boundary face.

    Cartesian elements are connected by Cartesian connections; deformed elements by deformed connections in the
    standard orientation {{d, d}, {1, 0}}. With `deformed`, interior vertices are moved by seeded uniform
    jitter*h per component (boundary vertices stay put). Element order is row-major over the grid position
    (dimension 0 slowest), matching the test meshes of test_Solver.cpp:416-456.
    """
    n = int(n)
    h = (hi - lo) / n
    n_e = n ** nd
    grid = np.stack(np.unravel_index(np.arange(n_e), (n,) * nd), axis=1)  # [n_e][nd]
    strides = [n ** (nd - 1 - d) for d in range(nd)]
    # connections between grid neighbours: for each dim, elements with index < n-1 along d
    con_a, con_b, con_d = [], [], []
    for d in range(nd):
        a = np.nonzero(grid[:, d] < n - 1)[0]
        con_a.append(a)
        con_b.append(a + strides[d])
        con_d.append(np.full(len(a), d))
    con_a, con_b, con_d = np.concatenate(con_a), np.concatenate(con_b), np.concatenate(con_d)
    n_int = len(con_a)
    bnd_e, bnd_d, bnd_s = [], [], []
    for d in range(nd):
        for s in (0, 1):
            e = np.nonzero(grid[:, d] == (n - 1 if s else 0))[0]
            bnd_e.append(e)
            bnd_d.append(np.full(len(e), d))
            bnd_s.append(np.full(len(e), s))
    bnd_e, bnd_d, bnd_s = np.concatenate(bnd_e), np.concatenate(bnd_d), np.concatenate(bnd_s)
    n_bnd = len(bnd_e)
    cls = abi.HX_DEFORMED if deformed else abi.HX_CARTESIAN
    n_con_car = 0 if deformed else n_int
    n_con_def = (n_int if deformed else 0) + n_bnd
    n_slot = 2 * (n_int + n_bnd)
    n_nslot = 2 * n_con_def
    m = HostMesh(nd, rs, 0 if deformed else n_e, n_e if deformed else 0, n_slot, n_nslot, n_con_car, n_con_def,
                 with_av=with_av)
    # slots: interior connections first (2 each), then boundary connections (inside, ghost)
    int_slots = np.arange(2 * n_int, dtype=np.int32).reshape(n_int, 2)
    bnd_slots = (2 * n_int + np.arange(2 * n_bnd, dtype=np.int32)).reshape(n_bnd, 2)
    fs = m.face_slot[cls]
    fs[con_a, 2 * con_d + 1] = int_slots[:, 0]
    fs[con_b, 2 * con_d] = int_slots[:, 1]
    fs[bnd_e, 2 * bnd_d + bnd_s] = bnd_slots[:, 0]
    std_dir = np.array([pack_dir(d, d, 1, 0) for d in range(nd)], np.int32)
    bnd_dir = np.array([pack_dir(d, d, s, 1 - s) for d, s in zip(bnd_d, bnd_s)], np.int32)
    if deformed:
        m.con_slot[1][:n_int] = int_slots
        m.con_dir[1][:n_int] = std_dir[con_d]
        off = n_int
    else:
        m.con_slot[0][:] = int_slots
        m.con_dir[0][:] = std_dir[con_d]
        off = 0
    m.con_slot[1][off:] = bnd_slots
    m.con_dir[1][off:] = bnd_dir
    m.con_nslot[:] = np.arange(n_nslot, dtype=np.int32).reshape(-1, 2)
    m.bc_con = np.arange(off, off + n_bnd, dtype=np.int32)
    m.bc_serial = np.zeros(n_bnd, np.int32)
    kind, fs_state = bc[0], bc[1]
    m.bcs = [(kind, list(fs_state) if fs_state is not None else [0.0] * (nd + 2)) + tuple(bc[2:])]
    m.nom_size[cls][:] = h
    # geometry
    rng = np.random.default_rng(seed)
    vgrid = np.stack(np.unravel_index(np.arange((n + 1) ** nd), (n + 1,) * nd), axis=1)
    vpos = lo + vgrid * h
    if deformed and jitter > 0:
        interior = np.all((vgrid > 0) & (vgrid < n), axis=1)
        vpos[interior] += rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), nd))
    vstrides = np.array([(n + 1) ** (nd - 1 - d) for d in range(nd)])
    corner = np.stack(np.unravel_index(np.arange(2 ** nd), (2,) * nd), axis=1)  # vertex index -> offsets
    vid = (grid[:, None, :] + corner[None, :, :]) @ vstrides                 # [n_e][nvert] global vertex ids
    if deformed:
        m.def_vertex_pos[:] = vpos[vid]
        metrics, fn, vtss, qpos = deformed_metrics(nd, rs, vpos[vid], m.nom_size[1])
        m.metrics[:] = metrics
        m.qpos[1] = qpos
        # connection normals: average of the two element face normals with sign correction (Solver.cpp:327-352);
        # standard orientation needs no permutation and has sign +1 on both sides
        nrm = np.zeros((n_con_def, 2, nd, m.nfq))
        n0 = fn[con_a, 2 * con_d + 1]
        n1 = fn[con_b, 2 * con_d]
        avg = 0.5 * n0 + 0.5 * n1
        nrm[:n_int, 0] = avg
        nrm[:n_int, 1] = avg
        # boundary: ghost face holds a copy of the inside normal, both sides flipped iff the inside face is negative
        nin = fn[bnd_e, 2 * bnd_d + bnd_s]
        sgn = np.where(bnd_s == 0, -1.0, 1.0)[:, None, None]
        avgb = 0.5 * sgn * nin + 0.5 * sgn * nin
        nrm[n_int:, 0] = sgn * avgb
        nrm[n_int:, 1] = sgn * avgb
        m.normals[:] = nrm.reshape(n_nslot, nd, m.nfq)
        m.face_nslot[con_a, 2 * con_d + 1] = m.con_nslot[:n_int, 0]
        m.face_nslot[con_b, 2 * con_d] = m.con_nslot[:n_int, 1]
        m.face_nslot[bnd_e, 2 * bnd_d + bnd_s] = m.con_nslot[n_int:, 0]
        elem_vtss = vtss
    else:
        nodes = node_coords(nd, rs)
        m.qpos[0] = ((nodes[None, :, :] + grid[:, None, :]) * h + lo).transpose(0, 2, 1).copy()
        # boundary connections of Cartesian elements carry the unit normal of Element::set_jacobian (Element.cpp:99-112)
        unit = np.zeros((n_bnd, nd, m.nfq))
        unit[np.arange(n_bnd), bnd_d] = 1.0
        m.normals[:] = np.repeat(unit, 2, axis=0)
        elem_vtss = np.full((n_e, 2 ** nd), h / nd)  # Element.cpp:17
    # share vertex time step scale by minimum over the elements using each vertex (Solver.cpp:380)
    vmin = np.full((n + 1) ** nd, np.inf)
    np.minimum.at(vmin, vid.reshape(-1), elem_vtss.reshape(-1))
    m.vertex_tss[cls][:] = vmin[vid]
    m.n_vertex = (n + 1) ** nd
    m.vertex_ids[cls][:] = vid
    return m


def set_state(m, func):
    """Fill the state of every element from func(pos[nd][..]) -> list of nv arrays (Solver::initialize)."""
    for c in range(2):
        if m.n_elem[c] == 0:
            continue
        pos = m.qpos[c]
        vals = func([pos[:, d] for d in range(m.n_dim)])
        for v in range(m.nv):
            m.state[c][:, v] = vals[v]


# ---- batches of independent cases -----------------------------------------------------------------------------

def concat(meshes):
    """One HostMesh holding several independent meshes of the same n_dim and row size (a batch of cases), with
    `case_id[c]` the case index of every element of class c. Element order, slots, connections, normals, hanging
    faces and boundary conditions of each mesh follow those of the previous ones; nothing connects two cases."""
    m0 = meshes[0]
    nd, rs = m0.n_dim, m0.row_size
    assert all(m.n_dim == nd and m.row_size == rs for m in meshes), "a batch shares one n_dim and row size"
    n_elem = [sum(m.n_elem[c] for m in meshes) for c in range(2)]
    n_slot = sum(m.faces.shape[0] for m in meshes)
    n_nslot = sum(m.normals.shape[0] for m in meshes)
    n_con = [sum(len(m.con_dir[c]) for m in meshes) for c in range(2)]
    n_ref = sum(m.ref_slot.shape[0] for m in meshes)
    with_av = any(m.av[c] is not None for m in meshes for c in range(2))
    out = HostMesh(nd, rs, n_elem[0], n_elem[1], n_slot, n_nslot, n_con[0], n_con[1], n_ref, with_av=with_av)
    out.case_id = [np.zeros(n, np.int32) for n in n_elem]
    out.n_cases = len(meshes)

    def shifted(a, off):
        a = a.copy()
        a[a >= 0] += off
        return a
    e_off, c_off = [0, 0], [0, 0]
    s_off = ns_off = r_off = b_off = v_off = 0
    bc_con, bc_serial, qpos, hanging = [], [], [[], []], []
    for k, m in enumerate(meshes):
        for c in range(2):
            n, lo = m.n_elem[c], e_off[c]
            sl = slice(lo, lo + n)
            out.state[c][sl] = m.state[c]
            out.rcache[c][sl] = m.rcache[c]
            out.tss[c][sl] = m.tss[c]
            if m.av[c] is not None:
                out.av[c][sl] = m.av[c]
            out.nom_size[c][sl] = m.nom_size[c]
            out.vertex_tss[c][sl] = m.vertex_tss[c]
            out.vertex_ids[c][sl] = m.vertex_ids[c] + v_off
            out.face_slot[c][sl] = m.face_slot[c] + s_off
            out.case_id[c][sl] = k
            if n:
                qpos[c].append(m.qpos[c])
            nc = len(m.con_dir[c])
            out.con_slot[c][c_off[c]:c_off[c] + nc] = m.con_slot[c] + s_off
            out.con_dir[c][c_off[c]:c_off[c] + nc] = m.con_dir[c]
            if c == 1:
                out.con_nslot[c_off[1]:c_off[1] + nc] = shifted(m.con_nslot, ns_off)
                out.metrics[sl] = m.metrics
                out.def_vertex_pos[sl] = m.def_vertex_pos
                out.face_nslot[sl] = shifted(m.face_nslot, ns_off)
        out.faces[s_off:s_off + m.faces.shape[0]] = m.faces
        out.normals[ns_off:ns_off + m.normals.shape[0]] = m.normals
        nr = m.ref_slot.shape[0]
        out.ref_slot[r_off:r_off + nr] = shifted(m.ref_slot, s_off)
        out.ref_stretch[r_off:r_off + nr] = m.ref_stretch
        bc_con.append(m.bc_con + c_off[1])
        bc_serial.append(m.bc_serial + b_off)
        out.bcs += list(m.bcs)
        h = m.hanging.copy()
        for f in range(4):
            h[:, 5 + 2 * f] += np.where(h[:, 4 + 2 * f] == 1, e_off[1], e_off[0])
        hanging.append(h)
        v_off += m.n_vertex
        for c in range(2):
            e_off[c] += m.n_elem[c]
            c_off[c] += len(m.con_dir[c])
        s_off += m.faces.shape[0]
        ns_off += m.normals.shape[0]
        r_off += nr
        b_off += len(m.bcs)
    out.bc_con = np.concatenate(bc_con).astype(np.int32)
    out.bc_serial = np.concatenate(bc_serial).astype(np.int32)
    out.qpos = [np.concatenate(q) if q else None for q in qpos]
    out.n_vertex = v_off
    out.hanging = np.concatenate(hanging).astype(np.int32)
    return out


# ---- general builder (arbitrary connection directions) -------------------------------------------------------

def face_vertex_inds(nd, i_dim, face_sign):
    """Permutation of the face vertices of side 1 so they align with side 0 (connection.cpp:6-27)."""
    n_vert = 2 ** (nd - 1)
    inds = list(range(n_vert))
    fn = [face_sign[0] == False, face_sign[1] == True]  # noqa: E712  flip_normal(i_side)
    flip_t = (i_dim[0] != i_dim[1]) and (fn[0] == fn[1])
    transpose = (i_dim[0] == 0 and i_dim[1] == 2) or (i_dim[0] == 2 and i_dim[1] == 0)
    if flip_t:
        stride = 1
        if nd == 3:
            unused = 3 - i_dim[0] - i_dim[1]
            if unused > i_dim[0]:
                stride = 2
        inds = [v + stride * (1 - 2 * ((v // stride) % 2)) for v in inds]
    if transpose:
        inds[1], inds[2] = inds[2], inds[1]
    return inds


def vertex_inds(nd, i_dim, face_sign):
    """Vertices of both elements on the connected faces, ordered so that entries correspond (connection.cpp:29-50)."""
    n_vert = 2 ** (nd - 1)
    inds = [[], []]
    for s in range(2):
        stride = 2 ** (nd - i_dim[s] - 1)
        inds[s] = [v for v in range(2 * n_vert) if (v // stride) % 2 == int(face_sign[s])]
    perm = face_vertex_inds(nd, i_dim, face_sign)
    inds[1] = [inds[1][perm[k]] for k in range(n_vert)]
    return inds


class Builder:
    """Small meshes through the reference's mesh-building vocabulary (Mesh.hpp:24-230): elements by refinement level
    and nominal position, Cartesian / deformed / boundary connections. Vertices merge as Vertex::eat does
    (position = mean of the merged nominal positions); geometry then follows the box builder's formulas."""

    def __init__(self, nd, rs, root_size=1.0):
        self.nd, self.rs, self.root = nd, rs, root_size
        self.elems = []      # dict(deformed, level, pos)
        self.car_cons = []   # (a, b, i_dim)
        self.def_cons = []   # (a, b, i_dims, face_signs)
        self.bnd_cons = []   # (elem, i_dim, sign, bc)
        self.ref_cons = []   # (coarse, [fine...], i_dim, reverse)
        self.bcs = []

    def add_element(self, ref_level, deformed, pos):
        self.elems.append(dict(deformed=bool(deformed), level=int(ref_level), pos=list(pos) + [0] * (self.nd - len(pos))))
        return len(self.elems) - 1

    def add_bc(self, kind, state, **params):
        """params: the hx_bc fields a condition takes (pressure, thermal, thermal_value, emissivity,
        heat_transfer_coef, temperature, coercion)."""
        self.bcs.append((kind, list(state), dict(params)) if params else (kind, list(state)))
        return len(self.bcs) - 1

    def connect_cartesian(self, a, b, i_dim):
        """Cartesian connection: face +i_dim of a to face -i_dim of b (connection.hpp:29-36)."""
        self.car_cons.append((a, b, i_dim))

    def connect_deformed(self, a, b, i_dims, face_signs):
        if i_dims[0] == i_dims[1] and face_signs[0] == face_signs[1]:
            raise RuntimeError("attempt to connect faces of same sign along same dimension which is forbidden")
        self.def_cons.append((a, b, tuple(i_dims), tuple(bool(s) for s in face_signs)))

    def connect_boundary(self, elem, i_dim, sign, bc):
        self.bnd_cons.append((elem, i_dim, int(sign), bc))

    def connect_refined(self, coarse, fine, i_dim, reverse=False):
        """Cartesian hanging-face connection between a coarse element and the 2^(nd-1) elements one level finer
        (Refined_connection, connection.hpp:168-286). `fine` lists the fine elements in row-major order over the
        tangential dimensions. Not reversed: the coarse element's +i_dim face meets the fine elements' -i_dim faces;
        reversed: the fine elements come first (their +i_dim faces). Each fine element gets a Cartesian connection to
        a mortar face; Prolong_refined / Restrict_refined move data between the coarse face and the mortar faces."""
        if len(fine) != 2 ** (self.nd - 1):
            raise RuntimeError("wrong number of elements in `Refined_connection`")
        for e in [coarse] + list(fine):
            if self.elems[e]["deformed"]:
                raise RuntimeError("this builder makes Cartesian refined connections only")
        self.ref_cons.append((coarse, list(fine), i_dim, bool(reverse)))

    def build(self, with_av=False, jitter=None):
        nd, rs = self.nd, self.rs
        nvert = 2 ** nd
        cls_of = [1 if e["deformed"] else 0 for e in self.elems]
        idx_of, counts = [], [0, 0]
        for c in cls_of:
            idx_of.append(counts[c])
            counts[c] += 1
        n_int_def = len(self.def_cons)
        n_bnd = len(self.bnd_cons)
        n_fine = sum(len(f) for _, f, _, _ in self.ref_cons)
        n_car_con = len(self.car_cons) + n_fine
        n_ref = len(self.ref_cons)
        m = HostMesh(nd, rs, counts[0], counts[1], 2 * (n_car_con + n_int_def + n_bnd) + n_ref,
                     2 * (n_int_def + n_bnd), n_car_con, n_int_def + n_bnd, n_ref=n_ref, with_av=with_av)
        # nominal vertex positions and union-find merge
        parent = list(range(len(self.elems) * nvert))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        def union(x, y):
            rx, ry = find(x), find(y)
            if rx != ry:
                parent[ry] = rx
        nom = np.zeros((len(self.elems), nvert, nd))
        for i, e in enumerate(self.elems):
            h = self.root / 2 ** e["level"]
            for v in range(nvert):
                for d in range(nd):
                    nom[i, v, d] = (e["pos"][d] + ((v >> (nd - 1 - d)) & 1)) * h
        for a, b, d in self.car_cons:
            ia, ib = vertex_inds(nd, (d, d), (True, False))
            for va, vb in zip(ia, ib):
                union(a * nvert + va, b * nvert + vb)
        for a, b, dims, signs in self.def_cons:
            ia, ib = vertex_inds(nd, dims, signs)
            for va, vb in zip(ia, ib):
                union(a * nvert + va, b * nvert + vb)
        for c, fine, d, rev in self.ref_cons:
            # the coarse face's corners eat the matching corner of each fine element (connection.hpp:241-247)
            vi = vertex_inds(nd, (d, d), (True, False))
            for i, f in enumerate(fine):
                union(c * nvert + vi[int(rev)][i], f * nvert + vi[int(not rev)][i])
        roots = np.array([find(x) for x in range(len(parent))])
        uniq, vid = np.unique(roots, return_inverse=True)
        vid = vid.reshape(len(self.elems), nvert)
        sums = np.zeros((len(uniq), nd))
        cnt = np.zeros(len(uniq))
        np.add.at(sums, vid.reshape(-1), nom.reshape(-1, nd))
        np.add.at(cnt, vid.reshape(-1), 1)
        vpos = sums / cnt[:, None]
        if jitter is not None:
            vpos = jitter(vpos)
        # element data
        for i, e in enumerate(self.elems):
            m.nom_size[cls_of[i]][idx_of[i]] = self.root / 2 ** e["level"]
        def_ids = [i for i in range(len(self.elems)) if cls_of[i] == 1]
        car_ids = [i for i in range(len(self.elems)) if cls_of[i] == 0]
        fn = None
        if def_ids:
            m.def_vertex_pos[:] = vpos[vid[def_ids]]
            metrics, fn, vtss_def, qpos = deformed_metrics(nd, rs, vpos[vid[def_ids]], m.nom_size[1])
            m.metrics[:] = metrics
            m.qpos[1] = qpos
        if car_ids:
            nodes = node_coords(nd, rs)
            pos = np.array([self.elems[i]["pos"] for i in car_ids], dtype=float)
            h = m.nom_size[0]
            m.qpos[0] = ((nodes[None] + pos[:, None, :]) * h[:, None, None]).transpose(0, 2, 1).copy()
        # slots
        slot = 0
        for k, (a, b, d) in enumerate(self.car_cons):
            m.con_slot[0][k] = (slot, slot + 1)
            m.con_dir[0][k] = pack_dir(d, d, 1, 0)
            m.face_slot[cls_of[a]][idx_of[a], 2 * d + 1] = slot
            m.face_slot[cls_of[b]][idx_of[b], 2 * d] = slot + 1
            slot += 2
        j = len(self.car_cons)
        for k, (c, fine, d, rev) in enumerate(self.ref_cons):
            m.ref_slot[k, 0] = slot
            m.face_slot[cls_of[c]][idx_of[c], 2 * d + (0 if rev else 1)] = slot
            slot += 1
            for i, f in enumerate(fine):
                # fine connection i: the mortar face on the coarse side, the fine element's face on the other
                mortar, fface = (slot + 1, slot) if rev else (slot, slot + 1)
                m.con_slot[0][j] = (slot, slot + 1)
                m.con_dir[0][j] = pack_dir(d, d, 1, 0)
                m.face_slot[cls_of[f]][idx_of[f], 2 * d + (1 if rev else 0)] = fface
                m.ref_slot[k, 1 + i] = mortar
                slot += 2
                j += 1
        nrm = np.zeros((n_int_def + n_bnd, 2, nd, m.nfq))
        for k, (a, b, dims, signs) in enumerate(self.def_cons):
            if cls_of[a] != 1 or cls_of[b] != 1:
                raise RuntimeError("deformed connections join deformed elements")
            packed = pack_dir(dims[0], dims[1], signs[0], signs[1])
            m.con_slot[1][k] = (slot, slot + 1)
            m.con_dir[1][k] = packed
            fa, fb = 2 * dims[0] + signs[0], 2 * dims[1] + signs[1]
            m.face_slot[1][idx_of[a], fa] = slot
            m.face_slot[1][idx_of[b], fb] = slot + 1
            m.con_nslot[k] = (2 * k, 2 * k + 1)
            m.face_nslot[idx_of[a], fa] = 2 * k
            m.face_nslot[idx_of[b], fb] = 2 * k + 1
            # shared normal (Solver.cpp:327-352)
            perm = face_permutation_index(nd, rs, packed)
            _, face_sign = unpack_dir(packed)
            sg = [1 - 2 * flip_normal(face_sign, s) for s in range(2)]
            n0 = fn[idx_of[a], fa]
            n1 = fn[idx_of[b], fb][:, perm]
            avg = 0.5 * sg[0] * n0 + 0.5 * sg[1] * n1
            nrm[k, 0] = sg[0] * avg
            back = np.zeros_like(avg)
            back[:, perm] = sg[1] * avg
            nrm[k, 1] = back
            slot += 2
        m.bc_con = np.arange(n_int_def, n_int_def + n_bnd, dtype=np.int32)
        m.bc_serial = np.zeros(n_bnd, np.int32)
        for j, (e, d, s, bc) in enumerate(self.bnd_cons):
            k = n_int_def + j
            m.con_slot[1][k] = (slot, slot + 1)
            m.con_dir[1][k] = pack_dir(d, d, s, 1 - s)
            m.con_nslot[k] = (2 * k, 2 * k + 1)
            m.face_slot[cls_of[e]][idx_of[e], 2 * d + s] = slot
            m.bc_serial[j] = bc
            if cls_of[e] == 1:
                m.face_nslot[idx_of[e], 2 * d + s] = 2 * k
                nin = fn[idx_of[e], 2 * d + s]
                sg = -1.0 if s == 0 else 1.0
                avg = 0.5 * sg * nin + 0.5 * sg * nin
                nrm[k, 0] = sg * avg
                nrm[k, 1] = sg * avg
            else:
                nrm[k, :, d, :] = 1.0
            slot += 2
        m.normals[:] = nrm.reshape(-1, nd, m.nfq)
        m.bcs = list(self.bcs)
        # vertex time step scale, shared by minimum
        per_elem = np.zeros((len(self.elems), nvert))
        for j, i in enumerate(car_ids):
            per_elem[i] = m.nom_size[0][j] / nd
        for j, i in enumerate(def_ids):
            per_elem[i] = vtss_def[j]
        vmin = np.full(len(uniq), np.inf)
        np.minimum.at(vmin, vid.reshape(-1), per_elem.reshape(-1))
        for j, i in enumerate(car_ids):
            m.vertex_tss[0][j] = vmin[vid[i]]
        for j, i in enumerate(def_ids):
            m.vertex_tss[1][j] = vmin[vid[i]]
        self.vertex_positions = vpos
        self.vertex_ids = vid
        m.n_vertex = len(uniq)
        for j, i in enumerate(car_ids):
            m.vertex_ids[0][j] = vid[i]
        for j, i in enumerate(def_ids):
            m.vertex_ids[1][j] = vid[i]
        # one hanging-vertex matcher per hanging face: the fine elements' face is on the side face_sign[!rev] of the
        # Cartesian direction {{d, d}, {1, 0}} (connection.hpp:35, 231)
        hang = np.zeros((len(self.ref_cons), 12), np.int32)
        for k, (c, fine, d, rev) in enumerate(self.ref_cons):
            hang[k, 0], hang[k, 1] = d, int(bool(rev))
            for f, e in enumerate(fine):
                hang[k, 4 + 2 * f], hang[k, 5 + 2 * f] = cls_of[e], idx_of[e]
        m.hanging = hang
        return m
