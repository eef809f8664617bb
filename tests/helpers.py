"""Shared test utilities: seeded states, oracle runs, comparisons."""
import ctypes
import sys

import numpy as np

from hexed_amd import abi, mesh

import oracle

GAMMA = 1.4


def freestream(nd, mach=0.5, rho=1.0, p=None):
    """Conserved freestream state with sound speed 1 (p = rho/gamma): momentum..., density, energy."""
    p = rho / GAMMA if p is None else p
    u = [mach] + [0.0] * (nd - 1)
    ke = 0.5 * rho * sum(x * x for x in u)
    return [rho * x for x in u] + [rho, p / (GAMMA - 1) + ke]


def smooth_state(m, mach=0.5, amp=0.05, seed=0):
    """Freestream plus seeded sinusoidal perturbations of density, velocity and pressure."""
    rng = np.random.default_rng(seed)
    nd = m.n_dim
    k = rng.uniform(0.5, 3.0, size=(4, nd))
    ph = rng.uniform(0, 2 * np.pi, size=4)

    def f(pos):
        def wave(i):
            arg = ph[i] + sum(k[i, d] * 2 * np.pi * pos[d] for d in range(nd))
            return np.sin(arg)
        rho = 1.0 + amp * wave(0)
        vel = [mach * (1 + amp * wave(1))] + [amp * mach * wave(2 + min(d, 1)) for d in range(1, nd)]
        p = (1.0 / GAMMA) * (1 + amp * wave(3))
        ke = 0.5 * rho * sum(v * v for v in vel)
        return [rho * v for v in vel] + [rho, p / (GAMMA - 1) + ke]
    mesh.set_state(m, f)


def oracle_lib():
    return oracle.load()


def o_call(name, *args):
    lib = oracle.load()
    oracle.check(getattr(lib, name)(*args))


def init_faces(m):
    """Solver::initialize's compute_write_face + compute_prolong (Solver.cpp:400-401), run by the oracle on the host
    mesh."""
    s = m.struct()
    for c in (0, 1):
        o_call("oracle_write_face", ctypes.byref(s), c, None)
    o_call("oracle_prolong_refined", ctypes.byref(s), None, 0, 0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / nb)


def update_folds(m, mode=-1):
    """Whether hx_update folds the face fluxes into the element kernels for mesh m (hx_mesh_set_face_fold `mode`)."""
    return mode == 1 or (mode == -1 and m.row_size >= 2 * m.n_dim)


def faces_rel_l2(got, ref, m, ldg=True):
    """Relative L2 of two face buffers [n_slot][face_sz]. ldg False leaves out the LDG blocks, which a folded
    Navier-Stokes update leaves holding the viscous-flux extrapolations instead of the reconciled jumps (values the
    reference overwrites before it reads them again)."""
    if ldg:
        return rel_l2(got, ref)
    k = m.nv * m.nfq
    keep = np.ones(got.shape[1], bool)
    keep[k:2 * k] = False
    return rel_l2(got[:, keep], ref[:, keep])


def per_var_rel_l2(a, b):
    """Relative L2 per conserved variable of [n][nv][nq] arrays."""
    return [rel_l2(a[:, v], b[:, v]) for v in range(a.shape[1])]


def say(*msg):
    print(*msg, flush=True)
    sys.stdout.flush()


def set_vertices(m):
    """Hand the oracle the vertex connectivity of `m` (shared vertex ids and hanging-vertex matchers), which
    fix_admissibility needs and hx_host_mesh does not carry."""
    s = m.struct()
    vid = [np.ascontiguousarray(m.vertex_ids[c], dtype=np.int32) for c in range(2)]
    hang = np.ascontiguousarray(m.hanging, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    o_call("oracle_set_vertices", ctypes.byref(s), int(m.n_vertex), vid[0].ctypes.data_as(ip),
           vid[1].ctypes.data_as(ip), int(hang.shape[0]), hang.ctypes.data_as(ip))
