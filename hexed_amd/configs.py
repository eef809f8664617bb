"""Workload generators for the benchmark configurations (synthetic, seeded).

Random numbers come from splitmix64 (Steele, Lea & Flood 2014) mapped to [0, 1) by the top 53 bits, so a
configuration is the same on every platform and library version.

  vortex    2D Euler isentropic vortex, 16x16 Cartesian quads on [-5, 5]^2, row size 6
  euler32   3D Euler, 32^3 Cartesian hexes on [0, 1]^3, row size 6, Mach 0.8 + 1% Fourier perturbations
  ns48      3D Navier-Stokes, 48^3 deformed hexes (interior vertices jittered +-10% of spacing), row size 6,
            Re 1e4 on unit length, Mach 0.3, Sutherland air, constant artificial viscosity
  amr16     3D Euler on a seeded 2:1-balanced tree with hanging faces (16^3 base), row size 6 or 4
  batch64   a batch of 64 independent 3D Euler boxes of log-uniform size, row size 4-6, Mach 0.3-0.9
"""
import math

import numpy as np

from . import abi, mesh

GAMMA = 1.4
R_AIR = 287.05287  # constants.hpp:44


def splitmix64(seed, n):
    """n uniform doubles in [0, 1) from the splitmix64 sequence started at `seed`."""
    mask = (1 << 64) - 1
    x = seed & mask
    out = np.empty(n)
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def splitmix64_np(seed, n):
    """Vectorised splitmix64: the i-th output uses state seed + (i+1)*golden (identical to `splitmix64`)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & ((1 << 64) - 1)) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def freestream_state(nd, rho, vel, p):
    ke = 0.5 * rho * sum(v * v for v in vel)
    return [rho * v for v in vel] + [rho, p / (GAMMA - 1) + ke]


def fourier_perturbation(pos, seed, n_modes=8, amplitude=0.01):
    """Sum over axes and modes 1..n_modes of seeded amplitude/phase sinusoids on the unit box, scaled so the
    sum of |amplitudes| is `amplitude` (the perturbation never exceeds it)."""
    nd = len(pos)
    r = splitmix64_np(seed, 2 * nd * n_modes).reshape(nd, n_modes, 2)
    amp = r[:, :, 0] + 0.5
    amp *= amplitude / amp.sum()
    phase = 2 * math.pi * r[:, :, 1]
    cp, sp = np.cos(phase), np.sin(phase)
    flat = [np.ascontiguousarray(p).reshape(-1) for p in pos]
    out = np.zeros(flat[0].size)
    chunk = 1 << 16  # keep the temporaries cache resident
    for lo in range(0, out.size, chunk):
        acc = out[lo:lo + chunk]
        for d in range(nd):
            # sin(k t + p) = sin(k t) cos p + cos(k t) sin p, sin/cos(k t) from the angle-addition recurrence
            t = 2 * math.pi * flat[d][lo:lo + chunk]
            s1, c1 = np.sin(t), np.cos(t)
            sk, ck = s1, c1
            for k in range(n_modes):
                acc += amp[d, k] * (sk * cp[d, k] + ck * sp[d, k])
                sk, ck = sk * c1 + ck * s1, ck * c1 - sk * s1
    return out.reshape(np.shape(pos[0]))


def euler_box(n=32, rs=6, mach=0.8, seed=0, nd=3, deformed=False, jitter=0.0):
    """Config 2: freestream rho=1, p=1/gamma (sound speed 1), velocity Mach*x, with 1% seeded Fourier
    perturbations of density and each velocity component; freestream boundary conditions."""
    fs = freestream_state(nd, 1.0, [mach] + [0.0] * (nd - 1), 1.0 / GAMMA)
    m = mesh.box(nd, rs, n, 0.0, 1.0, deformed=deformed, jitter=jitter, seed=seed,
                 bc=(abi.HX_BC_FREESTREAM, fs))

    def init(pos):
        rho = 1.0 * (1 + fourier_perturbation(pos, seed * 16 + 1))
        vel = [(mach if d == 0 else 0.0) + mach * fourier_perturbation(pos, seed * 16 + 2 + d) for d in range(nd)]
        p = 1.0 / GAMMA
        return [rho * v for v in vel] + [rho, p / (GAMMA - 1) + 0.5 * rho * sum(v * v for v in vel)]
    mesh.set_state(m, init)
    return m


def sutherland_air():
    """Sutherland viscosity and conductivity of air (Case.hil:83-90)."""
    return abi.sutherland(1.716e-5, 273., 111.), abi.sutherland(0.0241, 273., 194.)


def ns_freestream(mach=0.3, reynolds=1e4, length=1.0, temperature=288.15):
    """Dimensional freestream with the requested Mach and Reynolds numbers: returns (rho, U, p, mu)."""
    c = math.sqrt(GAMMA * R_AIR * temperature)
    u = mach * c
    mu = 1.716e-5 * (temperature / 273.) ** 1.5 * (273. + 111.) / (temperature + 111.)
    rho = reynolds * mu / (u * length)
    p = rho * R_AIR * temperature
    return rho, u, p, mu


def ns_deformed(n=48, rs=6, mach=0.3, reynolds=1e4, jitter=0.1, seed=0, art_visc=1e-3, nd=3):
    """Config 3: deformed box with interior vertices jittered by U(+-jitter*h), Navier-Stokes freestream with
    1% perturbations, freestream boundaries, constant bulk artificial viscosity `art_visc` [m^2/s]
    (Solver::set_art_visc_constant, Solver.cpp:415-426)."""
    rho, u, p, _ = ns_freestream(mach, reynolds)
    fs = freestream_state(nd, rho, [u] + [0.0] * (nd - 1), p)
    m = mesh.box(nd, rs, n, 0.0, 1.0, deformed=True, jitter=jitter, seed=seed, bc=(abi.HX_BC_FREESTREAM, fs),
                 with_av=True)

    def init(pos):
        r = rho * (1 + fourier_perturbation(pos, seed * 16 + 1))
        vel = [(u if d == 0 else 0.0) + u * fourier_perturbation(pos, seed * 16 + 2 + d) for d in range(nd)]
        return [r * v for v in vel] + [r, p / (GAMMA - 1) + 0.5 * r * sum(v * v for v in vel)]
    mesh.set_state(m, init)
    for c in range(2):
        if m.n_elem[c]:
            m.av[c][:, 0, :] = art_visc
    return m


def isentropic_vortex(n=16, rs=6, mach=0.5, strength=5.0, half_width=5.0):
    """Config 1 (2D): Cartesian quads on [-w, w]^2 with freestream boundaries. Shu's isentropic vortex with
    rho_inf = p_inf = 1 (sound speed sqrt(gamma)), freestream speed mach*sqrt(gamma) along x:
    T = p/rho = 1 - (gamma-1) beta^2/(8 gamma pi^2) e^(1-r^2), rho = T^(1/(gamma-1)), perturbation velocity
    beta/(2 pi) e^((1-r^2)/2) (-y, x). The exact solution is the initial vortex translated with the freestream.
    Returns (mesh, exact(pos, t) -> list of conserved fields)."""
    u_inf = mach * math.sqrt(GAMMA)
    fs = freestream_state(2, 1.0, [u_inf, 0.0], 1.0)
    m = mesh.box(2, rs, n, -half_width, half_width, bc=(abi.HX_BC_FREESTREAM, fs))

    def exact(pos, t):
        x = pos[0] - u_inf * t
        y = pos[1]
        r2 = x * x + y * y
        du = strength / (2 * math.pi) * np.exp(0.5 * (1 - r2))
        temp = 1 - (GAMMA - 1) * strength ** 2 / (8 * GAMMA * math.pi ** 2) * np.exp(1 - r2)
        rho = temp ** (1 / (GAMMA - 1))
        pr = rho * temp
        vx = u_inf - du * y
        vy = du * x
        return [rho * vx, rho * vy, rho, pr / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)]
    mesh.set_state(m, lambda pos: exact(pos, 0.0))
    return m, exact


def tree_leaves(base=16, levels=3, frac=0.085, seed=0, nd=3):
    """Leaves of a seeded, 2:1 face-balanced refinement tree over base^nd level-0 cells: at each level below
    `levels`, every leaf of that level is refined with probability `frac` (a splitmix64 uniform per leaf in a fixed
    order), then leaves are split until no leaf has a face neighbour more than one level finer. With the defaults
    (16^3, 3 levels, 0.085, seed 0) the balancing leaves 31% of the base elements refined: 19,461 leaves over levels
    0-3 (2835, 9355, 5663, 1608) and 7,691 hanging faces. Returns a dict {(level, integer position tuple): True}."""
    leaves = {(0, p): True for p in np.ndindex(*([base] * nd))}

    def split(key):
        lvl, p = key
        del leaves[key]
        for c in np.ndindex(*([2] * nd)):
            leaves[(lvl + 1, tuple(2 * x + y for x, y in zip(p, c)))] = True

    for lvl in range(levels):
        cur = sorted(k for k in leaves if k[0] == lvl)
        u = splitmix64(seed * 1009 + lvl, len(cur)) if cur else []
        for k, r in zip(cur, u):
            if r < frac:
                split(k)
    # 2:1 balance across faces: a leaf at level L needs every face neighbour covered by leaves of level >= L - 1
    changed = True
    while changed:
        changed = False
        for (lvl, p) in sorted(leaves, key=lambda k: -k[0]):
            if (lvl, p) not in leaves or lvl < 2:
                continue
            for d in range(nd):
                for s in (-1, 1):
                    q = list(p)
                    q[d] += s
                    if not 0 <= q[d] < base * 2 ** lvl:
                        continue
                    # the leaf covering the same-level neighbour cell, if it is coarser
                    for up in range(1, lvl + 1):
                        anc = (lvl - up, tuple(x >> up for x in q))
                        if anc in leaves:
                            if up >= 2:
                                split(anc)
                                changed = True
                            break
    return leaves


def tree_mesh(base=16, rs=6, levels=3, frac=0.085, seed=0, nd=3, bc=None):
    """Cartesian tree mesh on the unit cube from `tree_leaves`: same-level neighbours get Cartesian connections,
    a leaf whose face neighbour is refined gets a hanging-face (refined) connection to the 2^(nd-1) finer leaves
    on that face (connection.hpp:168-286), domain faces get boundary connections. Root size 1/base."""
    leaves = tree_leaves(base, levels, frac, seed, nd)
    b = mesh.Builder(nd, rs, root_size=1.0 / base)
    keys = sorted(leaves, key=lambda k: (k[0], k[1]))
    index = {}
    for k in keys:
        index[k] = b.add_element(k[0], False, list(k[1]))
    if bc is None:
        bc = (abi.HX_BC_FREESTREAM, freestream_state(nd, 1.0, [0.8] + [0.0] * (nd - 1), 1.0 / GAMMA))
    bc_id = b.add_bc(*bc)
    for (lvl, p) in keys:
        e = index[(lvl, p)]
        n_cells = base * 2 ** lvl
        for d in range(nd):
            for s in (0, 1):
                q = list(p)
                q[d] += 1 if s else -1
                if not 0 <= q[d] < n_cells:
                    b.connect_boundary(e, d, s, bc_id)
                    continue
                q = tuple(q)
                if (lvl, q) in leaves:
                    if s:
                        b.connect_cartesian(e, index[(lvl, q)], d)
                    continue
                # neighbour refined: its children on the shared face, row-major over the tangential dimensions
                kids = []
                tang = [t for t in range(nd) if t != d]
                for c in np.ndindex(*([2] * (nd - 1))):
                    child = [2 * x for x in q]
                    child[d] += 0 if s else 1
                    for t, ct in zip(tang, c):
                        child[t] += ct
                    kids.append((lvl + 1, tuple(child)))
                if all(k in leaves for k in kids):
                    b.connect_refined(e, [index[k] for k in kids], d, reverse=not s)
                # otherwise the neighbour is coarser and owns the connection
    m = b.build()
    m.tree_counts = {lvl: sum(1 for k in keys if k[0] == lvl) for lvl in range(levels + 1)}
    return m


def euler_tree(base=16, rs=6, levels=3, frac=0.085, seed=0, mach=0.8, nd=3):
    """Config 4: 3D Euler on a seeded 2:1-balanced tree (hanging faces through Prolong/Restrict_refined), freestream
    rho=1, p=1/gamma, velocity Mach*x with 1% seeded Fourier perturbations, freestream boundaries."""
    fs = freestream_state(nd, 1.0, [mach] + [0.0] * (nd - 1), 1.0 / GAMMA)
    m = tree_mesh(base, rs, levels, frac, seed, nd, bc=(abi.HX_BC_FREESTREAM, fs))

    def init(pos):
        rho = 1.0 * (1 + fourier_perturbation(pos, seed * 16 + 1))
        vel = [(mach if d == 0 else 0.0) + mach * fourier_perturbation(pos, seed * 16 + 2 + d) for d in range(nd)]
        p = 1.0 / GAMMA
        return [rho * v for v in vel] + [rho, p / (GAMMA - 1) + 0.5 * rho * sum(v * v for v in vel)]
    mesh.set_state(m, init)
    return m


def batch_cases(n_cases=64, seed=0, min_elem=512, max_elem=8192, row_sizes=(4, 5, 6), mach=(0.3, 0.9)):
    """Config 5 case list: for each case an element count drawn log-uniform in [min_elem, max_elem] and rounded to a
    cube n^3, a row size drawn uniformly from `row_sizes` and a freestream Mach number uniform in `mach`; returns
    dicts {n, rs, mach, seed}."""
    r = splitmix64_np(seed * 7919 + 12345, 3 * n_cases).reshape(n_cases, 3)
    out = []
    for k in range(n_cases):
        count = math.exp(math.log(min_elem) + r[k, 0] * (math.log(max_elem) - math.log(min_elem)))
        out.append(dict(n=max(2, int(round(count ** (1.0 / 3.0)))), rs=row_sizes[min(int(r[k, 1] * len(row_sizes)),
                                                                                   len(row_sizes) - 1)],
                        mach=mach[0] + r[k, 2] * (mach[1] - mach[0]), seed=seed * 1000 + k + 1))
    return out


def euler_batch(cases):
    """Config 5 meshes: one batch mesh per row size (mesh.concat of the cases of that row size, each an
    euler_box of its own size, Mach number and perturbation seed). Returns {rs: (mesh, [case indices])}."""
    groups = {}
    for k, c in enumerate(cases):
        groups.setdefault(c["rs"], []).append(k)
    return {rs: (mesh.concat([euler_box(cases[k]["n"], rs, cases[k]["mach"], cases[k]["seed"]) for k in ks]), ks)
            for rs, ks in sorted(groups.items())}
