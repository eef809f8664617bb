"""Gauss-Legendre basis tables for row sizes 2..8, generated from closed forms.

The reference generates these tables at build time (script/auto_generate.py:61-280 driving
script/Basis.py:10-204) and emits each entry as a 20-significant-digit decimal literal which the C++
compiler rounds to double. This module computes the same quantities from their definitions with mpmath
at 50 digits, rounds each to 20 significant digits exactly as that emission does, and then to double.

Tables (all on the reference interval [0, 1]):
  node[i]            Gauss-Legendre node                                   (Basis.py:58-62)
  weight[i]          quadrature weight, sums to 1                          (Basis.py:64-68)
  diff_mat[i][j]     d(ell_j)/dx at node i                                 (Basis.py:70-87)
  boundary[s][j]     ell_j(s), s in {0, 1}                                 (Basis.py:89-102)
  orthogonal[k][i]   Legendre P_k(2x-1) at node i, unit norm under weights (Basis.py:104-123)
  filter[i][j]       sum_k phi_k(i) 0.5^k phi_k(j) w_j                     (Basis.py:152-156)
  prolong[h][i][j]   ell_j((node_i + h)/2)                                 (Basis.py:129-139)
  restrict[h][i][j]  w_j/2 * ell_i((node_j + h)/2) / w_i                   (Basis.py:141-150)
  eigenvals[2]       min real part of the eigenvalues of the 1D periodic advection and diffusion
                     operators on 16 elements                              (Basis.py:158-204)

Run as a script to regenerate hexed_amd/csrc/basis_tables.inc and hexed_amd/basis_tables.json.
"""
import json
import os

import mpmath
import numpy as np
import sympy

CALC_DPS = 50
REPR_DIGITS = 20
MIN_ROW_SIZE = 2
MAX_ROW_SIZE = 8
N_ELEM_STABILITY = 16


def _to_double(x):
    """Round an mpf to 20 significant digits the way the reference's generator prints it, then to double."""
    text = str(sympy.Float(mpmath.nstr(x, CALC_DPS + 5, strip_zeros=False), CALC_DPS))
    return float(str(sympy.Float(text, REPR_DIGITS)))


def _legendre(n, x):
    """(P_n(x), P_n'(x)) by the three-term recurrence."""
    p0, p1 = mpmath.mpf(1), x
    if n == 0:
        return p0, mpmath.mpf(0)
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def gauss_legendre_nodes(n):
    """Nodes and weights on [0, 1] (weights sum to 1), ascending, by Newton iteration on P_n."""
    with mpmath.workdps(CALC_DPS + 20):
        xs, ws = [], []
        for i in range(n):
            x = mpmath.cos(mpmath.pi * (i + mpmath.mpf(3) / 4) / (n + mpmath.mpf(1) / 2))
            for _ in range(200):
                p, dp = _legendre(n, x)
                dx = p / dp
                x -= dx
                if abs(dx) < mpmath.mpf(10) ** (-(CALC_DPS + 15)):
                    break
            p, dp = _legendre(n, x)
            xs.append(x)
            ws.append(2 / ((1 - x * x) * dp * dp))
        order = sorted(range(n), key=lambda i: xs[i])
        nodes = [(xs[i] + 1) / 2 for i in order]
        weights = [ws[i] / 2 for i in order]
    return nodes, weights


def gauss_lobatto_nodes(n):
    """Gauss-Lobatto nodes on [0, 1], ascending: endpoints plus the roots of P'_{n-1}."""
    with mpmath.workdps(CALC_DPS + 20):
        m = n - 1
        inner = []
        for i in range(1, m):
            x = mpmath.cos(mpmath.pi * i / m)
            for _ in range(200):
                # roots of P'_m: Newton on P'_m using P''_m = (2x P'_m - m(m+1) P_m)/(1-x^2)
                p, dp = _legendre(m, x)
                d2p = (2 * x * dp - m * (m + 1) * p) / (1 - x * x)
                dx = dp / d2p
                x -= dx
                if abs(dx) < mpmath.mpf(10) ** (-(CALC_DPS + 15)):
                    break
            inner.append(x)
        xs = sorted([mpmath.mpf(-1), mpmath.mpf(1)] + inner)
        return [(x + 1) / 2 for x in xs]


def _lagrange(nodes, j, pos):
    r = mpmath.mpf(1)
    for n, xn in enumerate(nodes):
        if n != j:
            r *= (pos - xn) / (nodes[j] - xn)
    return r


def _lagrange_deriv(nodes, i, j):
    """Derivative of the j-th Lagrange polynomial at node i (Basis.py:70-87)."""
    rs = len(nodes)
    if i == j:
        return sum(1 / (nodes[i] - nodes[n]) for n in range(rs) if n != i)
    dp = mpmath.mpf(1)
    for n in range(rs):
        if n != j:
            dp /= nodes[j] - nodes[n]
            if n != i:
                dp *= nodes[i] - nodes[n]
    return dp


def basis_exact(rs):
    """All tables for row size `rs` as 50-digit mpf values."""
    with mpmath.workdps(CALC_DPS):
        nodes, weights = gauss_legendre_nodes(rs)
        nodes = [+x for x in nodes]
        weights = [+w for w in weights]
        diff = [[_lagrange_deriv(nodes, i, j) for j in range(rs)] for i in range(rs)]
        bound = [[_lagrange(nodes, j, mpmath.mpf(s)) for j in range(rs)] for s in (0, 1)]
        ortho = []
        for k in range(rs):
            vals = [mpmath.legendre(k, 2 * x - 1) for x in nodes]
            norm = mpmath.sqrt(sum(v * v * w for v, w in zip(vals, weights)))
            ortho.append([v / norm for v in vals])
        filt = [[sum(ortho[k][i] * mpmath.mpf(0.5) ** k * ortho[k][j] * weights[j] for k in range(rs))
                 for j in range(rs)] for i in range(rs)]
        prolong = [[[_lagrange(nodes, j, (nodes[i] + h) / 2) for j in range(rs)] for i in range(rs)]
                   for h in (0, 1)]
        restrict = [[[weights[j] / 2 * _lagrange(nodes, i, (nodes[j] + h) / 2) / weights[i]
                      for j in range(rs)] for i in range(rs)] for h in (0, 1)]
    return dict(node=nodes, weight=weights, diff_mat=diff, boundary=bound, orthogonal=ortho,
                filter=filt, prolong=prolong, restrict=restrict)


def _round_tree(t):
    if isinstance(t, dict):
        return {k: _round_tree(v) for k, v in t.items()}
    if isinstance(t, list):
        return [_round_tree(x) for x in t]
    return _to_double(t)


def stability_operators(rs, n_elem=N_ELEM_STABILITY):
    """Discrete -d/dx and d^2/dx^2 on a periodic 1D mesh of `n_elem` unit elements (Basis.py:158-196).

    Entries are formed the way the reference forms them: the 20-digit table values (derivative,
    boundary interpolant) combined with full-precision weights, accumulated in float64.
    """
    with mpmath.workdps(CALC_DPS):
        nodes, weights = gauss_legendre_nodes(rs)
        nodes = [+x for x in nodes]
        weights = [+w for w in weights]
        d20 = [[_to_double(_lagrange_deriv(nodes, i, j)) for j in range(rs)] for i in range(rs)]
        # interpolate() returns 20-digit values; dividing by the 50-digit weight happens at high precision
        b20 = [[mpmath.mpf(str(sympy.Float(mpmath.nstr(_lagrange(nodes, j, mpmath.mpf(s)), CALC_DPS + 5), REPR_DIGITS)))
                for j in range(rs)] for s in (0, 1)]
        n = n_elem * rs
        local_grad = np.zeros((n, n))
        avrg = np.zeros((n, n))
        jump = np.zeros((n, n))
        for e in range(n_elem):
            for i in range(rs):
                for j in range(rs):
                    local_grad[e * rs + i, e * rs + j] = d20[i][j]
                    for si in (0, 1):
                        for sj in (0, 1):
                            row = ((e + si) % n_elem) * rs + i
                            col = ((e + sj) % n_elem) * rs + j
                            # the reference adds a high-precision term to a float64 entry and rounds the sum once
                            bnd = b20[1 - sj][j] * b20[1 - si][i] / weights[i]
                            avrg[row, col] = float(mpmath.mpf(avrg[row, col]) + mpmath.mpf(0.5 * (2 * sj - 1)) * bnd)
                            jump[row, col] = float(mpmath.mpf(jump[row, col]) + mpmath.mpf(0.5 * (1 - 2 * (sj == si))) * bnd)
    advection = -local_grad + -avrg + jump
    diffusion = (local_grad + avrg) @ (local_grad + avrg)
    return advection, diffusion


def eigenvals(rs):
    return [float(np.linalg.eigvals(m).real.min()) for m in stability_operators(rs)]


def generate():
    tables = {}
    for rs in range(MIN_ROW_SIZE, MAX_ROW_SIZE + 1):
        t = _round_tree(basis_exact(rs))
        t["eigenvals"] = eigenvals(rs)
        tables[rs] = t
    return tables


def _flat(t):
    if isinstance(t, list):
        out = []
        for x in t:
            out.extend(_flat(x))
        return out
    return [t]


def write_outputs(tables, here=os.path.dirname(os.path.abspath(__file__))):
    js = {str(rs): {k: [float.hex(v) for v in _flat(t[k])] for k in t} for rs, t in tables.items()}
    with open(os.path.join(here, "basis_tables.json"), "w") as f:
        json.dump(js, f, indent=1)
        f.write("\n")
    lines = ["// Gauss-Legendre basis tables, row sizes 2..8; written by hexed_amd/basis_gen.py -- do not edit.",
             "// Row-major: diff_mat[i][j], boundary[s][j], orthogonal[k][i], filter[i][j],",
             "// prolong[h][i][j], restrict[h][i][j]; eigenvals = {min_eig_convection, min_eig_diffusion}.",
             "#pragma once", "namespace hx_basis_tables {"]
    keys = ["node", "weight", "diff_mat", "boundary", "orthogonal", "filter", "prolong", "restrict", "eigenvals"]
    for rs, t in tables.items():
        for k in keys:
            vals = _flat(t[k])
            body = ", ".join(float.hex(v) for v in vals)
            lines.append(f"static const double {k}{rs}[{len(vals)}] = {{{body}}};")
    for k in keys:
        refs = ", ".join(f"{k}{rs}" for rs in tables)
        lines.append(f"static const double* const {k}s[{len(tables)}] = {{{refs}}};")
    lines.append("}  // namespace hx_basis_tables")
    with open(os.path.join(here, "csrc", "basis_tables.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")


def load_tables(path=None):
    """Read the committed JSON tables into numpy arrays keyed by row size."""
    if path is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_tables.json")
    with open(path) as f:
        js = json.load(f)
    out = {}
    for rs_s, t in js.items():
        rs = int(rs_s)
        a = {k: np.array([float.fromhex(v) for v in vals]) for k, vals in t.items()}
        a["diff_mat"] = a["diff_mat"].reshape(rs, rs)
        a["boundary"] = a["boundary"].reshape(2, rs)
        a["orthogonal"] = a["orthogonal"].reshape(rs, rs)
        a["filter"] = a["filter"].reshape(rs, rs)
        a["prolong"] = a["prolong"].reshape(2, rs, rs)
        a["restrict"] = a["restrict"].reshape(2, rs, rs)
        out[rs] = a
    return out


if __name__ == "__main__":
    write_outputs(generate())
