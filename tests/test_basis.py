"""Basis tables (hexed_amd/basis_gen.py -> basis_tables.json / csrc/basis_tables.inc) pinned three ways:
the invariants of the reference's test_Basis.cpp:8-154, an independent library routine (numpy's Gauss-Legendre),
and the eigenvalue fixture written by the reference's own generator (tests/golden/reference_eigenvals.json)."""
import json
import os

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from hexed_amd import basis_gen

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = basis_gen.load_tables()
ROW_SIZES = list(range(2, 9))


def poly_vals(coefs, x):
    return sum(c * x ** k for k, c in enumerate(coefs))


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_nodes_weights_match_library_gauss_legendre(rs):
    t = TABLES[rs]
    x, w = npleg.leggauss(rs)
    np.testing.assert_allclose(t["node"], (x + 1) / 2, rtol=0, atol=2e-15)
    np.testing.assert_allclose(t["weight"], w / 2, rtol=0, atol=2e-15)
    assert np.all(np.diff(t["node"]) > 0) and t["node"][0] > 0 and t["node"][-1] < 1
    assert abs(t["weight"].sum() - 1) < 4e-16


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_quadrature_exact_to_degree_2rs_minus_1(rs):
    t = TABLES[rs]
    for k in range(2 * rs):
        assert abs(np.dot(t["weight"], t["node"] ** k) - 1 / (k + 1)) < 1e-15


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_diff_mat(rs):
    """Row sums vanish and the derivative is exact for polynomials of degree < rs (test_Basis.cpp)."""
    t = TABLES[rs]
    D, x = t["diff_mat"], t["node"]
    assert np.abs(D.sum(axis=1)).max() < 1e-12 * rs * rs
    rng = np.random.default_rng(rs)
    c = rng.normal(size=rs)
    dc = [k * c[k] for k in range(1, rs)]
    np.testing.assert_allclose(D @ poly_vals(c, x), poly_vals(dc, x), atol=1e-12 * rs ** 3)


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_boundary_interpolation(rs):
    t = TABLES[rs]
    B, x = t["boundary"], t["node"]
    rng = np.random.default_rng(10 + rs)
    c = rng.normal(size=rs)
    np.testing.assert_allclose(B @ poly_vals(c, x), [poly_vals(c, 0.0), poly_vals(c, 1.0)], atol=1e-13 * rs)
    np.testing.assert_allclose(B.sum(axis=1), [1, 1], atol=1e-14)


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_orthogonal_is_normalized_legendre(rs):
    t = TABLES[rs]
    O, x, w = t["orthogonal"], t["node"], t["weight"]
    np.testing.assert_allclose(O @ np.diag(w) @ O.T, np.eye(rs), atol=1e-14)
    for k in range(rs):
        ref = npleg.legval(2 * x - 1, [0] * k + [1]) * np.sqrt(2 * k + 1)
        np.testing.assert_allclose(O[k], ref, atol=1e-13)


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_filter_scales_modes_by_powers_of_half(rs):
    t = TABLES[rs]
    F, O = t["filter"], t["orthogonal"]
    for k in range(rs):
        np.testing.assert_allclose(F @ O[k], 0.5 ** k * O[k], atol=1e-14)


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_prolong_restrict(rs):
    """Prolongation interpolates exactly; restriction is its L2 pseudo-inverse: R0 P0 + R1 P1 = I (test_Basis.cpp:94-119)."""
    t = TABLES[rs]
    x = t["node"]
    rng = np.random.default_rng(20 + rs)
    c = rng.normal(size=rs)
    for h in (0, 1):
        np.testing.assert_allclose(t["prolong"][h] @ poly_vals(c, x), poly_vals(c, (x + h) / 2), atol=1e-13)
    RP = t["restrict"][0] @ t["prolong"][0] + t["restrict"][1] @ t["prolong"][1]
    np.testing.assert_allclose(RP, np.eye(rs), atol=1e-13)


def test_eigenvalues_match_reference_generator():
    """min_eig_convection / min_eig_diffusion equal the values the reference's script/Basis.py:197-204 produces."""
    with open(os.path.join(HERE, "golden", "reference_eigenvals.json")) as f:
        golden = json.load(f)
    for rs in ROW_SIZES:
        ref = [float.fromhex(h) for h in golden[str(rs)]["as_double_hex"]]
        mine = list(TABLES[rs]["eigenvals"])
        for a, b in zip(mine, ref):
            assert abs(a - b) <= 1e-12 * abs(b), (rs, a, b)


def test_max_cfl_and_step_ratio_formulas():
    """Basis::max_cfl = -2*0.9/min_eig_convection, step_ratio = 0.5/0.9 (Basis.cpp:6-14); row size 2 has the
    periodic-advection eigenvalue -6 (closed form for the two-point Gauss-Legendre upwind operator)."""
    assert abs(TABLES[2]["eigenvals"][0] + 6) < 1e-13


def test_regeneration_reproduces_committed_tables():
    tables = basis_gen.generate()
    for rs in ROW_SIZES:
        for k, v in tables[rs].items():
            np.testing.assert_array_equal(np.ravel(v), np.ravel(TABLES[rs][k]), err_msg=f"{rs} {k}")


@pytest.mark.parametrize("rs", ROW_SIZES)
def test_tables_exactly_mirror_symmetric(rs):
    """The FP64 element kernels take the mirrored half of D, B and L from the first half with a sign flip for even
    row sizes (same products, same order). That is exact only if the committed tables are bit-for-bit
    (anti)symmetric, which Gauss-Legendre nodes symmetric about 1/2 give in exact arithmetic."""
    t = TABLES[rs]
    d, b, w = t["diff_mat"], t["boundary"], t["weight"]
    lift0 = -(b[0] * (1.0 / w))
    lift1 = b[1] * (1.0 / w)
    assert np.array_equal(b[1], b[0][::-1])
    assert np.array_equal(lift1, -lift0[::-1])
    if rs % 2 == 0:
        assert np.array_equal(d, -d[::-1, ::-1])
