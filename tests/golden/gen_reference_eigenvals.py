"""Writes tests/golden/reference_eigenvals.json from the reference's own build-time basis generator.

`Basis.eigenvals()` (script/Basis.py:197-204) is the one basis quantity with no closed form: numpy
eigenvalues of the discrete 1D periodic advection and diffusion operators of `discretizations(16)`
(script/Basis.py:158-196). The reference's generator prints them with `{:.20f}` into its C++ tables
(script/auto_generate.py:115-118), so the double each C++ literal parses to is recorded here, together
with the raw float64 numpy returned. Nodes and weights are obtained the way script/auto_generate.py:68-71
does (sympy Gauss-Legendre at 50 digits mapped to [0, 1]).

Run once in a container that holds the reference at /root/reference; the JSON is the committed fixture.
"""
import json
import os
import sys

REFERENCE_SCRIPT_DIR = "/root/reference/script"


def main():
    sys.path.insert(0, REFERENCE_SCRIPT_DIR)
    import matplotlib
    matplotlib.use("Agg")
    from Basis import Basis  # noqa: E402  (reference build-time generator)
    from sympy.integrals.quadrature import gauss_legendre

    out = {}
    for row_size in range(2, 9):
        nodes, weights = gauss_legendre(row_size, 50)
        nodes = [(node + 1) / 2 for node in nodes]
        weights = [weight / 2 for weight in weights]
        basis = Basis(nodes, weights, calc_digits=50)
        vals = basis.eigenvals()
        printed = [f"{v:.20f}" for v in vals]
        out[str(row_size)] = {
            "printed": printed,
            "as_double_hex": [float.hex(float(p)) for p in printed],
            "numpy_hex": [float.hex(float(v)) for v in vals],
        }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_eigenvals.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
