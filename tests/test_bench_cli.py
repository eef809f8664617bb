"""bench.py without a GPU: it refuses more than one GPU before touching one, and the HBM traffic it reports for each
workload's dominant kernel comes from the committed rocprofv3 summary (profiles/traffic.json) under the kernel name
it derives, in both arithmetic forms."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       cwd=ROOT, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


# (workload, dominant timing category, n_dim, row size)
DOMINANT = [("ns48", "deformed/local/navier_stokes", 3, 6), ("euler32", "cartesian/local/euler", 3, 6),
            ("amr16", "cartesian/local/euler", 3, 6), ("amr16_rs4", "cartesian/local/euler", 3, 4)]


@pytest.mark.parametrize("form", ["fp32_mfma", "fp64"])
@pytest.mark.parametrize("config,category,nd,rs", DOMINANT, ids=[d[0] for d in DOMINANT])
def test_traffic_known_for_dominant_kernel(bench, config, category, nd, rs, form):
    symbol = bench.kernel_symbol(category, nd, rs, form, bench.folds_faces(nd, rs, -1))
    traffic = bench.measured_traffic(config, symbol)
    assert traffic is not None, f"{config}: no measured traffic for {symbol}"
    # at least the compulsory bytes of one launch
    m_elem = {"ns48": 48 ** 3, "euler32": 32 ** 3, "amr16": 19461, "amr16_rs4": 19461}[config]
    alg = bench.algorithmic_bytes(category, nd, rs ** nd, m_elem)
    assert alg is not None and traffic > alg


@pytest.mark.parametrize("form", ["Fp32_mfma_form", "Fp64_form"])
def test_traffic_known_for_fused_local(bench, form):
    symbol = f"k_fused_local_euler<3, {form}, false, false, 4, 5, 6>"
    assert bench.measured_traffic("batch64", symbol) is not None


def test_deformed_euler_local_runs_fp64_form(bench):
    assert "Fp64_form" in bench.kernel_symbol("deformed/local/euler", 3, 6, "fp32_mfma", True)
    assert "Fp32_mfma_form" in bench.kernel_symbol("cartesian/local/euler", 3, 6, "fp32_mfma", True)
    assert bench.folds_faces(3, 6, -1) and not bench.folds_faces(3, 4, -1) and bench.folds_faces(3, 4, 1)
