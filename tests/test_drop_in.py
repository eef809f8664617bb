"""The C++ drop-in (hexed_amd/host: the reference's kernel_factory / Spatial / compute_* interfaces over the C ABI),
exercised by tests/cpp/test_drop_in.cpp (built by __graft_entry__.build / make)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_drop_in")


def run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    p = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    return p


def test_drop_in_host_checks():
    """kernel_factory's range check and the Local assertions throw the reference's messages (no GPU needed)."""
    p = run("--host-only")
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASS" in p.stdout


@pytest.mark.gpu
def test_drop_in_against_oracle():
    """Solver::update, compute_write_face and kernel_factory-made Neighbor/Local through the adapter match the oracle
    to 1e-12 on Cartesian and deformed boxes, Euler and Navier-Stokes; unregistered Max_dt matches its closed form."""
    print("[gpu] drop-in adapter vs oracle", flush=True)
    p = run()
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASS" in p.stdout
