"""hx_update_cases (a batch of independent cases in one mesh) against hx_update on a mesh holding each
case alone: the same time step per case and bit-identical states, in both arithmetic forms, for Euler and
Navier-Stokes (single cases are themselves compared with the oracle in test_gpu_parity.py); plus the errors of
the registration."""
import numpy as np
import pytest

from hexed_amd import abi, configs, mesh
from hexed_amd.device import Context, Device_mesh, HxError

from helpers import init_faces, per_var_rel_l2, say

pytestmark = pytest.mark.gpu

STEPS = 3


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def euler_cases():
    return [configs.euler_box(2, 4, mach=0.3, seed=1), configs.euler_box(3, 4, mach=0.7, seed=2),
            configs.euler_box(2, 4, mach=0.5, seed=3, deformed=True, jitter=0.1),
            configs.euler_box(4, 4, mach=0.9, seed=4)]


def ns_cases():
    return [configs.ns_deformed(2, rs=5, seed=1), configs.ns_deformed(3, rs=5, mach=0.5, seed=2),
            configs.ns_deformed(2, rs=5, mach=0.2, jitter=0.0, seed=3)]


def run_batch_vs_single(ctx, cases, opts, precision, relabel=None):
    """relabel: a permutation of the case ids (case k registered as relabel[k]), so that the ids are not in order
    within a class and the update takes its general per-case reduction."""
    for m in cases:
        init_faces(m)
    b = mesh.concat(cases)
    dev = Device_mesh(ctx, b.copy())
    dev.set_precision(precision)
    perm = np.arange(len(cases)) if relabel is None else np.asarray(relabel)
    dev.set_cases(len(cases), [perm[c].astype(np.int32) for c in b.case_id])
    st = (abi.hx_update_status * len(cases))()
    singles = []
    for m in cases:
        d = Device_mesh(ctx, m.copy())
        d.set_precision(precision)
        d.set_face_fold(1)  # a batch update folds the face fluxes at every row size
        singles.append((d, abi.hx_update_status()))
    for it in range(STEPS):
        dev.update_cases(opts, st)
        for k, (d, s1) in enumerate(singles):
            d.update(opts, s1)
            sk = st[int(perm[k])]
            assert sk.time_step == s1.time_step, f"step {it} case {k}: dt {sk.time_step} vs {s1.time_step}"
            assert sk.flow_time == s1.flow_time and sk.iteration == s1.iteration == it + 1
    got = [dev.download("state", c, out=np.zeros_like(b.state[c])) for c in (0, 1)]
    off = [0, 0]
    for k, (d, _) in enumerate(singles):
        m = cases[k]
        for c in (0, 1):
            n = m.n_elem[c]
            if n:
                ref = d.download("state", c, out=np.zeros_like(m.state[c]))
                part = got[c][off[c]:off[c] + n]
                assert np.array_equal(part, ref), f"case {k} class {c}: rel L2 {per_var_rel_l2(part, ref)}"
            off[c] += n
        d.close()
    dev.close()


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
def test_batch_euler(ctx, precision):
    say(f"[gpu] batch euler precision {precision}")
    run_batch_vs_single(ctx, euler_cases(), abi.update_options(max_safety=0.5), precision)


@pytest.mark.parametrize("precision", [abi.HX_FP64, abi.HX_FP32_MFMA], ids=["fp64", "fp32_mfma"])
def test_batch_navier_stokes(ctx, precision):
    say(f"[gpu] batch navier-stokes precision {precision}")
    visc, cond = configs.sutherland_air()
    run_batch_vs_single(ctx, ns_cases(), abi.update_options(max_safety=0.5, visc=visc, cond=cond, use_ldg=True),
                        precision)


def test_batch_unordered_case_ids(ctx):
    say("[gpu] batch with case ids out of order")
    run_batch_vs_single(ctx, euler_cases(), abi.update_options(max_safety=0.5), abi.HX_FP64, relabel=[2, 0, 3, 1])


def test_batch_filter_and_chebyshev(ctx):
    say("[gpu] batch with filter, 2 chebyshev stages, max_time_step")
    opts = abi.update_options(max_safety=0.5, n_cheby_flow=2, use_filter=True, max_time_step=2e-3)
    run_batch_vs_single(ctx, euler_cases(), opts, abi.HX_FP64)


def test_batch_errors(ctx):
    say("[gpu] batch registration errors")
    one = configs.euler_box(2, 4, seed=1)
    init_faces(one)
    dev = Device_mesh(ctx, one)
    with pytest.raises(HxError, match="no cases registered"):
        dev.update_cases(abi.update_options(), (abi.hx_update_status * 1)())
    ids = np.zeros(one.n_elem[0], np.int32)
    ids[0] = 1  # element 0 and its neighbours in different cases
    with pytest.raises(HxError, match="joins two cases|shared by two cases"):
        dev.set_cases(2, [ids, None])
    with pytest.raises(HxError, match="out of range"):
        dev.set_cases(1, [ids, None])
    dev.close()
