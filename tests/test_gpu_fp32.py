"""FP32-MFMA form of the element kernels against the FP64 form on the same meshes and initial states.

The FP32-MFMA form computes volume fluxes, gradients and derivative contractions in FP32 (the contractions with
V_MFMA_F32_4X4X1_16B_F32) while the state, the RK update, face fluxes and jumps, and max_dt stay FP64. It is not
the reference's arithmetic, so the bar is a stated tolerance rather than parity: relative L2 per conserved
variable <= TOL32 after a few steps of smooth flow, and a freestream held to FP32 roundoff.
"""
import numpy as np
import pytest

from hexed_amd import abi, mesh
from hexed_amd.device import Context, Device_mesh

from helpers import freestream, init_faces, per_var_rel_l2, say, smooth_state

pytestmark = pytest.mark.gpu

TOL32 = 1e-7   # relative L2 per variable after STEPS steps vs the FP64 form (measured up to 2.4e-8)
STEPS = 3

# (n_dim, row_size, deformed, elements per axis, Navier-Stokes); includes row sizes whose workgroups are not whole
# wavefronts (rs 5: 2 x 125 points, rs 7: 343, 1D rs 5: 51 x 5)
CONFIGS = [
    (3, 6, True, 3, True),
    (3, 6, False, 3, False),
    (3, 6, True, 3, False),
    (3, 5, True, 2, True),
    (3, 7, False, 2, True),
    (2, 4, True, 4, True),
    (2, 8, False, 3, False),
    (1, 5, True, 6, True),
    (3, 2, True, 3, True),
]


def ids(c):
    return f"nd{c[0]}-rs{c[1]}-{'def' if c[2] else 'car'}-n{c[3]}-{'ns' if c[4] else 'euler'}"


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def sutherland_air():
    return abi.sutherland(1.716e-5, 273.0, 111.0), abi.sutherland(0.0241, 273.0, 194.0)


def run(ctx, host, opts, precision, steps):
    dev = Device_mesh(ctx, host.copy())
    dev.set_precision(precision)
    st = abi.hx_update_status()
    for _ in range(steps):
        dev.update(opts, st)
    out = [dev.download("state", c, out=np.zeros_like(host.state[c])) if host.n_elem[c] else None for c in (0, 1)]
    dev.close()
    return out


# every configuration without the modal filter, the first four with it
CASES = [(c, False) for c in CONFIGS] + [(c, True) for c in CONFIGS[:4]]


@pytest.mark.parametrize("cfg,use_filter", CASES, ids=[ids(c) + ("-filter" if f else "") for c, f in CASES])
def test_fp32_mfma_vs_fp64(ctx, cfg, use_filter):
    nd, rs, deformed, n, ns = cfg
    say(f"[gpu] fp32-mfma vs fp64 {ids(cfg)} filter={use_filter}")
    host = mesh.box(nd, rs, n, deformed=deformed, jitter=0.1 if deformed else 0.0, seed=3,
                    bc=(abi.HX_BC_FREESTREAM, freestream(nd)), with_av=ns)
    smooth_state(host, seed=3)
    if ns:
        for c in (0, 1):
            if host.n_elem[c]:
                host.av[c][:] = 1e-3
    init_faces(host)
    if ns:
        visc, cond = sutherland_air()
        opts = abi.update_options(max_safety=0.5, visc=visc, cond=cond, use_filter=use_filter)
    else:
        opts = abi.update_options(max_safety=0.5, use_filter=use_filter)
    ref = run(ctx, host, opts, abi.HX_FP64, STEPS)
    got = run(ctx, host, opts, abi.HX_FP32_MFMA, STEPS)
    for c in (0, 1):
        if host.n_elem[c] == 0:
            continue
        errs = per_var_rel_l2(got[c], ref[c])
        say(f"    class {c}: rel L2 per variable " + " ".join(f"{e:.2e}" for e in errs))
        assert max(errs) <= TOL32, f"class {c}: {errs}"


@pytest.mark.parametrize("nd,rs", [(3, 6), (2, 5)])
def test_fp32_mfma_freestream(ctx, nd, rs):
    """A uniform flow on a jittered mesh stays uniform to FP32 roundoff (free-stream preservation)."""
    say(f"[gpu] fp32-mfma freestream nd{nd} rs{rs}")
    fs = freestream(nd)
    host = mesh.box(nd, rs, 3, deformed=True, jitter=0.1, seed=1, bc=(abi.HX_BC_FREESTREAM, fs))
    mesh.set_state(host, lambda pos: fs)
    init_faces(host)
    got = run(ctx, host, abi.update_options(max_safety=0.5), abi.HX_FP32_MFMA, 2)
    scale = max(abs(x) for x in fs)
    devs = [float(np.abs(got[1][:, v] - fs[v]).max()) / scale for v in range(nd + 2)]
    say("    max deviation / freestream scale per variable " + " ".join(f"{d:.2e}" for d in devs))
    assert max(devs) <= 5e-8, f"max deviations {devs}"  # measured up to 4.7e-9


@pytest.mark.parametrize("nd,ns", [(3, False), (3, True), (2, True)], ids=["3d-euler", "3d-ns", "2d-ns"])
def test_fp32_mfma_vs_fp64_tree(ctx, nd, ns):
    """Seeded 2:1 tree meshes: the hanging faces' prolongation and restriction stay FP64 in both forms, the elements
    next to them run the FP32-MFMA contractions; same bar as the box meshes."""
    from hexed_amd import configs
    say(f"[gpu] fp32-mfma vs fp64 tree nd{nd} {'ns' if ns else 'euler'}")
    host = configs.euler_tree(base=3, rs=6, levels=2, frac=0.3, seed=5, nd=nd)
    if ns:
        host.av = [np.full((host.n_elem[0], 2, host.nq), 1e-3), None]
        opts = abi.update_options(max_safety=0.3, visc=abi.constant(2e-3), cond=abi.constant(2.8e-3))
    else:
        opts = abi.update_options(max_safety=0.5)
    init_faces(host)
    ref = run(ctx, host, opts, abi.HX_FP64, STEPS)
    got = run(ctx, host, opts, abi.HX_FP32_MFMA, STEPS)
    errs = per_var_rel_l2(got[0], ref[0])
    say("    rel L2 per variable " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) <= TOL32, errs
