#!/usr/bin/env python3
"""DOF-stage/s of hexed's explicit time step (Solver::update, two stages per step) on one MI355X.

A step is one Solver::update on the whole mesh: max_dt, then for each of the two stages the state boundary
conditions and the compute_navier_stokes (stage 0) / compute_euler (stage 1) kernel sequence. One DOF is one
quadrature point carrying all n_dim + 2 conserved variables; value = DOF x 2 stages x steps / wall seconds, with
the mesh and state resident in HBM before the timed region starts.

Both arithmetic forms are timed on the same mesh from the same initial state: FP64 (the reference's arithmetic)
and FP32-MFMA (volume fluxes, gradients and derivative contractions in FP32 on the matrix cores; state, RK update,
face fluxes and jumps, metrics scaling and max_dt in FP64). `value` is the form chosen with --precision (default
fp32_mfma, the headline); the line also carries the other form's rate, the speedup, and the FP32-MFMA form's relative
L2 error per conserved variable against the FP64 form after the same number of steps.

The timed region enqueues the steps with hx_update_async: each step's time step stays on the device and the
steady-state update is one launch of a captured HIP graph, so the host never waits inside the region. The same
steps are then run again through hx_update (the time step read back every step) with HIP events around every
kernel launch; that eager region gives the kernel breakdown and the roofline's launch time, and its rate is
reported as `eager`.

The JSON line also carries
  roofline:     the chosen form's dominant kernel: algorithmic HBM bytes per launch / its average launch time,
                measured with HIP events on the library's stream over the eager timed region, against the 8 TB/s
                HBM peak;
  cpu_baseline: the C++/OpenMP oracle (a loop-for-loop restatement of the reference CPU path) timed on this
                host's cores for a few steps of the same mesh;
  parity_vs_cpu: relative L2 error per conserved variable between the FP64 GPU form and the oracle after those steps.

The batch workload (batch64) holds 64 independent cases in one mesh per row size; the three meshes form one group
(hx_group), so a step is one hx_group_update: every kernel of an Euler stage launched once for all meshes, classes
and row sizes, every case with its own time step, with the fused Local's launch shape picked by the autotuner
(hx_group_autotune, run for each form before its timed region; its table is in the JSON line).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config ns48|euler32|amr16|amr16_rs4|batch64]
                       [--precision fp32_mfma|fp64] [--cpu-steps N] [--err-steps N]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from hexed_amd import abi, configs  # noqa: E402
from hexed_amd.device import Context, Device_group, Device_mesh  # noqa: E402

METRIC = "DOF-stage/s, 3D NS row_size 6: FP32-MFMA vs FP64 GPU at stated L2 err"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "ns48": dict(
        desc="3D Navier-Stokes, 48^3 deformed hexes (interior vertices jittered +-10% of spacing), row_size 6, "
             "Re 1e4 on unit length, Mach 0.3, Sutherland air, constant bulk artificial viscosity 1e-3 m^2/s, "
             "1% seeded Fourier perturbations, freestream boundaries, global time step at max_safety 0.7",
        build=lambda: configs.ns_deformed(48), visc=True, cpu_steps=20, bytes_per_dof_stage=224),
    "amr16": dict(
        desc="3D Euler on a seeded 2:1-balanced tree: base 16^3 Cartesian hexes, refinement probability 0.085 per "
             "level up to 3 levels (19,461 leaves, 31% of the base refined, 7,691 hanging faces through "
             "Prolong/Restrict_refined), row_size 6, Mach 0.8 + 1% seeded Fourier perturbations, freestream "
             "boundaries, global time step at max_safety 0.7",
        build=lambda: configs.euler_tree(), visc=False, cpu_steps=200, bytes_per_dof_stage=128),
    "amr16_rs4": dict(
        desc="3D Euler on the same seeded 2:1-balanced tree (base 16^3, 19,461 leaves over 4 levels, 7,691 hanging "
             "faces), row_size 4, Mach 0.8 + 1% seeded Fourier perturbations, freestream boundaries, global time "
             "step at max_safety 0.7",
        build=lambda: configs.euler_tree(rs=4), visc=False, cpu_steps=400, bytes_per_dof_stage=128),
    "euler32": dict(
        desc="3D Euler, 32^3 Cartesian hexes, row_size 6, Mach 0.8 + 1% seeded Fourier perturbations, "
             "freestream boundaries, global time step at max_safety 0.7",
        build=lambda: configs.euler_box(32), visc=False, cpu_steps=150, bytes_per_dof_stage=128),
    "batch64": dict(
        desc="64 independent 3D Euler boxes on the unit cube, element counts log-uniform in [512, 8192] rounded to "
             "cubes (142,551 elements), row_size uniform in {4, 5, 6}, Mach uniform in [0.3, 0.9] + 1% seeded "
             "Fourier perturbations, freestream boundaries, a global time step per case at max_safety 0.7; one "
             "mesh per row size, each kernel launched once per row size for all its cases",
        build=None, visc=False, cpu_steps=50, bytes_per_dof_stage=128),
}


def build_groups(config):
    """[(host mesh, n_cases or None, CPU-sample builder)] of a workload: one mesh, or for the batch one mesh of cases
    per row size. The CPU sample is a list of meshes the oracle steps one at a time: the mesh itself, or every case
    of the row size alone (the reference runs one Solver per case)."""
    if config == "batch64":
        cases = configs.batch_cases()
        out = []
        for rs, (m, ks) in configs.euler_batch(cases).items():
            out.append((m, len(ks), lambda ks=ks, rs=rs: [configs.euler_box(cases[k]["n"], rs, cases[k]["mach"],
                                                                            cases[k]["seed"]) for k in ks]))
        return out
    m = WORKLOADS[config]["build"]()
    return [(m, None, lambda: [m])]


def algorithmic_bytes(category, nd, nq, n_elem):
    """Compulsory HBM bytes of one launch of an element kernel (face buffers and other intermediates are not
    counted): the state read and written, the residual cache written (stage 0) or read (stage 1), the time step
    scale, the two artificial-viscosity coefficients for Navier-Stokes, and for deformed elements the nd*nd
    reference-level normals plus the Jacobian determinant."""
    nv = nd + 2
    parts = category.split("/")
    if len(parts) != 3 or parts[1] != "local":
        return None
    doubles = 3 * nv + 1
    if parts[2] == "navier_stokes":
        doubles += 2
    if parts[0] == "deformed":
        doubles += nd * nd + 1
    return 8.0 * doubles * nq * n_elem


FORMS = {"fp64": (abi.HX_FP64, "Fp64_form"), "fp32_mfma": (abi.HX_FP32_MFMA, "Fp32_mfma_form")}


def folds_faces(nd, rs, mode):
    """Whether hx_update folds the face fluxes into the Local (hx_mesh_set_face_fold)."""
    return mode == 1 or (mode == -1 and rs >= 2 * nd)


def kernel_symbol(category, nd, rs, form, fold):
    """The HIP kernel a timing category launches (the name rocprofv3 reports, without the namespace), for local
    categories. The deformed Euler Local runs the FP64 form under either precision."""
    cls, kernel, pde = category.split("/")
    if kernel != "local":
        return None
    name = "k_local_ns" if pde == "navier_stokes" else "k_local_euler"
    deformed = cls == "deformed"
    form_name = "Fp64_form" if deformed and pde == "euler" else FORMS[form][1]
    return (f"{name}<{nd}, {rs}, {'true' if deformed else 'false'}, {form_name}, "
            f"{'true' if fold else 'false'}>")


def measured_traffic(config, symbol):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, if one covers the kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if symbol is None or not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f).get(config, {})
    return table.get("bytes_per_launch", {}).get(symbol)


def time_artificial_viscosity(ctx, m, reps=3):
    """The Navier-Stokes workload's artificial-viscosity kernels on its own mesh, outside the time step (the workload
    itself runs a constant artificial viscosity): Solver::update_art_visc_smoothness (advection, Legendre projection,
    Chebyshev diffusion of the forcing, Solver.cpp:457-582) with the reference's defaults and advection length 0.1, and
    Solver::set_art_visc_admis (Stab_art_visc, vertex sharing, interpolation; Solver.cpp:636-659). Each is run once,
    then `reps` times with HIP events around every kernel launch. Returns milliseconds per call and the per-kernel
    breakdown."""
    rho, u, p, _ = configs.ns_freestream()
    host = m.copy().enable_av_generator()
    dev = Device_mesh(ctx, host)
    reset(dev, host)
    opts = abi.av_options(0.1, host.row_size, rho, p / 0.4 + 0.5 * rho * u * u)
    vd = host.vertex_data()
    char_speed = u + (1.4 * p / rho) ** 0.5
    out = {}
    for name, call in (("update_art_visc_smoothness", lambda: dev.update_art_visc_smoothness(opts)),
                       ("set_art_visc_admis", lambda: dev.set_art_visc_admis(char_speed, vd))):
        call()
        ctx.synchronize()
        ctx.set_timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps
        t = ctx.timings()
        ctx.set_timing(False)
        out[name] = {"ms_per_call": round(1e3 * wall, 3),
                     "kernels": {k: {"ms_per_call": round(v[0] / reps, 4), "launches": v[1] // reps}
                                 for k, v in sorted(t.items())}}
    dev.close()
    return out


def time_setup_paths(ctx, m, opts, reps=3):
    """The paths around the time step on the workload's mesh, outside the timed region: Solver::calc_jacobian (element
    metrics, face normals and vertex time step scales on the device; Solver.cpp:273-381), with the oracle's on the host
    cores beside it; Solver::is_admissible (Solver.cpp:921-959); and Solver::fix_admissibility (Solver.cpp:961-1095)
    on the initial state with the energy made negative at three quadrature points of three elements, with its
    iteration count. Milliseconds per call (the fix: one call)."""
    import oracle
    out = {}
    host = m.copy()
    dev = Device_mesh(ctx, host)
    reset(dev, host)
    vd = host.vertex_data()

    def timed(call, n):
        call()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            r = call()
        ctx.synchronize()
        return 1e3 * (time.perf_counter() - t0) / n, r

    ms, _ = timed(lambda: dev.calc_jacobian(vd), reps)
    lib = oracle.load()
    ref = m.copy()
    s, vdr = ref.struct(), ref.vertex_data()
    t0 = time.perf_counter()
    oracle.check(lib.oracle_calc_jacobian(ctypes.byref(s), ctypes.byref(vdr)))
    out["calc_jacobian"] = {"ms_per_call": round(ms, 3), "cpu_oracle_ms": round(1e3 * (time.perf_counter() - t0), 1),
                            "cpu_threads": lib.oracle_get_threads()}
    ms, ok = timed(dev.is_admissible, 10 * reps)
    out["is_admissible"] = {"ms_per_call": round(ms, 3), "admissible": ok}
    # three points of three elements made inadmissible (energy < 0), then one fix
    nv, nd = host.nv, host.n_dim
    c = 1 if host.n_elem[1] else 0
    bad = host.state[c].copy()
    n = host.n_elem[c]
    for k, e in enumerate((n // 7, n // 2, (9 * n) // 10)):
        bad[e, nd + 1, (17 * k + 5) % host.nq] = -0.02
    dev.upload("state", c, bad)
    dev.compute_write_face()
    ctx.synchronize()
    t0 = time.perf_counter()
    iters = dev.fix_admissibility(opts, 0.7)
    ctx.synchronize()
    out["fix_admissibility"] = {"ms_per_call": round(1e3 * (time.perf_counter() - t0), 3), "iterations": iters,
                                "admissible_after": dev.is_admissible()}
    dev.close()
    return out


def host_cpu():
    """The host's CPU model (lscpu), its logical CPU count (nproc) and the CPUs this process may run on."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc", "--all"], capture_output=True, text=True, timeout=10).stdout)
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = os.cpu_count()
    return {"model": model, "nproc_all": nproc, "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def run_oracle_steps(m, opts, steps):
    import oracle
    lib = oracle.load()
    s = m.struct()
    st = abi.hx_update_status()
    t0 = time.perf_counter()
    for _ in range(steps):
        oracle.check(lib.oracle_update(ctypes.byref(s), ctypes.byref(opts), ctypes.byref(st)))
    return time.perf_counter() - t0, oracle.load().oracle_get_threads()


def reset(dev, m):
    """Initial state on the device: state uploaded, residual cache and time step scale cleared, faces extrapolated and
    prolonged to the hanging-face mortars (Solver::initialize, Solver.cpp:400-401)."""
    for c in (0, 1):
        if m.n_elem[c]:
            dev.upload("state", c, m.state[c])
            dev.upload("rcache", c, np.zeros_like(m.rcache[c]))
            dev.upload("tss", c, np.ones_like(m.tss[c]))
    dev.compute_write_face()
    dev.prolong()


def download_state(dev, m):
    return [dev.download("state", c, out=np.zeros_like(m.state[c])) if m.n_elem[c] else None for c in (0, 1)]


def rel_l2_per_var(got, ref, nv):
    """Relative L2 error per conserved variable over all elements of all groups (momentum components, mass,
    energy); got/ref are lists of per-class state arrays."""
    errs = []
    for v in range(nv):
        pairs = [(g, r) for g, r in zip(got, ref) if r is not None and r.size]
        num = sum(float(np.linalg.norm(g[:, v] - r[:, v])) ** 2 for g, r in pairs)
        den = sum(float(np.linalg.norm(r[:, v])) ** 2 for _, r in pairs)
        errs.append((num / den) ** 0.5)
    return errs


def step(groups, opts, n=1, fused=None):
    """n time steps of every mesh: hx_update, hx_update_cases for a batch mesh, or one hx_group_update for all of them
    when they form a group (`fused`: (Device_group, status array))."""
    if fused is not None:
        for _ in range(n):
            fused[0].update(opts, fused[1])
        return
    for m, dev, nc, status in groups:
        for _ in range(n):
            if nc is None:
                dev.update(opts, status)
            else:
                dev.update_cases(opts, status)


def step_async(groups, opts, n, fused=None):
    """n time steps of every mesh enqueued with hx_update_async (batch meshes step as in `step`, reading their per-case
    time steps back)."""
    if fused is not None or any(nc is not None for _, _, nc, _ in groups):
        return step(groups, opts, n, fused)
    for m, dev, nc, status in groups:
        dev.update_async(opts, n)


def fresh_status(nc):
    return abi.hx_update_status() if nc is None else (abi.hx_update_status * nc)()


def timed_run(ctx, groups, opts, form, warmup, steps, group=None):
    """The timed region on the graph path (from the initial state, after `warmup` steps), then the same steps again
    from the initial state through hx_update with every kernel launch timed. A batch workload's group is autotuned
    first. Returns (graph seconds, eager seconds, per-kernel timings of the eager region, final status, autotune
    table or None)."""
    groups = [(m, dev, nc, fresh_status(nc)) for m, dev, nc, _ in groups]
    for m, dev, nc, _ in groups:
        dev.set_precision(FORMS[form][0])
        reset(dev, m)
    fused, table = None, None
    if group is not None:
        table = group.autotune(use_filter=bool(opts.use_filter), reps=5)
        for m, dev, nc, _ in groups:
            reset(dev, m)
        fused = (group, (abi.hx_update_status * group.n_cases)())
    step_async(groups, opts, warmup, fused)
    ctx.synchronize()
    t0 = time.perf_counter()
    step_async(groups, opts, steps, fused)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    for m, dev, nc, status in groups:
        if nc is None:
            dev.status_sync(status)
    st = fused[1] if fused else groups[0][3]
    status = st if (groups[0][2] is None and not fused) else st[0]
    groups = [(m, dev, nc, fresh_status(nc)) for m, dev, nc, _ in groups]
    for m, dev, nc, _ in groups:
        reset(dev, m)
    if fused:
        fused = (group, (abi.hx_update_status * group.n_cases)())
    step(groups, opts, warmup, fused)
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    step(groups, opts, steps, fused)
    ctx.synchronize()
    eager = time.perf_counter() - t0
    ctx.set_timing(False)
    return elapsed, eager, ctx.timings(), status, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="ns48", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32_mfma", choices=sorted(FORMS), help="form reported as `value`")
    ap.add_argument("--cpu-steps", type=int, default=None, help="oracle steps for the CPU baseline (0 to skip)")
    ap.add_argument("--err-steps", type=int, default=10, help="steps after which FP32-MFMA is compared with FP64")
    ap.add_argument("--av", type=int, default=1, help="ns48: also time the artificial-viscosity kernels (0 to skip)")
    ap.add_argument("--setup", type=int, default=1,
                    help="ns48: also time calc_jacobian, is_admissible and fix_admissibility (0 to skip)")
    ap.add_argument("--face-fold", type=int, default=-1, choices=(-1, 0, 1),
                    help="fold the face fluxes into the Local: 1 always, 0 never, -1 where 2 n_dim <= row_size")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: the benchmark runs on one GPU only; multi-GPU figures are not produced "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world})", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.config]
    cpu_steps = wl["cpu_steps"] if args.cpu_steps is None else args.cpu_steps

    t_build = time.perf_counter()
    hosts = build_groups(args.config)
    nd, nv = hosts[0][0].n_dim, hosts[0][0].nv
    n_elem = sum(sum(m.n_elem) for m, _, _ in hosts)
    n_dof = sum(m.n_dof for m, _, _ in hosts)
    if wl["visc"]:
        visc, cond = configs.sutherland_air()
        opts = abi.update_options(max_safety=0.7, visc=visc, cond=cond, use_ldg=True)
    else:
        opts = abi.update_options(max_safety=0.7)
    print(f"[bench] {args.config}: {n_elem} elements, {n_dof} DOF in {len(hosts)} mesh(es), built in "
          f"{time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)

    ctx = Context(0)
    copy_gbs = ctx.stream_copy(1 << 31, 10)  # achievable HBM bandwidth of a plain copy kernel on this box
    groups = []
    for m, nc, _ in hosts:
        dev = Device_mesh(ctx, m)
        dev.set_face_fold(args.face_fold)
        if nc is not None:
            dev.set_cases(nc, m.case_id)
        groups.append((m, dev, nc, None))
    group = Device_group(ctx, [dev for _, dev, _, _ in groups]) if args.config == "batch64" else None
    runs = {}
    for form in FORMS:
        elapsed, eager, timings, status, table = timed_run(ctx, groups, opts, form, args.warmup, args.steps, group)
        runs[form] = dict(elapsed=elapsed, eager=eager, timings=timings, value=n_dof * 2 * args.steps / elapsed,
                          autotune=table)
        print(f"[bench] GPU {form}: {args.steps} steps in {elapsed:.3f} s (graph), {eager:.3f} s (eager, "
              f"events), dt {status.time_step:.3e}", file=sys.stderr, flush=True)
    chosen = runs[args.precision]
    value, elapsed, timings = chosen["value"], chosen["elapsed"], chosen["timings"]

    # FP32-MFMA against FP64 from the same initial state after err_steps steps
    finals = {}
    for form in FORMS:
        gs = [(m, dev, nc, fresh_status(nc)) for m, dev, nc, _ in groups]
        for m, dev, nc, _ in gs:
            dev.set_precision(FORMS[form][0])
            reset(dev, m)
        step(gs, opts, args.err_steps, (group, (abi.hx_update_status * group.n_cases)()) if group else None)
        finals[form] = [a for m, dev, _, _ in gs for a in download_state(dev, m)]
    fp32_err = rel_l2_per_var(finals["fp32_mfma"], finals["fp64"], nv)
    print(f"[bench] FP32-MFMA vs FP64 after {args.err_steps} steps: rel L2 per variable "
          + " ".join(f"{e:.2e}" for e in fp32_err), file=sys.stderr, flush=True)
    for m, dev, _, _ in groups:
        dev.set_precision(abi.HX_FP64)

    aux = None
    if args.config == "ns48" and args.av:
        aux = time_artificial_viscosity(ctx, hosts[0][0])
        print("[bench] artificial viscosity: " + ", ".join(f"{k} {v['ms_per_call']} ms" for k, v in aux.items()),
              file=sys.stderr, flush=True)
    setup_paths = None
    if args.config == "ns48" and args.setup:
        setup_paths = time_setup_paths(ctx, hosts[0][0], opts)
        print("[bench] setup paths: " + ", ".join(f"{k} {v['ms_per_call']} ms" for k, v in setup_paths.items()),
              file=sys.stderr, flush=True)

    # dominant kernel of the chosen form and its roofline: algorithmic bytes of every launch in its category over the
    # time spent in them (for one mesh: bytes per launch over the average launch time)
    dominant = max(timings, key=lambda k: timings[k][0])
    ms, launches, units = timings[dominant]
    if dominant.startswith("fused/"):
        # one fused launch covers every mesh and class: its bytes are those of all of them
        per_round = [sum(algorithmic_bytes(f"{('cartesian', 'deformed')[c]}/local/euler", nd, m.nq, m.n_elem[c])
                         for c in (0, 1)) for m, _, _ in hosts]
        per_round = [sum(per_round)]
        launches_per_mesh = launches
    else:
        cls = abi.HX_DEFORMED if dominant.startswith("deformed") else abi.HX_CARTESIAN
        per_round = [algorithmic_bytes(dominant, nd, m.nq, m.n_elem[cls]) for m, _, _ in hosts]
        launches_per_mesh = launches / len(hosts)
    rs0 = hosts[0][0].row_size
    if dominant == "fused/local/euler":
        # the fused Local over the group's row sizes, in the launch shape the autotuner kept
        table = chosen.get("autotune") or []
        persist = bool(table) and min(table, key=lambda r: r[2])[0] != 2  # (order, workgroups per CU, ms)
        with_def = any(m.n_elem[1] for m, _, _ in hosts)
        symbol = (f"k_fused_local_euler<{nd}, {FORMS[args.precision][1]}, {'true' if with_def else 'false'}, "
                  f"{'true' if persist else 'false'}, 4, 5, 6>")
    else:
        symbol = (kernel_symbol(dominant, nd, rs0, args.precision, folds_faces(nd, rs0, args.face_fold))
                  if len(hosts) == 1 else None)
    roofline = {"bound": "hbm", "kernel": dominant, "symbol": symbol, "launch_ms": round(ms / launches, 5),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "traffic": measured_traffic(args.config, symbol) if symbol else None,
                "stream_copy_gbs": round(copy_gbs, 1)}
    if None not in per_round:
        alg = sum(per_round)  # one launch per mesh, or one fused launch for all
        achieved = alg * launches_per_mesh / (ms * 1e-3) / 1e9
        roofline.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": alg})
    step_gbs = value * wl["bytes_per_dof_stage"] / 1e9
    roofline["step"] = {"bytes_per_dof_stage": wl["bytes_per_dof_stage"], "achieved": round(step_gbs, 1),
                        "frac": round(step_gbs / HBM_PEAK_GBS, 4)}
    breakdown = {form: {k: {"ms_per_step": round(v[0] / args.steps, 4), "launches": v[1]}
                        for k, v in sorted(r["timings"].items())} for form, r in runs.items()}

    # CPU baseline and full-size parity: the oracle runs cpu_steps from the initial state (for the batch, on every case
    # alone, as one Solver per case would); the FP64 GPU form repeats them
    cpu = None
    parity = None
    if cpu_steps > 0:
        import oracle
        lib = oracle.load()
        cpu_s, cpu_dof, threads, got, ref, samples = 0.0, 0, 1, [], [], []
        for sample in (x for (_, _, make_sample) in hosts for x in make_sample()):
            samples.append(sample)
            init_state = [a.copy() for a in sample.state]
            s = sample.struct()
            for c in (0, 1):
                oracle.check(lib.oracle_write_face(ctypes.byref(s), c, None))
            oracle.check(lib.oracle_prolong_refined(ctypes.byref(s), None, 0, 0))
            t, threads = run_oracle_steps(sample, opts, cpu_steps)
            cpu_s += t
            cpu_dof += sample.n_dof
            ref += [a.copy() for a in sample.state]
            for c in (0, 1):
                sample.state[c][...] = init_state[c]
            d1 = Device_mesh(ctx, sample)
            reset(d1, sample)
            st = abi.hx_update_status()
            for _ in range(cpu_steps):
                d1.update(opts, st)
            got += download_state(d1, sample)
            d1.close()
        print(f"[bench] CPU oracle: {cpu_steps} steps in {cpu_s:.2f} s on {threads} threads", file=sys.stderr, flush=True)
        what = (f"the same {n_elem}-element mesh" if len(groups) == 1 else
                f"every case of the batch, one at a time ({cpu_dof} DOF)")
        cpu = {"value": round(cpu_dof * 2 * cpu_steps / cpu_s, 1), "unit": "DOF-stage/s", "cores": threads,
               "kind": "port", "host": host_cpu(),
               "sample": f"{cpu_steps} time steps ({2 * cpu_steps} stages) of {what}, FP64, OpenMP over {threads} "
                         f"threads, {cpu_s:.2f} s"}
        # the same oracle on one thread, over a shorter sample of the same meshes (timing only)
        steps_1t = max(1, cpu_steps // 20)
        prev_threads = lib.oracle_get_threads()
        lib.oracle_set_threads(1)
        try:
            t_1t = 0.0
            for sample in samples:
                t, _ = run_oracle_steps(sample, opts, steps_1t)
                t_1t += t
        finally:
            lib.oracle_set_threads(prev_threads)
        print(f"[bench] CPU oracle: {steps_1t} step(s) in {t_1t:.2f} s on 1 thread", file=sys.stderr, flush=True)
        cpu["value_1_thread"] = round(cpu_dof * 2 * steps_1t / t_1t, 1)
        cpu["sample_1_thread"] = f"{steps_1t} time step(s) ({2 * steps_1t} stages) of the same sample on 1 thread, {t_1t:.2f} s"
        errs = rel_l2_per_var(got, ref, nv)
        parity = {"steps": cpu_steps, "form": "fp64", "rel_l2_per_var": [float(f"{e:.3e}") for e in errs],
                  "max_rel_l2": max(errs)}

    other = "fp64" if args.precision == "fp32_mfma" else "fp32_mfma"
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "DOF-stage/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-mfma+f64" if args.precision == "fp32_mfma" else "f64", "data": "synthetic",
        "config": {"workload": wl["desc"], "name": args.config, "n_elem": n_elem, "dof": n_dof,
                   "row_size": sorted({m.row_size for m, _, _ in hosts}), "precision": args.precision,
                   "face_fold": args.face_fold},
        "forms": {
            args.precision: {"value": round(value, 1), "ms_per_step": round(1e3 * elapsed / args.steps, 4),
                             "eager": {"value": round(n_dof * 2 * args.steps / chosen["eager"], 1),
                                       "ms_per_step": round(1e3 * chosen["eager"] / args.steps, 4)}},
            other: {"value": round(runs[other]["value"], 1),
                    "ms_per_step": round(1e3 * runs[other]["elapsed"] / args.steps, 4),
                    "eager": {"value": round(n_dof * 2 * args.steps / runs[other]["eager"], 1),
                              "ms_per_step": round(1e3 * runs[other]["eager"] / args.steps, 4)}},
            "fp32_mfma_speedup_vs_fp64": round(runs["fp32_mfma"]["value"] / runs["fp64"]["value"], 3),
            "autotune": ({f: [{"order": o, "workgroups_per_cu": w, "ms": round(t, 4)} for o, w, t in runs[f]["autotune"]]
                          for f in FORMS} if group else None),
            "fp32_mfma_vs_fp64_rel_l2_per_var": {"steps": args.err_steps,
                                                  "vars": [f"momentum{i}" for i in range(nd)] + ["mass", "energy"],
                                                  "rel_l2": [float(f"{e:.3e}") for e in fp32_err],
                                                  "max": float(f"{max(fp32_err):.3e}")},
        },
        "roofline": roofline, "cpu_baseline": cpu, "parity_vs_cpu": parity, "kernel_breakdown": breakdown,
        "artificial_viscosity": aux, "setup_paths": setup_paths,
    }
    print(json.dumps(line))
    if group is not None:
        group.close()
    for m, dev, _, _ in groups:
        dev.close()
    ctx.close()


if __name__ == "__main__":
    main()
