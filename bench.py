"""Projection-step benchmark on one MI355X.

A step is one pass of the hot path over a device-resident scene:
RK2 semi-Lagrangian advection -> RHS -> pressure solve to the stated
relative residual -> gradient correction -> max|div u| gate -> commit.
Inputs are synthetic (seeded Fourier-mode spectrum), already resident in
HBM when the timed region starts.

Prints ONE JSON line (rank 0, N = 1).  `--gpus N` with N > 1 exits: this
build measures one GPU (a scene shards by whole scenes across GPUs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import rheon_amd as R  # noqa: E402
from rheon_amd import _native as N  # noqa: E402
from rheon_amd.scenes import GridSpec, centred_sphere, courant_dt, fourier_modes, serving_requests  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def mg_cycle_bytes(spec: GridSpec, s: int, m: int = 0, pre: int = 3, post: int = 1,
                   x0_implicit: bool = True, cpre: int = 0, cpost: int = 0) -> float:
    """Compulsory bytes of one defect-correction cycle (outer f64 residual +
    V(pre, post) in storage of s bytes): per finest cell
      outer residual  read p, b (f64) + write b0, x0           16 + 2s
      extra pre-sweeps  (pre-1) x (read x, b + write x)         3s each
      restrict  read x, b + write coarse b, x (1/8 cells)       2s + 2s/8
      prolong   read x + coarse x (1/8) + write x               2s + s/8
      post-sweeps (post-1) x 3s, last one adds to p (f64 r/w)   2s + 16
    and coarse levels l >= 1 the same minus the f64 terms with their own
    sweep counts (cpre, cpost; 0 = the finest level's), on N/8^l cells.
    x0_implicit (four-cell FP32 kernels, pre >= 2): the outer residual does
    not write x0 and the second pre-sweep reads b0 alone (s less each).
    With obstacles (m = 1) every finest stencil kernel also reads the 1-byte
    mask and every coarse stencil kernel the 4-byte face-fraction word."""
    n = spec.cell_len
    finest = (16 + 2 * s) + (pre - 1) * 3 * s + (2 * s + 2 * s / 8) + (2 * s + s / 8) + (post - 1) * 3 * s + (2 * s + 16)
    if x0_implicit and pre >= 2:
        finest -= 2 * s
    finest += m * (1 + (pre - 1) + 1 + post)  # obstacle mask reads
    cpre, cpost = cpre or pre, cpost or post
    coarse_per_cell = (cpre - 1) * 3 * s + (2 * s + 2 * s / 8) + (2 * s + s / 8) + cpost * 3 * s
    coarse_per_cell += 4 * m * (cpre + 1 + cpost)  # face-fraction words
    coarse = sum(coarse_per_cell * n / 8 ** l for l in range(1, 8))
    return finest * n + coarse


def coarse_sweeps(args) -> tuple:
    """--mg-coarse-sweeps as (pre, post); (0, 0) = the finest level's counts."""
    return tuple(int(v) for v in args.mg_coarse_sweeps.split(",")) if args.mg_coarse_sweeps else (0, 0)


def algorithmic_bytes(spec: GridSpec, face_bytes: int, method: str, solver_bytes: int, iterations: int,
                      m: int = 0, sweeps=(3, 1), x0_implicit: bool = False, coarse=(0, 0)) -> dict:
    """Compulsory HBM bytes of one step (per-kernel minimal-fusion model)."""
    n = spec.cell_len
    F = spec.face_total
    sv = solver_bytes
    if method in ("mixed-mg-dc", "mg-f64-dc"):
        per_iter = mg_cycle_bytes(spec, sv, m, *sweeps, x0_implicit=x0_implicit, cpre=coarse[0], cpost=coarse[1])
    else:
        # Krylov vectors in sv bytes, the solution accumulated in f64 (16 B/cell per update)
        per_iter = {"jacobi-pcg-v1": (9 * sv + 16) * n + m * n, "rbsgs-pcg": (15 * sv + 16) * n + 3 * m * n}.get(method, 0)
    return {
        "advect": 2 * face_bytes * F,
        "rhs": face_bytes * F + 8 * n,
        "solve": per_iter * iterations,
        "correct": 8 * n + 2 * face_bytes * F,
        "divmax": face_bytes * F,
    }


# Algorithmic bytes per launch of each profiled kernel, per cell (s = vector
# bytes, m = obstacle mask bytes per cell: 1 with obstacles, else 0; the
# first coarse level's face-fraction word is 4 bytes per coarse cell).
KERNEL_BYTES_PER_CELL = {
    "pcg_apply_dir": lambda s, m: 2 * s + m,      # read d (+mask), write q
    "pcg_update": lambda s, m: 4 * s + 16,        # read d, r, q (s), p (f64); write r (s), p (f64)
    "pcg_direction": lambda s, m: 3 * s,          # read r (or z), d; write d
    "pcg_sgs_black": lambda s, m: 3 * s + m,      # read r, z; write z (interleaved colours)
    "pcg_sgs_red": lambda s, m: 3 * s + m,        # read r, z; write z
    "pcg_true_residual": lambda s, m: 8 + 8 + s + m,  # read p (s), b (8), write product (s)
    "mg_outer_residual": lambda s, m: 16 + 2 * s + m,  # read p, b (f64) (+mask); write b0, x0
    "mg_smooth_first": lambda s, m: 2 * s + m,         # read b0 (+mask); write x1 (x0 recomputed)
    "mg_smooth": lambda s, m: 3 * s + m,               # read x, b (+mask); write x
    "mg_smooth_update": lambda s, m: 2 * s + 16 + m,   # read x, b (+mask); read+write p (f64)
    "mg_restrict": lambda s, m: 2 * s + 2 * s / 8 + m + 4 * m / 8,  # read x, b; write coarse b, x
    "mg_prolong": lambda s, m: 2 * s + s / 8 + 4 * m / 8,           # read x, coarse x; write x
}


# Algorithmic bytes per launch of the face-field kernels (f32 faces, F faces,
# N cells): advection reads three face arrays and writes three; the RHS pass
# reads the faces and writes f64 cells; the correction reads p and reads and
# writes the faces; the divergence/rate pass reads the faces.
FACE_KERNEL_BYTES = {
    "advect_velocity": lambda spec, m: 2 * 4 * spec.face_total,
    "build_rhs": lambda spec, m: 4 * spec.face_total + 8 * spec.cell_len,
    "correct": lambda spec, m: 8 * spec.cell_len + 2 * 4 * spec.face_total + m * spec.cell_len,
    "divergence_max": lambda spec, m: 4 * spec.face_total + m * spec.cell_len,
}
# Timed groups that are many kernels (not candidates for the dominant kernel).
AGGREGATES = ("mg_coarse_levels",)


def kernel_bytes_per_launch(name: str, spec: GridSpec, s: int, m: int) -> float:
    if name in FACE_KERNEL_BYTES:
        return FACE_KERNEL_BYTES[name](spec, m)
    return KERNEL_BYTES_PER_CELL[name](s, m) * spec.cell_len


def profiled_counters(kernel: str) -> dict:
    """`name=value` items of the committed instruction-counter summary line
    `counters_per_launch: kernel ...` (profiles/*_counters.md), or {}."""
    import glob
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*_counters.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith("counters_per_launch:"):
                    items = line.split(":", 1)[1].split()
                    if items and items[0] == kernel:
                        return {k: float(v) for k, _, v in (it.partition("=") for it in items[1:])}
    return {}


def profiled_traffic(kernel: str):
    """Calibrated HBM bytes per launch of `kernel` on the default workload,
    from the committed rocprofv3 --pmc summary (profiles/*_traffic.md, line
    `traffic_bytes_per_launch: name=bytes`), or None."""
    import glob
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*_traffic.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    for item in line.split(":", 1)[1].split():
                        name, _, value = item.partition("=")
                        if name == kernel:
                            return float(value)
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec: GridSpec, vel, dt: float, iterations_needed: int, solid=None) -> dict:
    """The oracle (C restatement of the reference, its parallel.rs
    decomposition: 2048-value chunks above 16 Ki values, serial Neumaier
    reductions) on this host's cores, on a bounded sample of the step: one
    advection, the RHS, Jacobi-PCG iterations (median of 3 repetitions after
    a warm-up) and the correction, extrapolated to the Jacobi-PCG iteration
    count the reference method needs for this step.  Also times one
    iteration on 1 thread, and the single-thread 64^3 iteration that the
    reference publishes (6.36 ms on 1 worker of an EPYC 9V74)."""
    import oracle as O
    from rheon_amd.scenes import close_obstacle_faces

    threads = min(os.cpu_count() or 1, 16)
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    ob = O.obstacle_mask(g, solid) if solid is not None else None
    t0 = time.perf_counter()
    adv = O.advect_velocity(g, vel, dt, threads=threads)
    if solid is not None:
        adv = close_obstacle_faces(spec, solid, adv)
    t_adv = time.perf_counter() - t0
    t0 = time.perf_counter()
    rhs = O.build_rhs(g, adv, dt)
    t_rhs = time.perf_counter() - t0

    def iteration_seconds(grid, b, th, k, reps=3, obstacle=None):
        st = O.Settings(0.0, 0.0, 0.0, k)
        run = (lambda: O.solve_rhs(grid, 1.0, b, dt, st, threads=th)) if obstacle is None else \
            (lambda: O.solve_rhs_masked(grid, 1.0, obstacle, b, dt, st, threads=th))
        run()  # warm-up
        times = []
        for _ in range(reps):
            t = time.perf_counter()
            r = run()
            times.append((time.perf_counter() - t) / k)
        return float(np.median(times)), r

    t_iter, r = iteration_seconds(g, rhs, threads, 2, obstacle=ob)
    t_iter1, _ = iteration_seconds(g, rhs, 1, 1, reps=3, obstacle=ob)
    t0 = time.perf_counter()
    if ob is None:
        out = O.correct(g, 1.0, r.pressure, dt, adv)
        O.actual_divergence_max(g, out)
    else:
        out = O.correct_masked(g, 1.0, ob, r.pressure, dt, adv)
        O.actual_divergence_max_masked(g, ob, out)
    t_corr = time.perf_counter() - t0
    # the reference's published single-worker figure: 64^3 Jacobi-PCG iteration
    s64 = GridSpec.unit_box(64)
    g64 = O.Grid(s64.counts, s64.spacing, s64.origin)
    v64 = fourier_modes(s64, 7, modes=8, kmax=4, dtype=np.float64)
    rhs64 = O.build_rhs(g64, v64, courant_dt(s64, v64))
    t64, _ = iteration_seconds(g64, rhs64, 1, 20)
    step = t_adv + t_rhs + t_corr + t_iter * iterations_needed
    return {
        "value": 1.0 / step,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{spec.counts[0]}x{spec.counts[1]}x{spec.counts[2]}: 1 advection ({t_adv:.2f}s), RHS, "
                   f"jacobi-pcg-v1 iterations ({t_iter * 1e3:.0f} ms each on {threads} threads, median of 3), "
                   f"correction+div; extrapolated to {iterations_needed} iterations (the reference method's "
                   f"count for this step, counted by the bit-identical device jacobi-pcg-v1)"),
        "seconds_per_step_estimate": step,
        "iteration_ms_all_threads": t_iter * 1e3,
        "iteration_ms_1_thread": t_iter1 * 1e3,
        "jpcg_64cubed_iteration_ms_1_thread": t64 * 1e3,
        "reference_64cubed_iteration_ms_1_worker": 6.36,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
    }


def advection_line(spec: GridSpec, kernels: dict, default_workload: bool) -> dict:
    """Advection is bound by f64 instruction issue, not HBM: its survey
    time, HBM fraction of the algorithmic 2 s_f F bytes, and (default
    workload) VALU instructions per face and issue fraction from the
    committed counter summary."""
    if "advect_velocity" not in kernels:
        return None
    n, sec = kernels["advect_velocity"]
    avg = sec / n
    out = {"avg_us": avg * 1e6,
           "algorithmic_bytes_per_launch": 8 * spec.face_total,
           "hbm_frac": 8 * spec.face_total / avg / 1e9 / HBM_PEAK_GBS}
    c = profiled_counters("advect_velocity") if default_workload else {}
    if c:
        out["valu_insts_per_face"] = c.get("valu_insts_per_face")
        out["valu_issue_frac"] = c.get("valu_issue_frac")
        out["counters_source"] = "profiles/advect_velocity_256_counters.md"
    return out


def jacobi_settings(st: R.PressureSettings) -> R.PressureSettings:
    return R.PressureSettings(st.relative_residual, st.absolute_residual, st.divergence_limit, 20000,
                              poll_interval=32)


def step_rhs(spec: GridSpec, scene: R.Scene, dt: float, solid) -> np.ndarray:
    """RHS of the next step from the scene's accepted faces (device advection,
    obstacle faces closed, device RHS)."""
    from rheon_amd.scenes import close_obstacle_faces

    adv = R.advect_velocity(spec, scene.get_velocity(), dt)
    if solid is not None:
        adv = close_obstacle_faces(spec, solid, adv)
    return R.build_rhs(spec, adv, dt)


def compare_methods(spec: GridSpec, scene: R.Scene, dt: float, solid, settings: R.PressureSettings) -> dict:
    """Time-to-tolerance of one pressure solve of the same RHS by each method
    (device time), and the pressure error of each against the f64
    jacobi-pcg-v1 solution (bit-identical to the oracle's on the sizes the
    parity tests cover)."""
    rhs = step_rhs(spec, scene, dt, solid)
    out, ref = {}, None
    cases = [("jacobi-pcg-v1", R.F64), ("jacobi-pcg-v1", R.F32), ("rbsgs-pcg", R.F64), ("mg-f64-dc", R.F64),
             ("mixed-mg-dc", R.F32)]
    for method, prec in cases:
        ws = R.PressureWorkspace(spec, method, prec)
        if solid is not None:
            ws.set_solid(solid)
        st = settings if method in ("mg-f64-dc", "mixed-mg-dc") else jacobi_settings(settings)
        if prec == R.F32 and method == "jacobi-pcg-v1":
            # f32 Krylov vectors stall far above 1e-10: capped, reported as reached
            st = R.PressureSettings(st.relative_residual, st.absolute_residual, st.divergence_limit, 3000,
                                    poll_interval=32)
        ws.solve_rhs(1.0, rhs, dt, st)  # warm (allocations, coarse inverse)
        rep = ws.solve_rhs(1.0, rhs, dt, st)
        p = ws.get_pressure()
        key = f"{method}/{'f64' if prec == R.F64 else 'f32'}"
        entry = {"status": R.STATUS_NAMES.get(rep.status), "iterations": rep.iterations,
                 "seconds": rep.seconds, "true_residual_l2": rep.true_residual_l2,
                 "predicted_div_max": rep.predicted_divergence_max}
        if ref is None:
            ref = p
        else:
            entry["pressure_rel_err_vs_jpcg_f64"] = float(np.max(np.abs(p - ref)) / np.max(np.abs(ref)))
        out[key] = entry
        ws.close()
    base = out["jacobi-pcg-v1/f64"]["seconds"]
    for k, v in out.items():
        v["speedup_vs_jpcg_f64"] = base / v["seconds"] if v["seconds"] > 0 else None
    return out


def copy_probe(mib: int) -> int:
    import ctypes as C
    lib = N.lib()
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    R._check(lib.rh_device_alloc(nbytes, C.byref(a)), "rh_device_alloc")
    R._check(lib.rh_device_alloc(nbytes, C.byref(b)), "rh_device_alloc")
    R._check(lib.rh_copy_device(b, a, nbytes), "rh_copy_device")
    lib.rh_device_free(a)
    lib.rh_device_free(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "kernel": "k_copy_bytes", "read_bytes": nbytes,
                      "write_bytes": nbytes}), flush=True)
    return 0


def run_dense3d(n: int = 256, reps: int = 5, cpu: bool = True) -> dict:
    """Dense3d frame of an n^3 accepted transport state (write_dense3d_frame,
    src/dense3d_sequence.rs:415-564): three f32 face arrays and the f32
    tracer printed with 10 significant digits, the f64 fraction and pressure
    with 18, every byte formatted on the device (rh_dense3d_frame) from
    device-resident fields into a device buffer.  Synthetic fields: normal
    faces and pressure, uniform tracer and fraction.  Timed: `reps` whole
    calls (preflight pass, count, scan, text pieces, write; host wall clock
    around the synchronous call) and the count and write kernels (HIP events).
    CPU baseline: the oracle's serialization (glibc's exact conversion, one
    thread, as the reference's writer) on a 64^3 frame, scaled per number."""
    import ctypes as C
    spec = GridSpec.unit_box(n)
    rng = np.random.default_rng(256)
    cells = n ** 3
    flen = [(n + (a == 0)) * (n + (a == 1)) * (n + (a == 2)) for a in range(3)]
    vel = [rng.standard_normal(flen[a], dtype=np.float32) for a in range(3)]
    host = (vel, rng.random(cells, dtype=np.float32), rng.random(cells), rng.standard_normal(cells))
    dev = [R.DeviceArray.from_host(a) for a in host[0]] + [R.DeviceArray.from_host(a) for a in host[1:]]
    state = R.Dense3dState(velocity=tuple(dev[:3]), tracer=dev[3], fraction=dev[4], pressure=dev[5],
                           carrier_stamp=(43, 0), liquid_stamp=(41, 0))
    try:
        R.dense3d_frame_device(spec, state, 0, None, None)
    except R.RheonError as e:
        needed = e.needed
    out = R.DeviceArray(needed, np.uint8)
    numbers = sum(flen) + 3 * cells
    lib = N.lib()
    R.dense3d_frame_device(spec, state, 0, None, out)  # warm-up
    lib.rh_profile_focus(None)
    lib.rh_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(reps):
        nbytes = R.dense3d_frame_device(spec, state, 0, None, out)
    wall = (time.perf_counter() - t0) / reps
    ks = {}
    for name in ("dense3d_count", "dense3d_write"):
        n_l, sec = C.c_int64(), C.c_double()
        if lib.rh_profile_query(name.encode(), C.byref(n_l), C.byref(sec)) == 0 and n_l.value:
            ks[name] = sec.value / reps
    lib.rh_profile_enable(0)
    head = out.download()[:64].tobytes().decode(errors="replace")
    # write kernel: reads every field value and writes the line's bytes
    read_bytes = 4 * (sum(flen) + cells) + 8 * 2 * cells
    write_s = ks.get("dense3d_write")
    res = {
        "value": 1.0 / wall, "unit": "frames/s",
        "workload": f"dense3d frame {n}^3 ({numbers} numbers: 3 f32 face arrays + f32 tracer at 10 digits, "
                    "f64 fraction + pressure at 18), device fields -> device line",
        "frame_bytes": nbytes, "ms_per_frame": wall * 1e3, "numbers_per_s": numbers / wall,
        "line_GBps": nbytes / wall / 1e9,
        "kernels_ms": {k: v * 1e3 for k, v in ks.items()},
        "head": head[:48],
    }
    if write_s:
        ach = (read_bytes + nbytes) / write_s / 1e9
        res["write_kernel_roofline"] = {"bound": "valu", "hbm_achieved_GBps": ach, "hbm_frac": ach / HBM_PEAK_GBS,
                                        "algorithmic_bytes": read_bytes + nbytes}
    if cpu:
        import oracle as O
        m = 64
        sflen = [(m + (a == 0)) * (m + (a == 1)) * (m + (a == 2)) for a in range(3)]
        srng = np.random.default_rng(64)
        sv = [srng.standard_normal(sflen[a], dtype=np.float32) for a in range(3)]
        sc = m ** 3
        args_ = (sv, srng.random(sc, dtype=np.float32), srng.random(sc), srng.standard_normal(sc))
        O.dense3d_frame((m, m, m), *args_, frame=0)  # warm-up
        times = []
        for _ in range(3):
            c0 = time.perf_counter()
            O.dense3d_frame((m, m, m), *args_, frame=0, threads=1)
            times.append(time.perf_counter() - c0)
        per_number = float(np.median(times)) / (sum(sflen) + 3 * sc)
        res["cpu_baseline"] = {"value": 1.0 / (per_number * numbers), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"{m}^3 frame (oracle: glibc exact conversion, one thread, median of 3), "
                                         f"scaled per number to {n}^3",
                               "numbers_per_s": 1.0 / per_number}
    for d in dev:
        d.free()
    out.free()
    return res


def bench_dense3d(args) -> int:
    r = run_dense3d(args.n, args.steps, cpu=not args.no_cpu_baseline)
    result = {"metric": "dense3d frames/s (write_dense3d_frame) at 256^3", "value": r["value"], "unit": "frames/s",
              "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": r["ms_per_frame"],
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64 -> u8",
              "data": "synthetic", "config": {k: v for k, v in r.items() if k not in ("value", "unit")}}
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0


def bench_batch(args) -> int:
    result = run_batch(args)
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if result.get("ok") else 1


def run_batch(args) -> dict:
    """Batched workload: 64 scenes, each axis log-uniform on [32, 128] and a
    multiple of 8 (cubic cells, longest side 1), Taylor-Green + noise with
    seed = scene id, 10 projection steps with the mixed-precision multigrid
    solve, as hipGraph segments per step (the cycles the scenes need) with
    autotuned launch shapes;
    against looping the one-scene path over the same scenes."""
    from rheon_amd.scenes import batch_workload

    specs, vels, dts = batch_workload(args.scenes)
    counts = [list(sp.counts) for sp in specs]
    cells = sum(sp.cell_len for sp in specs)
    st = R.PressureSettings(args.rel, 1e-14, 1e-3, 40)
    # graph segments: the cycles the last step needed, then one more at a
    # time while some scene is still solving, up to the solve's 40
    cycles = 40
    coarse = coarse_sweeps(args)
    batch = R.SceneBatch(specs, settings=st, use_graph=True, adaptive=True, actual_divergence_limit=1e-3,
                         max_cycles=cycles, mg_coarse_sweeps=coarse)
    for i in range(len(specs)):
        batch.scene(i).set_velocity(vels[i])
    batch.set_dt(dts)
    t0 = time.perf_counter()
    best, table = batch.autotune(steps=1)
    tune_s = time.perf_counter() - t0
    batch.step(1)  # graph of the chosen shape
    for i in range(len(specs)):
        batch.scene(i).set_velocity(vels[i])
    N.lib().rh_device_synchronize()
    t0 = time.perf_counter()
    reps = batch.step(args.steps)
    batch_s = time.perf_counter() - t0
    bad = [i for i, r in enumerate(reps) if r.status != R.OK]
    if bad:
        return {"ok": False, "failed": [(i, R.STATUS_NAMES.get(reps[i].status)) for i in bad]}
    # the one-scene path over the same scenes and steps
    singles = []
    for i, sp in enumerate(specs):
        sc = R.Scene(sp, 1.0, R.F32, "mixed-mg-dc", settings=st, actual_divergence_limit=1e-3,
                     mg_coarse_sweeps=coarse)
        sc.set_velocity(vels[i])
        sc.step(dts[i])  # warm (allocations, coarse inverse)
        sc.set_velocity(vels[i])
        singles.append(sc)
    N.lib().rh_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for i, sc in enumerate(singles):
            r = sc.step(dts[i])
            if r.status != R.OK:
                return {"ok": False, "failed_single": (i, R.STATUS_NAMES.get(r.status))}
    loop_s = time.perf_counter() - t0
    same = all(np.array_equal(batch.scene(i).get_pressure().view(np.uint64),
                              singles[i].get_pressure().view(np.uint64)) for i in range(len(specs)))
    value = cells * args.steps / batch_s
    result = {
        "metric": "batched projection steps: aggregate cell-updates/s (64 scenes x 10 steps, FP32 MG)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "f32 faces + f32 V-cycles, f64 outer residual/pressure",
        "data": "synthetic: Taylor-Green + LCG noise [-0.1, 0.1], seed = scene id",
        "config": {
            "workload": f"{len(specs)} scenes, axes log-uniform [32,128] multiple of 8, {args.steps} steps, "
                        f"mixed-mg-dc to rel {args.rel:g}",
            "total_cells": cells,
            "cycles_per_step": f"adaptive graph segments (at most {cycles})",
            "mg_coarse_sweeps": list(coarse) if args.mg_coarse_sweeps else "as finest",
            "shapes": counts,
            "batch_seconds": batch_s,
            "looped_seconds": loop_s,
            "looped_cell_updates_per_s": cells * args.steps / loop_s,
            "speedup_vs_looped": loop_s / batch_s,
            "batch_equals_looped_bitwise": same,
            "autotune": {"chosen": best, "candidates": table, "seconds": tune_s},
            "iterations_last_step": [r.pressure.iterations for r in reps],
        },
        "ok": bool(same),
    }
    for sc in singles:
        sc.close()
    batch.close()
    return result


def bench_serve(args) -> int:
    result = run_serve(args)
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if not result["config"]["failed"] else 1


def run_serve(args) -> dict:
    """Serving workload: Poisson arrivals at a fixed rate, continuous
    batching, admission by planned device bytes; per-step latency p50/p99
    and total steps/s at max|div u| <= 1e-6."""
    rng = np.random.default_rng(args.seed)
    t_gen = time.perf_counter()
    reqs = serving_requests(args.requests, args.seed)
    t_gen = time.perf_counter() - t_gen
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    server = R.ProjectionServer(settings=R.PressureSettings(1e-8, 1e-14, 5e-7, 40), actual_divergence_limit=1e-6,
                                hbm_budget=int(args.hbm_budget_gb * (1 << 30)), block_divisor=args.block_divisor)
    ids, submitted = [], 0
    ticks = []
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while submitted < len(reqs) and arrivals[submitted] <= now:
            spec, vel, dt, steps, solid, _ = reqs[submitted]
            ids.append(server.submit(spec, vel, dt, steps, solid))
            submitted += 1
        states = [server.state(i)["state"] for i in ids]
        if submitted == len(reqs) and all(s_ in ("done", "failed") for s_ in states):
            break
        if not any(s_ in ("queued", "running") for s_ in states):
            time.sleep(max(0.0, arrivals[submitted] - now))
            continue
        tk = server.tick()
        tk["t"] = time.perf_counter() - t0
        ticks.append(tk)
    makespan = time.perf_counter() - t0
    st = [server.state(i) for i in ids]
    failed = [(i, R.STATUS_NAMES.get(s_["status"])) for i, s_ in zip(ids, st) if s_["state"] != "done"]
    step_lat = np.array([tk["seconds"] for tk in ticks if tk["stepped"] > 0]) * 1e3

    def pct(key):
        v = np.array([tk[key] for tk in ticks if tk["stepped"] > 0]) * 1e3
        return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "max": float(v.max())}

    breakdown = {k: pct(k) for k in ("admit_seconds", "build_seconds", "wait_seconds", "device_seconds")}
    steps_total = int(sum(tk["stepped"] for tk in ticks))
    cells_total = int(sum(tk["cells"] for tk in ticks))
    done = [s_ for s_ in st if s_["state"] == "done"]
    wait = np.array([s_["admitted"] - s_["submitted"] for s_ in done]) * 1e3
    turn = np.array([s_["finished"] - s_["submitted"] for s_ in done])
    divs = [s_["actual_divergence_max"] for s_ in done]
    result = {
        "metric": "serving: per-step latency p50/p99 and total steps/s at max|div u| <= 1e-6",
        "value": steps_total / makespan,
        "unit": "scene-steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 faces + f32 V-cycles, f64 outer residual/pressure",
        "data": ("synthetic: 16 Fourier modes |k|<=8, max|u|=0.02, 0-3 spheres r in [0.05,0.2]; "
                 f"seed {args.seed}"),
        "config": {
            "workload": (f"{len(reqs)} requests, Poisson arrivals at {args.rate:g}/s, axes log-uniform [32,192] "
                         f"multiple of 8, run length U{{10..200}}, mixed-mg-dc rel 1e-8 + predicted div 5e-7, "
                         f"actual max|div u| gate 1e-6"),
            "hbm_budget_gb": args.hbm_budget_gb,
            "block_divisor": args.block_divisor,
            "latency_ms_p50": float(np.percentile(step_lat, 50)),
            "latency_ms_p99": float(np.percentile(step_lat, 99)),
            "latency_ms_max": float(step_lat.max()),
            "tick_breakdown_ms": breakdown,
            "cell_updates_per_s": cells_total / makespan,
            "makespan_s": makespan,
            "ticks": len(ticks),
            "mean_active": float(np.mean([tk["stepped"] for tk in ticks])),
            "max_active": int(max(tk["stepped"] for tk in ticks)),
            "max_bytes_in_use_gb": max(tk["bytes_in_use"] for tk in ticks) / (1 << 30),
            "queue_wait_ms_p50": float(np.percentile(wait, 50)) if len(wait) else None,
            "queue_wait_ms_p99": float(np.percentile(wait, 99)) if len(wait) else None,
            "turnaround_s_p50": float(np.percentile(turn, 50)) if len(turn) else None,
            "turnaround_s_p99": float(np.percentile(turn, 99)) if len(turn) else None,
            "max_div_u": max(divs) if divs else None,
            "failed": failed,
            "generation_s": t_gen,
            "spheres": int(sum(r[5] for r in reqs)),
            "scene_steps": steps_total,
        },
    }
    server.close()
    return result


def bench_config2(args) -> int:
    """128^3 FP32 drop-in provider workload (Taylor-Green + noise): Jacobi-PCG,
    red-black SGS-PCG and the 4-level FP32 multigrid, each to a relative
    residual of 1e-5; Jacobi iterates checked against the oracle's f64
    Jacobi-PCG iterates (1e-5 relative, the stated FP32 tolerance)."""
    import oracle as O
    from rheon_amd.scenes import taylor_green_noise

    spec = GridSpec.unit_box(128)
    vel = taylor_green_noise(spec, 6, 0.1, np.float32)
    dt = courant_dt(spec, vel, 0.5)
    rhs = R.build_rhs(spec, vel, dt)
    n = spec.cell_len
    out = {}
    bnorm = float(np.sqrt(rhs @ rhs))
    # the stated multigrid: 4 levels, weighted-Jacobi V(2,2); beside it the
    # default Chebyshev V(3,1) on the same 4 levels
    jac = {"mg_levels": 4, "mg_smoother": R.SMOOTHER_JACOBI, "mg_sweeps": (2, 2)}
    cheb = {"mg_levels": 4, "mg_smoother": R.SMOOTHER_CHEBYSHEV, "mg_sweeps": (3, 1)}
    cases = [("jacobi-pcg-v1", R.F32, {}, 9 * 4 + 16), ("rbsgs-pcg", R.F32, {}, 15 * 4 + 16),
             ("mixed-mg-dc", R.F32, jac, None), ("mixed-mg-dc", R.F32, cheb, None)]
    for method, prec, kw, bytes_per_cell in cases:
        ws = R.PressureWorkspace(spec, method, prec, **kw)
        entry = {}
        if bytes_per_cell:
            # The reference's termination (recursive residual below the
            # threshold -> true residual incl. the gauge row -> replace and
            # restart) stalls on this field at rel 1e-5, so the iteration
            # count is the first k whose true residual is <= 1e-5 ||b||,
            # found with the stopping tests off.
            def true_rel(k):
                r = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k, poll_interval=32))
                return r.true_residual_l2 / bnorm
            lo, hi = 1, 64
            while true_rel(hi) > 1e-5 and hi < 8192:
                lo, hi = hi, hi * 2
            while hi - lo > 1:
                mid = (lo + hi) // 2
                if true_rel(mid) > 1e-5:
                    lo = mid
                else:
                    hi = mid
            st = R.PressureSettings(0.0, 0.0, 0.0, hi, poll_interval=32)
            stalled = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-5, 1e-14, 1.0, 4 * hi, poll_interval=32))
            entry["reference_termination_at_rel_1e-5"] = {
                "status": R.STATUS_NAMES.get(stalled.status), "iterations": stalled.iterations,
                "true_rel_residual": stalled.true_residual_l2 / bnorm}
        else:
            st = R.PressureSettings(1e-5, 1e-14, 1.0, 100, poll_interval=2)
        ws.solve_rhs(1.0, rhs, dt, st)
        reps = [ws.solve_rhs(1.0, rhs, dt, st) for _ in range(3)]
        secs = float(np.median([r.seconds for r in reps]))
        it = reps[-1].iterations
        per_iter = bytes_per_cell * n if bytes_per_cell else mg_cycle_bytes(spec, 4, 0, *kw["mg_sweeps"],
                                                                             x0_implicit=True)
        entry.update({"iterations_to_rel_1e-5": it, "true_rel_residual": reps[-1].true_residual_l2 / bnorm,
                      "seconds": secs, "algorithmic_GBps": per_iter * it / secs / 1e9,
                      "cell_updates_per_s": n * it / secs})
        label = "" if not kw else ("/4-level-jacobi-V(2,2)" if kw is jac else "/4-level-chebyshev-V(3,1)")
        out[f"{method}/f32" + label] = entry
        ws.close()
    # Jacobi iterates against the oracle
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    orhs = O.build_rhs(g, vel, dt)
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F32)
    worst = 0.0
    for k in (1, 10, 50):
        ws.solve_rhs(1.0, orhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k))
        ref = O.solve_rhs(g, 1.0, orhs, dt, O.Settings(0.0, 0.0, 0.0, k), threads=min(os.cpu_count() or 1, 16))
        p = ws.get_pressure()
        worst = max(worst, float(np.max(np.abs(p - ref.pressure)) / np.max(np.abs(ref.pressure))))
    result = {
        "metric": "128^3 FP32 provider solves to rel 1e-5: seconds, iterations, HBM GB/s",
        "value": out["mixed-mg-dc/f32/4-level-jacobi-V(2,2)"]["seconds"],
        "unit": "s (4-level MG solve)",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f32 vectors, f64 scalars",
        "data": "synthetic: Taylor-Green + LCG noise 0.1, seed 6, 128^3 unit box",
        "config": {"workload": "128^3 pressure solves to rel 1e-5", "methods": out,
                   "jacobi_iterate_max_rel_err_vs_f64_oracle": worst, "iterates_checked": [1, 10, 50]},
    }
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if worst < 1e-5 else 1


def run_warm_start(spec, vel, solid, dt_requested, settings, sweeps, smoother, args, max_courant) -> dict:
    """The headline workload under AcceptedPreviousV1 (feature
    accepted-pressure-warm-start of the reference: every solve seeded with
    the last committed step's pressure): a fresh scene, the same untimed
    steps before the same number of timed ones as the cold headline.  The
    headline itself stays cold (the reference's default policy)."""
    scene = R.Scene(spec, density=1.0, precision=R.F32, method=args.method,
                    solver_precision=R.F32 if args.method == "mixed-mg-dc" else R.F64,
                    settings=settings, actual_divergence_limit=1e-3, max_courant=max_courant,
                    mg_sweeps=sweeps, mg_smoother=smoother, mg_coarse_sweeps=coarse_sweeps(args))
    if solid is not None:
        scene.set_solid(solid)
    scene.set_initialization(R.INIT_ACCEPTED_PREVIOUS)
    scene.set_velocity(vel)
    for _ in range(args.warmup + 1):  # the cold run's warm-up and survey steps
        if scene.step_requested(dt_requested).status != R.OK:
            return {"failed": "untimed step"}
    N.lib().rh_device_synchronize()
    iters, divs = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rep = scene.step_requested(dt_requested)
        if rep.status != R.OK:
            return {"failed": R.STATUS_NAMES.get(rep.status)}
        iters.append(rep.pressure.iterations)
        divs.append(rep.actual_divergence_max)
    N.lib().rh_device_synchronize()
    elapsed = time.perf_counter() - t0
    scene.close()
    return {"policy": "AcceptedPreviousV1", "steps_per_s": args.steps / elapsed,
            "ms_per_step": elapsed / args.steps * 1e3, "iterations": iters, "max_div_u": max(divs)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--method", default="mixed-mg-dc",
                    choices=["jacobi-pcg-v1", "rbsgs-pcg", "mixed-mg-dc", "mg-f64-dc"])
    ap.add_argument("--rel", type=float, default=1e-10)
    ap.add_argument("--sphere", type=float, default=0.2, help="radius of the centred solid sphere (0 = none)")
    ap.add_argument("--smoother", default="chebyshev", choices=["chebyshev", "jacobi"],
                    help="multigrid smoother (Chebyshev weights or weighted Jacobi)")
    ap.add_argument("--mg-sweeps", default="", metavar="PRE,POST",
                    help="V-cycle sweeps (default 3,1 for Chebyshev, 2,2 for Jacobi)")
    ap.add_argument("--mg-coarse-sweeps", default="", metavar="PRE,POST",
                    help="sweeps of the levels below the finest (default: the finest level's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--compare", action="store_true",
                    help="also time one solve of the step's RHS by every method and report pressure errors")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--headline-only", action="store_true",
                    help="default workload without the method comparison and the batch/serve summaries")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once with the plain copy kernel (the known-size "
                         "reference for hardware traffic counters) and exit")
    ap.add_argument("--workload", default="scene", choices=["scene", "batch", "serve", "config2", "dense3d"],
                    help="scene: the 256^3 projection step (default); batch: 64 batched scenes; "
                         "serve: the continuous-batching serving loop")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=8.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--hbm-budget-gb", type=float, default=16.0)
    ap.add_argument("--block-divisor", type=int, default=8)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py measures one GPU: a projection scene lives on one device and many scenes shard "
              "across GPUs by whole scenes (no collectives); run with --gpus 1.", file=sys.stderr)
        return 2
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "config2":
        return bench_config2(args)
    if args.workload == "dense3d":
        return bench_dense3d(args)
    if args.workload == "batch":
        if args.steps == 5 and "--steps" not in sys.argv:
            args.steps = 10
        if "--rel" not in sys.argv:
            args.rel = 1e-8
        return bench_batch(args)

    spec = GridSpec.unit_box(args.n)
    vel = fourier_modes(spec, seed=2026, modes=32, kmax=16, amplitude=1.0, dtype=np.float32)
    # Simulation::step's dt policy (src/simulation.rs:926-979): the recipe's
    # requested dt = 0.5 h / max|u|, clipped to 0.5 max_courant / rate with
    # rate = sum_d max|u_d| / h_d and max_courant = 1; the Courant gate then
    # checks every step's candidate against max_courant.
    max_courant = 1.0
    dt_requested = courant_dt(spec, vel, 0.5)
    rate0 = sum(float(np.max(np.abs(vel[d]))) / spec.spacing[d] for d in range(3))
    dt = min(dt_requested, 0.5 * max_courant / rate0)
    mg = args.method in ("mixed-mg-dc", "mg-f64-dc")
    s_vec = 4 if args.method == "mixed-mg-dc" else 8  # bytes per Krylov / V-cycle value
    settings = R.PressureSettings(args.rel, 1e-14, 1e-4, 200 if mg else 20000, poll_interval=2 if mg else 32)
    smoother = R.SMOOTHER_JACOBI if args.smoother == "jacobi" else R.SMOOTHER_CHEBYSHEV
    sweeps = tuple(int(v) for v in args.mg_sweeps.split(",")) if args.mg_sweeps else (
        (2, 2) if args.smoother == "jacobi" else (3, 1))
    scene = R.Scene(spec, density=1.0, precision=R.F32, method=args.method,
                    solver_precision=R.F32 if args.method == "mixed-mg-dc" else R.F64,
                    settings=settings, actual_divergence_limit=1e-3, max_courant=max_courant,
                    mg_sweeps=sweeps, mg_smoother=smoother, mg_coarse_sweeps=coarse_sweeps(args))
    solid = centred_sphere(spec, args.sphere) if args.sphere > 0 else None
    if solid is not None:
        scene.set_solid(solid)
    scene.set_velocity(vel)
    N.lib().rh_device_synchronize()

    for _ in range(args.warmup):
        rep = scene.step_requested(dt_requested)
        if rep.status != R.OK:
            print(f"warmup step failed: {R.STATUS_NAMES.get(rep.status)}", file=sys.stderr)
            return 1

    import ctypes as C
    lib = N.lib()

    def kernel_table():
        out = {}
        for name in list(KERNEL_BYTES_PER_CELL) + list(FACE_KERNEL_BYTES) + list(AGGREGATES):
            n_l, sec = C.c_int64(), C.c_double()
            if lib.rh_profile_query(name.encode(), C.byref(n_l), C.byref(sec)) == 0 and n_l.value:
                out[name] = (n_l.value, sec.value)
        return out

    # Survey step (untimed): every kernel bracketed by HIP events, to find
    # the dominant kernel.  Event pairs cost the stream a few microseconds
    # each, so the timed steps bracket only the dominant kernel.
    lib.rh_profile_focus(None)
    lib.rh_profile_enable(1)
    rep = scene.step_requested(dt_requested)
    if rep.status != R.OK:
        print(f"survey step failed: {R.STATUS_NAMES.get(rep.status)}", file=sys.stderr)
        return 1
    kernels = kernel_table()
    lib.rh_profile_enable(0)
    # the dominant kernel: the most device time per step among single kernels
    dom = max((k for k in kernels if k not in AGGREGATES), key=lambda k: kernels[k][1])

    lib.rh_profile_focus(dom.encode())
    lib.rh_profile_enable(1)
    lib.rh_device_synchronize()
    lat, iters, divs, dts, solve_s = [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        rep = scene.step_requested(dt_requested)
        lat.append(time.perf_counter() - s0)
        if rep.status != R.OK:
            print(f"step failed: {R.STATUS_NAMES.get(rep.status)}", file=sys.stderr)
            return 1
        iters.append(rep.pressure.iterations)
        solve_s.append(rep.pressure.seconds)
        divs.append(rep.actual_divergence_max)
        dts.append((rep.dt, rep.courant))
    lib.rh_device_synchronize()
    elapsed = time.perf_counter() - t0
    launches, secs = kernel_table()[dom]
    lib.rh_profile_enable(0)
    lib.rh_profile_focus(None)

    avg = secs / launches
    m_bytes = 1 if solid is not None else 0
    # four-cell FP32 path with >= 2 pre-sweeps: the outer residual does not store x0
    x0_implicit = args.method == "mixed-mg-dc" and sweeps[0] >= 2 and spec.counts[0] % 4 == 0

    def launch_bytes(name):
        b = kernel_bytes_per_launch(name, spec, s_vec, m_bytes)
        return b - s_vec * spec.cell_len if name == "mg_outer_residual" and x0_implicit else b

    per_launch = launch_bytes(dom)
    total_bytes = per_launch * launches
    first_launch = None
    if dom == "mg_outer_residual" and x0_implicit:
        # each solve's first residual does not read p (p = 0 is known): 8 B/cell less
        first_launch = per_launch - 8 * spec.cell_len
        total_bytes -= 8 * spec.cell_len * args.steps
    achieved = total_bytes / secs / 1e9
    # measured HBM traffic (committed counter summary) applies to the default workload only
    default_workload = (args.n == 256 and args.sphere == 0.2 and args.method == "mixed-mg-dc"
                        and args.smoother == "chebyshev" and not args.mg_sweeps and not args.mg_coarse_sweeps)
    traffic = profiled_traffic(dom) if default_workload else None

    mean_iters = float(np.mean(iters))
    ab = algorithmic_bytes(spec, 4, args.method, s_vec, int(round(mean_iters)), m_bytes, sweeps, x0_implicit,
                           coarse_sweeps(args))
    step_bytes = sum(ab.values())
    ms = elapsed / args.steps * 1e3

    result = {
        "metric": "projection steps/s + HBM GB/s at 256^3 to fixed max|div u|; p50/p99 latency",
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 faces + f32 V-cycles, f64 outer residual/pressure" if args.method == "mixed-mg-dc"
                  else "f32 faces, f64 pressure"),
        "data": ("synthetic: 32 seeded Fourier modes |k| in [1,16], amplitude ~|k|^-5/3, max|u| = 1"
                 + (f", solid sphere r={args.sphere:g} at the box centre" if solid is not None else "")),
        "config": {
            "workload": f"projection step {args.n}^3, {args.method} to rel {args.rel:g}",
            "grid": list(spec.counts),
            "method": args.method,
            "mg_smoother": (f"{args.smoother} V({sweeps[0]},{sweeps[1]})" if mg else None),
            "mg_coarse_sweeps": (list(coarse_sweeps(args)) if args.mg_coarse_sweeps else "as finest") if mg else None,
            "relative_residual": args.rel,
            "obstacle": (f"sphere r={args.sphere:g} at (0.5,0.5,0.5), {int(solid.sum())} solid cells"
                         if solid is not None else "none"),
            "dt_policy": {"requested": dt_requested, "max_courant": max_courant,
                          "dt_mean": float(np.mean([d for d, _ in dts])),
                          "courant_max": float(max(c for _, c in dts))},
            "iterations_mean": mean_iters,
            "max_div_u": max(divs),
            "latency_ms_p50": float(np.percentile(np.array(lat) * 1e3, 50)),
            "latency_ms_p99": float(np.percentile(np.array(lat) * 1e3, 99)),
            "algorithmic_GBps": step_bytes / (elapsed / args.steps) / 1e9,
            # the pressure solve alone: its compulsory bytes over its device time
            # (the solve report's seconds, HIP events around the solve's launches)
            "solve_ms_mean": float(np.mean(solve_s)) * 1e3,
            "solve_algorithmic_GBps": ab["solve"] / float(np.mean(solve_s)) / 1e9,
            "solve_hbm_frac": ab["solve"] / float(np.mean(solve_s)) / 1e9 / HBM_PEAK_GBS,
            "cell_updates_per_s": spec.cell_len * args.steps / elapsed,
            "kernels_survey_step": {
                k: {"launches": v[0], "avg_us": v[1] / v[0] * 1e6, "total_us": v[1] * 1e6,
                    **({"hbm_frac": launch_bytes(k) / (v[1] / v[0]) / 1e9 / HBM_PEAK_GBS}
                       if k not in AGGREGATES else {})}
                for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][1])},
            "advection": advection_line(spec, kernels, default_workload),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": per_launch,
            # each solve's first launch, when it reads no p (achieved weights it in)
            "algorithmic_bytes_first_launch": first_launch,
            "avg_launch_us": avg * 1e6,
            "timed_launches": launches,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the reference method's iteration count for this step (device jacobi-pcg-v1 is bit-identical)
        if args.method == "jacobi-pcg-v1":
            need = int(round(mean_iters))
        else:
            ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
            if solid is not None:
                ws.set_solid(solid)
            rhs = step_rhs(spec, scene, dt, solid)
            need = ws.solve_rhs(1.0, rhs, dt, jacobi_settings(settings)).iterations
            ws.close()
        result["cpu_baseline"] = cpu_baseline(spec, scene.get_velocity(), dt, need, solid)
    if args.compare or (default_workload and not args.headline_only):
        result["config"]["compare"] = compare_methods(spec, scene, dt, solid, settings)
    if default_workload and not args.headline_only:
        scene.close()
        result["config"]["warm_start"] = run_warm_start(spec, vel, solid, dt_requested, settings, sweeps, smoother,
                                                        args, max_courant)
        # the batched and serving workloads, summarised (each has its own
        # --workload line with every detail)
        b = run_batch(argparse.Namespace(scenes=64, steps=10, rel=1e-8, mg_coarse_sweeps="", json_out=None))
        c = b.get("config", {})
        result["batch"] = {k: v for k, v in {
            "value": b.get("value"), "unit": b.get("unit"), "workload": c.get("workload"),
            "looped_cell_updates_per_s": c.get("looped_cell_updates_per_s"),
            "speedup_vs_looped": c.get("speedup_vs_looped"),
            "batch_equals_looped_bitwise": c.get("batch_equals_looped_bitwise"),
            "autotune_chosen": c.get("autotune", {}).get("chosen"), "ok": b.get("ok"),
            "failed": b.get("failed")}.items() if v is not None}
        keys = ("latency_ms_p50", "latency_ms_p99", "latency_ms_max", "tick_breakdown_ms", "cell_updates_per_s",
                "queue_wait_ms_p50", "queue_wait_ms_p99", "max_bytes_in_use_gb", "hbm_budget_gb", "max_div_u",
                "failed", "scene_steps", "makespan_s")
        sv = run_serve(argparse.Namespace(requests=40, rate=8.0, seed=5, hbm_budget_gb=16.0, block_divisor=8,
                                          json_out=None))
        c = sv["config"]
        result["serve"] = {"value": sv["value"], "unit": sv["unit"], "workload": c["workload"],
                           **{k: c[k] for k in keys}}
        # the same stream under an HBM budget of 1.25x its largest request's
        # plan, so admission by device bytes binds (requests queue for memory)
        server = R.ProjectionServer()
        biggest = max(server.plan_bytes(r[0]) for r in serving_requests(40, 5))
        server.close()
        sv = run_serve(argparse.Namespace(requests=40, rate=8.0, seed=5, hbm_budget_gb=1.25 * biggest / (1 << 30),
                                          block_divisor=8, json_out=None))
        c = sv["config"]
        result["serve_budget_bound"] = {"value": sv["value"], "unit": sv["unit"], **{k: c[k] for k in keys}}
        # dataset output: one dense3d frame of a 256^3 state (--workload dense3d)
        d3 = run_dense3d(256, 5, cpu=not args.no_cpu_baseline)
        result["dense3d"] = {k: v for k, v in d3.items() if k != "head"}
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
