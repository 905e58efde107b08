"""Throughput benchmark of the D3Q19 wind-tunnel step on one MI355X.

Workload (default): the plain FP32 form at 512x256x256 cells (33.5 M), finned rocket of ~20k triangles
(cylinder L=160, D=24 cells, tangent-ogive nose 2.5 D, 4 trapezoidal fins, seeded +-2 % jitter),
inflow u = 0.08, Re = 2e4 (nu = u D / Re), Smagorinsky Cs = 0.17, seeded 1e-4 velocity noise.
A "step" is one full solver step over the grid: the bulk stream-collide kernel, the special-cell kernel
(face rows and cells next to the body), then the mass-anchor reduction. Prints one JSON line.

The roofline object describes the dominant kernel (the bulk stream-collide kernel): algorithmic bytes per
launch = bytes per cell update (152 B FP32, 76 B FP16: 19 populations read and written) x the cells that
kernel updates, divided by its mean duration from HIP events recorded on the context's stream around it.
`traffic` is the HBM bytes per launch of that kernel from the rocprofv3 FETCH_SIZE and WRITE_SIZE
counters, scaled by the counters' reading for a plain copy of known size (`--copy-calibration` runs that
copy), when the profile summary for the workload is present.

`--readouts` runs no throughput measurement: it times the derived-field readouts on the same workload after
`--steps` steps (the last with a readout), each repeated 20 times: the Q-criterion without the host copy of its
fields (maxima only) and with it, the RGBA slice on all three axes, and a 4096-point probe. It prints one JSON
object; for per-kernel durations run it under `rocprofv3 --kernel-trace --stats`.

`--cd-ensemble SEEDS` runs no throughput measurement: it runs the workload for 20 000 steps per noise seed in
FP32 and FP16 storage and prints the mean Cd of each run over the developed-wake window, with the ensemble
statistics of the FP16 - FP32 difference.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {0: 152.0, 1: 76.0}  # 19 populations read + written, FP32 / FP16 storage


def rocket_workload(nx, ny, nz, scale, seed=7):
    import aerosim_amd as A

    D, L, Ln = 24.0 * scale, 160.0 * scale, 60.0 * scale
    tip = (100.0 * scale - 0.5 * nx, 0.0, 0.0)
    tris = A.mesh_rocket(tip, L, D, Ln, 0.02, seed, 128)
    return tris, D


def dcd_vs_fp64(storage_name):
    """|dCd| of the GPU run against the FP64 oracle's committed Cd (tests/golden): the plain FP32 workload's
    128x64x64 downscale, or for FP16 storage the precision form's 256x128x128 comparison; mean over the last
    400 of 2000 steps. Runs after the timed region."""
    import aerosim_amd as A

    name = "rocket256_cd_fp64.json" if storage_name == "fp16" else "rocket128_cd_fp64.json"
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        fx = json.load(fh)
    c, refs = fx["case"], fx["refs"]
    tris, D = rocket_workload(c["nx"], c["ny"], c["nz"], c["scale"])  # the mesh the fixture's oracle run used
    storage = A.AERO_STORE_FP16 if storage_name == "fp16" else A.AERO_STORE_FP32
    ctx = A.Context(c["nx"], c["ny"], c["nz"], storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(c["nx"]), float(c["nz"]), float(c["ny"])))
    ctx.set_params(A.SolverParams(u_inlet=c["u"], nu=c["u"] * D / c["re"], smagorinsky_cs=c["cs"]))
    ctx.reset(c["seed"], c["amp"])
    ctx.step(c["steps"] - c["window"])
    cd = 0.0
    for _ in range(c["window"]):
        ctx.step(1)
        cd += ctx.read_forces(refs)["cd"]
    ctx.close()
    cd /= c["window"]
    ref = fx["mean"]["cd"]
    return {"value": round(abs(cd - ref) / abs(ref) * 100, 4), "unit": "%", "gpu_cd": round(cd, 6),
            "fp64_cd": round(ref, 6),
            "case": f"rocket {c['nx']}x{c['ny']}x{c['nz']}, {storage_name} populations, mean Cd of steps "
                    f"{c['steps'] - c['window'] + 1}-{c['steps']} against tests/golden/{name}"}


def cpu_baseline(seconds=15.0):
    """The oracle in reference-faithful FP32-storage mode on the host cores, on the 128x64x64 downscale of
    the same rocket case; bounded to about `seconds` of stepping."""
    import numpy as np  # noqa: F401
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    tris, D = rocket_workload(128, 64, 64, 0.25)
    vox = O.Voxels(tris, 64.0, (128, 64, 64))
    u = 0.08
    s = O.Solver(vox, "ref32", u_inlet=u, nu=u * D / 2e4, cs=0.17)
    s.reset(7, 1e-4)
    s.step(2)  # warm caches and thread pool
    cells = 128 * 64 * 64
    steps, t0 = 0, time.perf_counter()
    while True:
        s.step(5)
        steps += 5
        dt = time.perf_counter() - t0
        if dt >= seconds or steps >= 2000:
            break
    threads = O.max_threads()
    value = cells * steps / dt / 1e6
    # one thread, as the reference's single-threaded JavaScript solver runs (a third of the time budget)
    O.set_threads(1)
    s.step(1)
    steps1, t0 = 0, time.perf_counter()
    while True:
        s.step(1)
        steps1 += 1
        dt1 = time.perf_counter() - t0
        if dt1 >= seconds / 3 or steps1 >= 400:
            break
    O.set_threads(threads)
    return {"value": value, "unit": "MLUPS", "cores": threads, "kind": "port",
            "value_1_thread": cells * steps1 / dt1 / 1e6, "cpu_model": cpu_model(),
            "sample": f"oracle ref32 (reference-faithful FP32 storage, C++ restatement of lbmCpuReference.ts), "
                      f"rocket 128x64x64 downscale, {steps} steps in {dt:.1f} s on {threads} threads (OpenMP over "
                      f"z-planes), {steps1} steps in {dt1:.1f} s on 1 thread"}


def copy_calibration(mib=1024, reps=5):
    """Plain device-to-device copies of `mib` MiB (the byte counters' calibration kernel)."""
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4, "reps": reps}))


def cd_ensemble(seeds, steps, storage_names=("fp32", "fp16"), block=2000, every=5):
    """Mean Cd of the 512x256x256 workload over a long developed-wake window, for several noise seeds per
    storage type: Cd is sampled every `every` steps and averaged per `block` steps; a run's mean is over the
    last 60 % of its blocks. Prints one JSON line with every run's block means, each storage type's ensemble
    mean and run-to-run standard deviation, and the FP16 - FP32 difference of the ensemble means with its
    standard error."""
    import numpy as np

    import aerosim_amd as A

    nx, ny, nz = 512, 256, 256
    tris, D = rocket_workload(nx, ny, nz, 1.0)
    st = A.mesh_stats(tris)
    u = 0.08
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=float(nz),
                tunnel_height=float(ny), dx=1.0, u_ref=u)
    nblk = steps // block
    first = nblk - (6 * nblk) // 10
    out = {"grid": "512x256x256", "steps": nblk * block, "block": block, "sample_every": every,
           "window": [first * block + 1, nblk * block], "runs": {}, "ensemble": {}}
    for name in storage_names:
        storage = A.AERO_STORE_FP16 if name == "fp16" else A.AERO_STORE_FP32
        means = []
        for seed in seeds:
            ctx = A.Context(nx, ny, nz, storage=storage)
            ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
            ctx.set_params(A.SolverParams(u_inlet=u, nu=u * D / 2e4, smagorinsky_cs=0.17))
            ctx.reset(seed, 1e-4)
            blocks = []
            for _ in range(nblk):
                cds = []
                for _ in range(block // every):
                    ctx.step(every)
                    cds.append(ctx.read_forces(refs)["cd"])
                blocks.append(float(np.mean(cds)))
            ctx.step(1, readout=True)
            ms = ctx.mass_stats()
            ctx.close()
            if ms["has_nan"]:
                raise RuntimeError(f"{name} seed {seed}: NaN in the populations")
            means.append(float(np.mean(blocks[first:])))
            out["runs"][f"{name}_seed{seed}"] = [round(b, 6) for b in blocks]
            print(f"{name} seed {seed}: window mean Cd {means[-1]:.5f}", file=sys.stderr, flush=True)
        m = np.array(means)
        out["ensemble"][name] = {"mean": float(m.mean()), "sd": float(m.std(ddof=1)) if len(m) > 1 else None}
    if len(storage_names) == 2 and len(seeds) > 1:
        a, b = (np.array([np.mean(out["runs"][f"{n}_seed{s}"][first:]) for s in seeds]) for n in storage_names)
        ref = a.mean()
        out["dcd_fp16_vs_fp32"] = float((b.mean() - ref) / ref)
        out["dcd_standard_error"] = float(np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b)) / ref)
    print(json.dumps(out))


# Algorithmic HBM bytes per cell of the Q-criterion kernel: ux, uy, uz and the flag read once (13 B), Q, |omega|
# and the 3-float vorticity written (20 B).
Q_BYTES_PER_CELL = 33


def readout_timings(nx, ny, nz, steps, repeat=20):
    import numpy as np

    import aerosim_amd as A

    def timed(fn, reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps * 1e3

    tris, D = rocket_workload(nx, ny, nz, nx / 512.0)
    ctx = A.Context(nx, ny, nz)
    ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
    u = 0.08
    ctx.set_params(A.SolverParams(u_inlet=u, nu=u * D / 2e4, smagorinsky_cs=0.17))
    ctx.reset(7, 1e-4)
    ctx.step(steps, readout=True)
    ctx.sync()
    print(f"stepped {steps} on {nx}x{ny}x{nz}", file=sys.stderr, flush=True)
    N = nx * ny * nz
    res = {"mode": "readouts", "grid": [nx, ny, nz], "cells": N, "steps": steps, "repeat": repeat}
    q_ms = timed(lambda: ctx.compute_q_criterion(1.0, fields=False), repeat)
    res["qcriterion_ms"] = q_ms
    res["qcriterion_gbps_wall"] = N * Q_BYTES_PER_CELL / (q_ms * 1e-3) / 1e9
    res["qcriterion_with_copy_ms"] = timed(lambda: ctx.compute_q_criterion(1.0, fields=True), max(2, repeat // 5))
    r = ctx.compute_q_criterion(1.0, fields=False)
    res["max_q"], res["max_omega"] = r["max_q"], r["max_omega"]
    for axis in ("z", "y", "x"):
        res[f"slice_{axis}_ms"] = timed(lambda: ctx.extract_slice_rgba(axis, 0.5, "vorticity", "cpu", u), repeat)
    pts = np.random.default_rng(1).uniform(-0.5, 0.5, (4096, 3)) * np.array([nx, ny, nz])
    res["probe_4096_ms"] = timed(lambda: ctx.probe(pts, (-0.5 * nx, -0.5 * ny, -0.5 * nz), 1.0, 25.0, u), repeat)
    ctx.close()
    print(json.dumps(res), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measure_form(args, storage_name, with_frames=False):
    """Times `args.steps` steps of the production path (aero_step) for one storage form of the workload, after
    `args.warmup` untimed ones, with the bulk kernel's roofline from HIP events. Returns (line fields, context)."""
    import aerosim_amd as A

    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    scale = nx / 512.0
    storage = A.AERO_STORE_FP16 if storage_name == "fp16" else A.AERO_STORE_FP32
    tris, D = rocket_workload(nx, ny, nz, scale)
    u, Re, cs = 0.08, 2e4, 0.17
    ctx = A.Context(nx, ny, nz, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
    vox_ms = ctx.read_voxels(float(ny)).voxelization_time_ms if nx * ny * nz <= 64 * 1024 * 1024 else None
    ctx.set_params(A.SolverParams(u_inlet=u, nu=u * D / Re, smagorinsky_cs=cs))
    ctx.reset(7, 1e-4)
    if args.cells_per_thread:
        ctx.set_cells_per_thread(args.cells_per_thread)
    split = ctx.split_info()
    N = nx * ny * nz
    if args.warmup:
        ctx.step(args.warmup)
    ctx.sync()
    # the timed region: exactly `steps` steps of the production path (aero_step), synchronised on both sides
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    mlups = N / (ms_step * 1e-3) / 1e6
    # per-kernel durations from HIP events on the library's stream (a separate, shorter run)
    tm = ctx.time_steps(min(args.steps, 100))
    bulk_ms = tm["bulk_ms"]
    bpc = BYTES_PER_CELL[storage]
    bulk_bytes = bpc * split["bulk_cells"]
    achieved = bulk_bytes / (bulk_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"step_bulk_{args.grid}_{storage_name}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "value": round(mlups, 1),
        "ms_per_step": round(ms_step, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_step_bulk", "bytes_per_cell": bpc, "cells_per_launch": split["bulk_cells"],
                     "algorithmic_bytes_per_launch": bulk_bytes, "kernel_ms": round(bulk_ms, 4)},
        # the timed region replays captured step pairs; the eager_* figures come from separately launched,
        # event-timed steps, where the listed cells finish long before the bulk kernel and the time after it
        # is the side stream's join, the reduction kernels and the gaps between launches
        "step_breakdown_ms": {"bulk": round(bulk_ms, 4), "outside_bulk": round(ms_step - bulk_ms, 4),
                              "eager_step": round(tm["step_ms"], 4),
                              "eager_bulk_end_to_side_join": round(tm["after_bulk_ms"], 4),
                              "eager_reduction_and_gaps": round(tm["step_ms"] - tm["kernel_ms"], 4),
                              "special_cells": split["special_cells"], "face_cells": split["face_cells"],
                              "cells_per_thread": split["cells_per_thread"]},
        "step_hbm_gbs": round(bpc * N / (ms_step * 1e-3) / 1e9, 1),
        "wall_s_timed_region": round(wall, 3),
        "voxelize_ms": None if vox_ms is None else round(vox_ms, 2),
        "triangles": int(tris.shape[0]),
    }
    if with_frames:
        out["frame_loop"] = frame_loop(ctx, N, ms_step)
    return out, ctx


def frame_loop(ctx, N, ms_step, frames=25):
    """The reference's frame (main.ts:490-512): 4 substeps, the last writing the macro fields, then the force
    readout and the slice image. Times frames against 4 plain steps, each synchronised, and the step with the
    macro readout against a plain step."""
    def run(fn):
        for _ in range(2):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(frames):
            fn()
        return (time.perf_counter() - t0) * 1e3 / frames

    def plain():
        ctx.step(4)
        ctx.sync()

    def macros():
        ctx.step(4, macros=True)
        ctx.sync()

    def frame():
        ctx.step(4, macros=True)
        ctx.read_forces()
        ctx.extract_slice_rgba("z", 0.5, "vorticity")

    t_plain, t_mac, t_frame = run(plain), run(macros), run(frame)
    readout_ms = ms_step + (t_mac - t_plain)
    return {"substeps": 4, "ms_4_plain_steps": round(t_plain, 4), "ms_4_steps_macro_readout": round(t_mac, 4),
            "ms_frame_with_forces_and_slice": round(t_frame, 4),
            "readout_step_vs_plain_step": round(readout_ms / ms_step, 4),
            "frame_mlups": round(N * 4 / (t_frame * 1e-3) / 1e6, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--storage", choices=["fp32", "fp16"], default="fp32")
    ap.add_argument("--grid", default="512x256x256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cells-per-thread", type=int, default=0, help="bulk kernel cells per thread (0: default)")
    ap.add_argument("--copy-calibration", action="store_true", help="only run the counter-calibration copy")
    ap.add_argument("--no-dcd", action="store_true", help="skip the |dCd| runs against the FP64 fixtures")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-store form beside the FP32 line")
    ap.add_argument("--no-frames", action="store_true", help="skip the frame-loop timing")
    ap.add_argument("--readouts", action="store_true", help="only time the derived-field readouts")
    ap.add_argument("--cd-ensemble", default="", metavar="SEEDS",
                    help="only run the long-window Cd ensemble, FP32 and FP16, for these comma-separated seeds")
    ap.add_argument("--cd-ensemble-steps", type=int, default=20000)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.readouts:
        nx, ny, nz = (int(v) for v in args.grid.split("x"))
        readout_timings(nx, ny, nz, args.steps)
        return
    if args.cd_ensemble:
        cd_ensemble([int(v) for v in args.cd_ensemble.split(",")], args.cd_ensemble_steps)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver runs each tunnel on a single MI355X; splitting a "
              "tunnel into slabs over several GPUs is designed but not implemented.", file=sys.stderr)
        sys.exit(2)

    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    N = nx * ny * nz
    main_form, ctx = measure_form(args, args.storage, with_frames=not args.no_frames)
    # sanity: a NaN field would make the throughput meaningless
    ctx.step(1, macros=True)
    stats = ctx.mass_stats()
    ctx.close()
    line = {
        "metric": "MLUPS & % HBM roofline, FP32 vs FP16-store D3Q19; |dCd| vs FP64 oracle",
        "value": main_form.pop("value"),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_form.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else "f16-store/f32-math",
        "data": "synthetic (generated rocket mesh, seeded 1e-4 velocity noise)",
        "config": {"workload": f"D3Q19 rocket tunnel {args.grid} cells, {args.storage} populations",
                   "cells": N, "u_inlet": 0.08, "Re": 2e4, "smagorinsky_cs": 0.17,
                   "triangles": main_form.pop("triangles"), "parallelism": "single GPU"},
    }
    line.update(main_form)
    line["finite_after_run"] = not stats["has_nan"]
    line["mass_stats_after_run"] = {"mean_rho": round(stats["mean_rho"], 8), "max_speed": round(stats["max_speed"], 6)}
    if not args.no_dcd:
        line["dcd_vs_fp64"] = dcd_vs_fp64(args.storage)
    if args.storage == "fp32" and not args.no_fp16:
        # the FP16 step varies more from launch to launch (its listed cells on the side stream sometimes end after
        # the bulk kernel), so its line times at least 200 steps
        a16 = argparse.Namespace(**vars(args))
        a16.steps, a16.warmup = max(args.steps, 200), max(args.warmup, 20)
        f16, c16 = measure_form(a16, "fp16")
        f16["steps"], f16["warmup"] = a16.steps, a16.warmup
        c16.step(1, macros=True)
        s16 = c16.mass_stats()
        c16.close()
        f16.pop("triangles")
        f16["finite_after_run"] = not s16["has_nan"]
        f16["vs_fp32"] = round(f16["value"] / line["value"], 4)
        f16["dtype"] = "f16-store/f32-math"
        if not args.no_dcd:
            f16["dcd_vs_fp64"] = dcd_vs_fp64("fp16")
        line["fp16"] = f16
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
