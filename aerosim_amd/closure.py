"""Learned-closure workflow on the GPU: training data from filtered 2x-resolution runs that the solver itself
generates, training with the matrix-core kernels, and the evaluation (coarse learned vs coarse
Smagorinsky against the fine-grid plain run, the reference for both).

Training has two parts. A priori: the network (a factor on the Smagorinsky eddy viscosity) is fitted to the
sub-filter stresses of the filtered fine fields of the training buildings. A posteriori: its output bias is
then offset so that coarse learned runs of the same training buildings reproduce their fine-grid Cd on
average. The evaluation uses a held-out building (the nominal size, not among the seeded training sizes).

Workload (ground boundary layer around a block building): the lattice's vertical axis is y (floor at
iy = 0), as in the reference (voxelizer.ts:101, lbmCpuReference.ts:375-401); a workload quoted as
nx x lateral x height maps to a grid (nx, height, lateral). Inflow is the log law
u(y) = u_ref ln((y + 1/2)/z0) / ln(z_ref/z0) with a no-slip floor.
"""
from dataclasses import dataclass

import numpy as np

from . import lbm as L


@dataclass
class BuildingCase:
    """Coarse-grid description; the fine grid doubles every length."""
    nx: int = 512            # streamwise
    height: int = 128        # vertical (lattice y)
    lateral: int = 256       # spanwise (lattice z)
    bx: float = 40.0         # building streamwise size (cells)
    bz: float = 20.0         # building lateral size
    bh: float = 60.0         # building height
    x_front: float = 0.25    # building front at this fraction of nx
    u_ref: float = 0.08
    z0: float = 0.5
    z_ref: float = 64.0
    re: float = 1e5
    cs: float = 0.17

    def scaled(self, f):
        return BuildingCase(int(self.nx * f), int(self.height * f), int(self.lateral * f), self.bx * f, self.bz * f,
                            self.bh * f, self.x_front, self.u_ref, self.z0 * f, self.z_ref * f, self.re, self.cs)

    @property
    def nu(self):
        return self.u_ref * self.bh / self.re

    def mesh(self, jitter=0.0, seed=0):
        """Box building standing on the floor (its base sinks one cell into the floor)."""
        rng = np.random.default_rng(seed)
        j = 1.0 + jitter * (2.0 * rng.random(3) - 1.0) if jitter > 0 else np.ones(3)
        bx, bz, bh = self.bx * j[0], self.bz * j[1], self.bh * j[2]
        x0 = self.x_front * self.nx - 0.5 * self.nx
        floor = -0.5 * self.height
        return L.mesh_box((x0, floor - 1.0, -0.5 * bz), (x0 + bx, floor + bh, 0.5 * bz)), (bx, bz, bh)

    def context(self, storage=L.AERO_STORE_FP32, jitter=0.0, seed=0, closure=None):
        tris, dims = self.mesh(jitter, seed)
        ctx = L.Context(self.nx, self.height, self.lateral, storage=storage)
        ctx.set_mesh(tris, L.Tunnel(float(self.nx), float(self.lateral), float(self.height)))
        if closure is not None:
            ctx.set_closure(closure)
        ctx.set_params(L.SolverParams(u_inlet=self.u_ref, nu=self.nu, smagorinsky_cs=self.cs, ground_noslip=True,
                                      inflow_profile=L.AERO_PROFILE_LOGLAW, z0=self.z0, z_ref=self.z_ref,
                                      closure=L.AERO_CLOSURE_MLP if closure is not None else L.AERO_CLOSURE_SMAGORINSKY))
        ctx.reset(seed, 1e-4)
        st = L.mesh_stats(tris)
        refs = dict(frontal_area=dims[1] * dims[2], char_length=dims[0], tunnel_width=float(self.lateral),
                    tunnel_height=float(self.height), dx=1.0, u_ref=self.u_ref)
        return ctx, refs, st


def generate_samples(closure, case, sizes=4, snapshots_per_size=50, coarse_steps=3000, window=600, stride=4,
                     jitter=0.1, seed=100, log=print):
    """Runs the 2x-resolution plain (Smagorinsky) case for `sizes` seeded building sizes (2 coarse_steps fine
    steps each, the evaluation's protocol) and appends, from `snapshots_per_size` snapshots spread over the last two
    thirds of each run, samples filtered onto the coarse grid. Returns the number of samples added and, per
    run, (seed, mean fine Cd over its last 2 window steps)."""
    fine = case.scaled(2)
    total_steps, fwin = 2 * coarse_steps, 2 * window
    first = total_steps // 3
    snaps = [first + (total_steps - first) * (k + 1) // snapshots_per_size for k in range(snapshots_per_size)]
    forces = list(range(total_steps - fwin + 2, total_steps + 1, 2))
    events = sorted(set(snaps) | set(forces))
    total, runs = 0, []
    for s in range(sizes):
        ctx, refs, _ = fine.context(jitter=jitter, seed=seed + s)
        t, cds = 0, []
        for k, e in enumerate(events):
            snap = e in snaps
            ctx.step(e - t, readout=snap)
            t = e
            if snap:
                total += closure.sample_fine(ctx, stride, seed * 7919 + s * 131 + k, case.nu, case.cs)
            if e in forces:
                cds.append(ctx.read_forces(refs)["cd"])
        ctx.close()
        runs.append((seed + s, float(np.mean(cds))))
        log(f"fine run {s + 1}/{sizes}: {total} samples, fine Cd {runs[-1][1]:.4f}")
    return total, runs


def calibrate_bias(closure, case, runs, coarse_steps=3000, window=600, jitter=0.1, log=print, max_evals=8):
    """A-posteriori step of the training: finds an offset of the network's output bias b3 at which the coarse
    learned runs of the training buildings `runs` [(seed, fine Cd)] match their fine-grid Cd on average (mean
    signed relative Cd error zero), by bracketing and regula falsi on the offset (the error grows with it:
    a larger output means a larger eddy viscosity). Leaves the calibrated parameters in the closure."""
    p0 = closure.params()
    history = []

    def err(delta):
        p = p0.copy()
        p[-1] += delta
        closure.set_params(p)
        e = []
        for sd, cd_f in runs:
            cd_l, _ = run_coefficients(case, coarse_steps, window, closure=closure, seed=sd, jitter=jitter)
            e.append((cd_l - cd_f) / cd_f)
        history.append((float(delta), float(np.mean(e))))
        log(f"bias offset {delta:+.4f}: mean signed Cd error {np.mean(e) * 100:+.3f} %")
        return float(np.mean(e))

    a, ea = 0.0, err(0.0)
    step = 1.5 if ea < 0 else -1.5
    b, eb = a + step, err(a + step)
    while ea * eb > 0 and len(history) < max_evals:
        a, ea = b, eb
        b, eb = b + step, err(b + step)
    while ea * eb <= 0 and len(history) < max_evals and abs(eb) > 1e-3:
        c = b - eb * (b - a) / (eb - ea) if eb != ea else 0.5 * (a + b)
        ec = err(c)
        if ea * ec <= 0:
            b, eb = c, ec
        else:
            a, ea = c, ec
    best = min(history, key=lambda h: abs(h[1]))
    p = p0.copy()
    p[-1] += best[0]
    closure.set_params(p)
    return {"offset": best[0], "train_mean_signed_err": best[1], "history": history}


def shuffle_samples(closure, seed=0):
    X, T = closure.samples()
    p = np.random.default_rng(seed).permutation(len(T))
    closure.clear_samples()
    closure.add_samples(X[p], T[p])
    return X[p], T[p]


def run_coefficients(case, steps, window, closure=None, seed=0, jitter=0.0, sample_every=1):
    """Runs a case and returns mean Cd and Cl (the uplift coefficient) over the last `window` steps."""
    ctx, refs, _ = case.context(jitter=jitter, seed=seed, closure=closure)
    ctx.step(steps - window)
    cd, cl = [], []
    for k in range(0, window, sample_every):
        ctx.step(sample_every)
        f = ctx.read_forces(refs)
        cd.append(f["cd"])
        cl.append(f["cl"])
    ctx.close()
    return float(np.mean(cd)), float(np.mean(cl))


def evaluate(case, closure, coarse_steps, window, seed=999, jitter=0.0):
    """Cd/Cl of the fine plain run (the reference) and of the coarse Smagorinsky and coarse learned runs."""
    fine = case.scaled(2)
    cd_f, cl_f = run_coefficients(fine, 2 * coarse_steps, 2 * window, seed=seed, jitter=jitter, sample_every=2)
    cd_s, cl_s = run_coefficients(case, coarse_steps, window, seed=seed, jitter=jitter)
    cd_l, cl_l = run_coefficients(case, coarse_steps, window, closure=closure, seed=seed, jitter=jitter)
    return {"fine": {"cd": cd_f, "cl": cl_f}, "smagorinsky": {"cd": cd_s, "cl": cl_s}, "learned": {"cd": cd_l, "cl": cl_l},
            "err_cd_smagorinsky": abs(cd_s - cd_f) / abs(cd_f), "err_cd_learned": abs(cd_l - cd_f) / abs(cd_f),
            "err_cl_smagorinsky": abs(cl_s - cl_f) / max(1e-9, abs(cl_f)),
            "err_cl_learned": abs(cl_l - cl_f) / max(1e-9, abs(cl_f))}


def cost(case, closure, steps=100):
    """Milliseconds per coarse step with the Smagorinsky closure and with the fused learned closure."""
    out = {}
    for name, cl in (("smagorinsky", None), ("learned", closure)):
        ctx, _, _ = case.context(closure=cl)
        ctx.step(10)
        t = ctx.time_steps(steps)
        out[name] = {"ms_per_step": t["step_ms"], "mlups": case.nx * case.height * case.lateral / (t["step_ms"] * 1e-3) / 1e6}
        ctx.close()
    return out


def protocol_check(case, coarse_steps=24000, every=20, buildings=((999, 0.0), (100, 0.1)), log=print):
    """Coarse and fine Smagorinsky runs of the given (seed, jitter) buildings for `coarse_steps` coarse steps (twice
    as many fine steps), Cd sampled every `every` coarse-equivalent steps, and the coarse/fine mean Cd over the
    last 3000, 6000 and 12000 steps. This sizes the evaluation's averaging: a flow-through of the 512-cell domain is
    6400 coarse steps, and the building sheds on the order of a few thousand."""
    out = {}
    for seed, jitter in buildings:
        rec = {}
        for tag, cs, steps, ev in (("coarse", case, coarse_steps, every), ("fine", case.scaled(2), 2 * coarse_steps,
                                                                           2 * every)):
            ctx, refs, _ = cs.context(jitter=jitter, seed=seed)
            cd = []
            for k in range(steps // ev):
                ctx.step(ev)
                cd.append(ctx.read_forces(refs)["cd"])
                if (k + 1) % max(1, steps // ev // 4) == 0:
                    log(f"{tag} building {seed}: {(k + 1) * ev} steps, Cd {cd[-1]:.4f}")
            ctx.close()
            rec[tag] = np.array(cd)
        windows = {}
        for w in (3000, 6000, 12000):
            n = w // every
            c, f = float(rec["coarse"][-n:].mean()), float(rec["fine"][-n:].mean())
            windows[str(w)] = {"coarse": c, "fine": f, "err": c / f - 1.0}
        blocks = [[float(rec[t][-(4 - i) * (3000 // every):len(rec[t]) - (3 - i) * (3000 // every)].mean())
                   for t in ("coarse", "fine")] for i in range(4)]
        out[str(seed)] = {"jitter": jitter, "windows": windows, "blocks_of_3000": blocks}
        log(f"building {seed}: " + ", ".join(f"last {w}: coarse {v['coarse']:.4f} fine {v['fine']:.4f} "
                                              f"({v['err'] * 100:+.2f} %)" for w, v in windows.items()))
    return out


def main():
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="Learned-closure workflow: samples from filtered 2x runs, training "
                                             "on the GPU, and the comparison against the fine-grid plain run.")
    ap.add_argument("--scale", type=float, default=1.0, help="linear scale of the 512x256x128 building case")
    ap.add_argument("--sizes", type=int, default=4, help="seeded building sizes of the fine runs")
    ap.add_argument("--snapshots", type=int, default=50, help="snapshots per fine run")
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--coarse-steps", type=int, default=3000)
    ap.add_argument("--window", type=int, default=600)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--protocol-check", type=int, default=0, metavar="STEPS",
                    help="only run the long coarse/fine Smagorinsky statistics over STEPS coarse steps")
    ap.add_argument("--cost", type=int, default=0, metavar="STEPS",
                    help="only time STEPS coarse steps with Smagorinsky and with the fused closure (a seeded "
                         "random network: the cost does not depend on the weights)")
    args = ap.parse_args()
    case = BuildingCase().scaled(args.scale) if args.scale != 1.0 else BuildingCase()
    log = lambda m: print(m, flush=True)  # noqa: E731
    if args.protocol_check:
        print(json.dumps(protocol_check(case, args.protocol_check, log=log)))
        return
    if args.cost:
        rng = np.random.default_rng(args.seed)
        net = L.Closure(params=(rng.standard_normal(L.MLP_NPARAM) * 0.3).astype(np.float32))
        print(json.dumps({"case": f"{case.nx}x{case.height}x{case.lateral}", "steps": args.cost,
                          "cost": cost(case, net, steps=args.cost)}))
        return
    closure = L.Closure()
    t0 = time.perf_counter()
    n, runs = generate_samples(closure, case, sizes=args.sizes, snapshots_per_size=args.snapshots,
                               coarse_steps=args.coarse_steps, window=args.window, seed=args.seed, log=log)
    t1 = time.perf_counter()
    X, T = shuffle_samples(closure, args.seed)
    losses = closure.train(args.iters, batch=4096, lr=args.lr)
    log(f"trained (a priori): loss {float(losses[0]):.4g} -> {float(np.mean(losses[-50:])):.4g}")
    a_priori = closure.params()
    cal = calibrate_bias(closure, case, runs, args.coarse_steps, args.window, log=log)
    t2 = time.perf_counter()
    res = evaluate(case, closure, args.coarse_steps, args.window)
    # the a-priori network alone on the held-out building, for the record
    calibrated = closure.params()
    closure.set_params(a_priori)
    cd_a, cl_a = run_coefficients(case, args.coarse_steps, args.window, closure=closure, seed=999)
    closure.set_params(calibrated)
    res["a_priori_only"] = {"cd": cd_a, "cl": cl_a, "err_cd": abs(cd_a - res["fine"]["cd"]) / abs(res["fine"]["cd"])}
    t3 = time.perf_counter()
    c = cost(case, closure)
    out = {"case": {"nx": case.nx, "height": case.height, "lateral": case.lateral, "building": [case.bx, case.bz, case.bh],
                    "re": case.re, "u_ref": case.u_ref, "z0": case.z0, "z_ref": case.z_ref},
           "samples": n, "snapshots": args.sizes * args.snapshots, "target_mean": float(np.mean(T)),
           "target_std": float(np.std(T)), "loss_first": float(losses[0]), "loss_last50": float(np.mean(losses[-50:])),
           "iters": args.iters, "training_runs": [{"seed": sd, "fine_cd": cd} for sd, cd in runs],
           "calibration": cal, "evaluation": res, "cost": c,
           "seconds": {"samples": t1 - t0, "training": t2 - t1, "evaluation": t3 - t2}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
