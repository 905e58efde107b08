"""Learned-closure workflow on the GPU: training data from filtered 2x-resolution runs that the solver itself
generates, training with the matrix-core kernels, and the evaluation (coarse learned vs coarse
Smagorinsky against the fine-grid plain run, the reference for both).

Training: the network (a factor on the Smagorinsky eddy viscosity) is fitted to the sub-filter stresses of the
filtered fine fields of seeded training buildings. The evaluation uses the nominal 40 x 20 x 60 building, which is
not among the training sizes.

Protocol. Runs are long enough for the wake to develop: a flow-through of the 512-cell domain is 6400 coarse steps,
and the coefficients are averaged over the last 9600 of 16 000 coarse steps (twice as many fine steps). Building
sizes are whole cells (the lateral size an even number, the building being centred) so that every face lies on a
cell boundary of both grids, as the nominal building's do: a face between cells changes the Bouzidi distances and,
measured, the sign of the coarse-grid error.

Workload (ground boundary layer around a block building): the lattice's vertical axis is y (floor at iy = 0), as in
the reference (voxelizer.ts:101, lbmCpuReference.ts:375-401); a workload quoted as nx x lateral x height maps to a
grid (nx, height, lateral). Inflow is the log law u(y) = u_ref ln((y + 1/2)/z0) / ln(z_ref/z0) with a no-slip floor.
"""
import dataclasses
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

    def mesh(self):
        """Box building standing on the floor (its base sinks one cell into the floor)."""
        bx, bz, bh = self.bx, self.bz, self.bh
        x0 = self.x_front * self.nx - 0.5 * self.nx
        floor = -0.5 * self.height
        return L.mesh_box((x0, floor - 1.0, -0.5 * bz), (x0 + bx, floor + bh, 0.5 * bz)), (bx, bz, bh)

    def context(self, storage=L.AERO_STORE_FP32, seed=0, closure=None):
        """A context holding the case, reset with 1e-4 velocity noise of `seed`; also the force references (frontal
        area and length of the building) and the mesh statistics."""
        tris, dims = self.mesh()
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


COARSE_STEPS = 16000   # 2.5 flow-throughs of the 512-cell domain
WINDOW = 9600          # coarse steps averaged at the end of a run
NOMINAL = (40, 20, 60)


def seeded_sizes(n, seed=100, jitter=0.1, base=NOMINAL):
    """n distinct building sizes (streamwise, lateral, height) within +-jitter of `base`, each face on a cell
    boundary of both grids (whole cells; an even lateral size), none equal to `base`."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        j = 1.0 + jitter * (2.0 * rng.random(3) - 1.0)
        size = (int(round(base[0] * j[0])), 2 * int(round(base[1] * j[1] / 2)), int(round(base[2] * j[2])))
        if size != tuple(base) and size not in out:
            out.append(size)
    return out


def sized(case, size):
    """The case with building size (streamwise, lateral, height) in coarse cells."""
    return dataclasses.replace(case, bx=float(size[0]), bz=float(size[1]), bh=float(size[2]))


def generate_samples(closure, case, sizes, snapshots_per_size=50, coarse_steps=COARSE_STEPS, window=WINDOW,
                     stride=4, log=print):
    """Runs the 2x-resolution plain (Smagorinsky) case for each building size of `sizes` (2 coarse_steps fine steps,
    the evaluation's protocol) and appends, from `snapshots_per_size` snapshots spread over the last two thirds of
    each run, samples filtered onto the coarse grid. Returns the number of samples added and, per run, the size and
    the mean fine Cd and Cl over its last 2 window steps (sampled every 40 fine steps)."""
    total_steps, fwin = 2 * coarse_steps, 2 * window
    first = total_steps // 3
    snaps = [first + (total_steps - first) * (k + 1) // snapshots_per_size for k in range(snapshots_per_size)]
    forces = list(range(total_steps - fwin + 40, total_steps + 1, 40))
    events = sorted(set(snaps) | set(forces))
    total, runs = 0, []
    for s, size in enumerate(sizes):
        fine = sized(case, size).scaled(2)
        ctx, refs, _ = fine.context(seed=1)
        t, cds, cls = 0, [], []
        for k, e in enumerate(events):
            snap = e in snaps
            ctx.step(e - t, readout=snap)
            t = e
            if snap:
                total += closure.sample_fine(ctx, stride, 7919 * s + k, case.nu, case.cs)
            if e in forces:
                f = ctx.read_forces(refs)
                cds.append(f["cd"])
                cls.append(f["cl"])
            if (k + 1) % max(1, len(events) // 4) == 0:
                log(f"fine run {s + 1}/{len(sizes)} {size}: step {e}, {total} samples")
        ctx.close()
        runs.append({"size": list(size), "fine_cd": float(np.mean(cds)), "fine_cl": float(np.mean(cls))})
        log(f"fine run {s + 1}/{len(sizes)} {size}: fine Cd {runs[-1]['fine_cd']:.4f} Cl {runs[-1]['fine_cl']:.4f}")
    return total, runs


def shuffle_samples(closure, seed=0):
    X, T = closure.samples()
    p = np.random.default_rng(seed).permutation(len(T))
    closure.clear_samples()
    closure.add_samples(X[p], T[p])
    return X[p], T[p]


def run_coefficients(case, steps=COARSE_STEPS, window=WINDOW, closure=None, seed=1, sample_every=20):
    """Runs a case and returns mean Cd and Cl (the uplift coefficient) over the last `window` steps, sampled every
    `sample_every` steps."""
    ctx, refs, _ = case.context(seed=seed, closure=closure)
    ctx.step(steps - window)
    cd, cl = [], []
    for k in range(0, window, sample_every):
        ctx.step(sample_every)
        f = ctx.read_forces(refs)
        cd.append(f["cd"])
        cl.append(f["cl"])
    ctx.close()
    return float(np.mean(cd)), float(np.mean(cl))


def evaluate(case, closure, coarse_steps=COARSE_STEPS, window=WINDOW, seed=1):
    """Cd/Cl of the fine plain run (the reference) and of the coarse Smagorinsky and coarse learned runs."""
    fine = case.scaled(2)
    cd_f, cl_f = run_coefficients(fine, 2 * coarse_steps, 2 * window, seed=seed, sample_every=40)
    cd_s, cl_s = run_coefficients(case, coarse_steps, window, seed=seed)
    cd_l, cl_l = run_coefficients(case, coarse_steps, window, closure=closure, seed=seed)
    return {"fine": {"cd": cd_f, "cl": cl_f}, "smagorinsky": {"cd": cd_s, "cl": cl_s}, "learned": {"cd": cd_l, "cl": cl_l},
            "err_cd_smagorinsky": cd_s / cd_f - 1.0, "err_cd_learned": cd_l / cd_f - 1.0,
            "err_cl_smagorinsky": cl_s / cl_f - 1.0, "err_cl_learned": cl_l / cl_f - 1.0}


def factor_params(g):
    """Parameters of the network whose output is the constant factor g on the Smagorinsky eddy viscosity: all
    weights zero, output bias softplus^-1(g) - ln(e - 1)."""
    p = np.zeros(L.MLP_NPARAM, dtype=np.float32)
    p[-1] = np.log(np.expm1(g)) - np.log(np.e - 1.0)
    return p


def factor_closure(g):
    return L.Closure(params=factor_params(g))


def gate_params(g_lo, c, width=0.1):
    """Parameters of the network that gates the factor by the Smagorinsky strength of the cell (feature
    x5 = ln(1 + nu_S/nu)/4): factor g_lo where x5 < c - width, 1 (Smagorinsky) where x5 > c + width, the output y
    linear in x5 between. Two first-layer units relu(x5 - c -+ width), one second-layer unit holding their
    difference, and the output bias."""
    H, IN = L.MLP_HIDDEN, L.MLP_INPUTS
    p_w1, p_b1 = 0, H * IN
    p_w2 = p_b1 + H
    p_w3 = p_w2 + H * H + H
    p = np.zeros(L.MLP_NPARAM, dtype=np.float32)
    y_lo = np.log(np.expm1(g_lo)) - np.log(np.e - 1.0)
    p[p_w1 + 0 * IN + 5], p[p_b1 + 0] = 1.0, -(c - width)
    p[p_w1 + 1 * IN + 5], p[p_b1 + 1] = 1.0, -(c + width)
    p[p_w2 + 0], p[p_w2 + 1] = 1.0, -1.0
    p[p_w3] = -y_lo / (2.0 * width)
    p[-1] = y_lo
    return p


def gate_sweep(case, gates, fine_cd, seed=1, log=print):
    """Coarse Cd of each building of `fine_cd` {(L, W, H): fine Cd} with each strain gate (g_lo, c) of `gates`, as
    signed errors against the fine-grid Cd, and the mean absolute error per gate."""
    out = {}
    for g_lo, c in gates:
        errs = {}
        for size, fcd in fine_cd.items():
            net = L.Closure(params=gate_params(g_lo, c))
            cd, _ = run_coefficients(sized(case, size), closure=net, seed=seed)
            net.close()
            errs["x".join(map(str, size))] = cd / fcd - 1.0
        mean_abs = float(np.mean(np.abs(list(errs.values()))))
        out[f"{g_lo}:{c}"] = {"errors": errs, "mean_abs": mean_abs}
        log(f"gate g_lo {g_lo} c {c}: " + " ".join(f"{k} {v * 100:+.1f}%" for k, v in errs.items()) +
            f"; mean |error| {mean_abs * 100:.1f}%")
    return out


def factor_sweep(case, factors, seed=1, log=print):
    """Coarse Cd and Cl of the case with a constant factor g on the Smagorinsky eddy viscosity, for each g."""
    out = {}
    for g in factors:
        cl = factor_closure(g)
        cd, cll = run_coefficients(case, closure=cl, seed=seed)
        cl.close()
        out[str(g)] = {"cd": cd, "cl": cll}
        log(f"factor {g}: coarse Cd {cd:.4f} Cl {cll:.4f}")
    return out


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


def protocol_check(case, sizes, coarse_steps=COARSE_STEPS, every=20, log=print):
    """Coarse and fine Smagorinsky runs of the given building sizes for `coarse_steps` coarse steps (twice as many fine
    steps), Cd sampled every `every` coarse-equivalent steps: per size the coarse and fine mean Cd over the last
    WINDOW steps and the means of its four quarters (a measure of the averaging noise)."""
    out = {}
    for size in sizes:
        rec = {}
        for tag, cs, steps, ev in (("coarse", sized(case, size), coarse_steps, every),
                                   ("fine", sized(case, size).scaled(2), 2 * coarse_steps, 2 * every)):
            ctx, refs, _ = cs.context(seed=1)
            cd = []
            for k in range(steps // ev):
                ctx.step(ev)
                cd.append(ctx.read_forces(refs)["cd"])
                if (k + 1) % max(1, steps // ev // 4) == 0:
                    log(f"{tag} {size}: {(k + 1) * ev} steps, Cd {cd[-1]:.4f}")
            ctx.close()
            w = np.array(cd[-(WINDOW // every):])
            rec[tag] = {"cd": float(w.mean()), "quarters": [float(q.mean()) for q in np.array_split(w, 4)]}
        rec["err"] = rec["coarse"]["cd"] / rec["fine"]["cd"] - 1.0
        out["x".join(map(str, size))] = rec
        log(f"building {size}: coarse {rec['coarse']['cd']:.4f} fine {rec['fine']['cd']:.4f} ({rec['err'] * 100:+.2f} %)")
    return out


def _sizes_arg(text):
    return [tuple(int(v) for v in item.split("x")) for item in text.split(",")]


def main():
    import argparse
    import json
    import os
    import time

    ap = argparse.ArgumentParser(description="Learned-closure workflow: samples from filtered 2x runs, training "
                                             "on the GPU, and the comparison against the fine-grid plain run.")
    ap.add_argument("--sizes", type=int, default=4, help="seeded building sizes of the fine training runs")
    ap.add_argument("--snapshots", type=int, default=50, help="snapshots per fine run")
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--coarse-steps", type=int, default=COARSE_STEPS)
    ap.add_argument("--window", type=int, default=WINDOW)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--out", default=None, help="directory for the trained parameters (closure_params.npy)")
    ap.add_argument("--protocol-check", default=None, metavar="SIZES",
                    help="only the coarse/fine Smagorinsky statistics of buildings LxWxH[,LxWxH...]")
    ap.add_argument("--factor-sweep", default=None, metavar="G",
                    help="only the nominal building's coarse coefficients with constant factors G[,G...] on the "
                         "Smagorinsky eddy viscosity")
    ap.add_argument("--gate-sweep", default=None, metavar="G:C",
                    help="only the coarse Cd errors of strain-gated factors G:C[,G:C...] (factor G below the gate "
                         "threshold C of x5, Smagorinsky above) on the buildings of --fine")
    ap.add_argument("--fine", default=None, metavar="SIZE=CD",
                    help="fine-grid Cd of the buildings for --gate-sweep, LxWxH=CD[,LxWxH=CD...]")
    ap.add_argument("--cost", type=int, default=0, metavar="STEPS",
                    help="only time STEPS coarse steps with Smagorinsky and with the fused closure (a seeded "
                         "random network: the cost does not depend on the weights)")
    args = ap.parse_args()
    case = BuildingCase()
    log = lambda m: print(m, flush=True)  # noqa: E731
    if args.protocol_check:
        print(json.dumps(protocol_check(case, _sizes_arg(args.protocol_check), args.coarse_steps, log=log)))
        return
    if args.factor_sweep:
        print(json.dumps(factor_sweep(case, [float(g) for g in args.factor_sweep.split(",")], log=log)))
        return
    if args.gate_sweep:
        gates = [tuple(float(v) for v in item.split(":")) for item in args.gate_sweep.split(",")]
        fine_cd = {tuple(int(v) for v in k.split("x")): float(cd)
                   for k, cd in (item.split("=") for item in args.fine.split(","))}
        print(json.dumps(gate_sweep(case, gates, fine_cd, log=log)))
        return
    if args.cost:
        rng = np.random.default_rng(args.seed)
        net = L.Closure(params=(rng.standard_normal(L.MLP_NPARAM) * 0.3).astype(np.float32))
        print(json.dumps({"case": f"{case.nx}x{case.height}x{case.lateral}", "steps": args.cost,
                          "cost": cost(case, net, steps=args.cost)}))
        return
    closure = L.Closure()
    t0 = time.perf_counter()
    sizes = seeded_sizes(args.sizes, args.seed)
    n, runs = generate_samples(closure, case, sizes, snapshots_per_size=args.snapshots,
                               coarse_steps=args.coarse_steps, window=args.window, log=log)
    t1 = time.perf_counter()
    X, T = shuffle_samples(closure, args.seed)
    loss_smag = closure.grad(X[:262144], T[:262144])[0]  # the zero network is the Smagorinsky closure
    losses = closure.train(args.iters, batch=4096, lr=args.lr)
    loss_net = closure.grad(X[:262144], T[:262144])[0]
    factor = closure.predict(X[:262144])
    log(f"trained: loss on 262144 samples {loss_smag:.5g} (Smagorinsky) -> {loss_net:.5g}; network output "
        f"mean {float(np.mean(factor)):.4f} sd {float(np.std(factor)):.4f}")
    if args.out:
        os.makedirs(args.out, exist_ok=True)
        np.save(os.path.join(args.out, "closure_params.npy"), closure.params())
    t2 = time.perf_counter()
    for r in runs:  # the training buildings' own coarse errors
        c = sized(case, r["size"])
        r["coarse_smagorinsky_cd"], r["coarse_smagorinsky_cl"] = run_coefficients(c, args.coarse_steps, args.window)
        r["coarse_learned_cd"], r["coarse_learned_cl"] = run_coefficients(c, args.coarse_steps, args.window,
                                                                          closure=closure)
        log(f"training building {r['size']}: Cd fine {r['fine_cd']:.4f} Smagorinsky {r['coarse_smagorinsky_cd']:.4f} "
            f"learned {r['coarse_learned_cd']:.4f}")
    res = evaluate(case, closure, args.coarse_steps, args.window)
    log(f"nominal building (held out): Cd fine {res['fine']['cd']:.4f} Smagorinsky {res['smagorinsky']['cd']:.4f} "
        f"({res['err_cd_smagorinsky'] * 100:+.2f} %) learned {res['learned']['cd']:.4f} "
        f"({res['err_cd_learned'] * 100:+.2f} %)")
    t3 = time.perf_counter()
    out = {"case": {"nx": case.nx, "height": case.height, "lateral": case.lateral, "building": list(NOMINAL),
                    "re": case.re, "u_ref": case.u_ref, "z0": case.z0, "z_ref": case.z_ref,
                    "coarse_steps": args.coarse_steps, "window": args.window},
           "samples": n, "snapshots": len(sizes) * args.snapshots, "target_mean": float(np.mean(T)),
           "target_std": float(np.std(T)), "loss_smagorinsky": loss_smag, "loss_trained": loss_net,
           "network_output_mean": float(np.mean(factor)), "network_output_std": float(np.std(factor)),
           "iters": args.iters, "training_runs": runs, "evaluation": res,
           "seconds": {"samples": t1 - t0, "training": t2 - t1, "evaluation": t3 - t2}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
