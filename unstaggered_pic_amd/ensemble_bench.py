"""Batched-ensemble benchmark: independent P3 runs of uneven sizes, drawn
from a seed (grid side in {64, 128, 256}, particles per cell in {16, 64,
256}, thermal speed in [0.02, 0.2] c, perturbation amplitude in [1e-3,
1e-2]; Landau loads with quadratic splines at dt = 0.1 and Weibel loads with
cubic splines at dt = 0.05), mixed precision with the physics-predictor
guess and the selective particle refinement on, each on its own context
with graph-replayed field solves.

The same members are timed four times from the same state: stepped one
after the other with the host-driven outer loop (graphs off), stepped one
after the other with each step its own device-driven graph, advanced
together by upic_ensemble_step with host threads (graphs off), and advanced
together by upic_ensemble_step with graphs on (every ensemble step one
hipGraph launch).  Then the mixed and the
FP64 forms of every member run --fp64-steps steps from the same initial
state for the per-member energy error.  Prints one JSON line: aggregate
particle-steps/s of each, the ratio of the graph ensemble to the
sequential steps, the largest relative total-energy change of any member
over the timed steps and the per-member errors against FP64.

  python bench.py --form ensemble [--members 16] [--steps 5] [--warmup 2]
"""
import json
import sys
import time

import numpy as np


def member_specs(n, seed=2024):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        specs.append({"kind": "landau" if i % 2 == 0 else "weibel", "side": int(rng.choice([64, 128, 256])),
                      "ppc": int(rng.choice([16, 64, 256])), "vth": float(rng.uniform(0.02, 0.2)),
                      "amp": float(rng.uniform(1e-3, 1e-2)), "seed": int(rng.integers(1, 2 ** 31))})
    return specs


def build(lib, spec, mixed=True):
    from unstaggered_pic_amd.abi import (CODE_P3, LOAD_BIMAXWELL, LOAD_LANDAU, make_load, make_params,
                                         with_mixed_precision)
    s = spec["side"]
    if spec["kind"] == "landau":
        L = 4.0 * np.pi * s / 64.0
        p = make_params(code=CODE_P3, nx=s, ny=s, nz=1, Lx=L, Ly=L, Lz=1.0, dt=0.1, spline_order=2,
                        orbit_quad_order=3, nyquist_projection=1, nonlinear_rtol=1e-11, nonlinear_atol=1e-13,
                        nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12, local_atol=1e-14,
                        local_max_iter=64, sort_interval=1, selective_particle_fallback=1)
        load = make_load(LOAD_LANDAU, spec["ppc"], seed=spec["seed"], alpha=spec["amp"],
                         vth=(spec["vth"],) * 3)
    else:
        L = 20.0 * np.pi * s / 256.0
        p = make_params(code=CODE_P3, nx=s, ny=s, nz=1, Lx=L, Ly=L, Lz=1.0, dt=0.05, spline_order=3,
                        orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11, nonlinear_atol=1e-13,
                        nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12, local_atol=1e-14,
                        local_max_iter=64, sort_interval=1, selective_particle_fallback=1)
        load = make_load(LOAD_BIMAXWELL, spec["ppc"], seed=spec["seed"],
                         vth=(0.25 * spec["vth"], spec["vth"], spec["vth"]))
    if mixed:
        p = with_mixed_precision(p)
        p.learned_guess = 1
    n = s * s * spec["ppc"]
    x, v, q, m = lib.load_particles(p, load, n)
    return p, n, (x, v, q, m)


class _Host:
    def __init__(self, n, N):
        self.x, self.v = np.zeros(3 * n), np.zeros(3 * n)
        self.v_prev, self.P, self.vbar_last = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        self.q, self.m = np.zeros(n), np.zeros(n)
        self.phi, self.psi, self.rho = np.zeros(N), np.zeros(N), np.zeros(N)
        self.A, self.U, self.J = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)


def run(args):
    """bench.py --form ensemble: the measurement as the bench's one JSON line."""
    print(json.dumps(measure(args.members, args.steps, args.warmup, args.baseline_steps, args.fp64_steps,
                             guess=getattr(args, "ensemble_guess", "predictor"))))


def train_per_kind(lib, specs, loads, collect_steps=40, rows_per_step=4096, epochs=15):
    """One network per load kind (Landau, Weibel): rows from collect_steps
    steps of the kind's largest member, trained on the device; returns
    {kind: parameters} and the seconds spent."""
    t0 = time.perf_counter()
    nets = {}
    for kind in ("landau", "weibel"):
        ks = [k for k, sp in enumerate(specs) if sp["kind"] == kind]
        if not ks:
            continue
        k = max(ks, key=lambda j: loads[j][1])
        p, n, (x, v, q, m) = loads[k]
        ctx = lib.Context(p)
        sim = lib.Simulation(ctx, n)
        h = _Host(n, p.nx * p.ny)
        h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
        sim.upload(h)
        sim.init_fields(1.0)
        stride = max(1, n // rows_per_step)
        per = (n + stride - 1) // stride
        sim.collect_guess_data(stride, per * collect_steps)
        for _ in range(collect_steps):
            sim.step()
        sim.collect_guess_data(0, 0)
        sim.train_guess(epochs=epochs, batch=4096, lr=2e-3, seed=5)
        nets[kind] = sim.guess_params()
        sim.close()
        ctx.close()
        print(f"# trained the {kind} network on member {k} ({n} particles)", file=sys.stderr, flush=True)
    return nets, time.perf_counter() - t0


def measure(members, steps_timed, warmup, baseline_steps, fp64_steps, guess="predictor"):
    """guess: "predictor" (the physics predictor as every member's inner
    start) or "learned" (predictor + network: one network per load kind,
    trained on that kind's largest member and installed in all of them)."""
    from unstaggered_pic_amd import lib
    specs = member_specs(members)
    loads = []
    for i, sp in enumerate(specs):
        loads.append(build(lib, sp))
        if (i + 1) % 8 == 0 or i + 1 == len(specs):
            print(f"# loaded {i + 1}/{len(specs)} members", file=sys.stderr, flush=True)
    total = sum(n for _, n, _ in loads)
    nets, train_s = train_per_kind(lib, specs, loads) if guess == "learned" else ({}, 0.0)

    def fresh(graphs=True, fp64=False):
        ms = []
        for k, (p, n, (x, v, q, m)) in enumerate(loads):
            if fp64:
                p = build(lib, specs[k], mixed=False)[0]
            ctx = lib.Context(p)
            ctx.set_graphs(graphs)
            sim = lib.Simulation(ctx, n)
            h = _Host(n, p.nx * p.ny)
            h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
            sim.upload(h)
            sim.init_fields(1.0)
            if nets and not fp64:
                sim.set_guess_params(nets[specs[k]["kind"]])
                sim.set_guess_mode(2)
            ms.append((ctx, sim))
        return ms

    out = {}
    for mode in ("sequential_host_loop", "sequential", "threads", "ensemble"):
        ms = fresh(graphs=mode in ("sequential", "ensemble"))
        print(f"# {mode}: {len(ms)} members on the GPU", file=sys.stderr, flush=True)
        sims = [s for _, s in ms]
        one_by_one = mode.startswith("sequential")
        for _ in range(warmup):
            [s.step() for s in sims] if one_by_one else lib.ensemble_step(sims, threads=16)
        e0 = [s.diagnostics(1).total_energy for s in sims]
        for c, _ in ms:
            c.sync()
        steps = steps_timed if mode == "ensemble" else min(steps_timed, baseline_steps)
        t0 = time.perf_counter()
        launched = 0
        for k in range(steps):
            if not one_by_one:
                launched += sum(1 for i in lib.ensemble_step(sims, threads=16) if i.graph_launch == 2)
            else:
                for s in sims:
                    s.step()
            if (k + 1) % 100 == 0:
                print(f"# {mode}: step {k + 1}", file=sys.stderr, flush=True)
        for c, _ in ms:
            c.sync()
        el = time.perf_counter() - t0
        drifts = [abs(s.diagnostics(1).total_energy - a) / abs(a) for s, a in zip(sims, e0)]
        print(f"# {mode}: {1e3 * el / steps:.1f} ms per ensemble step", file=sys.stderr, flush=True)
        out[mode] = {"value": total * steps / el, "ms_per_step": 1e3 * el / steps, "steps": steps,
                     "max_rel_energy_change": max(drifts), "graph_launched": launched}
        if mode == "ensemble":
            out[mode]["rel_energy_change_per_member"] = [float(f"{d:.3e}") for d in drifts]
        for c, s in ms:
            s.close()
            c.close()
    # per-member energy error against the FP64 form of the same member: both
    # ensembles from the same initial state for fp64_steps steps
    hist = {}
    for form in ("mixed", "fp64"):
        ms = fresh(graphs=True, fp64=form == "fp64")
        sims = [s for _, s in ms]
        h = [[(d.total_energy, d.field_energy)] for d in (s.diagnostics(1) for s in sims)]
        for _ in range(fp64_steps):
            lib.ensemble_step(sims, threads=16)
            for k, s in enumerate(sims):
                d = s.diagnostics(1)
                h[k].append((d.total_energy, d.field_energy))
        hist[form] = np.array(h)
        print(f"# {form}: {fp64_steps} steps for the energy comparison", file=sys.stderr, flush=True)
        for c, s in ms:
            s.close()
            c.close()
    hm, hf = hist["mixed"], hist["fp64"]
    te_err = np.max(np.abs(hm[:, :, 0] - hf[:, :, 0]), axis=1) / np.abs(hf[:, 0, 0])
    fe_err = np.max(np.abs(hm[:, :, 1] - hf[:, :, 1]) / np.abs(hf[:, :, 1]), axis=1)
    line = {
        "metric": "aggregate particle-steps/s of a batched ensemble of independent P3 runs on 1 MI355X",
        "value": out["ensemble"]["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": steps_timed,
        "graph_launched_member_steps": out["ensemble"]["graph_launched"],
        "warmup": warmup, "ms_per_step": out["ensemble"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
        "data": "synthetic (seeded Landau and bi-Maxwellian loads)",
        "config": {"workload": f"ensemble_{members}_members_p3_mixed_{guess}", "particles": total,
                   "guess": guess if guess == "predictor" else
                   "learned: one network per load kind, trained on its largest member (40 collection steps, "
                   f"15 epochs; {train_s:.1f} s outside the timed region)",
                   "members": specs},
        "sequential": out["sequential"], "ensemble_vs_sequential": out["ensemble"]["value"] / out["sequential"]["value"],
        "max_rel_energy_change": out["ensemble"]["max_rel_energy_change"],
        "rel_energy_change_per_member": out["ensemble"]["rel_energy_change_per_member"],
        "threads": out["threads"], "sequential_host_loop": out["sequential_host_loop"],
        "ensemble_vs_sequential_host_loop": out["ensemble"]["value"] / out["sequential_host_loop"]["value"],
        "energy_error_vs_fp64": {
            "steps": fp64_steps,
            "definition": "per member: max over steps of |E_mixed - E_fp64| / |E_fp64(0)| (total energy), "
                          "max over steps of |W_mixed - W_fp64| / |W_fp64| (field energy); FP64 = the same member "
                          "stepped by the FP64 GPU path from the same initial state",
            "total_energy_max": float(te_err.max()), "field_energy_max": float(fe_err.max()),
            "total_energy_per_member": [float(f"{e:.3e}") for e in te_err],
            "field_energy_per_member": [float(f"{e:.3e}") for e in fe_err]},
    }
    return line
