"""Long-run check of the mixed-precision + learned-guess step against the
plain FP64 step on the Weibel workload (256 x 256 on [0, 20 pi)^2, 256
particles per cell, cubic splines, dt = 0.05; the deck of bench.py).

Both forms start from the same seeded load; every step records the total,
field and magnetic energies and the |Bz(k_y = 1)| mode.  Prints a progress
line every 100 steps and one JSON summary at the end:
  * relative total-energy drift of each form, max_k |TE_k - TE_0| / TE_0;
  * field-energy history of the mixed form against the FP64 form,
    max_k |FE_mixed - FE_fp64| / FE_fp64 (and the same for the magnetic
    energy);
  * the magnetic energy of the mixed form against the FP64 form, relative
    to the FP64 maximum (the field starts at B = 0);
  * the magnetic-energy growth rate of each form: half the slope of
    log(W_B - plateau) on the interval where W_B rises from 3x its
    thermal-noise plateau (the median over the first quarter of the time to
    the maximum) to 1/5 of its maximum, and their relative difference.
The histories go to <out>_fp64.csv and <out>_mixed.csv.

Usage: python -m unstaggered_pic_amd.validate --steps 2000 --out DIR/weibel
       (--fp64-history / --mixed-history reuse earlier histories)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args, prec, learned):
    sys.path.insert(0, ROOT)
    import bench
    from unstaggered_pic_amd import lib
    p = bench.workload_params(args.nx, args.ppc, prec)
    N = p.nx * p.ny
    n = N * args.ppc
    x, v, q, m = lib.load_particles(p, bench.workload_load(args.ppc), n)
    ctx = lib.Context(p)
    ctx.set_graphs(args.graphs)
    sim = lib.Simulation(ctx, n)
    h = bench.HostArrays(n, N)
    h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
    sim.upload(h)
    del h, x, v, q, m
    sim.init_fields(1.0)
    dV = (p.Lx / p.nx) * (p.Ly / p.ny) * p.Lz
    rows = []

    def record(k, t):
        d = sim.diagnostics(1)
        wb = 0.5 / p.sigma2 * d.B_l2 ** 2 * N * dV
        rows.append((k, t, d.total_energy, d.field_energy, wb, d.Bz_y_mode_abs, d.kinetic_energy))

    record(0, 0.0)
    tag = "mixed+learned" if learned else ("mixed" if prec else "fp64")
    t0 = time.perf_counter()
    k = 0
    if learned:
        sim.set_guess_mode(1)  # the collection steps start from the physics predictor
        stride = max(1, n // 16384)
        per = (n + stride - 1) // stride
        sim.collect_guess_data(stride, per * args.collect_steps)
        for _ in range(args.collect_steps):
            sim.step()
            k += 1
            record(k, k * p.dt)
        sim.collect_guess_data(0, 0)
        sim.train_guess(epochs=15, batch=4096, lr=2e-3, seed=5)
    while k < args.steps:
        sim.step()
        k += 1
        record(k, k * p.dt)
        if k % 100 == 0:
            r = rows[-1]
            print(f"{tag} step {k} TE {r[2]:.15e} FE {r[3]:.6e} WB {r[4]:.6e} "
                  f"{time.perf_counter() - t0:.1f}s", flush=True)
    el = time.perf_counter() - t0
    sim.close()
    ctx.close()
    return np.array(rows), el


def growth_rate(t, wb):
    imax = int(np.argmax(wb))
    if imax < 8:
        return None, None
    plateau = float(np.median(wb[1:imax // 4]))  # thermal-noise level of W_B
    lo = 3.0 * plateau
    if not np.any(wb > lo):
        return None, None
    i0 = int(np.argmax(wb > lo))
    hi = np.max(wb) / 5.0
    i1 = i0 + int(np.argmax(wb[i0:] > hi)) if np.any(wb[i0:] > hi) else len(wb) - 1
    if i1 - i0 < 5:
        return None, None
    # the unstable mode's energy grows on top of the noise plateau
    sl = np.polyfit(t[i0:i1], np.log(wb[i0:i1] - plateau), 1)[0]
    return 0.5 * sl, (float(t[i0]), float(t[i1]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--collect-steps", type=int, default=200,
                    help="steps whose final sweeps write the network's training rows (mixed + learned)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="1: each step's outer solve as one device-driven graph launch (0: host-driven loop)")
    ap.add_argument("--out", default="weibel_history")
    ap.add_argument("--fp64-history", default=None,
                    help="reuse the FP64 history CSV of an earlier run of the same deck instead of rerunning it")
    ap.add_argument("--mixed-history", default=None, help="the same for the mixed + learned history")
    args = ap.parse_args()
    hdr = "step,time,total_energy,field_energy,magnetic_energy,Bz_y_mode_abs,kinetic_energy"
    if args.fp64_history:
        a = np.loadtxt(args.fp64_history, delimiter=",", skiprows=1)[: args.steps + 1]
        el_a = None
    else:
        a, el_a = run(args, 0, False)
        np.savetxt(args.out + "_fp64.csv", a, delimiter=",", header=hdr, comments="", fmt="%.17g")
    if args.mixed_history:
        b = np.loadtxt(args.mixed_history, delimiter=",", skiprows=1)[: args.steps + 1]
        el_b = None
    else:
        b, el_b = run(args, 1, True)
        np.savetxt(args.out + "_mixed.csv", b, delimiter=",", header=hdr, comments="", fmt="%.17g")
    print(json.dumps(summary(args, a, b, el_a, el_b)), flush=True)


def summary(args, a, b, el_a, el_b):
    ga, wa = growth_rate(a[:, 1], a[:, 4])
    gb, wb = growth_rate(b[:, 1], b[:, 4])
    out = {
        "workload": f"weibel_p3_{args.nx}x{args.nx}_ppc{args.ppc}_cubic_dt0.05", "steps": args.steps,
        "fp64": {"drift": float(np.max(np.abs(a[:, 2] - a[0, 2])) / abs(a[0, 2])), "seconds": el_a,
                 "growth_rate": ga, "fit_window": wa},
        "mixed_learned": {"drift": float(np.max(np.abs(b[:, 2] - b[0, 2])) / abs(b[0, 2])), "seconds": el_b,
                          "growth_rate": gb, "fit_window": wb},
        "field_energy_max_rel_diff": float(np.max(np.abs(b[:, 3] - a[:, 3]) / np.abs(a[:, 3]))),
        "magnetic_energy_max_diff_rel_to_max": float(np.max(np.abs(b[:, 4] - a[:, 4])) / np.max(a[:, 4])),
        "growth_rate_rel_diff": (abs(gb - ga) / abs(ga)) if (ga and gb) else None,
        "magnetic_energy_growth_factor": float(np.max(a[:, 4]) / a[1, 4]),
    }
    return out


if __name__ == "__main__":
    main()
