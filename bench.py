"""Benchmark of the per-timestep PIC hot path on one MI355X.

Workload (the Weibel configuration the headline metric is quoted on):
P3 form (relativistic GM-HC-CN, outer Anderson on the mesh current, inner
per-particle Picard), 2-D periodic 256 x 256 grid on [0, 20 pi)^2, 256
electrons per cell (16,777,216 particles), bi-Maxwellian momenta
u_th = (0.05, 0.2, 0.2) c, uniform density, cubic B-splines, omega_p dt =
0.05, Nyquist projection on, tolerances of the reference's Weibel manuscript
deck (Project3/inputs/manuscript/weibel_main_dt02_T100.txt).  Particles are
generated by the library's seeded loader (synthetic, seed 12345).

Forms: the plain FP64 step, and the mixed-precision step (FP32 particle
state, push and deposit tiles; FP64 mesh, FFT and convergence tests, inner
and outer relative tolerances 1e-6 and 3e-6) with the learned initial guess of the
inner solve, whose network is trained on the device on rows the first
--collect-steps steps of the same run produce.  The default runs both and
reports the mixed + learned form as `value`, the plain FP64 form beside it.

One step = one upic_step_p3 call (a full nonlinear time step).  The timed
region covers exactly --steps steps bracketed by device synchronisation;
inputs are resident in HBM before it starts; data collection and training
happen before it and are reported separately.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector, vendor figure
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector, vendor figure


WORKLOADS = {
    # name: (grid side, particles per cell) defaults
    "weibel": (256, 256),
    "beams": (1024, 64),
    "landau": (64, 100),
}


def workload_params(nx=256, ppc=256, precision=0, workload="weibel"):
    """The P3 decks: Weibel (the headline: [0, 20 pi)^2, cubic, dt 0.05),
    counter-streaming beams ([0, 40 pi)^2, quadratic, dt 0.1) and Landau
    damping ([0, 4 pi)^2, quadratic, dt 0.1); precision 1 selects the mixed
    mode with its FP32 tolerance floors (inner relative 1e-6, outer 3e-6)."""
    from unstaggered_pic_amd.abi import CODE_P3, make_params, with_mixed_precision
    if workload == "weibel":
        L, order, nq, dt = 20.0 * np.pi, 3, 4, 0.05
    elif workload == "beams":
        L, order, nq, dt = 40.0 * np.pi, 2, 3, 0.1
    else:
        L, order, nq, dt = 4.0 * np.pi, 2, 3, 0.1
    p = make_params(code=CODE_P3, nx=nx, ny=nx, nz=1, Lx=L, Ly=L, Lz=1.0, dt=dt, spline_order=order,
                       orbit_quad_order=nq, split_orbit_at_knots=1, nyquist_projection=1, c=1.0, sigma1=1.0,
                       sigma2=1.0, nonlinear_solver_picard=0, nonlinear_max_iter=80, nonlinear_rtol=1e-11,
                       nonlinear_atol=1e-13, anderson_memory=4, anderson_start=2, anderson_damping=0.5,
                       anderson_regularization=1e-12, anderson_growth_limit=1.2, local_rtol=1e-12,
                       local_atol=1e-14, local_max_iter=64, local_damping=1.0, precision=0,
                       sort_interval=1)
    return with_mixed_precision(p) if precision == 1 else p


def workload_load(ppc, seed=12345, workload="weibel"):
    from unstaggered_pic_amd.abi import LOAD_BEAMS, LOAD_BIMAXWELL, LOAD_LANDAU, make_load
    if workload == "beams":
        return make_load(LOAD_BEAMS, ppc, seed=seed, n0=1.0, v0=0.2, noise=1e-3, vth=(0.01, 0.01, 0.01))
    if workload == "landau":
        return make_load(LOAD_LANDAU, ppc, seed=seed, n0=1.0, alpha=0.01, mode=1, vth=(0.05, 0.05, 0.05))
    return make_load(LOAD_BIMAXWELL, ppc, seed=seed, n0=1.0, vth=(0.05, 0.2, 0.2))


def workload_name(args):
    if args.workload == "weibel":
        return f"weibel_p3_{args.nx}x{args.nx}_ppc{args.ppc}_cubic_dt0.05"
    if args.workload == "beams":
        return f"beams_p3_{args.nx}x{args.nx}_ppc{args.ppc}_quadratic_dt0.1"
    return f"landau_p3_{args.nx}x{args.nx}_ppc{args.ppc}_quadratic_dt0.1"


class HostArrays:
    def __init__(self, np_, N):
        self.x, self.v = np.zeros(3 * np_), np.zeros(3 * np_)
        self.v_prev, self.P, self.vbar_last = np.zeros(3 * np_), np.zeros(3 * np_), np.zeros(3 * np_)
        self.q, self.m = np.zeros(np_), np.zeros(np_)
        self.phi, self.psi, self.rho = np.zeros(N), np.zeros(N), np.zeros(N)
        self.A, self.U, self.J = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)


def cpu_baseline(nx, dt_budget=30.0, workload="weibel"):
    """The reference's own FP64 step (compiled from its sources into
    oracle/_ref/libref_p3.so) -- or, where that is absent, the oracle port --
    timed on the host for one step of the same deck at 3 particles per cell
    (about 15 s of serial work at 256 x 256)."""
    from unstaggered_pic_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, State, ptr
    ppc = 3 if nx <= 256 else 1
    p = workload_params(nx=nx, ppc=ppc, workload=workload)
    np_ = nx * nx * ppc
    x, v, q, m = lib.load_particles(p, workload_load(ppc, workload=workload), np_)
    orc = Oracle()
    st = orc.init_fields(p, x, v, q, m, 1.0)
    refso = os.path.join(ROOT, "oracle", "_ref", "libref_p3.so")
    kind = "port"
    t0 = time.perf_counter()
    if os.path.exists(refso):
        L = C.CDLL(refso)
        out = State(st.np, st.N)
        it = C.c_int32()
        info = np.zeros(3)
        t0 = time.perf_counter()
        rc = L.ref_p3_step(C.byref(p), C.byref(st.host()), C.byref(out.host()), C.byref(it), ptr(info))
        el = time.perf_counter() - t0
        if rc == 0:
            kind = "reference"
        else:
            el = None
    else:
        el = None
    if kind == "port":
        t0 = time.perf_counter()
        orc.step(p, st)
        el = time.perf_counter() - t0
    return {"value": np_ / el, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"one step of the same deck on the {nx}x{nx} grid at {ppc} particles/cell ({np_} particles), "
                      f"serial FP64, {el:.1f} s"}


def measured_traffic(args, form):
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) for this workload and form,
    kept in profiles/; None for other workloads."""
    path = os.path.join(ROOT, "profiles", f"traffic_weibel{args.nx}_ppc{args.ppc}.json")
    if args.workload != "weibel" or args.nx != 256 or args.ppc != 256 or not os.path.exists(path):
        return None
    d = json.load(open(path))["forms"].get(form)
    return d["bytes_per_launch"] if d else None


def run_form(args, prec, learned):
    """Load, initialise and time one form of the step; returns a dict."""
    from unstaggered_pic_amd import lib
    p = workload_params(args.nx, args.ppc, prec, args.workload)
    N = p.nx * p.ny
    np_ = N * args.ppc
    x, v, q, m = lib.load_particles(p, workload_load(args.ppc, workload=args.workload), np_)
    ctx = lib.Context(p)
    ctx.set_graphs(args.graphs)
    sim = lib.Simulation(ctx, np_)
    h = HostArrays(np_, N)
    h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
    t0 = time.perf_counter()
    sim.upload(h)  # host arrays -> HBM (pageable host memory)
    ctx.sync()
    t_upload = time.perf_counter() - t0
    del h, x, v, q, m
    sim.init_fields(1.0)
    d0 = sim.diagnostics(1)
    extra = {}
    if learned:
        # training rows from the first collect_steps steps of the run itself
        stride = max(1, np_ // args.rows_per_step)
        per_step = (np_ + stride - 1) // stride
        sim.collect_guess_data(stride, per_step * args.collect_steps)
        t0 = time.perf_counter()
        for _ in range(args.collect_steps):
            sim.step()
        ctx.sync()
        t_collect = time.perf_counter() - t0
        sim.collect_guess_data(0, 0)
        t0 = time.perf_counter()
        losses = sim.train_guess(epochs=args.train_epochs, batch=4096, lr=2e-3, seed=5)
        t_train = time.perf_counter() - t0
        extra = {"collect_steps": args.collect_steps, "collect_seconds": t_collect, "training_rows":
                 per_step * args.collect_steps, "train_epochs": args.train_epochs, "train_seconds": t_train,
                 "train_loss_first_last": [float(losses[0]), float(losses[-1])]}
    for _ in range(args.warmup):
        sim.step()
    ctx.set_timing(not args.no_timing)
    d_start = sim.diagnostics(1)
    infos = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(sim.step())
    ctx.sync()
    el = time.perf_counter() - t0
    kname = "p3_map" if prec == 0 else "p3_map_mixed"
    launches, kms = ctx.timing(kname)
    dep = ctx.timing("deposit_mixed") if prec == 1 else (0, 0.0)
    d1 = sim.diagnostics(1)
    dep_choice = sim.deposit_variant() if prec == 1 else None
    sim.close()
    ctx.close()
    return {"p": p, "np": np_, "N": N, "elapsed": el, "infos": infos, "launches": launches, "kms": kms, "dep": dep,
            "kname": kname, "d0": d0, "d_start": d_start, "d1": d1, "extra": extra, "dep_choice": dep_choice,
            "upload_seconds": t_upload}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="weibel",
                    help="P3 deck; sets the grid side and particles per cell unless --nx/--ppc are given")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ppc", type=int, default=None)
    ap.add_argument("--members", type=int, default=16, help="ensemble members (--form ensemble)")
    ap.add_argument("--form", choices=["both", "fp64", "mixed", "learned", "ensemble"], default="both",
                    help="both: plain FP64 and mixed + learned guess; value is the last")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None, help="alias of --form fp64|mixed")
    ap.add_argument("--collect-steps", type=int, default=40)
    ap.add_argument("--rows-per-step", type=int, default=16384)
    ap.add_argument("--train-epochs", type=int, default=15)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graphs", type=int, default=1, help="replay the field solves from hipGraphs (1) or not (0)")
    ap.add_argument("--no-timing", action="store_true", help="skip the per-kernel event timing")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a known-size kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.precision:
        args.form = args.precision
    args.nx = args.nx or WORKLOADS[args.workload][0]
    args.ppc = args.ppc or WORKLOADS[args.workload][1]
    if args.calibrate_copy_mib > 0:
        from unstaggered_pic_amd import lib
        n = args.calibrate_copy_mib * (1 << 20) // 8
        a, b = lib.DeviceArray(n), lib.DeviceArray(n)
        lib.lib().upic_memcpy_htod(a.p, np.zeros(n).ctypes.data_as(C.c_void_p), 8 * n)
        if lib.lib().upic_memcpy_dtod(b.p, a.p, 8 * n) != 0:
            raise RuntimeError("device copy failed")
        del a, b
    if args.form == "ensemble":
        from unstaggered_pic_amd import ensemble_bench
        ensemble_bench.run(args)
        return
    forms = {"both": [(0, False), (1, True)], "fp64": [(0, False)], "mixed": [(1, False)],
             "learned": [(1, True)]}[args.form]
    runs = []
    for prec, learned in forms:
        runs.append(run_form(args, prec, learned))
        r = runs[-1]
        print(f"# form prec={prec} learned={learned}: {1e3 * r['elapsed'] / args.steps:.2f} ms/step",
              file=sys.stderr, flush=True)
    r = runs[-1]
    prec = 1 if forms[-1][0] == 1 else 0
    learned = forms[-1][1]
    np_, N, infos = r["np"], r["N"], r["infos"]
    value = np_ * args.steps / r["elapsed"]
    launches, kms = r["launches"], r["kms"]
    # algorithmic HBM bytes of one map launch.  FP64: reads x, P, v (72 B),
    # q, m (16 B) per particle, writes x, P, v, vbar_macro (96 B) once per
    # step, reads A_n, A_end, grad phi_mid (9 N doubles), writes J (3 N).
    # Mixed: reads cell (4 B), offsets, P, v (3 x 12 B), q, m (8 B) and, with
    # the learned guess, the previous sweep's vbar and A^n at the start point
    # (2 x 12 B, the first sweep of a step writes A^n instead); every sweep
    # writes the end state cell, offsets, P, v, vbar_macro (52 B) and the
    # converged vbar for the deposit (12 B); reads the FP32 node records
    # (32 N B).
    maps = sum(i.map_evaluations for i in infos)
    if prec == 0:
        bytes_per_launch = np_ * 88.0 + (np_ * 96.0 * args.steps) / max(maps, 1) + 12.0 * 8 * N
    else:
        later = 1.0 - args.steps / max(maps, 1)  # share of sweeps after a step's first
        bytes_per_launch = np_ * (48.0 + ((12.0 + 12.0 * later) if learned else 0.0) + 64.0) + 32.0 * N
    avg_ms = kms / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if launches else None
    # operations of one inner iteration at r = 3, nq = 4 (counted from the
    # kernel source: per Gauss point 2 axis stencils + 16 nodes x (3 gather +
    # 12 D_A) FMAs; per iteration the end-point gather and the HC push).
    segs = np.mean([i.path.total_segments / max(i.path.n_particles, 1) for i in infos])
    flops_iter = segs * 4 * (2 * 4 * 24 + 16 * (2 * 15 + 6 * 2 + 3)) + 16 * 6 + 250
    local_total = sum(i.local_iterations_total for i in infos)
    peak_c = FP64_PEAK_TFLOPS if prec == 0 else FP32_PEAK_TFLOPS
    compute_tflops = (flops_iter * local_total / args.steps) / (kms * 1e-3 / args.steps) / 1e12 if launches else None

    def phys(rr):
        d0, ds, d1 = rr["d0"], rr["d_start"], rr["d1"]
        return {"rel_total_energy_change_timed": (d1.total_energy - ds.total_energy) / abs(ds.total_energy),
                "rel_total_energy_change_since_init": (d1.total_energy - d0.total_energy) / abs(d0.total_energy),
                "outer_iterations": [i.iterations for i in rr["infos"]],
                "map_evaluations": [i.map_evaluations for i in rr["infos"]],
                "mean_inner_iterations": float(np.mean([i.local_iterations_mean for i in rr["infos"]])),
                # summed over the step's sweeps (the guess acts in the first)
                "inner_iterations_per_particle_step": float(np.mean([i.local_iterations_total for i in rr["infos"]]))
                / rr["np"],
                # constraints on the final state: Gauss's law and the Lorenz gauge residuals (FP64 diagnostics)
                "gauss_rms": d1.gauss_rms, "rho_rms": d1.rho_rms, "gauge_rms": d1.gauge_rms,
                "divA_rms": d1.divA_rms, "psi_over_c2_rms": d1.psi_over_c2_rms}

    form_name = {(0, False): "plain FP64", (1, False): "mixed precision", (1, True): "mixed precision + learned guess"}
    out = {
        "metric": "particle-steps/s on 1 MI355X, mixed+learned vs plain FP64; rel. energy drift",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32/f64",
        "data": "synthetic (seeded bi-Maxwellian load, splitmix64 seed 12345)",
        "config": {"workload": workload_name(args), "particles": np_,
                   "grid": [args.nx, args.nx], "spline_order": r["p"].spline_order,
                   "orbit_quad_order": r["p"].orbit_quad_order,
                   "form": form_name[(prec, learned)],
                   "solver": "P3 nested solve (Anderson outer, per-particle Picard inner)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(args, form_name[(prec, learned)]),
                     "kernel": "k_p3_map" if prec == 0 else "k_p3_map_mixed", "avg_launch_ms": avg_ms,
                     "launches": launches, "bytes_per_launch": bytes_per_launch,
                     "deposit_kernel_avg_ms": (r["dep"][1] / r["dep"][0]) if r["dep"][0] else None,
                     "deposit_variant_sorted_unsorted": r["dep_choice"]},
        "compute_roofline": {"bound": "fp64_valu" if prec == 0 else "fp32_valu", "achieved": compute_tflops,
                             "peak": peak_c, "unit": "TFLOP/s",
                             "frac": (compute_tflops / peak_c) if compute_tflops else None,
                             "flops_per_inner_iteration": flops_iter},
        "physics": phys(r),
    }
    out["physics"]["avg_path_segments"] = float(segs)
    # the state upload over PCIe (not part of `value`: inputs resident in HBM)
    out["host_upload_seconds"] = r["upload_seconds"]
    if r["extra"]:
        out["learned_guess"] = r["extra"]
    if len(runs) > 1:
        r0 = runs[0]
        v0 = r0["np"] * args.steps / r0["elapsed"]
        out["plain_fp64"] = {"value": v0, "ms_per_step": 1e3 * r0["elapsed"] / args.steps,
                             "avg_launch_ms": r0["kms"] / max(r0["launches"], 1), "physics": phys(r0)}
        out["speedup_vs_plain_fp64"] = value / v0
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.nx, workload=args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
