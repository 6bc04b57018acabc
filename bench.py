"""Benchmark of the per-timestep PIC hot path on one MI355X.

Headline workload (the Weibel configuration the metric is quoted on): P3
form (relativistic GM-HC-CN, outer Anderson on the mesh current, inner
per-particle Picard), 2-D periodic 256 x 256 grid on [0, 20 pi)^2, 256
electrons per cell (16,777,216 particles), bi-Maxwellian momenta
u_th = (0.05, 0.2, 0.2) c, uniform density, cubic B-splines, omega_p dt =
0.05, Nyquist projection on, tolerances of the reference's Weibel manuscript
deck (Project3/inputs/manuscript/weibel_main_dt02_T100.txt).  Particles are
generated by the library's seeded loader (synthetic, seed 12345).

Legs of the default run (--form all), each from the same initial state and
timing the same steps (--spinup start-up steps, --warmup, then --steps; the
first steps of a run are a start-up transient -- a Taylor-started history
velocity, a first step without the extrapolated current, and in the mixed
form a pause of the vbar start of later sweeps -- whose cost per step is
reported as spinup_ms_per_step):
  * plain FP64: the reference's algorithm and tolerances;
  * mixed precision, guess mode 0: FP32 particle state, push and deposit
    tiles, FP64 mesh, FFT and convergence tests at the FP32 tolerance floors
    (inner/outer relative 1e-6 / 3e-6), inner solves started from v^n as the
    reference does -- the precision change alone;
  * mixed, guess mode 1: the physics predictor as the inner start;
  * mixed, guess mode 2 (`value`): predictor + the learned network, trained
    on the device on rows that a separate 200-step run of the same workload
    produced (collection and training are outside the timed region and are
    reported in `learned_guess`).
After each leg's timed steps the same simulation runs --ref-steps steps in
the reference's scope ("as_reference": compute_diagnostics and
compute_step_diagnostics after every step, P3:8835-8850).
Beside the headline: the 64 x 64 Landau workload in the P1 form (plain
FP64, 409,600 particles, quadratic, dt 0.1, Picard tolerance 1e-10), and the
CPU baselines (the reference's own P3 step, compiled from its sources,
serial and with OpenMP on the host's cores, and its serial P1 step, each on
a bounded sample of the same workload), and the batched ensemble
(`ensemble`: --ensemble-members uneven Landau/Weibel runs drawn from a seed,
mixed precision + predictor, every ensemble step one graph launch, timed
over --ensemble-steps steps, with the same members stepped one by one for
the ratio and each member's energies against its own FP64 run over 5 steps;
unstaggered_pic_amd/ensemble_bench.py, which `--form ensemble` runs alone
at any step count).

One step = one upic_step_p3 (upic_step_p1) call.  The timed region covers
exactly --steps steps bracketed by device synchronisation; inputs are
resident in HBM before it starts.  With --graphs 1 (default) each P3 step's
outer solve runs as one device-driven hipGraph (upic_ctx_set_graphs).

Roofline (HBM, the north-star bookkeeping): B_alg per particle-step is one
read and one write of the persistent particle state plus 192 N_g / N_p
bytes of mesh traffic (72.75 B mixed P3, 128.75 B FP64 P3, 177.9 B FP64 P1
on these grids).  `roofline.achieved` = B_alg x the particles one launch of
the dominant kernel processes / that kernel's mean launch time (HIP events
on the context stream, over --timing-steps steps run right after the timed
ones; event-bracketed launches take the host-driven outer loop, whose
kernels are the same); `roofline.step_level` = B_alg x particle-steps/s.
`traffic` is the rocprofv3-measured HBM bytes per launch of that kernel
(profiles/); `traffic_model` is the kernel's own byte count (it also moves
the sweep-to-sweep scratch).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
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
    "landau_p1": (64, 100),
}
# persistent particle state, bytes per particle (x, v, P [, v_prev]; mixed:
# int32 cell + FP32 offsets and vectors)
STATE_BYTES = {"fp64_p3": 64.0, "mixed_p3": 36.0, "fp64_p1": 88.0, "mixed_p1": 48.0}


def b_alg(form, N, np_):
    """Algorithmic HBM bytes per particle-step: one read and one write of the
    persistent state, plus 12 FP64 mesh fields read and written once."""
    return 2.0 * STATE_BYTES[form] + 192.0 * N / np_


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


def p1_params(nx=64):
    """The Landau workload in the P1 form: [0, 4 pi)^2, quadratic splines,
    3 Gauss points per orbit segment (exact for its polynomial integrands),
    dt 0.1, Picard tolerance 1e-10 with the final consistency sweep
    (Project1/examples/input_landau_weak.txt:35-37)."""
    from unstaggered_pic_amd.abi import CODE_P1, make_params
    return make_params(code=CODE_P1, nx=nx, ny=nx, nz=1, Lx=4 * np.pi, Ly=4 * np.pi, Lz=1.0, dt=0.1,
                       spline_order=2, orbit_quad_order=3, split_orbit_at_knots=1, c=1.0, eps0=1.0, mu0=1.0,
                       picard_tol=1e-10, picard_max_iter=24, final_consistency_sweep=1, sort_interval=1)


def workload_load(ppc, seed=12345, workload="weibel"):
    from unstaggered_pic_amd.abi import LOAD_BEAMS, LOAD_BIMAXWELL, LOAD_LANDAU, make_load
    if workload == "beams":
        return make_load(LOAD_BEAMS, ppc, seed=seed, n0=1.0, v0=0.2, noise=1e-3, vth=(0.01, 0.01, 0.01))
    if workload in ("landau", "landau_p1"):
        return make_load(LOAD_LANDAU, ppc, seed=seed, n0=1.0, alpha=0.01, mode=1, vth=(0.05, 0.05, 0.05))
    return make_load(LOAD_BIMAXWELL, ppc, seed=seed, n0=1.0, vth=(0.05, 0.2, 0.2))


def workload_name(workload, nx, ppc):
    if workload == "weibel":
        return f"weibel_p3_{nx}x{nx}_ppc{ppc}_cubic_dt0.05"
    if workload == "beams":
        return f"beams_p3_{nx}x{nx}_ppc{ppc}_quadratic_dt0.1"
    if workload == "landau_p1":
        return f"landau_p1_{nx}x{nx}_ppc{ppc}_quadratic_dt0.1"
    return f"landau_p3_{nx}x{nx}_ppc{ppc}_quadratic_dt0.1"


class HostArrays:
    def __init__(self, np_, N):
        self.x, self.v = np.zeros(3 * np_), np.zeros(3 * np_)
        self.v_prev, self.P, self.vbar_last = np.zeros(3 * np_), np.zeros(3 * np_), np.zeros(3 * np_)
        self.q, self.m = np.zeros(np_), np.zeros(np_)
        self.phi, self.psi, self.rho = np.zeros(N), np.zeros(N), np.zeros(N)
        self.A, self.U, self.J = np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N)


# ---------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------
def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def _ref_p3_time(so, p, st):
    from oracle_bind import State, ptr
    L = C.CDLL(so)
    out = State(st.np, st.N)
    it = C.c_int32()
    info = np.zeros(3)
    t0 = time.perf_counter()
    rc = L.ref_p3_step(C.byref(p), C.byref(st.host()), C.byref(out.host()), C.byref(it), ptr(info))
    el = time.perf_counter() - t0
    return el if rc == 0 else None


def cpu_baseline_p3(nx, workload="weibel"):
    """The reference's own P3 step (nonlinear_step_orbit, compiled from
    Project3/ec_pic_nyquist_projected.cpp into oracle/_ref) on the host, for
    one step of the same deck on the same grid at reduced particles per
    cell: serially (libref_p3.so) and with OpenMP on the host's cores
    (libref_p3_omp.so, -fopenmp as Project3/run_case.sb:38).  Where oracle/_ref
    is absent, the oracle port is timed serially instead (kind "port")."""
    from unstaggered_pic_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle
    orc = Oracle()
    cpu = host_cpu()
    out = {}
    for leg, ppc, so in (("serial", 3 if nx <= 256 else 1, "libref_p3.so"),
                         ("openmp", 16 if nx <= 256 else 2, "libref_p3_omp.so")):
        p = workload_params(nx=nx, ppc=ppc, workload=workload)
        np_ = nx * nx * ppc
        x, v, q, m = lib.load_particles(p, workload_load(ppc, workload=workload), np_)
        st = orc.init_fields(p, x, v, q, m, 1.0)
        path = os.path.join(ROOT, "oracle", "_ref", so)
        threads = 1
        if os.path.exists(path):
            kind = "reference"
            if leg == "openmp":
                gomp = C.CDLL("libgomp.so.1")
                threads = int(gomp.omp_get_max_threads())
            el = _ref_p3_time(path, p, st)
        else:
            if leg == "openmp":
                continue
            kind = "port"
            t0 = time.perf_counter()
            orc.step(p, st)
            el = time.perf_counter() - t0
        if el is None:
            continue
        out[leg] = {"value": np_ / el, "unit": "particle-steps/s", "cores": threads, "kind": kind,
                    "sample": f"one step of the same deck on the {nx}x{nx} grid at {ppc} particles/cell "
                              f"({np_} particles), FP64, {threads} thread(s), {el:.1f} s",
                    "cpu_model": cpu["model"], "nproc": cpu["nproc"]}
    return out


def cpu_baseline_p1(nx=64, ppc=12, steps=2):
    """The reference's own serial P1 step (picard_step_orbit, compiled from
    Project1/ec_pic_serial_demo.cpp) on a bounded sample of the P1 Landau
    workload (same grid and deck, fewer particles per cell)."""
    from unstaggered_pic_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import State, ptr
    path = os.path.join(ROOT, "oracle", "_ref", "libref_p1.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    p = p1_params(nx)
    np_ = nx * nx * ppc
    x, v, q, m = lib.load_particles(p, workload_load(ppc, workload="landau_p1"), np_)
    st = State(np_, nx * nx)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    if L.ref_p1_init_from_particles(C.byref(p), C.c_double(1.0), C.byref(st.host())) != 0:
        return None
    el = 0.0
    for _ in range(steps):
        out = State(np_, nx * nx)
        it = C.c_int32()
        errs = np.zeros(2)
        t0 = time.perf_counter()
        rc = L.ref_p1_step(C.byref(p), C.byref(st.host()), C.byref(out.host()), C.byref(it), ptr(errs))
        el += time.perf_counter() - t0
        if rc != 0:
            return None
        st = out
    cpu = host_cpu()
    return {"value": np_ * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} steps of the P1 Landau deck on the {nx}x{nx} grid at {ppc} particles/cell "
                      f"({np_} particles), serial FP64, {el:.1f} s", "cpu_model": cpu["model"], "nproc": cpu["nproc"]}


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------
def measured_traffic(workload, nx, ppc, form):
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) for this workload and form,
    kept in profiles/; None where no measurement is kept."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}{nx}_ppc{ppc}.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))["forms"].get(form)
    return d["bytes_per_launch"] if d else None


def load_sim(lib, p, np_, ppc, workload, graphs):
    x, v, q, m = lib.load_particles(p, workload_load(ppc, workload=workload), np_)
    ctx = lib.Context(p)
    ctx.set_graphs(graphs)
    sim = lib.Simulation(ctx, np_)
    h = HostArrays(np_, p.nx * p.ny)
    h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
    t0 = time.perf_counter()
    sim.upload(h)  # host arrays -> HBM (pageable host memory)
    ctx.sync()
    t_upload = time.perf_counter() - t0
    sim.init_fields(1.0)
    return ctx, sim, t_upload


def train_network(args, lib):
    """Collect (features, converged correction) rows from --collect-steps
    steps of the workload in the mixed mode and train the network on the
    device; returns its parameters and the training record."""
    p = workload_params(args.nx, args.ppc, 1, args.workload)
    np_ = p.nx * p.ny * args.ppc
    ctx, sim, _ = load_sim(lib, p, np_, args.ppc, args.workload, args.graphs)
    # the collection run starts its inner solves from the physics predictor:
    # the rows (features, converged correction) do not depend on the start,
    # and the steps cost half those from v^n
    sim.set_guess_mode(1)
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
    P = sim.guess_params()
    sim.close()
    ctx.close()
    rec = {"collect_steps": args.collect_steps, "collect_seconds": t_collect, "training_rows": per_step * args.collect_steps,
           "train_epochs": args.train_epochs, "train_seconds": t_train,
           "train_loss_first_last": [float(losses[0]), float(losses[-1])],
           "outside_timed_region": True}
    return P, rec


def run_leg(args, lib, code, prec, mode, net=None):
    """Load, initialise and time one form; returns a dict of raw results."""
    if code == 1:
        p = p1_params(args.p1_nx)
        ppc, workload = args.p1_ppc, "landau_p1"
    else:
        p = workload_params(args.nx, args.ppc, prec, args.workload)
        ppc, workload = args.ppc, args.workload
    N = p.nx * p.ny
    np_ = N * ppc
    ctx, sim, t_upload = load_sim(lib, p, np_, ppc, workload, args.graphs)
    if prec == 1:
        if mode == 2:
            sim.set_guess_params(net)
        sim.set_guess_mode(mode)
    d0 = sim.diagnostics(1)
    # start-up steps of the run (Taylor-started history, first step without
    # the extrapolated current), then the warmup
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.spinup):
        sim.step()
    ctx.sync()
    t_spin = time.perf_counter() - t0
    for _ in range(args.warmup):
        sim.step()
    d_start = sim.diagnostics(1)
    infos = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(sim.step())
    ctx.sync()
    el = time.perf_counter() - t0
    d1 = sim.diagnostics(1)
    dep_choice = sim.deposit_variant() if prec == 1 else None
    # per-kernel HIP-event timing on the context stream over a few more steps
    # (event-bracketed launches run the host-driven outer loop; the kernels
    # are the same)
    kname = "p1_push" if code == 1 else ("p3_map" if prec == 0 else "p3_map_mixed")
    launches, kms, dep, tsteps, tinfos = 0, 0.0, (0, 0.0), 0, []
    if not args.no_timing and args.timing_steps > 0:
        ctx.set_timing(True)
        for _ in range(args.timing_steps):
            tinfos.append(sim.step())
        tsteps = args.timing_steps
        launches, kms = ctx.timing(kname)
        dep = ctx.timing("deposit_mixed") if prec == 1 else (0, 0.0)
        ctx.set_timing(False)
    # the reference's scope: compute_diagnostics and compute_step_diagnostics
    # after every step (P3:8835-8850, P1:1847-1848)
    asref = None
    if args.ref_steps > 0:
        sim.keep_step_start(True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.ref_steps):
            sim.step()
            sim.diagnostics(1)
            sim.step_diagnostics()
        ctx.sync()
        el_ref = time.perf_counter() - t0
        sim.keep_step_start(False)
        asref = {"ms_per_step": 1e3 * el_ref / args.ref_steps, "value": np_ * args.ref_steps / el_ref,
                 "steps": args.ref_steps,
                 "includes": "the step, compute_diagnostics and compute_step_diagnostics after every step"}
    sim.close()
    ctx.close()
    return {"p": p, "np": np_, "N": N, "elapsed": el, "infos": infos, "launches": launches, "kms": kms, "dep": dep,
            "timing_steps": tsteps, "timing_infos": tinfos,
            "kname": kname, "d0": d0, "d_start": d_start, "d1": d1, "dep_choice": dep_choice,
            "upload_seconds": t_upload, "as_reference": asref, "code": code, "prec": prec, "mode": mode,
            "spinup_ms_per_step": 1e3 * t_spin / args.spinup if args.spinup else None,
            "workload": workload, "ppc": ppc}


def physics(r):
    d0, ds, d1 = r["d0"], r["d_start"], r["d1"]
    infos = r["infos"]
    out = {"rel_total_energy_change_timed": (d1.total_energy - ds.total_energy) / abs(ds.total_energy),
           "rel_total_energy_change_since_init": (d1.total_energy - d0.total_energy) / abs(d0.total_energy),
           "outer_iterations": [i.iterations for i in infos],
           "gauss_rms": d1.gauss_rms, "rho_rms": d1.rho_rms, "gauge_rms": d1.gauge_rms,
           "divA_rms": d1.divA_rms, "psi_over_c2_rms": d1.psi_over_c2_rms,
           "avg_path_segments": float(np.mean([i.path.total_segments / max(i.path.n_particles, 1) for i in infos]))}
    if r["code"] == 3:
        out["map_evaluations"] = [i.map_evaluations for i in infos]
        # steps run as one device-driven graph launch (the others took the
        # host-driven loop: deposit-autotuner timings, refinement)
        out["graph_launched_steps"] = sum(1 for i in infos if i.graph_launch > 0)
        out["mean_inner_iterations"] = float(np.mean([i.local_iterations_mean for i in infos]))
        # summed over the step's sweeps (the guess acts in the first)
        out["inner_iterations_per_particle_step"] = float(np.mean([i.local_iterations_total for i in infos])) / r["np"]
    else:
        out["picard_iterations"] = [i.iterations for i in infos]
    return out


STATE_FORM = {(3, 0): "fp64_p3", (3, 1): "mixed_p3", (1, 0): "fp64_p1"}
KERNEL_NAME = {"p3_map": "k_p3_map", "p3_map_mixed": "k_p3_map_mixed", "p1_push": "k_p1_push"}


def roofline(r, args, traffic_form):
    """HBM roofline of the dominant kernel (see the module docstring)."""
    np_, N = r["np"], r["N"]
    balg = b_alg(STATE_FORM[(r["code"], r["prec"])], N, np_)
    avg_ms = r["kms"] / max(r["launches"], 1)
    per_launch = balg * np_
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if r["launches"] else None
    value = np_ * args.steps / r["elapsed"]
    step_gbs = balg * value / 1e9
    maps = sum(getattr(i, "map_evaluations", 0) for i in r["infos"])
    # the kernel's own byte count per launch (particle state, sweep-to-sweep
    # scratch and node records; see the kernel sources)
    if r["code"] == 1:
        model = np_ * (88.0 + 24.0 + 24.0) + 12.0 * 8 * N
    elif r["prec"] == 0:
        model = np_ * 88.0 + (np_ * 96.0 * args.steps) / max(maps, 1) + 12.0 * 8 * N
    else:
        later = 1.0 - args.steps / max(maps, 1)
        model = np_ * (48.0 + ((12.0 + 12.0 * later) if r["mode"] >= 1 else 0.0) + 64.0) + 32.0 * N
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
           "traffic": measured_traffic(r["workload"], r["p"].nx, r["ppc"], traffic_form) if traffic_form else None,
           "kernel": KERNEL_NAME[r["kname"]], "avg_launch_ms": avg_ms, "launches": r["launches"],
           "particles_per_launch": np_, "b_alg_bytes_per_particle_step": balg, "bytes_per_launch": per_launch,
           "step_level": {"achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                          "definition": "B_alg x particle-steps/s"},
           "traffic_model": {"bytes_per_launch": model,
                             "definition": "bytes the kernel itself reads and writes per launch, scratch included"}}
    return out


def compute_roofline(r, args):
    """FP64/FP32 VALU roofline of the sweep: operations of one inner
    iteration at r = 3, nq = 4 (counted from the kernel source: per Gauss
    point 2 axis stencils + 16 nodes x (3 gather + 12 D_A) FMAs; per
    iteration the end-point gather and the HC push) x measured iterations."""
    if r["code"] != 3 or not r["launches"]:
        return None
    infos = r["infos"]
    segs = np.mean([i.path.total_segments / max(i.path.n_particles, 1) for i in infos])
    flops_iter = segs * 4 * (2 * 4 * 24 + 16 * (2 * 15 + 6 * 2 + 3)) + 16 * 6 + 250
    tinfos = r["timing_infos"]
    local_total = sum(i.local_iterations_total for i in tinfos)
    peak = FP64_PEAK_TFLOPS if r["prec"] == 0 else FP32_PEAK_TFLOPS
    tflops = (flops_iter * local_total) / (r["kms"] * 1e-3) / 1e12
    return {"bound": "fp64_valu" if r["prec"] == 0 else "fp32_valu", "achieved": tflops, "peak": peak,
            "unit": "TFLOP/s", "frac": tflops / peak, "flops_per_inner_iteration": flops_iter}


def leg_summary(r, args, traffic_form=None):
    value = r["np"] * args.steps / r["elapsed"]
    out = {"value": value, "ms_per_step": 1e3 * r["elapsed"] / args.steps,
           "spinup_ms_per_step": r["spinup_ms_per_step"],
           "avg_launch_ms": r["kms"] / max(r["launches"], 1), "physics": physics(r),
           "roofline": roofline(r, args, traffic_form)}
    cr = compute_roofline(r, args)
    if cr:
        out["compute_roofline"] = cr
    if r["as_reference"]:
        out["as_reference"] = r["as_reference"]
    if r["prec"] == 1:
        out["deposit_kernel_avg_ms"] = (r["dep"][1] / r["dep"][0]) if r["dep"][0] else None
        out["deposit_variant_sorted_unsorted"] = r["dep_choice"]
    return out


FORM_NAME = {(3, 0, 0): "plain FP64", (3, 1, 0): "mixed precision", (3, 1, 1): "mixed precision + predictor",
             (3, 1, 2): "mixed precision + learned guess", (1, 0, 0): "plain FP64 (P1 form)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spinup", type=int, default=20,
                    help="start-up steps of each leg's run before the warmup (reported as spinup_ms_per_step)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="weibel",
                    help="deck; sets the grid side and particles per cell unless --nx/--ppc are given")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ppc", type=int, default=None)
    ap.add_argument("--members", type=int, default=16, help="ensemble members (--form ensemble)")
    ap.add_argument("--baseline-steps", type=int, default=20,
                    help="timed steps of the sequential and threaded ensemble baselines (--form ensemble)")
    ap.add_argument("--fp64-steps", type=int, default=10,
                    help="steps of the per-member energy comparison against FP64 (--form ensemble)")
    ap.add_argument("--form", choices=["all", "both", "fp64", "mixed", "predictor", "learned", "ensemble"],
                    default="all", help="all: the FP64 and three mixed legs, the P1 leg and the CPU legs; "
                                        "both: FP64 and mixed + learned; value is the last P3 leg")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None, help="alias of --form fp64|mixed")
    ap.add_argument("--collect-steps", type=int, default=200)
    ap.add_argument("--rows-per-step", type=int, default=4096)
    ap.add_argument("--train-epochs", type=int, default=15)
    ap.add_argument("--ref-steps", type=int, default=3, help="steps timed in the reference's scope per leg")
    ap.add_argument("--p1-nx", type=int, default=64)
    ap.add_argument("--p1-ppc", type=int, default=100)
    ap.add_argument("--no-p1", action="store_true", help="skip the P1 Landau leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the batched-ensemble sub-object of the default run (0: none)")
    ap.add_argument("--ensemble-guess", choices=["predictor", "learned"], default="predictor",
                    help="the ensemble members' inner start (learned: one network per load kind)")
    ap.add_argument("--ensemble-steps", type=int, default=20,
                    help="timed ensemble steps of that sub-object (after 18 warmup steps; the sequential "
                         "baselines: 5, the per-member FP64 comparison: 5)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="1: each P3 step's outer solve as one device-driven hipGraph (0: host-driven loop)")
    ap.add_argument("--timing-steps", type=int, default=4,
                    help="steps after the timed ones whose kernels are timed with HIP events")
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
    from unstaggered_pic_amd import lib
    if args.calibrate_copy_mib > 0:
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
    if args.workload == "landau_p1":
        args.p1_nx, args.p1_ppc = args.nx, args.ppc
        legs = [(1, 0, 0)]
    else:
        legs = {"all": [(3, 0, 0), (3, 1, 0), (3, 1, 1), (3, 1, 2)], "both": [(3, 0, 0), (3, 1, 2)],
                "fp64": [(3, 0, 0)], "mixed": [(3, 1, 0)], "predictor": [(3, 1, 1)],
                "learned": [(3, 1, 2)]}[args.form]
    net, train_rec = None, None
    if any(m == 2 for _, _, m in legs):
        print(f"# collecting {args.collect_steps} steps of training rows", file=sys.stderr, flush=True)
        net, train_rec = train_network(args, lib)
    runs = {}
    for leg in legs:
        runs[leg] = run_leg(args, lib, *leg, net=net)
        r = runs[leg]
        print(f"# {FORM_NAME[leg]}: {1e3 * r['elapsed'] / args.steps:.2f} ms/step", file=sys.stderr, flush=True)
    head = legs[-1]
    r = runs[head]
    code, prec = head[0], head[1]
    value = r["np"] * args.steps / r["elapsed"]
    summ = leg_summary(r, args, FORM_NAME[head])
    workload = "landau_p1" if code == 1 else args.workload
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
        "data": f"synthetic (seeded {'bi-Maxwellian' if workload == 'weibel' else workload} load, splitmix64 seed 12345)",
        "config": {"workload": workload_name(workload, r["p"].nx, r["ppc"]), "particles": r["np"],
                   "grid": [r["p"].nx, r["p"].ny], "spline_order": r["p"].spline_order,
                   "orbit_quad_order": r["p"].orbit_quad_order, "form": FORM_NAME[head],
                   "solver": "P1 global Picard" if code == 1 else
                   "P3 nested solve (Anderson outer, per-particle Picard inner)"},
        "roofline": summ["roofline"],
        "physics": summ["physics"],
        "host_upload_seconds": r["upload_seconds"],
        "spinup_steps": args.spinup, "spinup_ms_per_step": r["spinup_ms_per_step"],
    }
    for k in ("compute_roofline", "as_reference", "deposit_kernel_avg_ms", "deposit_variant_sorted_unsorted"):
        if k in summ:
            out[k] = summ[k]
    if train_rec:
        run_s = train_rec["collect_seconds"] + train_rec["train_seconds"]
        train_rec["share_of_2000_step_run"] = run_s / (run_s + 2000 * out["ms_per_step"] * 1e-3)
        out["learned_guess"] = train_rec
    legs_out = {}
    for leg in legs[:-1]:
        legs_out[FORM_NAME[leg]] = leg_summary(runs[leg], args, FORM_NAME[leg])
    if legs_out:
        out["legs"] = legs_out
    if (3, 0, 0) in runs and head != (3, 0, 0):
        out["speedup_vs_plain_fp64"] = value / legs_out["plain FP64"]["value"]
    if code == 3 and args.form == "all" and not args.no_p1:
        r1 = run_leg(args, lib, 1, 0, 0)
        print(f"# {FORM_NAME[(1, 0, 0)]}: {1e3 * r1['elapsed'] / args.steps:.2f} ms/step", file=sys.stderr,
              flush=True)
        s1 = leg_summary(r1, args)
        s1["config"] = {"workload": workload_name("landau_p1", r1["p"].nx, r1["ppc"]), "particles": r1["np"]}
        out["p1_landau"] = s1
    if code == 3 and args.form == "all" and args.ensemble_members > 0:
        from unstaggered_pic_amd import ensemble_bench
        # 18 warmup steps: each member's deposit autotuner times its eight
        # variants twice (one timing per step: the later evaluations of a
        # step deposit incrementally) on the host-driven path before its
        # steps join the ensemble graph
        e = ensemble_bench.measure(args.ensemble_members, args.ensemble_steps, 18, 5, 5)
        out["ensemble"] = {k: e[k] for k in (
            "metric", "value", "unit", "steps", "graph_launched_member_steps", "ms_per_step", "ensemble_vs_sequential",
            "ensemble_vs_sequential_host_loop", "max_rel_energy_change", "energy_error_vs_fp64")}
        out["ensemble"]["config"] = {"workload": e["config"]["workload"], "particles": e["config"]["particles"],
                                     "members": len(e["config"]["members"]), "member_draw": "seed 2024",
                                     "form": "mixed precision + predictor, one graph launch per ensemble step"}
        out["ensemble"]["sequential_ms_per_step"] = e["sequential"]["ms_per_step"]
        out["ensemble"]["energy_error_vs_fp64"].pop("definition", None)
    if not args.no_cpu_baseline:
        if code == 3:
            cpu = cpu_baseline_p3(args.nx, workload=args.workload)
            main_leg = cpu.get("openmp") or cpu.get("serial")
            if main_leg:
                out["cpu_baseline"] = dict(main_leg)
                if "serial" in cpu and main_leg is not cpu["serial"]:
                    out["cpu_baseline"]["serial"] = cpu["serial"]
        if code == 1 or (args.form == "all" and not args.no_p1):
            c1 = cpu_baseline_p1(args.p1_nx)
            if c1:
                if code == 1:
                    out["cpu_baseline"] = c1
                else:
                    out.setdefault("p1_landau", {})["cpu_baseline"] = c1
    print(json.dumps(out))


if __name__ == "__main__":
    main()
