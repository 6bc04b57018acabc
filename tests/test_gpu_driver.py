"""The deck-driven driver (unstaggered_pic_amd/upic_run, the replacement of
the reference codes' main loops) run on reference decks.

P1: the 300-step Landau-weak case rebuilt from the reference's config echo
(Project1/output/landau_weak_config_echo.txt): the driver's
landau_weak_diagnostics.csv matches the reference's own CSV column by column
-- total energy to 1e-12 relative, field energy to 1e-9, Picard iteration
counts exactly, Gauss RMS at roundoff, path statistics exactly -- and its
config echo matches line by line.

P3: a small Weibel-type deck in FP64 and in the mixed mode: the driver
writes the reference's 62-column diagnostics schema, every step converges and
the total energy is conserved (1e-12 FP64, 1e-6 mixed).
"""
import csv
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "unstaggered_pic_amd", "upic_run")

LANDAU_WEAK_DECK = """# deck rebuilt from landau_weak_config_echo.txt
test_case = landau_weak
nx = 32
ny = 1
nz = 1
Lx = 12.566370614359172
Ly = 1.0
Lz = 1.0
n_steps = 300
dt = 0.1
particles_per_cell = 32
spline_order = 1
orbit_quad_order = 12
split_orbit_at_knots = true
c = 1.0
eps0 = 1.0
mu0 = 1.0
n0 = 1.0
thermal_velocity = 1.0
landau_alpha = 0.01
perturbation_mode = 1
theory_omega = 1.4156618886
theory_gamma = -0.1533594669
picard_tol = 1.0e-10
picard_max_iter = 24
final_consistency_sweep = true
run_unit_tests = false
output_prefix = landau_weak
"""


def run_driver(tmp_path, code, deck):
    path = tmp_path / "deck.txt"
    path.write_text(deck)
    r = subprocess.run([EXE, code, str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return r


def test_driver_p1_landau_weak_matches_reference_csv(tmp_path):
    run_driver(tmp_path, "p1", LANDAU_WEAK_DECK)
    got = list(csv.DictReader(open(tmp_path / "landau_weak_diagnostics.csv")))
    ref = list(csv.DictReader(open(os.path.join(HERE, "golden", "landau_weak_diagnostics.csv"))))
    assert len(got) == len(ref) == 301
    assert list(got[0].keys()) == list(ref[0].keys())
    for k, (a, b) in enumerate(zip(got, ref)):
        assert abs(float(a["total_energy"]) - float(b["total_energy"])) / float(b["total_energy"]) < 1e-12, k
        assert abs(float(a["field_energy"]) - float(b["field_energy"])) / float(b["field_energy"]) < 1e-9, k
        assert float(a["gauss_rms"]) < 1e-13
        assert a["picard_iterations"] == b["picard_iterations"], k
        if k > 0:
            assert a["crossing_particles"] == b["crossing_particles"], k
            assert a["max_path_segments"] == b["max_path_segments"], k
    echo = (tmp_path / "landau_weak_config_echo.txt").read_text().splitlines()
    ref_echo = open(os.path.join(HERE, "golden", "landau_weak_config_echo.txt")).read().splitlines()
    assert echo == ref_echo
    sample = list(csv.DictReader(open(tmp_path / "landau_weak_particle_sample.csv")))
    assert len(sample) == 1024 and list(sample[0].keys()) == ["particle", "x", "y", "z", "vx", "vy", "vz", "q", "m"]
    # <prefix>_step_diagnostics.csv (compute_step_diagnostics, P1:1384-1439):
    # energy changes and works as the reference's own file, path statistics
    # exact; the chain-rule defect and the work residuals are rounding noise
    # in both (P1's E and J are exactly conjugate)
    got = list(csv.DictReader(open(tmp_path / "landau_weak_step_diagnostics.csv")))
    ref = list(csv.DictReader(open(os.path.join(HERE, "golden", "landau_weak_step_diagnostics.csv"))))
    assert len(got) == len(ref) == 300
    assert list(got[0].keys()) == list(ref[0].keys())
    for k, (a, b) in enumerate(zip(got, ref)):
        for col in ("delta_K", "delta_W", "particle_work", "grid_work"):
            assert abs(float(a[col]) - float(b[col])) <= 1e-9 * abs(float(b[col])) + 1e-14, (k, col)
        for col in ("crossing_particles", "total_crossings", "max_path_segments", "avg_path_segments"):
            assert float(a[col]) == float(b[col]), (k, col)
        assert float(a["A_chain_abs_rms"]) < 1e-15 and float(a["A_chain_rel_rms"]) < 1e-12, k
        for col in ("delta_total_energy", "deposit_gather_work_residual", "particle_energy_residual",
                    "field_energy_residual"):
            assert abs(float(a[col])) < 1e-12, (k, col)


P3_DECK = """test_case = bimaxwell
nx = 32
ny = 32
nz = 1
Lx = 7.853981633974483
Ly = 7.853981633974483
Lz = 1.0
n_steps = 6
dt = 0.05
particles_per_cell = 16
spline_order = 3
orbit_quad_order = 4
vth_x = 0.05
vth_y = 0.2
vth_z = 0.2
nonlinear_rtol = 1.0e-11
nonlinear_atol = 1.0e-13
nonlinear_max_iter = 80
anderson_damping = 0.5
particle_local_rtol = 1.0e-12
particle_local_atol = 1.0e-14
particle_local_max_iter = 64
precision = {prec}
nyquist_projection = true
output_prefix = weibel_small
"""


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_driver_p3_small_weibel(tmp_path, prec, tol):
    run_driver(tmp_path, "p3", P3_DECK.format(prec=prec))
    rows = list(csv.DictReader(open(tmp_path / "weibel_small_diagnostics.csv")))
    assert len(rows) == 7 and len(rows[0]) == 62
    te = np.array([float(r["total_energy"]) for r in rows])
    assert np.max(np.abs(te - te[0])) / te[0] < tol
    assert all(r["nonlinear_converged"] == "1" for r in rows[1:])
    assert all(float(r["gauss_rms"]) < 1e-12 for r in rows)


# The reference's cold relativistic Weibel eigenmode deck (its initialiser,
# P3:3961-4108, and the linear-theory columns of its CSV, P3:8427-8433 and
# 8576-8594) on a small 2-D grid.
P3_WEIBEL_DECK = """test_case = cold_relativistic_weibel
nx = 8
ny = 32
nz = 1
Lx = 6.283185307179586
Ly = 6.283185307179586
Lz = 1.0
n_steps = 20
dt = 0.1
particles_per_cell_pair = 4
spline_order = 2
orbit_quad_order = 6
v0 = 0.9
weibel_B0 = 1.0e-4
nonlinear_rtol = 1.0e-11
nonlinear_atol = 1.0e-13
nonlinear_max_iter = 80
anderson_damping = 0.5
particle_local_rtol = 1.0e-12
particle_local_atol = 1.0e-14
particle_local_max_iter = 64
nyquist_projection = true
output_prefix = weibel_eig
"""


def test_driver_p3_weibel_eigenmode_follows_theory(tmp_path):
    from unstaggered_pic_amd import lib as U
    from unstaggered_pic_amd.abi import P3_WEIBEL, make_params
    out = run_driver(tmp_path, "p3", P3_WEIBEL_DECK)
    rows = list(csv.DictReader(open(tmp_path / "weibel_eig_diagnostics.csv")))
    assert len(rows) == 21 and len(rows[0]) == 62
    p = make_params(code=3, nx=8, ny=32, nz=1, Lx=6.283185307179586, Ly=6.283185307179586, Lz=1.0,
                    spline_order=2)
    rate = U.p3_initial_state(p, U.p3_case(P3_WEIBEL, particles_per_cell_pair=4, v0=0.9,
                                           weibel_B0=1e-4))["growth_rate"]
    assert all(float(r["theory_growth_rate"]) == pytest.approx(rate, rel=1e-15) for r in rows)
    ratio = np.array([float(r["mode_over_theory"]) for r in rows])
    print("weibel |B_z(k_y)| / theory:", " ".join(f"{x:.5f}" for x in ratio), flush=True)
    assert ratio[0] == pytest.approx(1.0, rel=1e-14)
    assert np.max(np.abs(ratio - 1.0)) < 0.005
    te = np.array([float(r["total_energy"]) for r in rows])
    assert np.max(np.abs(te - te[0])) / te[0] < 1e-12
    assert "linear theory growth/damping rate" in out.stdout
    # <prefix>_step_diagnostics.csv (P3:8462-8473): the work identities of
    # every step close to rounding of the total energy
    # <prefix>_config_echo.txt with the reference's keys (P3:8217-8300), as
    # Project3/summarize_hpc_run.py and pic_analysis.py read it
    echo = dict(line.split("=", 1) for line in (tmp_path / "weibel_eig_config_echo.txt").read_text().splitlines())
    assert echo["test_case"] == "cold_relativistic_weibel" and int(echo["n_steps"]) == 20
    assert float(echo["target_time"]) == pytest.approx(2.0) and int(echo["particles"]) == 2 * 8 * 32 * 4
    assert echo["adaptive_dt"] == "false" and echo["uniform_particle_subcycling"] == "false"
    assert float(echo["auto_theory_growth_rate"]) == pytest.approx(rate, rel=1e-15)
    assert len(echo) == 80
    sample = list(csv.DictReader(open(tmp_path / "weibel_eig_particle_sample.csv")))
    assert len(sample) == 2 * 8 * 32 * 4 and list(sample[0].keys()) == ["particle", "x", "y", "z", "vx", "vy", "vz",
                                                                        "q", "m"]
    steps = list(csv.DictReader(open(tmp_path / "weibel_eig_step_diagnostics.csv")))
    assert len(steps) == 20 and len(steps[0]) == 46
    for r in steps:
        for k in ("particle_energy_residual", "field_energy_residual", "deposit_gather_work_residual"):
            assert abs(float(r[k])) < 1e-13 * te[0], k
        assert float(r["A_chain_abs_rms"]) < 1e-15


# The key values of the reference's manuscript two-stream deck
# (Project3/inputs/manuscript/tsi_main_dt01_T60.txt), cut from 600 steps to
# the first 100 (t = 10, the linear phase).
P3_TSI_DECK = """test_case = cold_relativistic_two_stream
nx = 32
ny = 32
nz = 32
Lx = 25.132741228718345
Ly = 25.132741228718345
Lz = 25.132741228718345
n_steps = 100
dt = 0.1
particles_per_cell_pair = 10
spline_order = 2
orbit_quad_order = 8
split_orbit_at_knots = true
kappa = 1.0
sigma1 = 1.0
sigma2 = 1.0
rho0 = 1.0
v0 = 0.9
perturbation = 0.0
two_stream_density_alpha = 0.005
two_stream_eigenmode = true
two_stream_finite_grid_theory = true
perturbation_mode = 1
nonlinear_solver = anderson_quasi_newton
nonlinear_rtol = 1.0e-11
nonlinear_atol = 1.0e-13
nonlinear_max_iter = 80
anderson_memory = 4
anderson_start = 2
anderson_damping = 0.5
anderson_regularization = 1.0e-12
anderson_growth_limit = 1.2
selective_particle_fallback = false
particle_local_rtol = 1.0e-11
particle_local_atol = 1.0e-13
particle_local_max_iter = 64
particle_local_damping = 1.0
nyquist_projection = true
output_prefix = tsi_main
"""


def test_driver_p3_manuscript_two_stream_linear_phase(tmp_path):
    from unstaggered_pic_amd import lib as U
    from unstaggered_pic_amd.abi import P3_TWO_STREAM, make_params
    run_driver(tmp_path, "p3", P3_TSI_DECK)
    rows = list(csv.DictReader(open(tmp_path / "tsi_main_diagnostics.csv")))
    assert len(rows) == 101
    L = 25.132741228718345
    p = make_params(code=3, nx=32, ny=32, nz=32, Lx=L, Ly=L, Lz=L, spline_order=2)
    rate = U.p3_initial_state(p, U.p3_case(P3_TWO_STREAM, particles_per_cell_pair=10, v0=0.9,
                                           two_stream_density_alpha=0.005, two_stream_eigenmode=1,
                                           two_stream_finite_grid_theory=1))["growth_rate"]
    assert rate == pytest.approx(0.0910913804, rel=1e-9)
    assert all(float(r["theory_growth_rate"]) == pytest.approx(rate, rel=1e-15) for r in rows)
    # |E_x(k_x)| / (E_0 exp(Gamma t)) over t = 0..10: measured within 5e-4
    ratio = np.array([float(r["mode_over_theory"]) for r in rows])
    print("two-stream |E_x(k)| / theory:", " ".join(f"{x:.5f}" for x in ratio[::10]), flush=True)
    assert np.max(np.abs(ratio - 1.0)) < 2e-3
    te = np.array([float(r["total_energy"]) for r in rows])
    assert np.max(np.abs(te - te[0])) / te[0] < 1e-12


# The reference's P2 example decks (Project2/examples/*.txt; their key
# values): the driver's runs against Project2/docs/validation_summary.json.
P2_DECK = """# {name}
example = {example}
method = {method}
charge_update = {update}
nx = 32
ny = 32
Lx = {L}
Ly = {L}
n_steps = {steps}
dt = {dt}
particles_per_cell = 4
cloud_particles = 4096
cloud_sigma = 1.0
cloud_drift = 0.01
cloud_thermal = 0.005
c = 1.0
eps0 = 1.0
mu0 = 1.0
n0 = 1.0
weibel_vx = 0.5
weibel_vy_max = 0.01
weibel_sine_perturb = 0.002
perturbation_mode = 1
seed = 12345
output_prefix = {name}
"""


@pytest.mark.parametrize("example", ["weibel", "cloud"])
@pytest.mark.parametrize("method", ["bdf1", "bdf2"])
@pytest.mark.parametrize("update", ["conserving", "naive"])
def test_driver_p2_matches_validation_summary(tmp_path, example, method, update):
    """The P2 deck driver on each example deck: final total energy and
    |B_z| against the reference's validation summary (1e-10 / 1e-8), the
    largest speed (1e-10), the conserving runs' Gauss/gauge/continuity at
    roundoff and the naive runs' violations to 1e-6."""
    import json
    name = f"{example}_{method}_{update}"
    L = 6.283185307179586 if example == "weibel" else 16.0
    steps, dt = (200, 0.02) if (example, method) == ("cloud", "bdf2") else (160, 0.05)
    out = run_driver(tmp_path, "p2", P2_DECK.format(name=name, example=example, method=method, update=update, L=L,
                                                   steps=steps, dt=dt))
    assert f"wrote {name}_diagnostics.csv" in out.stdout
    ref = json.load(open(os.path.join(HERE, "golden", "p2_validation_summary.json")))["runs"][name]
    rows = list(csv.DictReader(open(tmp_path / f"{name}_diagnostics.csv")))
    assert len(rows) == steps + 1 and len(rows[0]) == 18
    assert rows[0]["active_method"] == "initial" and rows[1]["active_method"] == method
    col = lambda k: np.array([float(r[k]) for r in rows])
    assert abs(col("total_energy")[-1] - ref["final_total_energy"]) <= 1e-10 * ref["final_total_energy"]
    assert abs(col("Bz_l2")[-1] - ref["final_Bz_l2"]) <= 1e-8 * ref["final_Bz_l2"]
    # the summary's maxima skip the t = 0 row (Project2/tools/validate_against_paper.py:59-60)
    assert abs(col("max_speed")[1:].max() - ref["max_speed"]) <= 1e-10 * ref["max_speed"]
    for k in ("continuity_rms", "gauge_rms", "gauss_rms"):
        got, want = col(k)[1:].max(), ref["max_" + k]
        if update == "conserving":
            assert got < 1e-13, k
        else:
            assert abs(got - want) <= 1e-6 * want, (k, got, want)
    echo = (tmp_path / f"{name}_config_echo.txt").read_text()
    assert f"static_ions = {4096 if example == 'cloud' else 0}" in echo
    sample = list(csv.DictReader(open(tmp_path / f"{name}_particle_sample.csv")))
    assert len(sample) == 2000 and list(sample[0].keys())[-1] == "gamma"


# ---------------------------------------------------------------------------
# The driver against the reference program itself: the reference's main
# (compiled from Project3/ec_pic_nyquist_projected.cpp into oracle/_ref,
# run in-process on the CPU) and upic_run p3 on the same deck write the same
# <prefix>_diagnostics.csv and <prefix>_step_diagnostics.csv.
# ---------------------------------------------------------------------------
# Project3/tests/test_solver.cpp:48-60: one inner iteration at rtol 2e-3
# forces the selective particle refinement
P3_FORCED_REFINEMENT_DECK = """test_case = cold_relativistic_weibel
nx = 4
ny = 4
nz = 4
Lx = 6.283185307179586
Ly = 6.283185307179586
Lz = 6.283185307179586
n_steps = 2
dt = 1.0
particles_per_cell_pair = 1
spline_order = 1
orbit_quad_order = 8
v0 = 0.9
weibel_B0 = 0.2
weibel_eigenmode = false
nonlinear_rtol = 1e-9
nonlinear_atol = 1e-11
nonlinear_max_iter = 32
anderson_damping = 0.5
uniform_local_rtol = 2e-3
uniform_local_atol = 1e-10
uniform_local_max_iter = 1
selective_particle_fallback = true
"""
# the reference's manuscript decks (inputs/manuscript/tsi_main_dt01_T60.txt,
# weibel_main_dt02_T100.txt) reduced as its validate.py does (4^3, two
# particles per cell pair, six steps)
P3_MANUSCRIPT_TSI_REDUCED = """test_case = cold_relativistic_two_stream
nx = 4
ny = 4
nz = 4
Lx = 25.132741228718345
Ly = 25.132741228718345
Lz = 25.132741228718345
n_steps = 6
dt = 0.1
particles_per_cell_pair = 2
spline_order = 2
orbit_quad_order = 8
split_orbit_at_knots = true
kappa = 1.0
sigma1 = 1.0
sigma2 = 1.0
rho0 = 1.0
v0 = 0.9
perturbation = 0.0
two_stream_density_alpha = 0.005
two_stream_eigenmode = true
two_stream_finite_grid_theory = true
perturbation_mode = 1
nonlinear_solver = anderson_quasi_newton
nonlinear_rtol = 1.0e-11
nonlinear_atol = 1.0e-13
nonlinear_max_iter = 80
anderson_memory = 4
anderson_start = 2
anderson_damping = 0.5
anderson_regularization = 1.0e-12
anderson_growth_limit = 1.2
selective_particle_fallback = false
particle_local_rtol = 1.0e-11
particle_local_atol = 1.0e-13
particle_local_max_iter = 64
particle_local_damping = 1.0
nyquist_projection = true
"""
P3_MANUSCRIPT_WEIBEL_REDUCED = """test_case = cold_relativistic_weibel
nx = 4
ny = 4
nz = 4
Lx = 6.283185307179586
Ly = 6.283185307179586
Lz = 6.283185307179586
n_steps = 6
dt = 0.2
particles_per_cell_pair = 2
spline_order = 2
orbit_quad_order = 8
split_orbit_at_knots = true
kappa = 1.0
sigma1 = 1.0
sigma2 = 1.0
rho0 = 1.0
v0 = 0.9
weibel_B0 = 1.0e-4
weibel_eigenmode = true
weibel_finite_grid_theory = true
perturbation_mode = 1
nonlinear_solver = anderson_quasi_newton
nonlinear_rtol = 1.0e-11
nonlinear_atol = 1.0e-13
nonlinear_max_iter = 80
anderson_memory = 4
anderson_start = 2
anderson_damping = 0.5
anderson_regularization = 1.0e-12
anderson_growth_limit = 1.2
uniform_particle_subcycling = false
selective_particle_fallback = false
selective_particle_max_substeps = 64
particle_local_rtol = 1.0e-12
particle_local_atol = 1.0e-14
particle_local_max_iter = 64
particle_local_damping = 1.0
adaptive_dt = false
nyquist_projection = true
"""
# columns: exact (integers, strings), to 1e-11 relative (state quantities),
# rounding-level residuals (absolute 1e-10 of the column's scale), skipped
EXACT = {"step", "nonlinear_iterations", "nonlinear_converged", "anderson_restarts", "anderson_steps",
         "picard_fallback_steps", "convergence_method", "backtracked_steps", "minimum_update_alpha",
         "local_particle_iterations_max", "uniform_substeps_used", "selective_fallback_used",
         "direct_attempt_iterations", "selective_refinement_restarts", "selective_refined_particles",
         "selective_total_particle_substeps", "selective_work_ratio", "selective_n1", "selective_n2", "selective_n4",
         "selective_n8", "selective_n16", "selective_n32", "selective_n64", "crossing_particles",
         "max_path_segments", "total_crossings", "avg_path_segments", "dt", "time"}
NOISE = {"gauss_rms", "gauss_max", "gauss_relative", "gauge_rms", "gauge_max", "gauge_relative",
         "nonlinear_rms_residual", "nonlinear_max_particle_residual", "nonlinear_absolute_rms_residual",
         "observed_convergence_rate", "estimated_map_contraction", "mean_rho", "A_chain_abs_rms", "A_chain_rel_rms",
         "A_chain_max_norm", "A_chain_mean_norm", "A_chain_work_defect", "delta_total_energy",
         "deposit_gather_work_residual", "particle_energy_residual", "field_energy_residual", "rho_mode",
         "E_mode_im", "Bz_y_mode_re", "Bz_y_mode_im", "mode_over_theory"}
SKIP = {"elapsed_wall_seconds"}
# a rounding-level residual over its normalisation: compared where the
# normalisation is resolved (a charge-neutral load has only rounding noise
# in rho, and the ratio is then noise over noise)
RATIO = {"gauss_relative": "gauss_normalization_rms", "gauge_relative": "gauge_normalization_rms"}


# decided at the inner solves' tolerance (a particle accepting its vbar one
# iteration earlier or later moves them): the same order of magnitude
TOL_LEVEL = {"nonlinear_rms_residual", "nonlinear_max_particle_residual", "nonlinear_absolute_rms_residual",
             "observed_convergence_rate", "estimated_map_contraction"}


# energy changes and works of a step: their rounding and the inner
# tolerance's effect are set by the total energy
ENERGY_CHANGE = {"delta_K", "delta_W", "delta_total_energy", "particle_work", "grid_work",
                 "deposit_gather_work_residual", "particle_energy_residual", "field_energy_residual"}


def compare_csv(got_path, want_path, energy_scale=0.0):
    got = list(csv.DictReader(open(got_path)))
    want = list(csv.DictReader(open(want_path)))
    assert len(got) == len(want)
    assert list(got[0].keys()) == list(want[0].keys())
    cols = list(want[0].keys())
    scale = {c: max([abs(float(r[c])) for r in want if c not in EXACT | {"convergence_method"}
                     and np.isfinite(float(r[c]))] or [0.0]) for c in cols if c not in ("convergence_method",)}
    bad = []
    for k, (a, b) in enumerate(zip(got, want)):
        for c in cols:
            if c in SKIP:
                continue
            if c == "convergence_method":
                if a[c] != b[c]:
                    bad.append((k, c, a[c], b[c]))
                continue
            x, y = float(a[c]), float(b[c])
            if c in EXACT:
                ok = x == pytest.approx(y, rel=1e-14, abs=0)
            elif c in RATIO:
                ok = float(b[RATIO[c]]) <= 1e-10 or abs(x - y) <= 1e-9
            elif not np.isfinite(y):
                ok = str(x) == str(y)
            elif c in TOL_LEVEL:
                ok = abs(x - y) <= 1e-9 or (y != 0.0 and 0.01 <= x / y <= 100.0)
            elif c in ENERGY_CHANGE:
                ok = abs(x - y) <= 1e-12 * energy_scale
            else:
                # 1e-11 relative, 1e-13 absolute (rounding-level entries
                # such as a Gauss residual or a vanishing mode); the
                # rounding-level residual columns to 1e-9 of their largest
                # value
                ok = abs(x - y) <= 1e-11 * abs(y) + 1e-13 + (1e-9 * scale[c] if c in NOISE else 0.0)
            if not ok:
                bad.append((k, c, x, y))
    assert not bad, bad[:20]


@pytest.mark.parametrize("name,deck", [("forced", P3_FORCED_REFINEMENT_DECK), ("tsi", P3_MANUSCRIPT_TSI_REDUCED),
                                       ("weibel", P3_MANUSCRIPT_WEIBEL_REDUCED)], ids=["forced", "tsi", "weibel"])
def test_driver_p3_matches_reference_program(tmp_path, name, deck):
    """Diagnostics and step-diagnostics CSVs of the driver equal the
    reference program's on the same deck: iteration counts, method strings,
    refinement columns and path statistics exactly; every other column to
    1e-11 relative plus 1e-13 absolute (rounding-level entries), the
    rounding-level residual columns to 1e-9 of their largest value, the
    step's energy changes and works to 1e-12 of the total energy, and the
    outer residual norms (set at the inner solves' tolerance) to a factor of
    100."""
    import ctypes as C
    from oracle_bind import Ref
    ref = Ref()
    if ref.p3 is None:
        pytest.skip("reference harness (oracle/_ref) not built")
    tail = "run_unit_tests = false\nfield_diagnostics = false\nselective_failure_diagnostics = false\n"
    (tmp_path / "ref").mkdir()
    (tmp_path / "ref" / "deck.txt").write_text(deck + tail + f"output_prefix = {tmp_path / 'ref' / name}\n")
    argv = (C.c_char_p * 3)(b"ref", str(tmp_path / "ref" / "deck.txt").encode(), None)
    assert ref.p3.ref_p3_main(2, argv) == 0
    run_driver(tmp_path, "p3", deck + tail + f"output_prefix = {name}\n")
    compare_csv(tmp_path / f"{name}_diagnostics.csv", tmp_path / "ref" / f"{name}_diagnostics.csv")
    rows = list(csv.DictReader(open(tmp_path / "ref" / f"{name}_diagnostics.csv")))
    te = max(abs(float(r["total_energy"])) for r in rows)
    compare_csv(tmp_path / f"{name}_step_diagnostics.csv", tmp_path / "ref" / f"{name}_step_diagnostics.csv", te)
    echo = (tmp_path / f"{name}_config_echo.txt").read_text().splitlines()
    ref_echo = (tmp_path / "ref" / f"{name}_config_echo.txt").read_text().splitlines()
    assert [l for l in echo if not l.startswith("output_prefix")] == [l for l in ref_echo if not l.startswith("output_prefix")]
