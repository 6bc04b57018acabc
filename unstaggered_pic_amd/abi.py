"""ctypes mirror of the plain-data structs in include/upic.h.

Only data layouts live here; the compute entry points are bound in
:mod:`unstaggered_pic_amd.lib`.
"""
import ctypes as C

UPIC_OK, UPIC_EINVAL, UPIC_ENONLINEAR, UPIC_EREFINE, UPIC_EDEVICE, UPIC_ENOMEM = range(6)
STATUS_NAMES = {0: "UPIC_OK", 1: "UPIC_EINVAL", 2: "UPIC_ENONLINEAR", 3: "UPIC_EREFINE",
                4: "UPIC_EDEVICE", 5: "UPIC_ENOMEM"}
CODE_P1, CODE_P2, CODE_P3 = 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("code", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("dt", C.c_double),
        ("spline_order", C.c_int32), ("orbit_quad_order", C.c_int32),
        ("split_orbit_at_knots", C.c_int32), ("nyquist_projection", C.c_int32),
        ("c", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
        ("sigma1", C.c_double), ("sigma2", C.c_double),
        ("picard_tol", C.c_double), ("picard_max_iter", C.c_int32),
        ("final_consistency_sweep", C.c_int32),
        ("nonlinear_solver_picard", C.c_int32), ("nonlinear_max_iter", C.c_int32),
        ("nonlinear_rtol", C.c_double), ("nonlinear_atol", C.c_double),
        ("anderson_memory", C.c_int32), ("anderson_start", C.c_int32),
        ("anderson_damping", C.c_double), ("anderson_regularization", C.c_double),
        ("anderson_growth_limit", C.c_double),
        ("local_rtol", C.c_double), ("local_atol", C.c_double),
        ("local_max_iter", C.c_int32), ("local_damping", C.c_double),
        ("selective_particle_fallback", C.c_int32),
        ("selective_particle_max_substeps", C.c_int32),
        ("bdf_order", C.c_int32), ("charge_conserving", C.c_int32),
        ("precision", C.c_int32), ("sort_interval", C.c_int32), ("learned_guess", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class PathStats(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("crossing_particles", C.c_int64),
                ("total_segments", C.c_int64), ("max_segments", C.c_int32), ("pad", C.c_int32)]


class StepInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32),
        ("current_error", C.c_double), ("velocity_error", C.c_double),
        ("hit_max_iter", C.c_int32), ("anderson_restarts", C.c_int32),
        ("anderson_steps", C.c_int32), ("picard_fallback_steps", C.c_int32),
        ("backtracked_steps", C.c_int32), ("map_evaluations", C.c_int32),
        ("minimum_update_alpha", C.c_double), ("final_rms_residual", C.c_double),
        ("final_max_particle_residual", C.c_double), ("final_absolute_rms_residual", C.c_double),
        ("estimated_map_contraction", C.c_double),
        ("local_iterations_max", C.c_int32), ("failed_particles", C.c_int32),
        ("local_iterations_mean", C.c_double), ("local_scaled_residual_max", C.c_double),
        ("local_iterations_total", C.c_int64),
        ("path", PathStats),
        ("update_method", C.c_int32), ("uniform_substeps_used", C.c_int32),
        ("selective_fallback_used", C.c_int32), ("direct_attempt_iterations", C.c_int32),
        ("selective_refinement_restarts", C.c_int32), ("graph_launch", C.c_int32),
        ("selective_refined_particles", C.c_int64), ("selective_total_particle_substeps", C.c_int64),
        ("selective_work_ratio", C.c_double), ("selective_level_counts", C.c_int64 * 7),
        ("observed_convergence_rate", C.c_double),
    ]


# upic_update_method -> the reference's convergence_method suffix (P3:5898, 6122-6130)
UPDATE_METHOD_NAMES = ("picard_fallback", "picard", "current_anderson", "current_anderson_backtracked",
                       "current_picard", "current_picard_backtracked")


class Diagnostics(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gauss_rms", "gauss_max", "gauge_rms", "gauge_max",
        "divE_rms", "rho_rms", "psi_over_c2_rms", "divA_rms",
        "field_energy", "kinetic_energy", "total_energy",
        "rho_mode", "E_mode_re", "E_mode_im", "E_mode_abs",
        "Bz_y_mode_re", "Bz_y_mode_im", "Bz_y_mode_abs", "mean_rho",
        "continuity_rms", "Bz_l2", "B_l2", "E_l2", "rho_l2", "J_l2", "max_speed", "mean_gamma")]


class HostState(C.Structure):
    _fields_ = [("np", C.c_int64)] + [(n, C.POINTER(C.c_double)) for n in (
        "x", "v", "v_prev", "P", "vbar_last", "q", "m", "phi", "psi", "rho", "A", "U", "J",
        "phi_hist", "A_hist", "rho_hist", "fixed_ion_rho")]


LOAD_LANDAU, LOAD_BIMAXWELL, LOAD_BEAMS = 1, 2, 3


class LoadSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ppc", C.c_int32), ("seed", C.c_uint64), ("n0", C.c_double),
                ("alpha", C.c_double), ("mode", C.c_int32), ("pad", C.c_int32), ("vth", C.c_double * 3),
                ("v0", C.c_double), ("noise", C.c_double)]


class StepDiagnostics(C.Structure):
    """upic_step_diagnostics (the reference's step-diagnostics CSV columns)."""
    _fields_ = [(n, C.c_double) for n in (
        "A_chain_abs_rms", "A_chain_rel_rms", "A_chain_max_norm", "A_chain_mean_norm", "A_chain_work_defect",
        "delta_K", "delta_W", "delta_total_energy", "particle_work", "grid_work", "deposit_gather_work_residual",
        "particle_energy_residual", "field_energy_residual")] + [
        ("crossing_particles", C.c_int64), ("total_crossings", C.c_int64), ("avg_path_segments", C.c_double),
        ("max_path_segments", C.c_int32), ("pad", C.c_int32)]


P3_TWO_STREAM, P3_WEIBEL, P3_LANDAU = 1, 2, 3  # upic_p3_case_kind


class P3Case(C.Structure):
    """upic_p3_case: the deck keys of the reference's P3 initialisers."""
    _fields_ = [("kind", C.c_int32), ("particles_per_cell_pair", C.c_int32), ("perturbation_mode", C.c_int32),
                ("tsi_transverse_mode", C.c_int32), ("two_stream_eigenmode", C.c_int32),
                ("two_stream_finite_grid_theory", C.c_int32), ("weibel_eigenmode", C.c_int32),
                ("weibel_finite_grid_theory", C.c_int32), ("n0", C.c_double), ("v0", C.c_double),
                ("perturbation", C.c_double), ("two_stream_density_alpha", C.c_double), ("weibel_B0", C.c_double),
                ("landau_alpha", C.c_double), ("thermal_velocity", C.c_double), ("tsi_transverse_seed", C.c_double)]


def make_load(kind, ppc, seed=12345, n0=1.0, alpha=0.0, mode=1, vth=(0.0, 0.0, 0.0), v0=0.0, noise=0.0):
    s = LoadSpec()
    s.kind, s.ppc, s.seed, s.n0, s.alpha, s.mode = kind, ppc, seed, n0, alpha, mode
    s.vth[:] = list(vth)
    s.v0, s.noise = v0, noise
    return s


def make_params(**kw):
    """Params with the reference's deck defaults (P1 Config 233-303,
    P3 Config 837-990), overridden by keyword arguments."""
    p = Params()
    defaults = dict(code=CODE_P1, nx=4, ny=4, nz=1, Lx=6.283185307179586, Ly=6.283185307179586,
                    Lz=1.0, dt=0.00625, spline_order=2, orbit_quad_order=16, split_orbit_at_knots=1,
                    nyquist_projection=0, c=1.0, eps0=1.0, mu0=1.0, sigma1=1.0, sigma2=1.0,
                    picard_tol=1e-11, picard_max_iter=16, final_consistency_sweep=1,
                    nonlinear_solver_picard=0, nonlinear_max_iter=32, nonlinear_rtol=1e-12,
                    nonlinear_atol=1e-13, anderson_memory=4, anderson_start=2, anderson_damping=0.7,
                    anderson_regularization=1e-12, anderson_growth_limit=1.2, local_rtol=1e-11,
                    local_atol=1e-13, local_max_iter=24, local_damping=1.0,
                    selective_particle_fallback=0, selective_particle_max_substeps=64,
                    bdf_order=1, charge_conserving=1, precision=PREC_FP64, sort_interval=1,
                    learned_guess=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


# Tolerance floors of the mixed-precision mode: an FP32 orbit map resolves
# the orbit-averaged velocity to ~1e-7 relative, so the inner fixed-point
# test needs a relative tolerance of about 1e-6 to be reachable; its noise is
# set by the size of the momentum terms (~0.2 c here), not by the particle's
# own speed, so a slow particle needs an absolute floor of 1e-7 c (with 1e-9
# the inner solve of a nearly stopped particle never converges).  The mesh
# current of an FP32 map carries ~2.5e-7 relative rms noise; an outer
# tolerance within a few times of that makes the Anderson iteration chase
# the noise (8-9 map evaluations per step on the 256^2 Weibel deck against 5
# at 3e-6, with the same energy conservation and field-energy history to
# 1e-6 of the FP64 run), so the outer floor is 3e-6.
MIXED_TOLERANCE_FLOORS = dict(local_rtol=1e-6, local_atol=1e-7, nonlinear_rtol=3e-6, nonlinear_atol=1e-9)


def with_mixed_precision(p):
    """Copy of Params `p` in the mixed-precision mode with its tolerances
    raised to the FP32 floors (values already above a floor are kept)."""
    q = Params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(Params))
    q.precision = PREC_MIXED
    for k, v in MIXED_TOLERANCE_FLOORS.items():
        setattr(q, k, max(getattr(q, k), v))
    return q
