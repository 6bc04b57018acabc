/*
 * upic.h -- C ABI of the MI355X-native per-timestep PIC hot path.
 *
 * Drop-in boundary for the three teaching codes of sgong11/Unstaggered_PIC:
 *   P1 = Project1/ec_pic_serial_demo.cpp       (non-relativistic EC, global Picard)
 *   P2 = Project2/spectral_pic_part3.cpp       (explicit BDF spectral GM push)
 *   P3 = Project3/ec_pic_nyquist_projected.cpp (relativistic GM-HC-CN, nested solve)
 *
 * The reference has no FFI layer: its call surface is a set of file-static C++
 * functions called by each code's step driver.  Every entry point below names
 * the reference function it replaces (file:line).  Conventions of this ABI:
 *   - no exceptions cross it; every call returns a upic_status;
 *   - particle arrays are structure-of-arrays, component-major:
 *       x[0..np) = x, x[np..2np) = y, x[2np..3np) = z   ("3*np" arrays);
 *   - mesh vector fields are component-major (3*N), mesh index (i*ny+j)*nz+k,
 *     exactly the reference's Grid::index (P1:381-383, P3:1450-1452);
 *   - "dev" pointers are device (HBM) pointers; entry points that take host
 *     arrays say so.  All work is enqueued on the context's HIP stream;
 *     results are complete after upic_ctx_sync() or any call that returns
 *     host data.
 *   - numbers are FP64 unless the name says otherwise.
 */
#ifndef UPIC_H
#define UPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mirror the reference error conventions:
 * std::runtime_error for bad input (P3:1339-1433) -> UPIC_EINVAL,
 * NonlinearSolveFailure (P3:199-202)           -> UPIC_ENONLINEAR,
 * SelectiveParticleRefinementNeeded (P3:252-260) -> UPIC_EREFINE. */
typedef enum upic_status {
    UPIC_OK = 0,
    UPIC_EINVAL = 1,
    UPIC_ENONLINEAR = 2,
    UPIC_EREFINE = 3,
    UPIC_EDEVICE = 4,
    UPIC_ENOMEM = 5
} upic_status;

/* Which reference code's conventions and step driver to use. */
typedef enum upic_code {
    UPIC_CODE_P1 = 1, /* grid origin -L/2, Nyquist symbol kept (P1:386-392)   */
    UPIC_CODE_P2 = 2, /* grid origin 0, Nyquist symbol zeroed (P2:320-330)    */
    UPIC_CODE_P3 = 3  /* grid origin -L/2, even Nyquist zeroed (P3:1485-1504) */
} upic_code;

/* Per-particle failure reasons of the P3 inner solve (P3:204-211). */
typedef enum upic_particle_failure {
    UPIC_FAIL_NONE = 0,
    UPIC_FAIL_NONFINITE_MAP = 1,
    UPIC_FAIL_SUPERLUMINAL_MAP = 2,
    UPIC_FAIL_NONFINITE_ITERATE = 3,
    UPIC_FAIL_SUPERLUMINAL_ITERATE = 4,
    UPIC_FAIL_MAXIMUM_ITERATIONS = 5
} upic_particle_failure;

/* Particle/grid precision of a simulation handle. */
typedef enum upic_precision {
    UPIC_PREC_FP64 = 0,  /* everything FP64 (the reference's arithmetic)            */
    UPIC_PREC_MIXED = 1  /* P3 on 2-D grids: particles held as an int32 cell index
                            plus FP32 in-cell offsets, FP32 momenta/velocities;
                            FP32 orbit push and LDS deposit tiles flushed into
                            FP64 mesh currents; FP64 grids, FFT, diagnostics
                            and convergence tests                              */
} upic_precision;

/* Numerical parameters of one run.  Field names follow the reference decks
 * (P1 Config 233-303, P2 Config 174-231, P3 Config 837-990). */
typedef struct upic_params {
    int32_t code;                 /* upic_code                                   */
    int32_t nx, ny, nz;           /* powers of two (P1:395-396, P3:1521-1523)     */
    double Lx, Ly, Lz;
    double dt;
    int32_t spline_order;         /* 1..4 (P1, P3); P2 is fixed quadratic         */
    int32_t orbit_quad_order;     /* Gauss points per orbit segment               */
    int32_t split_orbit_at_knots; /* 1 = split orbit at B-spline knots            */
    int32_t nyquist_projection;   /* P3 only (P3:885)                             */
    double c;                     /* P1/P2 light speed; P3 kappa                  */
    double eps0;                  /* P1/P2 eps0; P3: 1/sigma1 is NOT used -- see sigma1 */
    double mu0;                   /* P1/P2 mu0                                    */
    double sigma1, sigma2;        /* P3 source coefficients (P3:955-957)          */
    /* P1 global Picard (P1:280-283) */
    double picard_tol;
    int32_t picard_max_iter;
    int32_t final_consistency_sweep;
    /* P3 outer current solve (P3:891-901) */
    int32_t nonlinear_solver_picard; /* 0 = anderson_quasi_newton, 1 = picard     */
    int32_t nonlinear_max_iter;
    double nonlinear_rtol, nonlinear_atol;
    int32_t anderson_memory, anderson_start;
    double anderson_damping, anderson_regularization, anderson_growth_limit;
    /* P3 inner per-particle solve (P3:917-920) */
    double local_rtol, local_atol;
    int32_t local_max_iter;
    double local_damping;
    /* P3 selective refinement (P3:927-931) */
    int32_t selective_particle_fallback;
    int32_t selective_particle_max_substeps;
    /* P2 (P2:174-231) */
    int32_t bdf_order;            /* 1 or 2                                       */
    int32_t charge_conserving;    /* 1 = continuity rho, 0 = naive redeposit      */
    /* GPU-only controls (no reference counterpart) */
    int32_t precision;            /* upic_precision                               */
    int32_t sort_interval;        /* cell-sort particles every k steps (0 = never) */
    int32_t learned_guess;        /* P3 mixed: initial guess mode (upic_guess_mode) */
    int32_t reserved[5];
} upic_params;

/* Orbit splitting statistics (PathStats, P1:771-776, P3:2509-2514). */
typedef struct upic_path_stats {
    int64_t n_particles;
    int64_t crossing_particles;
    int64_t total_segments;
    int32_t max_segments;
    int32_t pad;
} upic_path_stats;

/* The update that produced the accepted outer iterate of a P3 step (the
 * reference's convergence_method, P3:5898, 6122-6130): its CSV string is
 * "nested_full_dt_" (or, after a selective refinement, "selective_") followed
 * by picard_fallback, picard, current_anderson, current_anderson_backtracked,
 * current_picard or current_picard_backtracked. */
typedef enum upic_update_method {
    UPIC_UPDATE_PICARD_FALLBACK = 0,  /* converged at the first evaluation (Anderson solver) */
    UPIC_UPDATE_PICARD = 1,           /* converged at the first evaluation (Picard solver)   */
    UPIC_UPDATE_CURRENT_ANDERSON = 2,
    UPIC_UPDATE_CURRENT_ANDERSON_BACKTRACKED = 3,
    UPIC_UPDATE_CURRENT_PICARD = 4,
    UPIC_UPDATE_CURRENT_PICARD_BACKTRACKED = 5
} upic_update_method;

/* Outcome of one time step (P1 PicardHistoryEntry 1503-1507, P3
 * NonlinearSolveInfo 5493-5519). */
typedef struct upic_step_info {
    int32_t iterations;              /* P1 picard_iterations / P3 nonlinear_iterations */
    int32_t converged;
    double current_error;            /* P1 relRMS(J - J_prev) of the last iterate      */
    double velocity_error;           /* P1 relRMS(v_next - v_guess)                    */
    /* P3 outer solve */
    int32_t hit_max_iter;
    int32_t anderson_restarts;
    int32_t anderson_steps;
    int32_t picard_fallback_steps;
    int32_t backtracked_steps;
    int32_t map_evaluations;
    double minimum_update_alpha;
    double final_rms_residual;
    double final_max_particle_residual;
    double final_absolute_rms_residual;
    double estimated_map_contraction;   /* |dG| / |dx| of the last two outer iterates (P3:5983-5996) */
    /* P3 inner solve */
    int32_t local_iterations_max;
    int32_t failed_particles;
    double local_iterations_mean;
    double local_scaled_residual_max;
    int64_t local_iterations_total;  /* summed over all map evaluations of the step   */
    upic_path_stats path;
    /* P3 outer update and selective refinement (NonlinearSolveInfo
     * P3:5508-5518, set at 6171-6185 and 6346-6356); without a refinement the
     * full-dt layout counts: one substep per particle, work ratio 1. */
    int32_t update_method;                   /* upic_update_method                          */
    int32_t uniform_substeps_used;           /* largest substep count of the accepted layout */
    int32_t selective_fallback_used;         /* 1 = the step needed the particle refinement  */
    int32_t direct_attempt_iterations;       /* outer iterations of the failed full-dt try   */
    int32_t selective_refinement_restarts;   /* substep doublings before the accepted layout */
    int32_t graph_launch;                    /* no reference counterpart: 0 = host-driven outer
                                                loop, 1 = the step's own device-driven graph,
                                                2 = one graph launch for the whole ensemble
                                                (upic_ctx_set_graphs, upic_ensemble_step)       */
    int64_t selective_refined_particles;     /* particles with more than one substep         */
    int64_t selective_total_particle_substeps;
    double selective_work_ratio;             /* total substeps / particles                    */
    int64_t selective_level_counts[7];       /* particles at 1, 2, 4, ..., 64 substeps         */
    double observed_convergence_rate;        /* (final / first rms residual)^(1/(iters-1)),
                                                P3:5535-5547                                    */
} upic_step_info;

/* Per-step diagnostics (P1 Diagnostics 1037-1055, P3 Diagnostics 3469-3499,
 * P2 Diagnostics 882-896). */
typedef struct upic_diagnostics {
    double gauss_rms, gauss_max, gauge_rms, gauge_max;
    double divE_rms, rho_rms, psi_over_c2_rms, divA_rms;
    double field_energy, kinetic_energy, total_energy;
    double rho_mode, E_mode_re, E_mode_im, E_mode_abs;
    double Bz_y_mode_re, Bz_y_mode_im, Bz_y_mode_abs;
    double mean_rho;
    /* P2 */
    double continuity_rms, Bz_l2, B_l2, E_l2, rho_l2, J_l2, max_speed, mean_gamma;
} upic_diagnostics;

/* ------------------------------------------------------------------------- */
/* Context: device, stream, grid symbols, quadrature.                        */
/* ------------------------------------------------------------------------- */
typedef struct upic_ctx upic_ctx;

/* Creates a context on the current HIP device (make_grid P1:394-418,
 * P3:1520-1546; gauss_legendre_01 P1:739-763).  Validates the deck values the
 * reference validates (P3:1339-1433). */
upic_status upic_ctx_create(const upic_params* params, upic_ctx** out);
upic_status upic_ctx_destroy(upic_ctx* ctx);
upic_status upic_ctx_sync(upic_ctx* ctx);
/* Last error message of this context (static storage, valid until next call). */
const char* upic_last_error(const upic_ctx* ctx);
/* Kernel timing with HIP events on the context stream (no reference
 * counterpart; the reference only times whole steps, P3:145-194).  When
 * enabled, each launch of a timed kernel family ("p1_push", "p1_push_final",
 * "p1_deposit_guess", "p3_map", "p3_map_mixed", "deposit_mixed", "p2_deposit",
 * "p2_push") is bracketed by
 * events and synchronised; upic_ctx_timing returns the accumulated launches
 * and milliseconds since the last upic_ctx_set_timing call. */
upic_status upic_ctx_set_timing(upic_ctx* ctx, int enable);
upic_status upic_ctx_timing(upic_ctx* ctx, const char* name, int64_t* launches, double* total_ms);
/* No reference counterpart: hipGraph replay.  A uniform-layout P3 step
 * (FP64 or mixed precision, Anderson memory <= 4) runs its whole outer solve
 * as one captured graph: device controller kernels make the outer loop's
 * decisions (Anderson, backtracking, restarts, convergence) between gated
 * map evaluations, so the step needs one launch and one synchronisation;
 * graphs are cached per buffer configuration.  A step the controller does
 * not finish within the unrolled evaluations continues two at a time, and
 * one it reports failed (particle failures, backtracking, nonfinite
 * currents) runs again on the host-driven path, as the reference would.
 * Other steps (P1, and the map evaluations of the host-driven path) replay
 * their field solves from graphs.  enable: 0 off (default), 1 on with the
 * number of unrolled evaluations adapted to the steps, n >= 2 on with n
 * unrolled evaluations.  Results agree with the host-driven path to
 * roundoff (the same decisions; sums in another order). */
upic_status upic_ctx_set_graphs(upic_ctx* ctx, int enable);
/* Deterministic mode, the counterpart of the reference's deposit into
 * per-thread meshes summed in a fixed order (deposit_current_to_mesh,
 * P3:2177-2216; the orbit deposit alike): with it on, an FP64 P3 step gives
 * the same bits on every run from the same state.  The orbit current is
 * accumulated as exact 64-bit fixed-point integer words (order-independent
 * sums, converted to FP64 once per map evaluation), the cell sort keeps the
 * particles of a cell in their storage order, and norms, Gram entries and
 * diagnostics sums run as one workgroup each.  Off by default (slower);
 * results agree with the default mode to roundoff.  The mixed-precision
 * sweep's FP32 tiles are not covered. */
upic_status upic_ctx_set_deterministic(upic_ctx* ctx, int enable);
/* Gauss-Legendre nodes/weights on [0,1] used by the context (host arrays of
 * orbit_quad_order entries). */
upic_status upic_quadrature(const upic_ctx* ctx, double* s, double* w);

/* ---- Device memory plumbing for FFI callers without a HIP allocator ------ */
/* (no reference counterpart) */
upic_status upic_device_count(int* count);
upic_status upic_device_alloc(size_t bytes, void** ptr);
upic_status upic_device_free(void* ptr);
upic_status upic_memcpy_htod(void* dst, const void* src, size_t bytes);
upic_status upic_memcpy_dtoh(void* dst, const void* src, size_t bytes);
upic_status upic_memcpy_dtod(void* dst, const void* src, size_t bytes);

/* ---- Group A: kernel-level entry points (device pointers) --------------- */

/* path_average_deposit_current (P1:883-906, P3:2784-2857).  J is overwritten
 * (3*N).  stats may be NULL (host pointer). */
upic_status upic_deposit_current_orbit(upic_ctx* ctx, const double* x0, const double* dx,
                                       const double* vbar, const double* q, int64_t np,
                                       double* J, upic_path_stats* stats);
/* path_average_gather_vector (P1:858-881, P3:2689-2713). out is 3*np. */
upic_status upic_gather_orbit(upic_ctx* ctx, const double* x0, const double* dx,
                              const double* field3, int64_t np, double* out);
/* orbit_discrete_gradient_A(_interval) (P1:908-934, P3:2958-2984).
 * theta(s) = theta0 + (theta1-theta0) s; D is 9*np, D[(a*3+b)*np + p] = dA_a/dx_b. */
upic_status upic_orbit_grad_A(upic_ctx* ctx, const double* x0, const double* dx,
                              const double* A_n, const double* A_np1, double theta0,
                              double theta1, int64_t np, double* D);
/* solve_local_particle_substep (P3:4930-4983) with
 * evaluate_local_particle_substep (4843-4884): the inner per-particle Picard
 * solve of one substep [theta0, theta1] of length h from `guess`, with A at
 * time theta interpolated between A_n and A_end and grad phi_mid = gpm
 * (P3 contexts; local tolerances, damping and iteration cap from the
 * parameters; on 2-D grids the z part of gpm is taken as zero, as for any
 * gradient there).  Device arrays: guess, x0, P0 (3*np), q, m (np), A_n, A_end,
 * gpm (3N) in; vbar, P_end, v_end (3*np; zero for failed particles),
 * iterations (np) and reasons (np, LocalParticleFailureReason P3:204-211)
 * out. */
upic_status upic_local_substep(upic_ctx* ctx, const double* guess, const double* x0, const double* P0,
                               const double* q, const double* m, double h, double theta0, double theta1,
                               const double* A_n, const double* A_end, const double* gpm, int64_t np,
                               double* vbar, double* P_end, double* v_end, int32_t* iterations, uint8_t* reasons);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): upic_local_substep with the same arguments and results,
 * calling no hipMalloc/hipFree.  Its 2-D node records (8 doubles per mesh
 * node) live in a static device array of the library, one per GPU, sized for
 * meshes of at most 2^18 nodes (512 x 512; 16 MiB); a 2-D context with more
 * nodes gets UPIC_EINVAL.  3-D contexts use no scratch.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (two streams, two
 * threads) race and must be serialised by the caller. */
upic_status upic_local_substep_noalloc(upic_ctx* ctx, const double* guess, const double* x0, const double* P0,
                                       const double* q, const double* m, double h, double theta0, double theta1,
                                       const double* A_n, const double* A_end, const double* gpm, int64_t np,
                                       double* vbar, double* P_end, double* v_end, int32_t* iterations,
                                       uint8_t* reasons);
/* gather_vector_at over particles (P1:666-678, P3:2267-2346). out is 3*np. */
upic_status upic_gather_point(upic_ctx* ctx, const double* pos, const double* field3,
                              int64_t np, double* out);
/* deposit_scalar_to_mesh (P1:617-623, P3:2117-2126); mesh (N) overwritten. */
upic_status upic_deposit_scalar_point(upic_ctx* ctx, const double* pos, const double* values,
                                      int64_t np, int divide_by_dV, double* mesh);
/* fft3d on interleaved complex data (2*N doubles) (P1:456-485, P3:1614-1656,
 * P2:406-424); inverse scales by 1/N. */
upic_status upic_fft(upic_ctx* ctx, double* data, int inverse);
/* spectral_gradient_scalar / spectral_divergence_vector / spectral_curl
 * (P1:505-546, P3:1721-1795). */
upic_status upic_spectral_gradient(upic_ctx* ctx, const double* f, double* out3);
upic_status upic_spectral_divergence(upic_ctx* ctx, const double* V3, double* out);
upic_status upic_spectral_curl(upic_ctx* ctx, const double* A3, double* out3);
/* continuity_charge_update (P1:995-1001, P3:3288-3296). */
upic_status upic_continuity(upic_ctx* ctx, const double* rho_n, const double* J3, double dt,
                            double* rho_out);
/* cn_potential_update (P1:977-993, P3:3249-3270).  P1 uses c/eps0/mu0,
 * P3 uses kappa/sigma1/sigma2 from the context parameters. */
upic_status upic_cn_potentials(upic_ctx* ctx, const double* phi, const double* psi,
                               const double* A3, const double* U3, const double* rho_mid,
                               const double* J3, double dt, double* phi_out, double* psi_out,
                               double* A_out3, double* U_out3);
/* project_nyquist (P3:2044-2078) on ncomp consecutive fields. */
upic_status upic_project_nyquist(upic_ctx* ctx, double* f, int ncomp);
/* solve_periodic_poisson_for_phi (P1:959-968, P3:3196-3206). */
upic_status upic_poisson(upic_ctx* ctx, const double* rho, double* phi);
/* solve_bdf_wave_scalar (P2:829-848), P2 contexts:
 * [b0/(c^2 dt^2) + k^2] u^ = S^ - sum_{l>=1} b_l u^_{l-1} / (c^2 dt^2) with
 * b = a * a (convolution) of the BDF1 {1, -1} or BDF2 {3/2, -2, 1/2}
 * coefficients (P2:769-785).  src (N); hist holds u^n, u^{n-1}, ... as
 * consecutive N-arrays (2 for BDF1, 4 for BDF2); out (N) may alias src. */
upic_status upic_bdf_wave(upic_ctx* ctx, const double* src, const double* hist, int order, double* out);

/* ---- Group B: per-step drivers on a device-resident simulation ---------- */
typedef struct upic_sim upic_sim;

/* Host-side particle/mesh state, reference State layout (P1:970-975,
 * P3:3208-3222; P2 ParticleSet/FieldState).  Particle arrays are 3*np SoA;
 * any pointer may be NULL where the code does not use it. */
typedef struct upic_host_state {
    int64_t np;
    double *x, *v, *v_prev, *P, *vbar_last; /* 3*np each                          */
    double *q, *m;                         /* np each                            */
    double *phi, *psi, *rho;               /* N each                             */
    double *A, *U, *J;                     /* 3*N each                           */
    /* P2 history (BDF, newest first): 5*N each for phi and rho, 5*3*N for A.
     * fixed_ion_rho is N.  (P2:787-810, 1126-1131) */
    double *phi_hist, *A_hist, *rho_hist, *fixed_ion_rho;
} upic_host_state;

upic_status upic_sim_create(upic_ctx* ctx, int64_t np, upic_sim** out);
upic_status upic_sim_destroy(upic_sim* sim);
/* Host arrays -> the resident state (NULL arrays keep their values).  Waits
 * for the context's stream first: work still queued there may touch the
 * same buffers. */
upic_status upic_sim_upload(upic_sim* sim, const upic_host_state* st);
upic_status upic_sim_download(upic_sim* sim, upic_host_state* st);
/* P1 picard_step_orbit (P1:1509-1552). */
upic_status upic_step_p1(upic_sim* sim, upic_step_info* info);
/* P3 nonlinear_step_orbit (P3:6477-6526) -> nonlinear_step_uniform_subcycled
 * (5878-6215) with the selective current map (5235-5431). */
upic_status upic_step_p3(upic_sim* sim, upic_step_info* info);
/* P2 one iteration of run_simulation's step loop (P2:1133-1189). */
upic_status upic_step_p2(upic_sim* sim, upic_step_info* info);
/* No reference counterpart (the reference runs one simulation per process):
 * one step of each of n independent simulations -- any codes, decks and
 * sizes, each on its own context (its own stream) -- advanced concurrently,
 * so that small members fill the GPU together.  When every member is a P3
 * simulation whose context has graphs on (upic_ctx_set_graphs), the
 * members that can take the device-driven step (deposit autotuner settled)
 * are one captured hipGraph, one branch per member, launched once and
 * waited on once (info.graph_launch = 2; a member that needs more
 * evaluations than its branch holds continues on its own stream, one that
 * needs the selective refinement takes the host loop), and the others take
 * the host-driven step; the members' host work before and after the graph
 * (step start, predictor, completion, host-driven steps) runs on
 * min(n, threads) host threads (threads <= 0: 16).  Otherwise that many host
 * threads step the members.
 * infos (n entries) and status (n entries) may be NULL.  Returns UPIC_OK
 * when every member stepped, else the first member status that is not
 * UPIC_OK (each member's error text is on its context); UPIC_EINVAL when
 * two members share a context. */
upic_status upic_ensemble_step(upic_sim* const* sims, int32_t n, int32_t threads, upic_step_info* infos,
                               upic_status* status);
/* compute_diagnostics (P1:1057-1087, P3:3521-3574, P2:898-962) on the current
 * state; mode = perturbation_mode. */
upic_status upic_sim_diagnostics(upic_sim* sim, int mode, upic_diagnostics* out);
/* Initial mesh state and particle momenta from the uploaded x, v, q, m: the
 * common tail of the reference initialisers (P1:1287-1303 / P3:3920-3937):
 * rho = deposit(q)/dV + n0 with the mean removed, J = point current deposit,
 * phi from Poisson, psi = A = U = 0, P = m v (P1) or m gamma v (P3), and the
 * history velocity v_prev by the Taylor start-up (P1:1296-1303,
 * P3:3423-3441). */
upic_status upic_sim_init_fields(upic_sim* sim, double n0);
/* upic_sim_init_fields with the mesh vector potential A and U (host arrays of
 * 3N, NULL = zero; P3 only), as the Weibel initialiser sets them
 * (P3:4081-4104): A and U are Nyquist-projected, E = -grad phi - U and
 * B = curl A drive the history velocity, and P = m gamma v + q A(x)
 * (project_initial_state, P3:3398-3405). */
upic_status upic_sim_init_fields_AU(upic_sim* sim, double n0, const double* A, const double* U);
/* ---- Step diagnostics (compute_step_diagnostics, P1:1384-1439,
 * P3:4501-4570): the orbit residual R_A of the vector-potential chain rule
 * (compute_RA_orbit_residual, P3:4242-4296) and the energy/work identities
 * of the last step, from the step's start state (kept on the device while
 * upic_sim_keep_step_start is enabled, after the step's cell sort) and the
 * current state.  vbar is (v^n + v^{n+1})/2 for P1 and the accepted
 * vbar_last for P3.  Fields in the order of the reference's
 * <prefix>_step_diagnostics.csv columns (P1:1818-1819). */
typedef struct upic_step_diagnostics {
    double A_chain_abs_rms, A_chain_rel_rms, A_chain_max_norm, A_chain_mean_norm, A_chain_work_defect;
    double delta_K, delta_W, delta_total_energy, particle_work, grid_work;
    double deposit_gather_work_residual, particle_energy_residual, field_energy_residual;
    int64_t crossing_particles, total_crossings;
    double avg_path_segments;
    int32_t max_path_segments, pad;
} upic_step_diagnostics;
/* Keep (enable != 0) or stop keeping each step's start state (P1, P3;
 * 6 np + 7 N doubles of HBM). */
upic_status upic_sim_keep_step_start(upic_sim* sim, int enable);
/* The step diagnostics of the last step (UPIC_EINVAL without a kept start).
 * After a P3 step that needed the selective per-particle refinement they are
 * compute_subcycled_step_diagnostics (P3:4363-4473, dispatched at
 * 4505-4512): every substep's chain-rule defect (A at the substep's times
 * theta0, theta1 = s/nsub, (s+1)/nsub), the works over the substep orbits
 * and the mesh work against the step's current. */
upic_status upic_sim_step_diagnostics(upic_sim* sim, upic_step_diagnostics* out);
/* The substep paths of the last P3 step when it was selectively refined
 * (State::subcycle_offset_last, subcycle_x0_last, subcycle_vbar_last,
 * P3:3220-3221, filled at 6211-6213): host arrays in the original particle
 * order -- offsets (np + 1; particle p's substeps are entries
 * offsets[p] .. offsets[p+1]-1), start points x0 and orbit-averaged
 * velocities vbar (3 * n_total each, component-major).  *n_total is set to
 * the number of substeps (0 when the last step was not refined); with
 * offsets == NULL only *n_total is set. */
upic_status upic_sim_substep_paths(upic_sim* sim, int64_t* n_total, int64_t* offsets, double* x0, double* vbar);
/* Device pointers of the resident state (for zero-copy callers).  With
 * UPIC_PREC_MIXED the particle pointers are an FP64 view rebuilt from the
 * FP32 state by this call; writes to them reach the simulation only through
 * upic_sim_upload. */
upic_status upic_sim_device_state(upic_sim* sim, upic_host_state* dev);
/* Per-particle failure reasons of the last P3 map evaluation (np bytes, host). */
upic_status upic_sim_failure_reasons(upic_sim* sim, uint8_t* reasons);

/* ---- Learned initial guess of the P3 inner solve (mixed precision) ----- */
/* No reference counterpart: the reference starts every per-particle Picard
 * solve (solve_local_particle_substep, P3:4930-4983) from v^n (P3:5318).
 * These entry points change only that starting iterate; the iteration to
 * the unchanged tolerance decides convergence.  Modes: v^n; the physics
 * predictor v^n + (h/2) q/(m gamma)(F - v (v.F)/c^2), F = E + v x B at the
 * start point; the predictor plus a 12 -> 64 -> 64 -> 3 SiLU network of the
 * features (v, E, B, (v . grad) E), BF16 on MFMA with FP32 accumulation,
 * fused into the push kernel.  The network is trained on the device on rows
 * (features, converged correction) the solver itself produces. */
#define UPIC_MLP_NPARAM 5214  /* [in_mean 12|in_scale 12|out_scale 3|W1 64x12|b1 64|W2 64x64|b2 64|W3 3x64|b3 3] */
typedef enum upic_guess_mode {
    UPIC_GUESS_PLAIN = 0,
    UPIC_GUESS_PREDICTOR = 1,
    UPIC_GUESS_LEARNED = 2
} upic_guess_mode;
/* Collect one training row per `stride` stored particles from the final map
 * evaluation of each following step, up to `capacity` rows (stride 0 stops). */
upic_status upic_sim_collect_guess_data(upic_sim* sim, int64_t stride, int64_t capacity);
/* Rows collected so far; copies min(n, max_rows) rows of 15 floats
 * (12 features, 3 target components) to `rows` when it is not NULL. */
upic_status upic_sim_guess_data(upic_sim* sim, float* rows, int64_t max_rows, int64_t* n_rows);
/* Train the network on the collected rows (Adam, minibatches of `batch`,
 * `epochs` passes), install it and switch to UPIC_GUESS_LEARNED.  Writes the
 * mean minibatch loss of each epoch to loss_per_epoch when not NULL. */
upic_status upic_sim_train_guess(upic_sim* sim, int32_t epochs, int32_t batch, double lr, uint64_t seed,
                                 double* loss_per_epoch);
/* Modes 1 and 2 apply to the first outer evaluation of a step; the later
 * evaluations (whose fields differ from it by the outer update only) start
 * each particle from the orbit-averaged velocity it converged to in the
 * previous evaluation.  That start remembers earlier evaluations within the
 * inner tolerance, which slows some outer solves (weak net currents, the
 * first steps of a run whose network guesses pass at their first
 * iteration): when two steps in a row that start from the extrapolated
 * current (every step but a run's first) need more than
 * four outer evaluations, every evaluation starts from mode 1 or 2 for the
 * next 16 steps, and the vbar start is then tried again (each relapse
 * doubles the pause, up to 1024 steps).  Mode 0 keeps the reference's v^n
 * throughout. */
upic_status upic_sim_set_guess_mode(upic_sim* sim, int32_t mode);
upic_status upic_sim_guess_params(upic_sim* sim, float* params);          /* UPIC_MLP_NPARAM floats out */
upic_status upic_sim_set_guess_params(upic_sim* sim, const float* params); /* and in (installs the network) */
/* ---- Deposit autotuner (mixed precision) --------------------------------- */
/* No reference counterpart (the reference deposits inside its per-particle
 * loop, P3:5380-5409).  The mixed-precision current deposit has eight
 * variants: 0 particle order with a wavefront reduction of the common
 * stencil window into an 8 x 32-node LDS tile, 1 one workgroup per 64
 * cells of a row, 2 eight workgroups per 64 cells of a row, 3 particle
 * order with per-particle atomics, 4 and 5 the wavefront reduction into
 * 4 x 64 and 16 x 16-node tiles, 6 and 7 as 1 and 2 on 8 x 8-cell blocks;
 * 1, 2, 6 and 7 need a freshly cell-sorted particle order (1 and 2:
 * ny % 64 == 0; 6 and 7: nx, ny multiples of 8 and at least 16).  By
 * default (variant -1) the
 * first map evaluations time every usable variant twice, one variant per
 * evaluation, and then keep the fastest, separately for sorted and unsorted
 * particle orders.  A variant >= 0 pins it (an unusable pin runs variant 0).
 * upic_sim_deposit_variant reports the choices (-1 while still tuning). */
upic_status upic_sim_set_deposit_variant(upic_sim* sim, int32_t variant);
upic_status upic_sim_deposit_variant(upic_sim* sim, int32_t* sorted_choice, int32_t* unsorted_choice);

/* Standalone network kernels on device buffers: the MFMA inference the push
 * kernel runs (X: n rows of 12 features -> Y: n rows of 3, scaled outputs)
 * and the FP32 training loss/gradient of rows of 15 floats (grad: UPIC_MLP_NPARAM
 * floats, entries before W1 left zero; loss: host double). */
upic_status upic_mlp_forward(upic_ctx* ctx, const float* params, const float* X, int64_t n, float* Y);
upic_status upic_mlp_loss_grad(upic_ctx* ctx, const float* params, const float* rows, int64_t n, float* grad,
                               double* loss);

/* ---- Particle loading for the synthetic benchmark workloads ------------- */
/* No reference counterpart: the reference's initialisers (P1:1151-1304,
 * P3:3689-4108) build quiet-start eigenmodes; the benchmark workloads need
 * seeded random loads.  Generation is host-side C++ (splitmix64 + Box-Muller)
 * and deterministic for a seed. */
typedef enum upic_load_kind {
    UPIC_LOAD_LANDAU = 1,   /* density 1 + alpha cos(k x) by inverse CDF, y uniform,
                               Maxwellian velocities of width vth[d] (P3:
                               Maxwellian momenta u = gamma v / c, below c)     */
    UPIC_LOAD_BIMAXWELL = 2,/* uniform density, Gaussian momenta u = gamma v / c of
                               widths vth[d] (stays sub-luminal)                 */
    UPIC_LOAD_BEAMS = 3     /* cell lattice + noise*dx jitter, two beams at +-v0 in x
                               with Maxwellian spread vth[d]                    */
} upic_load_kind;
typedef struct upic_load_spec {
    int32_t kind;           /* upic_load_kind                                  */
    int32_t ppc;            /* particles per cell                              */
    uint64_t seed;
    double n0, alpha;
    int32_t mode;           /* perturbation mode (k = 2 pi mode / Lx)          */
    int32_t pad;
    double vth[3];
    double v0, noise;
} upic_load_spec;
/* Fills st->x, v, q, m (host arrays of st->np = nx*ny*nz*ppc particles) with
 * q = -n0 Lx Ly Lz / np and m = |q|. */
upic_status upic_load_particles(const upic_params* params, const upic_load_spec* spec, upic_host_state* st);

/* ---- The reference's P3 initial conditions ------------------------------ */
typedef enum upic_p3_case_kind {
    UPIC_P3_TWO_STREAM = 1, /* initialize_two_stream_3d_linear_x (P3:3689-3851)     */
    UPIC_P3_WEIBEL = 2,     /* initialize_weibel_filamentation_3d (P3:3961-4108)    */
    UPIC_P3_LANDAU = 3      /* initialize_landau_damping_1d3d (P3:3853-3937)        */
} upic_p3_case_kind;
/* The deck keys the initialisers read (Config, P3:872-987; parser
 * P3:1201-1334), with the reference's defaults in brackets. */
typedef struct upic_p3_case {
    int32_t kind;                          /* upic_p3_case_kind                     */
    int32_t particles_per_cell_pair;       /* [4]                                   */
    int32_t perturbation_mode;             /* [1]                                   */
    int32_t tsi_transverse_mode;           /* [1]                                   */
    int32_t two_stream_eigenmode;          /* [1]                                   */
    int32_t two_stream_finite_grid_theory; /* [1]                                   */
    int32_t weibel_eigenmode;              /* [1]                                   */
    int32_t weibel_finite_grid_theory;     /* [1]                                   */
    double n0, v0, perturbation;
    double two_stream_density_alpha;       /* [-1]: < 0 = velocity seed only        */
    double weibel_B0;                      /* [-1]: < 0 = perturbation              */
    double landau_alpha;                   /* [-1]: < 0 = perturbation              */
    double thermal_velocity;               /* [1]                                   */
    double tsi_transverse_seed;            /* [0]                                   */
} upic_p3_case;
/* Fills upic_p3_case with the reference's defaults for a kind. */
void upic_p3_case_defaults(int32_t kind, upic_p3_case* cs);
/* The particle part of the initialiser: st->x, v, q, m for st->np particles
 * (two streams of nx*ny*nz*ppc each for two-stream and Weibel, nx*ny*nz*ppc
 * quiet-start samples for Landau), and, when st->A / st->U are not NULL, the
 * mesh A and U (3N each; nonzero for the Weibel eigenmode only).  With
 * st->x == NULL only st->np is set.  growth_rate (may be NULL) receives the
 * cold linear-theory rate (cold_two_stream_growth_rate P3:4160-4195,
 * cold_weibel_growth_rate 4197-4226; 0 for Landau).  The input checks and
 * their messages are the reference's (UPIC_EINVAL, upic_last_error(NULL)).
 * The field tail is upic_sim_init_fields_AU. */
upic_status upic_p3_initial_state(const upic_params* params, const upic_p3_case* cs, upic_host_state* st,
                                  double* growth_rate);

#ifdef __cplusplus
}
#endif
#endif /* UPIC_H */
