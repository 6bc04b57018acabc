// oracle/ref_p3.cpp -- extern "C" access to Project3/ec_pic_nyquist_projected.cpp.
// TEST INFRASTRUCTURE ONLY (see ref_common.hpp).
#define main upic_ref_p3_program_main
#include REF_P3_SOURCE
#undef main
#include "ref_common.hpp"

static Config p3_cfg(const upic_params* p) {
    Config c;
    c.nx = p->nx; c.ny = p->ny; c.nz = p->nz;
    c.Lx = p->Lx; c.Ly = p->Ly; c.Lz = p->Lz;
    c.dt = p->dt;
    c.spline_order = p->spline_order;
    c.orbit_quad_order = p->orbit_quad_order;
    c.split_orbit_at_knots = p->split_orbit_at_knots != 0;
    c.nyquist_projection = p->nyquist_projection != 0;
    c.nonlinear_solver = p->nonlinear_solver_picard ? "picard" : "anderson_quasi_newton";
    c.nonlinear_rtol = p->nonlinear_rtol; c.nonlinear_atol = p->nonlinear_atol;
    c.nonlinear_max_iter = p->nonlinear_max_iter;
    c.anderson_memory = p->anderson_memory; c.anderson_start = p->anderson_start;
    c.anderson_damping = p->anderson_damping; c.anderson_regularization = p->anderson_regularization;
    c.anderson_growth_limit = p->anderson_growth_limit;
    c.uniform_local_rtol = p->local_rtol; c.uniform_local_atol = p->local_atol;
    c.uniform_local_max_iter = p->local_max_iter; c.uniform_local_damping = p->local_damping;
    c.selective_particle_fallback = p->selective_particle_fallback != 0;
    c.selective_particle_max_substeps = p->selective_particle_max_substeps;
    c.kappa = p->c; c.sigma1 = p->sigma1; c.sigma2 = p->sigma2;
    c.output_prefix = "/tmp/upic_ref_p3";
    return c;
}
static State p3_load(const upic_host_state* s, int N, Real dt) {
    State st;
    st.x = ref_load3<Vec3>(s->x, s->np); st.v = ref_load3<Vec3>(s->v, s->np);
    st.v_prev = ref_load3<Vec3>(s->v_prev, s->np);
    st.vbar_last = s->vbar_last ? ref_load3<Vec3>(s->vbar_last, s->np) : st.v;
    st.P = ref_load3<Vec3>(s->P, s->np);
    st.q.assign(s->q, s->q + s->np); st.m.assign(s->m, s->m + s->np);
    st.phi.assign(s->phi, s->phi + N); st.psi.assign(s->psi, s->psi + N); st.rho.assign(s->rho, s->rho + N);
    st.A = ref_loadF3<VecField>(s->A, N); st.U = ref_loadF3<VecField>(s->U, N); st.J = ref_loadF3<VecField>(s->J, N);
    st.dt_last = dt;
    return st;
}
static void p3_store(const State& st, upic_host_state* s) {
    ref_store3(st.x, s->x); ref_store3(st.v, s->v); ref_store3(st.v_prev, s->v_prev);
    if (s->vbar_last && !st.vbar_last.empty()) ref_store3(st.vbar_last, s->vbar_last);
    ref_store3(st.P, s->P);
    std::copy(st.q.begin(), st.q.end(), s->q); std::copy(st.m.begin(), st.m.end(), s->m);
    std::copy(st.phi.begin(), st.phi.end(), s->phi); std::copy(st.psi.begin(), st.psi.end(), s->psi);
    std::copy(st.rho.begin(), st.rho.end(), s->rho);
    ref_storeF3(st.A, s->A); ref_storeF3(st.U, s->U); ref_storeF3(st.J, s->J);
}

extern "C" {
const char* ref_p3_last_error(void) { return ref_err.c_str(); }

int ref_p3_fft(const upic_params* p, double* data, int inverse) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c);
        CField z(g.N);
        for (int i = 0; i < g.N; ++i) z[i] = Complex(data[2 * i], data[2 * i + 1]);
        fft3d(z, g, inverse != 0);
        for (int i = 0; i < g.N; ++i) { data[2 * i] = z[i].real(); data[2 * i + 1] = z[i].imag(); }
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_bspline(int degree, int n, const double* r, double* w, double* dw) {
    try { for (int i = 0; i < n; ++i) { w[i] = centered_bspline(degree, r[i]); dw[i] = centered_bspline_derivative(degree, r[i]); } return 0; }
    catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_project_nyquist(const upic_params* p, double* f) {
    try { const Config c = p3_cfg(p); const Grid g = make_grid(c); Field x(f, f + g.N); project_nyquist(x, g); std::copy(x.begin(), x.end(), f); return 0; }
    catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_deposit_current_orbit(const upic_params* p, const double* x0, const double* dx, const double* vbar,
                                 const double* q, int64_t np, double* J, int64_t* stats4) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        PathStats ps;
        const VecField Jf = path_average_deposit_current(ref_load3<Vec3>(x0, np), ref_load3<Vec3>(dx, np),
            ref_load3<Vec3>(vbar, np), std::vector<Real>(q, q + np), g, c.spline_order, qd, c.split_orbit_at_knots, &ps);
        ref_storeF3(Jf, J);
        stats4[0] = ps.n_particles; stats4[1] = ps.crossing_particles; stats4[2] = ps.total_segments; stats4[3] = ps.max_segments;
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_orbit_grad_A_interval(const upic_params* p, const double* x0, const double* dx, const double* An,
                                 const double* Anp1, double th0, double th1, int64_t np, double* D) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        const auto X = ref_load3<Vec3>(x0, np), Dx = ref_load3<Vec3>(dx, np);
        const VecField A0 = ref_loadF3<VecField>(An, g.N), A1 = ref_loadF3<VecField>(Anp1, g.N);
        for (int64_t i = 0; i < np; ++i) {
            const Mat3 M = orbit_discrete_gradient_A_interval(X[i], Dx[i], A0, A1, th0, th1, g, c.spline_order, qd, c.split_orbit_at_knots);
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) D[(a * 3 + b) * np + i] = M[a][b];
        }
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_local_solve(const upic_params* p, const double* guess, const double* x0, const double* P0, const double* q,
                       const double* m, double h, double th0, double th1, const double* A_n, const double* A_end,
                       const double* gpm, int64_t np, double* vbar, double* P_end, double* v_end, int32_t* iters,
                       int32_t* reasons) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        State st; st.A = ref_loadF3<VecField>(A_n, g.N);
        const VecField Ae = ref_loadF3<VecField>(A_end, g.N), Gp = ref_loadF3<VecField>(gpm, g.N);
        const auto G = ref_load3<Vec3>(guess, np), X = ref_load3<Vec3>(x0, np), P0v = ref_load3<Vec3>(P0, np);
        std::vector<Vec3> vb(np, Vec3{0, 0, 0}), Pe(np, Vec3{0, 0, 0}), ve(np, Vec3{0, 0, 0});
        for (int64_t i = 0; i < np; ++i) {
            LocalParticleStep ls; LocalParticleSolveReport rep;
            solve_local_particle_substep(G[i], X[i], P0v[i], q[i], m[i], h, th0, th1, st, Ae, Gp, g, c, qd, vb[i], ls, rep);
            Pe[i] = ls.P_end; ve[i] = ls.v_end; iters[i] = rep.iterations; reasons[i] = static_cast<int32_t>(rep.failure_reason);
        }
        ref_store3(vb, vbar); ref_store3(Pe, P_end); ref_store3(ve, v_end);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// kind 0 = landau, 1 = two_stream (cold relativistic), 2 = weibel
int ref_p3_init(const upic_params* p, int kind, int ppc, double alpha, double vth, double v0, double pert, int mode,
                double n0, upic_host_state* st) {
    try {
        Config c = p3_cfg(p);
        c.test_case = kind == 0 ? "landau_weak" : (kind == 1 ? "cold_relativistic_two_stream" : "cold_relativistic_weibel");
        c.particles_per_cell_pair = ppc; c.landau_alpha = alpha; c.thermal_velocity = vth; c.v0 = v0;
        c.perturbation = pert; c.perturbation_mode = mode; c.n0 = n0;
        if (kind == 1) c.two_stream_density_alpha = alpha;
        if (kind == 2) c.weibel_B0 = alpha;
        const Grid g = make_grid(c);
        const State s = initialize_state(c, g);
        if (!st->x) { st->np = static_cast<int64_t>(s.x.size()); return 0; }
        p3_store(s, st);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// ref_p3_init with the eigenmode switch of the two-stream / Weibel loads
// (two_stream_eigenmode, weibel_eigenmode; P3:3689-4108)
int ref_p3_init_eigen(const upic_params* p, int kind, int ppc, double alpha, double vth, double v0, double pert,
                      int mode, double n0, int eigenmode, upic_host_state* st) {
    try {
        Config c = p3_cfg(p);
        c.test_case = kind == 0 ? "landau_weak" : (kind == 1 ? "cold_relativistic_two_stream" : "cold_relativistic_weibel");
        c.particles_per_cell_pair = ppc; c.landau_alpha = alpha; c.thermal_velocity = vth; c.v0 = v0;
        c.perturbation = pert; c.perturbation_mode = mode; c.n0 = n0;
        if (kind == 1) { c.two_stream_density_alpha = alpha; c.two_stream_eigenmode = eigenmode != 0; }
        if (kind == 2) { c.weibel_B0 = alpha; c.weibel_eigenmode = eigenmode != 0; }
        const Grid g = make_grid(c);
        const State s = initialize_state(c, g);
        if (!st->x) { st->np = static_cast<int64_t>(s.x.size()); return 0; }
        p3_store(s, st);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_step(const upic_params* p, const upic_host_state* in, upic_host_state* out, int32_t* iters, double* info3) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        std::vector<NonlinearHistoryEntry> hist;
        NonlinearSolveInfo info;
        const State s1 = nonlinear_step_orbit(p3_load(in, g.N, c.dt), g, c, qd, hist, &info, 0.0, nullptr);
        p3_store(s1, out);
        *iters = info.iterations;
        info3[0] = info.final_rms_residual; info3[1] = info.local_particle_iterations_mean;
        info3[2] = info.local_particle_iterations_max;
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// compute_step_diagnostics (P3:4501-4570) of s0 -> s1; out17 in the order
// of upic_step_diagnostics (13 doubles, crossing_particles, total_crossings,
// avg_path_segments, max_path_segments)
int ref_p3_step_diagnostics(const upic_params* p, const upic_host_state* s0, const upic_host_state* s1,
                            double* out17) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        const State a = p3_load(s0, g.N, c.dt), b = p3_load(s1, g.N, c.dt);
        const StepDiagnostics d = compute_step_diagnostics(a, b, g, c.dt, c.spline_order, qd, c.sigma1, c.sigma2,
                                                           c.kappa, c.split_orbit_at_knots);
        const double v[17] = {d.A_chain_abs_rms, d.A_chain_rel_rms, d.A_chain_max_norm, d.A_chain_mean_norm,
                              d.A_chain_work_defect, d.delta_K, d.delta_W, d.delta_total_energy, d.particle_work,
                              d.grid_work, d.deposit_gather_work_residual, d.particle_energy_residual,
                              d.field_energy_residual, (double)d.crossing_particles, (double)d.total_crossings,
                              d.average_path_segments, (double)d.max_path_segments};
        std::memcpy(out17, v, sizeof(v));
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
static void p3_step_diag17(const StepDiagnostics& d, double* out17) {
    const double v[17] = {d.A_chain_abs_rms, d.A_chain_rel_rms, d.A_chain_max_norm, d.A_chain_mean_norm,
                          d.A_chain_work_defect, d.delta_K, d.delta_W, d.delta_total_energy, d.particle_work,
                          d.grid_work, d.deposit_gather_work_residual, d.particle_energy_residual,
                          d.field_energy_residual, (double)d.crossing_particles, (double)d.total_crossings,
                          d.average_path_segments, (double)d.max_path_segments};
    std::memcpy(out17, v, sizeof(v));
}
// compute_step_diagnostics of s0 -> s1 where s1 carries substep paths
// (offsets np + 1; x0, vbar 3*n component-major): the reference dispatches
// to compute_subcycled_step_diagnostics (P3:4505-4512)
int ref_p3_step_diagnostics_paths(const upic_params* p, const upic_host_state* s0, const upic_host_state* s1,
                                  const int64_t* off, int64_t n, const double* x0, const double* vb, double* out17) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        const State a = p3_load(s0, g.N, c.dt);
        State b = p3_load(s1, g.N, c.dt);
        b.subcycle_offset_last.assign(off, off + s0->np + 1);
        b.subcycle_x0_last = ref_load3<Vec3>(x0, n);
        b.subcycle_vbar_last = ref_load3<Vec3>(vb, n);
        p3_step_diag17(compute_step_diagnostics(a, b, g, c.dt, c.spline_order, qd, c.sigma1, c.sigma2, c.kappa,
                                                c.split_orbit_at_knots), out17);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// One nonlinear_step_orbit of the reference (selective fallback as the
// parameters say) with its NonlinearSolveInfo (info16: iterations,
// converged, uniform_substeps_used, selective_fallback_used,
// direct_attempt_iterations, selective_refinement_restarts,
// selective_refined_particles, selective_total_particle_substeps,
// selective_work_ratio, selective_level_counts[7]) and the step diagnostics
// of in -> out computed from the reference's own end state (out17).
int ref_p3_step_refined(const upic_params* p, const upic_host_state* in, upic_host_state* out, double* info16,
                        double* out17) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        std::vector<NonlinearHistoryEntry> hist;
        NonlinearSolveInfo info;
        const State s0 = p3_load(in, g.N, c.dt);
        const State s1 = nonlinear_step_orbit(s0, g, c, qd, hist, &info, 0.0, nullptr);
        p3_store(s1, out);
        const double v[9] = {(double)info.iterations, info.converged ? 1.0 : 0.0, (double)info.uniform_substeps_used,
                             info.selective_fallback_used ? 1.0 : 0.0, (double)info.direct_attempt_iterations,
                             (double)info.selective_refinement_restarts, (double)info.selective_refined_particles,
                             (double)info.selective_total_particle_substeps, info.selective_work_ratio};
        std::memcpy(info16, v, sizeof(v));
        for (int k = 0; k < 7; ++k) info16[9 + k] = (double)info.selective_level_counts[k];
        p3_step_diag17(compute_step_diagnostics(s0, s1, g, c.dt, c.spline_order, qd, c.sigma1, c.sigma2, c.kappa,
                                                c.split_orbit_at_knots), out17);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// The reference program itself (its main, renamed where the source is
// included): argv[1] is the deck; its output files go where the deck's
// output_prefix points.
int ref_p3_main(int argc, char** argv) { return upic_ref_p3_program_main(argc, argv); }
// read_config + write_config_echo (P3:1178-1435, 8217-8300) of a deck, with
// the particle count of its initial state
int ref_p3_config_echo(const char* deck, const char* out) {
    try {
        const Config c = read_config(deck);
        const Grid g = make_grid(c);
        const State s = initialize_state(c, g);
        write_config_echo(out, c, g, s.x.size());
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p3_diagnostics(const upic_params* p, const upic_host_state* s, int mode, double* out10) {
    try {
        const Config c = p3_cfg(p); const Grid g = make_grid(c);
        const Diagnostics d = compute_diagnostics(p3_load(s, g.N, c.dt), g, c.sigma1, c.sigma2, c.kappa, mode);
        const double v[10] = {d.gauss_rms, d.gauss_max, d.gauge_rms, d.field_energy, d.kinetic_energy, d.total_energy,
                              d.E_mode_abs, d.rho_mode, d.Bz_y_mode_abs, d.mean_rho};
        std::memcpy(out10, v, sizeof(v));
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
}
