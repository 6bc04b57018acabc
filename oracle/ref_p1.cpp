// oracle/ref_p1.cpp -- extern "C" access to Project1/ec_pic_serial_demo.cpp.
// TEST INFRASTRUCTURE ONLY (see ref_common.hpp).
#define main upic_ref_p1_program_main
#include REF_P1_SOURCE
#undef main
#include "ref_common.hpp"

static Config p1_cfg(const upic_params* p) {
    Config c;
    c.nx = p->nx; c.ny = p->ny; c.nz = p->nz;
    c.Lx = p->Lx; c.Ly = p->Ly; c.Lz = p->Lz;
    c.dt = p->dt;
    c.spline_order = p->spline_order;
    c.orbit_quad_order = p->orbit_quad_order;
    c.split_orbit_at_knots = p->split_orbit_at_knots != 0;
    c.picard_tol = p->picard_tol;
    c.picard_max_iter = p->picard_max_iter;
    c.final_consistency_sweep = p->final_consistency_sweep != 0;
    c.c = p->c; c.eps0 = p->eps0; c.mu0 = p->mu0;
    return c;
}
static State p1_load(const upic_host_state* s, int N) {
    State st;
    st.x = ref_load3<Vec3>(s->x, s->np); st.v = ref_load3<Vec3>(s->v, s->np);
    st.v_prev = ref_load3<Vec3>(s->v_prev, s->np); st.P = ref_load3<Vec3>(s->P, s->np);
    st.q.assign(s->q, s->q + s->np); st.m.assign(s->m, s->m + s->np);
    st.phi.assign(s->phi, s->phi + N); st.psi.assign(s->psi, s->psi + N); st.rho.assign(s->rho, s->rho + N);
    st.A = ref_loadF3<VecField>(s->A, N); st.U = ref_loadF3<VecField>(s->U, N); st.J = ref_loadF3<VecField>(s->J, N);
    return st;
}
static void p1_store(const State& st, upic_host_state* s) {
    ref_store3(st.x, s->x); ref_store3(st.v, s->v); ref_store3(st.v_prev, s->v_prev); ref_store3(st.P, s->P);
    std::copy(st.q.begin(), st.q.end(), s->q); std::copy(st.m.begin(), st.m.end(), s->m);
    std::copy(st.phi.begin(), st.phi.end(), s->phi); std::copy(st.psi.begin(), st.psi.end(), s->psi);
    std::copy(st.rho.begin(), st.rho.end(), s->rho);
    ref_storeF3(st.A, s->A); ref_storeF3(st.U, s->U); ref_storeF3(st.J, s->J);
}

extern "C" {
const char* ref_p1_last_error(void) { return ref_err.c_str(); }

int ref_p1_deposit_current_orbit(const upic_params* p, const double* x0, const double* dx, const double* vbar,
                                 const double* q, int64_t np, double* J, int64_t* stats4) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        PathStats ps;
        const VecField Jf = path_average_deposit_current(ref_load3<Vec3>(x0, np), ref_load3<Vec3>(dx, np),
            ref_load3<Vec3>(vbar, np), std::vector<Real>(q, q + np), g, c.spline_order, qd, c.split_orbit_at_knots, &ps);
        ref_storeF3(Jf, J);
        stats4[0] = ps.n_particles; stats4[1] = ps.crossing_particles; stats4[2] = ps.total_segments; stats4[3] = ps.max_segments;
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_gather_orbit(const upic_params* p, const double* x0, const double* dx, const double* f3, int64_t np, double* out) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        ref_store3(path_average_gather_vector(ref_load3<Vec3>(x0, np), ref_load3<Vec3>(dx, np), ref_loadF3<VecField>(f3, g.N),
                                              g, c.spline_order, qd, c.split_orbit_at_knots), out);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_orbit_grad_A(const upic_params* p, const double* x0, const double* dx, const double* An, const double* Anp1,
                        int64_t np, double* D) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        const auto M = orbit_discrete_gradient_A(ref_load3<Vec3>(x0, np), ref_load3<Vec3>(dx, np), ref_loadF3<VecField>(An, g.N),
                                                 ref_loadF3<VecField>(Anp1, g.N), g, c.spline_order, qd, c.split_orbit_at_knots);
        for (int64_t i = 0; i < np; ++i) for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) D[(a * 3 + b) * np + i] = M[i][a][b];
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_fft(const upic_params* p, double* data, int inverse) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c);
        CField z(g.N);
        for (int i = 0; i < g.N; ++i) z[i] = Complex(data[2 * i], data[2 * i + 1]);
        fft3d(z, g, inverse != 0);
        for (int i = 0; i < g.N; ++i) { data[2 * i] = z[i].real(); data[2 * i + 1] = z[i].imag(); }
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// kind 0 = landau, 1 = two_stream; mirrors initialize_state (P1:1306-1318)
int ref_p1_init(const upic_params* p, int kind, int ppc, double alpha, double vth, double v0, double pert, int mode,
                double n0, upic_host_state* st) {
    try {
        Config c = p1_cfg(p);
        c.test_case = kind == 0 ? "landau_weak" : "two_stream";
        c.particles_per_cell_pair = ppc; c.landau_alpha = alpha; c.thermal_velocity = vth; c.v0 = v0;
        c.perturbation = pert; c.perturbation_mode = mode; c.n0 = n0;
        const Grid g = make_grid(c);
        const State s = initialize_state(c, g);
        if (!st->x) { st->np = static_cast<int64_t>(s.x.size()); return 0; }
        p1_store(s, st);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// The field and history tail of the reference's initialisers (P1:1279-1301,
// the same steps as the two-stream initialiser's) applied to the particles
// st->x, v, q, m given by the caller: rho, J, phi, psi, A, U, P and the
// Taylor start-up v_prev, composed from the reference's own functions.
int ref_p1_init_from_particles(const upic_params* p, double n0, upic_host_state* st) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c);
        State s;
        s.x = ref_load3<Vec3>(st->x, st->np); s.v = ref_load3<Vec3>(st->v, st->np);
        s.q.assign(st->q, st->q + st->np); s.m.assign(st->m, st->m + st->np);
        const std::size_t np = s.x.size();
        s.v_prev.assign(np, Vec3{0.0, 0.0, 0.0});
        s.P.assign(np, Vec3{0.0, 0.0, 0.0});
        Field rho_e = deposit_scalar_to_mesh(s.x, s.q, g, c.spline_order, true);
        s.rho.assign(g.N, 0.0);
        for (int id = 0; id < g.N; ++id) s.rho[id] = rho_e[id] + n0;
        subtract_mean(s.rho);
        s.J = deposit_current_to_mesh(s.x, s.v, s.q, g, c.spline_order);
        s.phi = solve_periodic_poisson_for_phi(s.rho, g, c.eps0);
        s.psi.assign(g.N, 0.0);
        s.A = zero_vecfield(g.N);
        s.U = zero_vecfield(g.N);
        for (std::size_t i = 0; i < np; ++i) s.P[i] = mul3(s.m[i], s.v[i]);
        State tmp = s;
        const VecField E0 = compute_E(tmp, g);
        const auto E0p = gather_vector_all(s.x, E0, g, c.spline_order);
        for (std::size_t i = 0; i < np; ++i) s.v_prev[i] = sub3(s.v[i], mul3(c.dt, mul3(s.q[i] / s.m[i], E0p[i])));
        p1_store(s, st);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_step(const upic_params* p, const upic_host_state* in, upic_host_state* out, int32_t* iters, double* errs) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        std::vector<PicardHistoryEntry> hist;
        const State s1 = picard_step_orbit(p1_load(in, g.N), g, c, qd, hist);
        p1_store(s1, out);
        *iters = static_cast<int32_t>(hist.size());
        errs[0] = hist.back().current_error; errs[1] = hist.back().velocity_error;
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_diagnostics(const upic_params* p, const upic_host_state* s, int mode, double* out8) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c);
        const Diagnostics d = compute_diagnostics(p1_load(s, g.N), g, c.eps0, c.mu0, c.c, mode);
        out8[0] = d.gauss_rms; out8[1] = d.gauss_max; out8[2] = d.gauge_rms; out8[3] = d.field_energy;
        out8[4] = d.kinetic_energy; out8[5] = d.total_energy; out8[6] = d.E_mode_abs; out8[7] = d.rho_mode;
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
int ref_p1_step_diagnostics(const upic_params* p, const upic_host_state* s0, const upic_host_state* s1, double* out17) {
    try {
        const Config c = p1_cfg(p); const Grid g = make_grid(c); const Quadrature qd = gauss_legendre_01(c.orbit_quad_order);
        const StepDiagnostics d = compute_step_diagnostics(p1_load(s0, g.N), p1_load(s1, g.N), g, c.dt, c.spline_order, qd,
                                                           c.eps0, c.mu0, c.split_orbit_at_knots);
        const double v[17] = {d.A_chain_abs_rms, d.A_chain_rel_rms, d.A_chain_max_norm, d.A_chain_mean_norm, d.A_chain_work_defect,
                              d.delta_K, d.delta_W, d.delta_total_energy, d.particle_work, d.grid_work,
                              d.deposit_gather_work_residual, d.particle_energy_residual, d.field_energy_residual,
                              static_cast<double>(d.crossing_particles), static_cast<double>(d.total_crossings),
                              d.average_path_segments, static_cast<double>(d.max_path_segments)};
        std::memcpy(out17, v, sizeof(v));
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
}
