// oracle/ref_p2.cpp -- extern "C" access to Project2/spectral_pic_part3.cpp.
// TEST INFRASTRUCTURE ONLY (see ref_common.hpp).
#define main upic_ref_p2_program_main
#include REF_P2_SOURCE
#undef main
#include "ref_common.hpp"
#include <fstream>

extern "C" {
const char* ref_p2_last_error(void) { return ref_err.c_str(); }
// Runs run_simulation (P2:1094-1196) from a deck written to a temporary file;
// the diagnostics CSV lands at <prefix>_diagnostics.csv.
int ref_p2_run_deck(const char* deck_path) {
    try { const Config c = read_config(deck_path); run_simulation(c); return 0; }
    catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
// Initial Weibel particles (P2:624-669) as SoA arrays: x (positions, z=0), v, P.
int ref_p2_init_weibel(int nx, int ny, double Lx, double Ly, int ppc, double vx, double vymax, double sine, int mode,
                       unsigned seed, double n0, double c, int64_t* np, double* x, double* v, double* P, double* q, double* m) {
    try {
        Config cfg; cfg.nx = nx; cfg.ny = ny; cfg.Lx = Lx; cfg.Ly = Ly; cfg.particles_per_cell = ppc;
        cfg.weibel_vx = vx; cfg.weibel_vy_max = vymax; cfg.weibel_sine_perturb = sine; cfg.perturbation_mode = mode;
        cfg.seed = seed; cfg.n0 = n0; cfg.c = c;
        const Grid g = make_grid(cfg);
        ParticleSet ps; initialize_weibel_particles(cfg, g, ps);
        const int64_t n = static_cast<int64_t>(ps.electrons.size());
        if (!x) { *np = n; return 0; }
        for (int64_t i = 0; i < n; ++i) {
            const Particle& pp = ps.electrons[i];
            x[i] = pp.x; x[n + i] = pp.y; x[2 * n + i] = 0.0;
            v[i] = pp.v.x; v[n + i] = pp.v.y; v[2 * n + i] = pp.v.z;
            P[i] = pp.P.x; P[n + i] = pp.P.y; P[2 * n + i] = pp.P.z;
            q[i] = pp.q; m[i] = pp.m;
        }
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
}
extern "C" int ref_p2_initial_fields(int nx, int ny, double Lx, double Ly, int64_t np, const double* x, const double* v,
                                     const double* q, double* rho0, double* phi0, double* J0) {
    try {
        Config cfg; cfg.nx = nx; cfg.ny = ny; cfg.Lx = Lx; cfg.Ly = Ly;
        const Grid g = make_grid(cfg);
        ParticleSet ps; ps.fixed_ion_rho.assign(g.N, cfg.n0);
        for (int64_t i = 0; i < np; ++i) { Particle pp; pp.x = x[i]; pp.y = x[np + i]; pp.v = Vec3{v[i], v[np + i], v[2 * np + i]}; pp.q = q[i]; pp.m = -q[i]; ps.electrons.push_back(pp); }
        const Field r = deposit_total_charge_naive(ps.electrons, ps, g);
        const auto J = deposit_current(ps.electrons, g);
        const Field ph = solve_initial_phi_from_poisson(r, cfg, g);
        std::copy(r.begin(), r.end(), rho0); std::copy(ph.begin(), ph.end(), phi0); ref_storeF3(J, J0);
        return 0;
    } catch (const std::exception& e) { ref_err = e.what(); return 1; }
}
