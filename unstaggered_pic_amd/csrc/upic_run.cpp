// upic_run.cpp -- deck-driven driver of the hot path: the replacement for the
// reference codes' main() time loops on top of the C ABI (include/upic.h).
//
//   upic_run p1 DECK      (Project1/ec_pic_serial_demo.cpp main, 1776-1892)
//   upic_run p2 DECK      (Project2/spectral_pic_part3.cpp main + run_simulation, 1093-1211)
//   upic_run p3 DECK      (Project3/ec_pic_nyquist_projected.cpp main, 8320-9058)
//
// Decks are `key = value` lines with `#` comments, parsed with the
// reference's key names and aliases (P1 read_config 320-367, P3 read_config
// 1178-1435); keys this driver does not use are reported once and ignored,
// as P1 does.  Outputs, with the reference's file names and column schemas:
//   <prefix>_diagnostics.csv    (P1 1811-1842; P3 8444-8460 and 8561-8624)
//   <prefix>_config_echo.txt    (P1 write_config_echo 1731-1770, P2 964-984,
//                                P3 8217-8300)
//   <prefix>_particle_sample.csv (P1 write_particle_sample 1716-1729, P2 986-1001,
//                                P3 9039)
// P1 and P3 also write <prefix>_step_diagnostics.csv (R_A and the work
// residuals of every step, P1:1818-1867, P3:8462-8890, the subcycled form
// after a selectively refined P3 step) from upic_sim_step_diagnostics (deck
// key step_diagnostics = false skips it for P3).  Not written: P3's
// transverse, field-probe, field-grid, particle-snapshot, selective-failure
// and particle-inner-history files; a deck that asks for them gets a
// warning.
//
// Initial conditions: P1 two_stream and landau_* are the reference's
// deterministic loads (lattice / quiet start, P1:1151-1304) computed here on
// the host, with the field tail (rho, J, phi, P, v_prev) on the device
// (upic_sim_init_fields).  P3 decks: the reference's two-stream, Weibel
// and Landau initial conditions (P3:3689-4141, upic_p3_initial_state +
// upic_sim_init_fields_AU) with its linear-theory growth-rate columns, or
// the seeded synthetic loads of upic_load_particles (test_case = bimaxwell |
// beams | landau_random).
//
// Exit codes as the reference: 0 ok, 1 error, 2 usage.
#include "../../include/upic.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using Real = double;
constexpr Real kPi = 3.14159265358979323846;
const Real kNaN = std::numeric_limits<Real>::quiet_NaN();

std::string trim(const std::string& s) {
    const auto a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const auto b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
std::string lower(std::string s) {
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
bool parse_bool(const std::string& v) {
    const std::string s = lower(trim(v));
    return s == "1" || s == "true" || s == "yes" || s == "on";
}

// One deck entry: its value and the line it was last set on.
struct DeckValue {
    std::string value;
    int line = 0;
};

// key -> value of a deck (later lines win, as in the reference parsers)
std::map<std::string, DeckValue> read_deck(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("could not open input file: " + path);
    std::map<std::string, DeckValue> kv;
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        ++n;
        const auto sharp = line.find('#');
        if (sharp != std::string::npos) line = line.substr(0, sharp);
        const auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        if (!key.empty()) kv[key] = DeckValue{trim(line.substr(eq + 1)), n};
    }
    return kv;
}

struct Deck {
    std::map<std::string, DeckValue> kv;
    std::set<std::string> used;
    bool has(std::initializer_list<const char*> keys) const {
        for (const char* k : keys)
            if (kv.count(k)) return true;
        return false;
    }
    // The value of the alias set last in the file: the reference parsers
    // apply the lines in order, so the latest line of any alias wins.
    const std::string* find(std::initializer_list<const char*> keys) {
        const DeckValue* r = nullptr;
        for (const char* k : keys) {
            auto it = kv.find(k);
            if (it == kv.end()) continue;
            used.insert(k);
            if (!r || it->second.line > r->line) r = &it->second;
        }
        return r ? &r->value : nullptr;
    }
    template <class T>
    void get(std::initializer_list<const char*> keys, T& out) {
        if (const std::string* v = find(keys)) {
            std::istringstream ss(*v);
            ss >> out;
        }
    }
    void get_bool(std::initializer_list<const char*> keys, bool& out) {
        if (const std::string* v = find(keys)) out = parse_bool(*v);
    }
    void get_str(std::initializer_list<const char*> keys, std::string& out) {
        if (const std::string* v = find(keys)) out = *v;
    }
    void warn_unused() const {
        for (const auto& kvp : kv)
            if (!used.count(kvp.first)) std::cerr << "warning: unknown input key ignored: " << kvp.first << "\n";
    }
};

void check(upic_status rc, upic_ctx* ctx, const char* what) {
    if (rc != UPIC_OK) {
        const char* msg = upic_last_error(ctx);
        throw std::runtime_error(std::string(what) + ": " + (msg && *msg ? msg : "failed"));
    }
}

double wall_since(const std::chrono::steady_clock::time_point& t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// wrap1 (P1:160-164): into [-L/2, L/2)
Real wrap1(Real x, Real L) {
    Real y = std::fmod(x + 0.5 * L, L);
    if (y < 0.0) y += L;
    return y - 0.5 * L;
}

std::vector<Real> cell_centered(int n, Real L) {  // cell_centered_grid (P1:1093-1101), centred
    std::vector<Real> x((size_t)n);
    const Real dx = L / (Real)n;
    for (int i = 0; i < n; ++i) x[(size_t)i] = ((Real)i + 0.5) * dx - 0.5 * L;
    return x;
}

// Inverse standard-normal CDF by Acklam's rational approximation (the
// published coefficients; P1:1103-1149 uses the same method).
Real normal_quantile(Real p) {
    if (!(p > 0.0 && p < 1.0)) throw std::runtime_error("normal_quantile requires 0<p<1");
    static const Real a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const Real b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
    static const Real c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const Real d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                             3.754408661907416e+00};
    const Real plow = 0.02425, phigh = 1.0 - plow;
    if (p < plow) {
        const Real q = std::sqrt(-2.0 * std::log(p));
        return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    if (p <= phigh) {
        const Real q = p - 0.5, r = q * q;
        return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
               (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
    }
    const Real q = std::sqrt(-2.0 * std::log(1.0 - p));
    return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
}

struct Particles {
    int64_t np = 0;
    std::vector<Real> x, v, q, m;  // SoA, component-major
    void resize(int64_t n) {
        np = n;
        x.assign(3 * (size_t)n, 0.0);
        v.assign(3 * (size_t)n, 0.0);
        q.assign((size_t)n, 0.0);
        m.assign((size_t)n, 0.0);
    }
};

// ---------------------------------------------------------------------------
// P1
// ---------------------------------------------------------------------------
struct P1Config {  // defaults of P1 Config (233-303)
    int nx = 4, ny = 4, nz = 4;
    Real Lx = 2.0 * kPi, Ly = 2.0 * kPi, Lz = 2.0 * kPi;
    int n_steps = 6;
    Real dt = 0.00625;
    std::string test_case = "two_stream";
    int ppc = 4, spline_order = 2, orbit_quad_order = 16;
    bool split = true, final_sweep = true, run_unit_tests = true;
    Real picard_tol = 1.0e-11;
    int picard_max_iter = 16;
    Real c = 1.0, eps0 = 1.0, mu0 = 1.0, n0 = 1.0, v0 = 0.3, perturbation = 0.02, landau_alpha = -1.0;
    Real thermal_velocity = 1.0;
    int perturbation_mode = 1;
    Real theory_omega = 0.0, theory_gamma = 0.0;
    unsigned seed = 12345;
    std::string output_prefix = "orbit_smoke";
    int precision = UPIC_PREC_FP64, sort_interval = 1;
};

P1Config read_p1(Deck& d) {
    P1Config c;
    d.get({"nx"}, c.nx); d.get({"ny"}, c.ny); d.get({"nz"}, c.nz);
    d.get({"Lx"}, c.Lx); d.get({"Ly"}, c.Ly); d.get({"Lz"}, c.Lz);
    d.get({"n_steps"}, c.n_steps); d.get({"dt"}, c.dt);
    if (const std::string* v = d.find({"test_case", "initial_condition", "example"})) c.test_case = lower(*v);
    d.get({"particles_per_cell_pair", "particles_per_cell", "velocity_samples_per_cell"}, c.ppc);
    d.get({"spline_order"}, c.spline_order); d.get({"orbit_quad_order"}, c.orbit_quad_order);
    d.get_bool({"split_orbit_at_knots"}, c.split);
    d.get({"picard_tol"}, c.picard_tol); d.get({"picard_max_iter"}, c.picard_max_iter);
    d.get_bool({"final_consistency_sweep"}, c.final_sweep); d.get_bool({"run_unit_tests"}, c.run_unit_tests);
    d.get({"c"}, c.c); d.get({"eps0"}, c.eps0); d.get({"mu0"}, c.mu0); d.get({"n0"}, c.n0);
    d.get({"v0"}, c.v0); d.get({"perturbation"}, c.perturbation);
    d.get({"landau_alpha", "landau_density_perturbation"}, c.landau_alpha);
    d.get({"thermal_velocity", "vth"}, c.thermal_velocity);
    d.get({"perturbation_mode"}, c.perturbation_mode);
    d.get({"theory_omega", "landau_theory_omega"}, c.theory_omega);
    d.get({"theory_gamma", "landau_theory_gamma"}, c.theory_gamma);
    d.get({"seed"}, c.seed);
    d.get_str({"output_prefix"}, c.output_prefix);
    d.get({"sort_interval"}, c.sort_interval);
    d.warn_unused();
    return c;
}

bool is_landau(const std::string& tc) {
    return tc == "landau" || tc == "landau_weak" || tc == "landau_strong" || tc == "weak_landau" ||
           tc == "strong_landau";
}

// initialize_two_stream_3d_linear_x (P1:1151-1206) and
// initialize_landau_damping_1d3d (P1:1208-1291): particle part
Particles p1_particles(const P1Config& c) {
    Particles P;
    const Real L[3] = {c.Lx, c.Ly, c.Lz};
    const Real kx = 2.0 * kPi * (Real)c.perturbation_mode / c.Lx;
    const Real volume = c.Lx * c.Ly * c.Lz;
    const std::string tc = lower(c.test_case);
    if (tc == "two_stream" || tc == "two-stream" || tc == "twostream") {
        const int nxp = c.nx * c.ppc;
        const auto xl = cell_centered(nxp, c.Lx), yl = cell_centered(c.ny, c.Ly), zl = cell_centered(c.nz, c.Lz);
        const int64_t nbase = (int64_t)nxp * c.ny * c.nz;
        P.resize(2 * nbase);
        int64_t b = 0;
        for (Real y : yl)
            for (Real z : zl)
                for (Real x : xl) {
                    const Real pos[3] = {x, y, z};
                    const Real seed_v = c.perturbation * std::sin(kx * x);
                    for (int d = 0; d < 3; ++d) {
                        P.x[(size_t)d * P.np + b] = wrap1(pos[d], L[d]);
                        P.x[(size_t)d * P.np + nbase + b] = wrap1(pos[d], L[d]);
                    }
                    P.v[b] = c.v0 + seed_v;
                    P.v[nbase + b] = -c.v0 + seed_v;
                    ++b;
                }
        const Real qm = -c.n0 * volume / (Real)P.np;
        for (int64_t p = 0; p < P.np; ++p) { P.q[p] = qm; P.m[p] = std::abs(qm); }
        return P;
    }
    if (is_landau(tc)) {
        const int nv = c.ppc;
        if (nv < 2) throw std::runtime_error("Landau examples need at least two velocity samples per cell");
        const Real alpha = c.landau_alpha >= 0.0 ? c.landau_alpha : c.perturbation;
        if (std::abs(alpha) >= 1.0)
            throw std::runtime_error("Landau density amplitude must satisfy |alpha| < 1 so all macro weights stay positive");
        const auto xl = cell_centered(c.nx, c.Lx), yl = cell_centered(c.ny, c.Ly), zl = cell_centered(c.nz, c.Lz);
        std::vector<Real> vq((size_t)nv);
        for (int j = 0; j < nv; ++j) vq[(size_t)j] = c.thermal_velocity * normal_quantile(((Real)j + 0.5) / (Real)nv);
        P.resize((int64_t)c.nx * c.ny * c.nz * nv);
        const Real q0 = -c.n0 * volume / (Real)P.np;
        int64_t p = 0;
        for (Real y : yl)
            for (Real z : zl)
                for (Real x : xl) {
                    const Real wgt = 1.0 + alpha * std::cos(kx * x);
                    const Real pos[3] = {x, y, z};
                    for (int j = 0; j < nv; ++j) {
                        for (int d = 0; d < 3; ++d) P.x[(size_t)d * P.np + p] = wrap1(pos[d], L[d]);
                        P.v[p] = vq[(size_t)j];
                        P.q[p] = q0 * wgt;
                        P.m[p] = std::abs(P.q[p]);
                        ++p;
                    }
                }
        return P;
    }
    throw std::runtime_error("unknown test_case/initial_condition: " + c.test_case);
}

void write_p1_echo(const P1Config& c, int64_t np) {  // write_config_echo (P1:1731-1770)
    std::ofstream out(c.output_prefix + "_config_echo.txt");
    const Real alpha = c.landau_alpha >= 0.0 ? c.landau_alpha : c.perturbation;
    out << "test_case=" << c.test_case << "\n";
    out << "nx=" << c.nx << "\nny=" << c.ny << "\nnz=" << c.nz << "\n";
    out << "Lx=" << std::setprecision(17) << c.Lx << "\n";
    out << "Ly=" << std::setprecision(17) << c.Ly << "\n";
    out << "Lz=" << std::setprecision(17) << c.Lz << "\n";
    out << "n_steps=" << c.n_steps << "\n";
    out << "dt=" << std::setprecision(17) << c.dt << "\n";
    out << "particles=" << np << "\n";
    out << "particles_per_cell_pair_or_velocity_samples=" << c.ppc << "\n";
    out << "spline_order=" << c.spline_order << "\norbit_quad_order=" << c.orbit_quad_order << "\n";
    out << "split_orbit_at_knots=" << (c.split ? "true" : "false") << "\n";
    out << "picard_tol=" << c.picard_tol << "\npicard_max_iter=" << c.picard_max_iter << "\n";
    out << "final_consistency_sweep=" << (c.final_sweep ? "true" : "false") << "\n";
    out << "c=" << c.c << "\neps0=" << c.eps0 << "\nmu0=" << c.mu0 << "\nn0=" << c.n0 << "\n";
    out << "v0=" << c.v0 << "\nperturbation=" << c.perturbation << "\nlandau_alpha=" << alpha << "\n";
    out << "thermal_velocity=" << c.thermal_velocity << "\nperturbation_mode=" << c.perturbation_mode << "\n";
    out << "theory_omega=" << c.theory_omega << "\ntheory_gamma=" << c.theory_gamma << "\n";
    out << "dV=" << (c.Lx / c.nx) * (c.Ly / c.ny) * (c.Lz / c.nz) << "\n";
}

struct Run {  // device handles and host buffers of one simulation
    upic_ctx* ctx = nullptr;
    upic_sim* sim = nullptr;
    int64_t np = 0;
    int N = 0;
    ~Run() {
        if (sim) upic_sim_destroy(sim);
        if (ctx) upic_ctx_destroy(ctx);
    }
};

void write_particle_sample(const std::string& path, Run& r, size_t max_particles) {  // P1:1716-1729
    std::vector<Real> x(3 * (size_t)r.np), v(3 * (size_t)r.np), q((size_t)r.np), m((size_t)r.np);
    upic_host_state hs{};
    hs.np = r.np;
    hs.x = x.data(); hs.v = v.data(); hs.q = q.data(); hs.m = m.data();
    check(upic_sim_download(r.sim, &hs), r.ctx, "download");
    std::ofstream out(path);
    out << "particle,x,y,z,vx,vy,vz,q,m\n";
    const size_t np = (size_t)r.np, nw = std::min(max_particles, np);
    for (size_t k = 0; k < nw; ++k) {
        const size_t p = (nw == 1) ? 0 : (k * (np - 1) / (nw - 1));
        out << p << "," << std::setprecision(17) << x[p] << "," << x[np + p] << "," << x[2 * np + p] << "," << v[p]
            << "," << v[np + p] << "," << v[2 * np + p] << "," << q[p] << "," << m[p] << "\n";
    }
}

int run_p1(const std::string& deck_path) {
    Deck d;
    d.kv = read_deck(deck_path);
    const P1Config c = read_p1(d);
    upic_params p{};
    p.code = UPIC_CODE_P1;
    p.nx = c.nx; p.ny = c.ny; p.nz = c.nz; p.Lx = c.Lx; p.Ly = c.Ly; p.Lz = c.Lz; p.dt = c.dt;
    p.spline_order = c.spline_order; p.orbit_quad_order = c.orbit_quad_order;
    p.split_orbit_at_knots = c.split ? 1 : 0;
    p.c = c.c; p.eps0 = c.eps0; p.mu0 = c.mu0; p.sigma1 = 1.0; p.sigma2 = 1.0;
    p.picard_tol = c.picard_tol; p.picard_max_iter = c.picard_max_iter;
    p.final_consistency_sweep = c.final_sweep ? 1 : 0;
    p.precision = c.precision; p.sort_interval = c.sort_interval;
    Particles P = p1_particles(c);
    Run r;
    r.np = P.np;
    r.N = c.nx * c.ny * c.nz;
    check(upic_ctx_create(&p, &r.ctx), nullptr, "context");
    check(upic_sim_create(r.ctx, P.np, &r.sim), r.ctx, "simulation");
    upic_host_state hs{};
    hs.np = P.np;
    std::vector<Real> zeros(3 * (size_t)P.np, 0.0), fz(3 * (size_t)r.N, 0.0);
    hs.x = P.x.data(); hs.v = P.v.data(); hs.q = P.q.data(); hs.m = P.m.data();
    hs.v_prev = zeros.data(); hs.P = zeros.data(); hs.vbar_last = zeros.data();
    hs.phi = fz.data(); hs.psi = fz.data(); hs.rho = fz.data(); hs.A = fz.data(); hs.U = fz.data(); hs.J = fz.data();
    check(upic_sim_upload(r.sim, &hs), r.ctx, "upload");
    check(upic_sim_init_fields(r.sim, c.n0), r.ctx, "initial fields");
    write_p1_echo(c, P.np);

    upic_diagnostics d0{};
    check(upic_sim_diagnostics(r.sim, c.perturbation_mode, &d0), r.ctx, "diagnostics");
    const Real theory_E0 = d0.E_mode_abs;
    const bool theory = (std::abs(c.theory_gamma) > 0.0 || std::abs(c.theory_omega) > 0.0) && theory_E0 > 0.0;
    std::ofstream diag(c.output_prefix + "_diagnostics.csv");
    diag << "step,time,gauss_rms,gauss_max,gauge_rms,field_energy,kinetic_energy,total_energy,"
         << "rho_mode,E_mode_re,E_mode_im,E_mode_abs,theory_E_abs,E_over_theory,mean_rho,"
         << "picard_iterations,picard_current_error,picard_velocity_error,"
         << "crossing_particles,avg_path_segments,max_path_segments\n";
    auto row = [&](int step, Real t, int iters, Real ce, Real ve, const upic_path_stats& ps) {
        upic_diagnostics g{};
        check(upic_sim_diagnostics(r.sim, c.perturbation_mode, &g), r.ctx, "diagnostics");
        Real th = kNaN, ratio = kNaN;
        if (theory) {
            th = theory_E0 * std::exp(c.theory_gamma * t);
            ratio = g.E_mode_abs / (th + 1.0e-300);
        }
        diag << step << "," << std::setprecision(17) << t << "," << g.gauss_rms << "," << g.gauss_max << ","
             << g.gauge_rms << "," << g.field_energy << "," << g.kinetic_energy << "," << g.total_energy << ","
             << g.rho_mode << "," << g.E_mode_re << "," << g.E_mode_im << "," << g.E_mode_abs << "," << th << ","
             << ratio << "," << g.mean_rho << "," << iters << "," << ce << "," << ve << ","
             << ps.crossing_particles << ","
             << (ps.n_particles ? (Real)ps.total_segments / (Real)ps.n_particles : 0.0) << "," << ps.max_segments
             << "\n";
        return g;
    };
    // <prefix>_step_diagnostics.csv (P1:1818-1819, 1859-1867)
    std::ofstream stepcsv(c.output_prefix + "_step_diagnostics.csv");
    stepcsv << "step,time,A_chain_abs_rms,A_chain_rel_rms,A_chain_max_norm,A_chain_mean_norm,A_chain_work_defect,"
            << "delta_K,delta_W,delta_total_energy,particle_work,grid_work,deposit_gather_work_residual,"
            << "particle_energy_residual,field_energy_residual,crossing_particles,total_crossings,"
            << "avg_path_segments,max_path_segments\n";
    check(upic_sim_keep_step_start(r.sim, 1), r.ctx, "step diagnostics");
    upic_path_stats ps0{};
    ps0.max_segments = 1;
    row(0, 0.0, 0, kNaN, kNaN, ps0);
    for (int n = 0; n < c.n_steps; ++n) {
        upic_step_info info{};
        check(upic_step_p1(r.sim, &info), r.ctx, "step");
        upic_step_diagnostics sd{};
        check(upic_sim_step_diagnostics(r.sim, &sd), r.ctx, "step diagnostics");
        const upic_diagnostics g = row(n + 1, (n + 1) * c.dt, info.iterations, info.current_error,
                                       info.velocity_error, info.path);
        stepcsv << (n + 1) << "," << std::setprecision(17) << (n + 1) * c.dt << "," << sd.A_chain_abs_rms << ","
                << sd.A_chain_rel_rms << "," << sd.A_chain_max_norm << "," << sd.A_chain_mean_norm << ","
                << sd.A_chain_work_defect << "," << sd.delta_K << "," << sd.delta_W << "," << sd.delta_total_energy
                << "," << sd.particle_work << "," << sd.grid_work << "," << sd.deposit_gather_work_residual << ","
                << sd.particle_energy_residual << "," << sd.field_energy_residual << "," << sd.crossing_particles
                << "," << sd.total_crossings << "," << sd.avg_path_segments << "," << sd.max_path_segments << "\n";
        if ((n + 1) % std::max(1, c.n_steps / 5) == 0 || n == c.n_steps - 1)
            std::cout << std::scientific << std::setprecision(3) << "step " << std::setw(5) << (n + 1) << "/"
                      << c.n_steps << " Gauss=" << g.gauss_rms << " Picard=" << info.iterations << "\n";
    }
    write_particle_sample(c.output_prefix + "_particle_sample.csv", r, 20000);
    upic_diagnostics fd{};
    check(upic_sim_diagnostics(r.sim, c.perturbation_mode, &fd), r.ctx, "diagnostics");
    std::cout << std::scientific << std::setprecision(6) << "Final Gauss RMS = " << fd.gauss_rms << "\n"
              << "Final gauge RMS = " << fd.gauge_rms << "\n"
              << "Final total energy = " << fd.total_energy << "\n";
    return 0;
}

// ---------------------------------------------------------------------------
// P3
// ---------------------------------------------------------------------------
int run_p3(const std::string& deck_path) {
    Deck d;
    d.kv = read_deck(deck_path);
    upic_params p{};
    p.code = UPIC_CODE_P3;
    // the defaults of the reference's Config (P3:837-990)
    p.nx = 4; p.ny = 4; p.nz = 4; p.Lx = 2.0 * kPi; p.Ly = 2.0 * kPi; p.Lz = 2.0 * kPi;
    p.dt = 0.00625; p.spline_order = 2; p.orbit_quad_order = 16; p.split_orbit_at_knots = 1;
    p.nyquist_projection = 0; p.c = 1.0; p.eps0 = 1.0; p.mu0 = 1.0; p.sigma1 = 1.0; p.sigma2 = 1.0;
    p.nonlinear_solver_picard = 0; p.nonlinear_max_iter = 32; p.nonlinear_rtol = 1e-12; p.nonlinear_atol = 1e-13;
    p.anderson_memory = 4; p.anderson_start = 2; p.anderson_damping = 0.7; p.anderson_regularization = 1e-12;
    p.anderson_growth_limit = 1.2; p.local_rtol = 1e-11; p.local_atol = 1e-13; p.local_max_iter = 24;
    p.local_damping = 1.0; p.selective_particle_fallback = 0; p.selective_particle_max_substeps = 64;
    p.precision = UPIC_PREC_FP64; p.sort_interval = 1; p.learned_guess = 0;
    int n_steps = 6, ppc = 4, mode = 1;
    std::string tc = "cold_relativistic_two_stream", prefix = "orbit_smoke", solver = "anderson_quasi_newton",
                prec = "fp64";
    bool split = true, proj = false, sel = false;
    Real rho0 = 1.0, v0 = 0.3, noise = 0.0, alpha = 0.02;
    Real vth[3] = {0.05, 0.2, 0.2};
    uint64_t seed = 12345;
    d.get({"nx"}, p.nx); d.get({"ny"}, p.ny); d.get({"nz"}, p.nz);
    d.get({"Lx"}, p.Lx); d.get({"Ly"}, p.Ly); d.get({"Lz"}, p.Lz);
    d.get({"n_steps"}, n_steps); d.get({"dt"}, p.dt);
    if (const std::string* v = d.find({"test_case", "initial_condition", "example"})) tc = lower(*v);
    d.get({"particles_per_cell_pair", "particles_per_cell", "velocity_samples_per_cell"}, ppc);
    d.get({"spline_order"}, p.spline_order); d.get({"orbit_quad_order"}, p.orbit_quad_order);
    d.get_bool({"split_orbit_at_knots"}, split);
    if (const std::string* v = d.find({"nonlinear_solver"})) solver = lower(*v);
    d.get({"nonlinear_rtol", "nonlinear_tol", "picard_tol"}, p.nonlinear_rtol);
    d.get({"nonlinear_atol"}, p.nonlinear_atol);
    d.get_bool({"nyquist_projection"}, proj);
    d.get({"nonlinear_max_iter", "picard_max_iter"}, p.nonlinear_max_iter);
    d.get({"anderson_memory"}, p.anderson_memory); d.get({"anderson_start"}, p.anderson_start);
    d.get({"anderson_damping"}, p.anderson_damping);
    d.get({"anderson_regularization"}, p.anderson_regularization);
    d.get({"anderson_growth_limit"}, p.anderson_growth_limit);
    d.get_bool({"selective_particle_fallback", "selective_subcycle_fallback"}, sel);
    d.get({"selective_particle_max_substeps", "selective_max_substeps"}, p.selective_particle_max_substeps);
    d.get({"particle_local_rtol", "uniform_local_rtol"}, p.local_rtol);
    d.get({"particle_local_atol", "uniform_local_atol"}, p.local_atol);
    d.get({"particle_local_max_iter", "uniform_local_max_iter"}, p.local_max_iter);
    d.get({"particle_local_damping", "uniform_local_damping"}, p.local_damping);
    // kappa, sigma1, sigma2 and their deprecated names, applied in file
    // order with the reference's warnings (P3:1308-1315)
    {
        struct Coef { int line; std::string key, value; };
        std::vector<Coef> cs;
        for (const char* k : {"kappa", "c", "sigma1", "sigma2", "eps0", "mu0"}) {
            auto it = d.kv.find(k);
            if (it == d.kv.end()) continue;
            d.used.insert(k);
            cs.push_back({it->second.line, k, it->second.value});
        }
        std::sort(cs.begin(), cs.end(), [](const Coef& a, const Coef& b) { return a.line < b.line; });
        for (const Coef& cf : cs) {
            std::istringstream ss(cf.value);
            if (cf.key == "kappa") ss >> p.c;
            else if (cf.key == "c") {
                ss >> p.c;
                std::cerr << "warning: c is deprecated; use kappa in the nondimensional paper notation\n";
            } else if (cf.key == "sigma1") ss >> p.sigma1;
            else if (cf.key == "sigma2") ss >> p.sigma2;
            else if (cf.key == "eps0") {
                Real eps0 = 1.0;
                ss >> eps0;
                p.sigma1 = 1.0 / eps0;
                std::cerr << "warning: eps0 is deprecated; use sigma1=1/eps0 in the nondimensional paper notation\n";
            } else {
                ss >> p.sigma2;
                std::cerr << "warning: mu0 is deprecated; use sigma2 in the nondimensional paper notation\n";
            }
        }
    }
    // keys the reference accepts and ignores (P3:1295-1306)
    for (const char* k : {"picard_relaxation", "picard_omega", "relaxed_picard_omega", "picard_relaxation_min",
                          "picard_omega_min", "picard_relaxation_shrink", "picard_relaxation_growth",
                          "picard_adaptive_relaxation", "adaptive_picard_relaxation", "picard_current_in_convergence",
                          "picard_converge_current", "final_consistency_sweep", "picard_require_converged",
                          "reject_unconverged_picard"})
        if (d.kv.count(k)) { d.used.insert(k); std::cerr << "warning: obsolete Picard control ignored: " << k << "\n"; }
    for (const char* k : {"adaptive_dt_max", "dt_max", "adaptive_max_accepted_steps", "max_accepted_steps",
                          "adaptive_energy_rtol", "energy_accept_rtol", "adaptive_finish_exactly"})
        if (d.kv.count(k)) {
            d.used.insert(k);
            std::cerr << "warning: unsupported adaptive time-step control ignored: " << k << "\n";
        }
    // controls of the reference's features outside this driver: read (and
    // echoed) as the reference reads them; the manuscript runs need the
    // adaptive dt and uniform subcycling off (P3:8344-8345), and outputs
    // this driver does not write are named in a warning
    bool adaptive = false, uniform_sub = false, uniform_refine = false, transverse = false, field_diag = true,
         field_grid = false, fail_diag = true;
    Real transverse_interval = 1.0, field_interval = 10.0, adaptive_dt_min = 0.0, adaptive_shrink = 0.5,
         adaptive_growth = 2.0, uniform_start = 0.0;
    int adaptive_growth_interval = 4, adaptive_attempts = 4, uniform_initial = 1, uniform_substeps = 10,
        uniform_max = 20, dynamic_chunk = 4, fail_limit = 16, probe_points = 32;
    std::string probe_axis = "auto", snapshot_times_raw;
    std::vector<Real> snapshot_times;
    d.get_bool({"adaptive_dt", "adaptive_time_step", "adaptive_timestep", "adaptive_reject_on_unconverged"}, adaptive);
    d.get({"adaptive_dt_min", "dt_min"}, adaptive_dt_min);
    d.get({"adaptive_dt_shrink", "dt_shrink"}, adaptive_shrink);
    d.get({"adaptive_dt_growth", "dt_growth"}, adaptive_growth);
    d.get({"adaptive_dt_growth_interval", "dt_growth_interval"}, adaptive_growth_interval);
    d.get({"adaptive_max_attempts", "max_step_attempts"}, adaptive_attempts);
    d.get_bool({"uniform_particle_subcycling", "uniform_subcycling"}, uniform_sub);
    d.get_bool({"uniform_refine_on_failure", "particle_refine_on_failure"}, uniform_refine);
    d.get({"uniform_initial_substeps", "particle_initial_substeps"}, uniform_initial);
    d.get({"uniform_substeps", "particle_substeps"}, uniform_substeps);
    d.get({"uniform_max_substeps", "particle_max_substeps"}, uniform_max);
    d.get({"uniform_subcycling_start_time", "particle_subcycling_start_time"}, uniform_start);
    d.get({"selective_particle_dynamic_chunk", "selective_dynamic_chunk"}, dynamic_chunk);
    d.get_bool({"selective_failure_diagnostics", "failed_particle_diagnostics"}, fail_diag);
    d.get({"selective_failure_diagnostic_limit", "failed_particle_diagnostic_limit"}, fail_limit);
    d.get_bool({"field_diagnostics", "write_field_diagnostics"}, field_diag);
    d.get({"field_diagnostics_interval", "field_output_interval"}, field_interval);
    d.get({"field_probe_points_per_cell", "field_samples_per_cell"}, probe_points);
    if (const std::string* v = d.find({"field_probe_axis"})) probe_axis = lower(*v);
    d.get_bool({"field_grid_snapshots", "write_full_field_snapshots"}, field_grid);
    d.get_bool({"transverse_diagnostics"}, transverse);
    d.get({"transverse_spectrum_interval"}, transverse_interval);
    if (const std::string* v = d.find({"particle_snapshot_times"})) {
        std::string times = *v;
        std::replace(times.begin(), times.end(), ',', ' ');
        std::istringstream values(times);
        Real t;
        while (values >> t) {
            if (!std::isfinite(t) || t < 0.0) throw std::runtime_error("particle_snapshot_times must be finite and nonnegative");
            snapshot_times.push_back(t);
        }
        if (!values.eof()) throw std::runtime_error("invalid particle_snapshot_times list");
        std::sort(snapshot_times.begin(), snapshot_times.end());
        snapshot_times.erase(std::unique(snapshot_times.begin(), snapshot_times.end()), snapshot_times.end());
    }
    if (adaptive || uniform_sub)
        throw std::runtime_error("manuscript runs require adaptive_dt=false, uniform_particle_subcycling=false");
    {
        std::vector<std::string> skipped;
        if (transverse) skipped.push_back("transverse_diagnostics (<prefix>_transverse_spectrum.csv, _velocity_moments.csv)");
        if (field_diag) skipped.push_back("field_diagnostics (<prefix>_field_probe*.csv, _knot_jumps.csv)");
        if (field_grid) skipped.push_back("field_grid_snapshots (<prefix>_field_grid_t*.csv)");
        if (fail_diag) skipped.push_back("selective_failure_diagnostics (<prefix>_selective_failures.csv)");
        if (!snapshot_times.empty()) skipped.push_back("particle_snapshot_times (<prefix>_particles_t*.csv)");
        for (const std::string& k : skipped)
            std::cerr << "warning: " << k << " is requested; upic_run does not write these files\n";
        if (d.has({"run_unit_tests"})) d.used.insert("run_unit_tests");
    }
    // the reference's initial conditions (is_two_stream_case / is_weibel_case
    // / is_landau_case, P3:1042-1105)
    int ref_kind = 0;
    if (tc == "two_stream" || tc == "two-stream" || tc == "twostream" || tc == "relativistic_two_stream" ||
        tc == "relativistic-two-stream" || tc == "cold_relativistic_two_stream" ||
        tc == "cold-relativistic-two-stream" || tc == "cold_two_stream" || tc == "cold-two-stream")
        ref_kind = UPIC_P3_TWO_STREAM;
    else if (tc == "weibel" || tc == "relativistic_weibel" || tc == "relativistic-weibel" ||
             tc == "cold_relativistic_weibel" || tc == "cold-relativistic-weibel" || tc == "filamentation" ||
             tc == "cold_filamentation" || tc == "cold-filamentation")
        ref_kind = UPIC_P3_WEIBEL;
    else if (is_landau(tc))
        ref_kind = UPIC_P3_LANDAU;
    upic_p3_case cs{};
    Real theory_gamma = 0.0, theory_omega = 0.0;
    d.get({"theory_gamma", "landau_theory_gamma"}, theory_gamma);
    d.get({"theory_omega", "landau_theory_omega"}, theory_omega);
    if (ref_kind) {
        upic_p3_case_defaults(ref_kind, &cs);
        cs.particles_per_cell_pair = ppc;
        d.get({"n0", "rho0", "background_density"}, cs.n0); d.get({"v0"}, cs.v0);
        d.get({"perturbation"}, cs.perturbation); d.get({"perturbation_mode"}, cs.perturbation_mode);
        d.get({"two_stream_density_alpha", "density_perturbation", "alpha_x"}, cs.two_stream_density_alpha);
        d.get({"weibel_B0", "magnetic_perturbation", "initial_B0"}, cs.weibel_B0);
        d.get({"landau_alpha", "landau_density_perturbation"}, cs.landau_alpha);
        d.get({"thermal_velocity", "vth"}, cs.thermal_velocity);
        d.get({"tsi_transverse_seed"}, cs.tsi_transverse_seed); d.get({"tsi_transverse_mode"}, cs.tsi_transverse_mode);
        // finite_grid_theory sets both theory switches; the line set last wins
        {
            bool b = cs.two_stream_finite_grid_theory != 0;
            d.get_bool({"two_stream_finite_grid_theory", "finite_grid_theory"}, b);
            cs.two_stream_finite_grid_theory = b;
            b = cs.weibel_finite_grid_theory != 0;
            d.get_bool({"weibel_finite_grid_theory", "finite_grid_theory"}, b);
            cs.weibel_finite_grid_theory = b;
        }
        bool b = cs.two_stream_eigenmode != 0;
        d.get_bool({"two_stream_eigenmode", "initialize_two_stream_eigenmode"}, b); cs.two_stream_eigenmode = b;
        b = cs.weibel_eigenmode != 0;
        d.get_bool({"weibel_eigenmode", "initialize_weibel_eigenmode"}, b); cs.weibel_eigenmode = b;
        rho0 = cs.n0;
        mode = cs.perturbation_mode;
    } else {
        d.get({"rho0", "n0", "background_density"}, rho0); d.get({"v0"}, v0); d.get({"perturbation_mode"}, mode);
        d.get({"perturbation", "landau_alpha"}, alpha);
        d.get({"vth_x"}, vth[0]); d.get({"vth_y"}, vth[1]); d.get({"vth_z"}, vth[2]);
        if (const std::string* v = d.find({"vth", "thermal_velocity"})) vth[0] = vth[1] = vth[2] = std::atof(v->c_str());
        d.get({"position_noise"}, noise); d.get({"seed"}, seed);
    }
    d.get_str({"output_prefix"}, prefix);
    d.get_str({"precision"}, prec);
    d.get({"learned_guess"}, p.learned_guess);
    d.get({"sort_interval"}, p.sort_interval);
    // <prefix>_step_diagnostics.csv every step, as the reference (false skips
    // its extra orbit pass per step; a key of this driver)
    bool step_diag = true;
    d.get_bool({"step_diagnostics"}, step_diag);
    d.warn_unused();
    p.split_orbit_at_knots = split ? 1 : 0;
    p.nyquist_projection = proj ? 1 : 0;
    p.selective_particle_fallback = sel ? 1 : 0;
    p.nonlinear_solver_picard = (lower(solver) == "picard") ? 1 : 0;
    if (lower(prec) == "mixed") {
        p.precision = UPIC_PREC_MIXED;
        // FP32 tolerance floors of the mixed mode
        p.local_rtol = std::max(p.local_rtol, 1e-6); p.local_atol = std::max(p.local_atol, 1e-7);
        p.nonlinear_rtol = std::max(p.nonlinear_rtol, 3e-6); p.nonlinear_atol = std::max(p.nonlinear_atol, 1e-9);
    }
    const int learned = p.learned_guess;
    if (learned == UPIC_GUESS_LEARNED) p.learned_guess = UPIC_GUESS_PREDICTOR;  // until the network is trained
    Run r;
    r.N = p.nx * p.ny * p.nz;
    Particles P;
    upic_host_state hs{};
    std::vector<Real> A0(3 * (size_t)r.N, 0.0), U0(3 * (size_t)r.N, 0.0);
    double growth = 0.0;
    if (ref_kind) {
        check(upic_p3_initial_state(&p, &cs, &hs, nullptr), nullptr, "initial condition");
        r.np = hs.np;
        P.resize(r.np);
        hs.x = P.x.data(); hs.v = P.v.data(); hs.q = P.q.data(); hs.m = P.m.data();
        hs.A = A0.data(); hs.U = U0.data();
        check(upic_p3_initial_state(&p, &cs, &hs, &growth), nullptr, "initial condition");
    } else {
        upic_load_spec ls{};
        ls.ppc = ppc; ls.seed = seed; ls.n0 = rho0; ls.alpha = alpha; ls.mode = mode;
        ls.vth[0] = vth[0]; ls.vth[1] = vth[1]; ls.vth[2] = vth[2]; ls.v0 = v0; ls.noise = noise;
        if (tc == "bimaxwell" || tc == "weibel_bimaxwell" || tc == "maxwellian") ls.kind = UPIC_LOAD_BIMAXWELL;
        else if (tc == "beams" || tc == "counter_streaming_beams") ls.kind = UPIC_LOAD_BEAMS;
        else if (tc == "landau_random") ls.kind = UPIC_LOAD_LANDAU;
        else throw std::runtime_error("unknown test_case/initial_condition: " + tc +
                                      " (two_stream, weibel, landau or the seeded bimaxwell, beams, landau_random)");
        r.np = (int64_t)r.N * ppc;
        P.resize(r.np);
        hs.np = r.np;
        hs.x = P.x.data(); hs.v = P.v.data(); hs.q = P.q.data(); hs.m = P.m.data();
        check(upic_load_particles(&p, &ls, &hs), nullptr, "particle load");
    }
    // linear-theory reference (P3:8427-8433): the selected growth rate times
    // the initial |E_x(k_x)| (|B_z(k_y)| for Weibel)
    const Real theory_growth = std::abs(theory_gamma) > 0.0 ? theory_gamma : growth;
    // the echo is written before any device call, as the reference writes it
    // before the time loop
    {  // write_config_echo (P3:8217-8300), every key as the deck set it
        if (!ref_kind) upic_p3_case_defaults(0, &cs);
        const Real wB0 = cs.weibel_B0 >= 0.0 ? cs.weibel_B0 : cs.perturbation;
        const Real lal = cs.landau_alpha >= 0.0 ? cs.landau_alpha : cs.perturbation;
        auto tf = [](bool b) { return b ? "true" : "false"; };
        std::ofstream out(prefix + "_config_echo.txt");
        out << std::setprecision(17);
        out << "test_case=" << tc << "\ntransverse_diagnostics=" << tf(transverse)
            << "\ntransverse_spectrum_interval=" << transverse_interval
            << "\ntsi_transverse_seed=" << cs.tsi_transverse_seed << "\ntsi_transverse_mode=" << cs.tsi_transverse_mode
            << "\nnyquist_projection=" << p.nyquist_projection << "\nnx=" << p.nx << "\nny=" << p.ny << "\nnz=" << p.nz
            << "\nLx=" << p.Lx << "\nLy=" << p.Ly << "\nLz=" << p.Lz << "\nn_steps=" << n_steps << "\ndt=" << p.dt
            << "\ntarget_time=" << (Real)n_steps * p.dt << "\nparticles=" << r.np
            << "\nparticles_per_cell_pair_or_velocity_samples=" << ppc
            << "\nspline_order=" << p.spline_order << "\norbit_quad_order=" << p.orbit_quad_order
            << "\nsplit_orbit_at_knots=" << tf(p.split_orbit_at_knots != 0) << "\nmacro_time_step=fixed"
            << "\nadaptive_dt=false\nadaptive_dt_min=" << (adaptive_dt_min > 0.0 ? adaptive_dt_min : p.dt / 16.0)
            << "\nadaptive_dt_shrink=" << adaptive_shrink << "\nadaptive_dt_growth=" << adaptive_growth
            << "\nadaptive_dt_growth_interval=" << adaptive_growth_interval
            << "\nadaptive_max_attempts=" << adaptive_attempts
            << "\nnonlinear_solver=" << solver << "\nnonlinear_rtol=" << p.nonlinear_rtol
            << "\nnonlinear_atol=" << p.nonlinear_atol << "\nnonlinear_max_iter=" << p.nonlinear_max_iter
            << "\nanderson_memory=" << p.anderson_memory << "\nanderson_start=" << p.anderson_start
            << "\nanderson_damping=" << p.anderson_damping << "\nanderson_regularization=" << p.anderson_regularization
            << "\nanderson_growth_limit=" << p.anderson_growth_limit
            << "\nuniform_particle_subcycling=false\nuniform_refine_on_failure=" << tf(uniform_refine)
            << "\nuniform_initial_substeps=" << uniform_initial << "\nuniform_substeps=" << uniform_substeps
            << "\nuniform_max_substeps=" << uniform_max << "\nuniform_subcycling_start_time=" << uniform_start
            << "\nparticle_local_rtol=" << p.local_rtol << "\nparticle_local_atol=" << p.local_atol
            << "\nparticle_local_max_iter=" << p.local_max_iter << "\nuniform_local_rtol=" << p.local_rtol
            << "\nuniform_local_atol=" << p.local_atol << "\nuniform_local_max_iter=" << p.local_max_iter
            << "\nuniform_local_damping=" << p.local_damping
            << "\nselective_particle_fallback=" << tf(p.selective_particle_fallback != 0)
            << "\nselective_particle_max_substeps=" << p.selective_particle_max_substeps
            << "\nselective_particle_dynamic_chunk=" << dynamic_chunk
            << "\nselective_failure_diagnostics=" << tf(fail_diag)
            << "\nselective_failure_diagnostic_limit=" << fail_limit << "\nparticle_snapshot_times=";
        for (Real t : snapshot_times) out << t << ' ';
        out << "\nfield_diagnostics=" << tf(field_diag) << "\nfield_diagnostics_interval=" << field_interval
            << "\nfield_probe_points_per_cell=" << probe_points << "\nfield_probe_axis=" << probe_axis
            << "\nfield_grid_snapshots=" << tf(field_grid) << "\nkappa=" << p.c << "\nsigma1=" << p.sigma1
            << "\nsigma2=" << p.sigma2
            << "\nsigma1_minus_kappa2_sigma2=" << (p.sigma1 - p.c * p.c * p.sigma2) << "\nrho0=" << rho0
            << "\nv0=" << (ref_kind ? cs.v0 : v0) << "\nperturbation=" << (ref_kind ? cs.perturbation : alpha)
            << "\ntwo_stream_density_alpha=" << cs.two_stream_density_alpha
            << "\ntwo_stream_eigenmode=" << tf(cs.two_stream_eigenmode != 0)
            << "\ntwo_stream_finite_grid_theory=" << tf(cs.two_stream_finite_grid_theory != 0)
            << "\nweibel_B0=" << wB0 << "\nweibel_eigenmode=" << tf(cs.weibel_eigenmode != 0)
            << "\nweibel_finite_grid_theory=" << tf(cs.weibel_finite_grid_theory != 0) << "\nlandau_alpha=" << lal
            << "\nthermal_velocity=" << cs.thermal_velocity << "\nperturbation_mode=" << mode
            << "\ntheory_omega=" << theory_omega << "\ntheory_gamma=" << theory_gamma
            << "\nauto_theory_growth_rate=" << theory_growth
            << "\ndV=" << (p.Lx / p.nx) * (p.Ly / p.ny) * (p.Lz / p.nz) << "\n";
    }
    check(upic_ctx_create(&p, &r.ctx), nullptr, "context");
    check(upic_sim_create(r.ctx, r.np, &r.sim), r.ctx, "simulation");
    std::vector<Real> zeros(3 * (size_t)r.np, 0.0), fz(3 * (size_t)r.N, 0.0);
    hs.v_prev = zeros.data(); hs.P = zeros.data(); hs.vbar_last = zeros.data();
    hs.phi = fz.data(); hs.psi = fz.data(); hs.rho = fz.data(); hs.A = fz.data(); hs.U = fz.data(); hs.J = fz.data();
    check(upic_sim_upload(r.sim, &hs), r.ctx, "upload");
    if (ref_kind == UPIC_P3_WEIBEL)
        check(upic_sim_init_fields_AU(r.sim, rho0, A0.data(), U0.data()), r.ctx, "initial fields");
    else
        check(upic_sim_init_fields(r.sim, rho0), r.ctx, "initial fields");
    upic_diagnostics d0{};
    check(upic_sim_diagnostics(r.sim, mode, &d0), r.ctx, "diagnostics");
    const bool use_B = ref_kind == UPIC_P3_WEIBEL;
    const Real theory_amp0 = use_B ? d0.Bz_y_mode_abs : d0.E_mode_abs;
    const bool has_theory = (std::abs(theory_growth) > 0.0 || std::abs(theory_omega) > 0.0) && theory_amp0 > 0.0;
    if (std::abs(theory_growth) > 0.0)
        std::cout << std::scientific << std::setprecision(6) << "linear theory growth/damping rate = " << theory_growth
                  << (use_B ? " using |B_z(k_y)|" : " using |E_x(k_x)|") << "\n";
    if (learned == UPIC_GUESS_LEARNED) {
        const int64_t stride = std::max<int64_t>(1, r.np / 16384);
        check(upic_sim_collect_guess_data(r.sim, stride, ((r.np + stride - 1) / stride) * 20), r.ctx, "collect");
    }
    const auto t0 = std::chrono::steady_clock::now();
    // <prefix>_step_diagnostics.csv (P3:8462-8473, rows 8835-8890): the
    // nonlinear solve's columns and compute_step_diagnostics
    std::ofstream stepcsv;
    if (step_diag) {
        check(upic_sim_keep_step_start(r.sim, 1), r.ctx, "step diagnostics");
        stepcsv.open(prefix + "_step_diagnostics.csv");
        stepcsv << "step,time,nonlinear_iterations,nonlinear_rms_residual,nonlinear_max_particle_residual,"
                << "anderson_restarts,convergence_method,observed_convergence_rate,"
                << "estimated_map_contraction,backtracked_steps,minimum_update_alpha,"
                << "local_particle_iterations_mean,local_particle_iterations_max,uniform_substeps_used,"
                << "selective_fallback_used,direct_attempt_iterations,selective_refinement_restarts,"
                << "selective_refined_particles,selective_total_particle_substeps,selective_work_ratio,"
                << "selective_n1,selective_n2,selective_n4,selective_n8,selective_n16,selective_n32,selective_n64,"
                << "A_chain_abs_rms,A_chain_rel_rms,A_chain_max_norm,"
                << "A_chain_mean_norm,A_chain_work_defect,delta_K,delta_W,delta_total_energy,particle_work,"
                << "grid_work,deposit_gather_work_residual,particle_energy_residual,field_energy_residual,"
                << "crossing_particles,total_crossings,avg_path_segments,max_path_segments,dt,elapsed_wall_seconds\n";
    }
    // the solve's convergence_method string (P3:6501-6502, 6185)
    auto method_name = [](const upic_step_info& in) {
        static const char* names[] = {"picard_fallback", "picard", "current_anderson", "current_anderson_backtracked",
                                      "current_picard", "current_picard_backtracked"};
        const int k = (in.update_method >= 0 && in.update_method < 6) ? in.update_method : 0;
        return std::string(in.selective_fallback_used ? "selective_" : "nested_full_dt_") + names[k];
    };
    // the NonlinearSolveInfo columns shared by both CSVs, from
    // local_particle_iterations_mean to selective_n64 (P3:8604-8620)
    auto solve_cols = [](std::ostream& o, const upic_step_info& in) {
        o << "," << in.local_iterations_mean << "," << in.local_iterations_max << "," << in.uniform_substeps_used << ","
          << in.selective_fallback_used << "," << in.direct_attempt_iterations << ","
          << in.selective_refinement_restarts << "," << in.selective_refined_particles << ","
          << in.selective_total_particle_substeps << "," << in.selective_work_ratio;
        for (int k = 0; k < 7; ++k) o << "," << in.selective_level_counts[k];
    };
    auto step_row = [&](int step, Real t, const upic_step_info& in, const upic_step_diagnostics& sd) {
        stepcsv << step << "," << std::setprecision(17) << t << "," << in.iterations << "," << in.final_rms_residual
                << "," << in.final_max_particle_residual << "," << in.anderson_restarts << "," << method_name(in)
                << "," << in.observed_convergence_rate << "," << in.estimated_map_contraction << ","
                << in.backtracked_steps << "," << in.minimum_update_alpha;
        solve_cols(stepcsv, in);
        stepcsv << "," << sd.A_chain_abs_rms << "," << sd.A_chain_rel_rms << "," << sd.A_chain_max_norm << ","
                << sd.A_chain_mean_norm << "," << sd.A_chain_work_defect << "," << sd.delta_K << "," << sd.delta_W
                << "," << sd.delta_total_energy << "," << sd.particle_work << "," << sd.grid_work << ","
                << sd.deposit_gather_work_residual << "," << sd.particle_energy_residual << ","
                << sd.field_energy_residual << "," << sd.crossing_particles << "," << sd.total_crossings << ","
                << sd.avg_path_segments << "," << sd.max_path_segments << "," << p.dt << "," << wall_since(t0) << "\n";
        stepcsv.flush();
    };
    std::ofstream diag(prefix + "_diagnostics.csv");
    diag << "step,time,gauss_rms,gauss_max,divE_rms,sigma1_rho_rms,"
         << "gauss_normalization_rms,gauss_relative,"
         << "gauge_rms,gauge_max,psi_over_kappa2_rms,divA_rms,"
         << "gauge_normalization_rms,gauge_relative,"
         << "field_energy,kinetic_energy,total_energy,"
         << "rho_mode,E_mode_re,E_mode_im,E_mode_abs,Bz_y_mode_re,Bz_y_mode_im,Bz_y_mode_abs,"
         << "theory_growth_rate,theory_mode_abs,mode_over_theory,mean_rho,"
         << "nonlinear_iterations,nonlinear_converged,nonlinear_rms_residual,"
         << "nonlinear_max_particle_residual,nonlinear_absolute_rms_residual,"
         << "anderson_restarts,anderson_steps,picard_fallback_steps,convergence_method,"
         << "observed_convergence_rate,estimated_map_contraction,backtracked_steps,minimum_update_alpha,"
         << "local_particle_iterations_mean,local_particle_iterations_max,uniform_substeps_used,"
         << "selective_fallback_used,direct_attempt_iterations,selective_refinement_restarts,"
         << "selective_refined_particles,selective_total_particle_substeps,selective_work_ratio,"
         << "selective_n1,selective_n2,selective_n4,selective_n8,selective_n16,selective_n32,selective_n64,"
         << "crossing_particles,avg_path_segments,max_path_segments,dt,elapsed_wall_seconds\n";
    // ps: the step's path statistics (crossing particles, average and
    // maximum segments; the step diagnostics' when they are computed,
    // P3:8839-8849)
    auto row = [&](int step, Real t, const upic_step_info& in, const std::string& method, Real crossing, Real avg_seg,
                   Real max_seg, Real step_dt) {
        upic_diagnostics g{};
        check(upic_sim_diagnostics(r.sim, mode, &g), r.ctx, "diagnostics");
        const Real gn = g.divE_rms + g.rho_rms, an = g.psi_over_c2_rms + g.divA_rms;
        Real th_abs = kNaN, th_ratio = kNaN;
        if (has_theory) {
            th_abs = theory_amp0 * std::exp(theory_growth * t);
            th_ratio = (use_B ? g.Bz_y_mode_abs : g.E_mode_abs) / (th_abs + 1.0e-300);
        }
        diag << step << "," << std::setprecision(17) << t << "," << g.gauss_rms << "," << g.gauss_max << ","
             << g.divE_rms << "," << g.rho_rms << "," << gn << "," << (gn > 0.0 ? g.gauss_rms / gn : 0.0) << ","
             << g.gauge_rms << "," << g.gauge_max << "," << g.psi_over_c2_rms << "," << g.divA_rms << "," << an << ","
             << (an > 0.0 ? g.gauge_rms / an : 0.0) << "," << g.field_energy << "," << g.kinetic_energy << ","
             << g.total_energy << "," << g.rho_mode << "," << g.E_mode_re << "," << g.E_mode_im << ","
             << g.E_mode_abs << "," << g.Bz_y_mode_re << "," << g.Bz_y_mode_im << "," << g.Bz_y_mode_abs << ","
             << theory_growth << "," << th_abs << "," << th_ratio << "," << g.mean_rho << "," << in.iterations << ","
             << in.converged << "," << in.final_rms_residual << "," << in.final_max_particle_residual << ","
             << in.final_absolute_rms_residual << "," << in.anderson_restarts << "," << in.anderson_steps << ","
             << in.picard_fallback_steps << "," << method << "," << in.observed_convergence_rate << ","
             << in.estimated_map_contraction << "," << in.backtracked_steps << "," << in.minimum_update_alpha;
        solve_cols(diag, in);
        diag << "," << crossing << "," << avg_seg << "," << max_seg << "," << step_dt << "," << wall_since(t0) << "\n";
        diag.flush();
        return g;
    };
    // row 0: the reference's initial NonlinearSolveInfo (P3:8626-8633)
    upic_step_info init{};
    init.converged = 1;
    init.final_rms_residual = init.final_max_particle_residual = init.final_absolute_rms_residual = kNaN;
    init.observed_convergence_rate = init.estimated_map_contraction = std::numeric_limits<Real>::infinity();
    init.minimum_update_alpha = 1.0;
    init.selective_work_ratio = 1.0;
    row(0, 0.0, init, "initial_state", 0, 0.0, 1, 0.0);
    for (int n = 0; n < n_steps; ++n) {
        upic_step_info info{};
        check(upic_step_p3(r.sim, &info), r.ctx, "step");
        const Real t = (n + 1) * p.dt;
        const std::string method = method_name(info);
        upic_diagnostics g{};
        if (step_diag) {
            upic_step_diagnostics sd{};
            check(upic_sim_step_diagnostics(r.sim, &sd), r.ctx, "step diagnostics");
            step_row(n + 1, t, info, sd);
            g = row(n + 1, t, info, method, (Real)sd.crossing_particles, sd.avg_path_segments,
                    (Real)sd.max_path_segments, p.dt);
        } else {
            const Real avg = info.path.n_particles ? (Real)info.path.total_segments / (Real)info.path.n_particles : 0.0;
            g = row(n + 1, t, info, method, (Real)info.path.crossing_particles, avg, (Real)info.path.max_segments, p.dt);
        }
        if (learned == UPIC_GUESS_LEARNED && n + 1 == 20) {
            check(upic_sim_collect_guess_data(r.sim, 0, 0), r.ctx, "collect");
            check(upic_sim_train_guess(r.sim, 15, 4096, 2e-3, 5, nullptr), r.ctx, "training");
        }
        if ((n + 1) % std::max(1, n_steps / 5) == 0 || n == n_steps - 1)
            std::cout << std::scientific << std::setprecision(3) << "step " << std::setw(5) << (n + 1) << "/"
                      << n_steps << " Gauss=" << g.gauss_rms << " TE=" << g.total_energy
                      << " outer=" << info.iterations << " inner=" << info.local_iterations_mean << "\n";
    }
    write_particle_sample(prefix + "_particle_sample.csv", r, 20000);  // P3:9039, same format as P1
    return 0;
}

// ---------------------------------------------------------------------------
// P2 (run_simulation, P2:1093-1189): the explicit spectral BDF code
// ---------------------------------------------------------------------------
struct P2Config {  // Config (P2:173-231) with its defaults
    std::string example = "weibel", method = "bdf1", charge_update = "conserving", output_prefix = "run";
    int nx = 32, ny = 32, n_steps = 100, particles_per_cell = 4, cloud_particles = 4096, perturbation_mode = 1;
    int diagnostic_stride = 1, particle_sample_count = 2000;
    Real Lx = 2.0 * kPi, Ly = 2.0 * kPi, dt = 0.05, c = 1.0, eps0 = 1.0, mu0 = 1.0, n0 = 1.0;
    Real electron_charge_sign = -1.0, ion_charge_sign = 1.0;
    Real weibel_vx = 0.5, weibel_vy_max = 0.01, weibel_sine_perturb = 0.002;
    Real cloud_sigma = 1.0, cloud_drift = 0.01, cloud_thermal = 0.005;
    unsigned int seed = 12345;
    bool run_unit_tests = true;
};

// read_config (P2:233-291): keys are case-insensitive
P2Config read_p2(const std::string& path) {
    Deck d;
    for (const auto& kv : read_deck(path)) {
        DeckValue& slot = d.kv[lower(kv.first)];
        if (kv.second.line >= slot.line) slot = kv.second;
    }
    P2Config c;
    if (const std::string* v = d.find({"example", "test_case"})) c.example = lower(*v);
    if (const std::string* v = d.find({"method"})) c.method = lower(*v);
    if (const std::string* v = d.find({"charge_update"})) c.charge_update = lower(*v);
    d.get({"nx"}, c.nx); d.get({"ny"}, c.ny); d.get({"lx"}, c.Lx); d.get({"ly"}, c.Ly);
    d.get({"n_steps"}, c.n_steps); d.get({"dt"}, c.dt);
    d.get({"particles_per_cell"}, c.particles_per_cell); d.get({"cloud_particles"}, c.cloud_particles);
    d.get({"c"}, c.c); d.get({"eps0"}, c.eps0); d.get({"mu0"}, c.mu0); d.get({"n0"}, c.n0);
    d.get({"electron_charge_sign"}, c.electron_charge_sign); d.get({"ion_charge_sign"}, c.ion_charge_sign);
    d.get({"weibel_vx"}, c.weibel_vx); d.get({"weibel_vy_max"}, c.weibel_vy_max);
    d.get({"weibel_sine_perturb"}, c.weibel_sine_perturb); d.get({"perturbation_mode"}, c.perturbation_mode);
    d.get({"cloud_sigma"}, c.cloud_sigma); d.get({"cloud_drift"}, c.cloud_drift);
    d.get({"cloud_thermal"}, c.cloud_thermal); d.get({"seed"}, c.seed);
    d.get({"diagnostic_stride"}, c.diagnostic_stride); d.get({"particle_sample_count"}, c.particle_sample_count);
    if (const std::string* v = d.find({"output_prefix"})) c.output_prefix = *v;
    d.get_bool({"run_unit_tests"}, c.run_unit_tests);
    d.warn_unused();
    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    if (!pow2(c.nx) || !pow2(c.ny))
        throw std::runtime_error("the internal radix-2 FFT requires nx and ny to be powers of two");
    if (c.method != "bdf1" && c.method != "bdf2")
        throw std::runtime_error("method must be bdf1 or bdf2 in this teaching code");
    if (c.charge_update != "conserving" && c.charge_update != "naive")
        throw std::runtime_error("charge_update must be conserving or naive");
    if (c.example != "weibel" && c.example != "cloud" && c.example != "drifting_cloud")
        throw std::runtime_error("example must be weibel or cloud");
    return c;
}

Real p2_gamma(Real vx, Real vy, Real vz, Real c) {  // gamma_from_v (P2:596-600)
    const Real beta2 = (vx * vx + vy * vy + vz * vz) / (c * c);
    return 1.0 / std::sqrt(std::max(1.0e-14, 1.0 - beta2));
}

Real wrap_periodic(Real x, Real L) {  // P2:160-166: into [0, L)
    x -= std::floor(x / L) * L;
    if (x >= L) x -= L;
    if (x < 0.0) x += L;
    return x;
}

// quadratic_weights + scatter_scalar (P2:537-562): nearest-node quadratic
// spline onto the [0, L) grid
void p2_scatter(std::vector<Real>& f, const P2Config& c, Real x, Real y, Real value) {
    const Real dx = c.Lx / c.nx, dy = c.Ly / c.ny, dA = dx * dy;
    int ix[3], iy[3];
    Real wx[3], wy[3];
    auto weights = [](Real pos, Real h, int n, int* idx, Real* w) {
        const Real s = pos / h;
        const int ic = (int)std::floor(s + 0.5);
        const Real r = s - (Real)ic;
        for (int k = 0; k < 3; ++k) idx[k] = (ic - 1 + k + n) % n;
        w[0] = 0.5 * (0.5 - r) * (0.5 - r);
        w[1] = 0.75 - r * r;
        w[2] = 0.5 * (0.5 + r) * (0.5 + r);
    };
    weights(x, dx, c.nx, ix, wx);
    weights(y, dy, c.ny, iy, wy);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) f[(size_t)ix[a] * c.ny + iy[b]] += value * wx[a] * wy[b] / dA;
}

// initialize_weibel_particles / initialize_cloud_particles (P2:624-726), with
// the same libstdc++ generators and draw order: x | v | P (SoA, 3 components,
// z = 0), q, m and the fixed ion density
struct P2Load {
    int64_t np = 0, ions = 0;
    std::vector<Real> x, v, P, q, m, ion_rho;
};
P2Load p2_load(const P2Config& c) {
    P2Load L;
    const Real dx = c.Lx / c.nx, dy = c.Ly / c.ny;
    std::mt19937 rng(c.seed);
    std::vector<Real> xs, ys, vx, vy;
    Real qe, me;
    L.ion_rho.assign((size_t)c.nx * c.ny, 0.0);
    if (c.example == "weibel") {
        const int nsub = std::max(1, (int)std::round(std::sqrt((Real)c.particles_per_cell)));
        const int total = 2 * c.nx * c.ny * nsub * nsub;
        const Real w = c.n0 * (c.Lx * c.Ly) / (Real)total;
        qe = c.electron_charge_sign * w;
        me = w;
        std::uniform_real_distribution<Real> uy(-c.weibel_vy_max, c.weibel_vy_max);
        for (int i = 0; i < c.nx; ++i)
            for (int j = 0; j < c.ny; ++j)
                for (int a = 0; a < nsub; ++a)
                    for (int b = 0; b < nsub; ++b) {
                        const Real x = ((Real)i + ((Real)a + 0.5) / nsub) * dx;
                        const Real y = ((Real)j + ((Real)b + 0.5) / nsub) * dy;
                        const Real seed = c.weibel_sine_perturb * c.c * std::sin(2.0 * kPi * c.perturbation_mode * x / c.Lx);
                        for (int beam = -1; beam <= 1; beam += 2) {
                            xs.push_back(x);
                            ys.push_back(y);
                            vx.push_back((Real)beam * c.weibel_vx * c.c);
                            vy.push_back(uy(rng) + seed);
                        }
                    }
        std::fill(L.ion_rho.begin(), L.ion_rho.end(), c.n0);
    } else {
        const int Np = c.cloud_particles;
        const Real w = c.n0 * (c.Lx * c.Ly) / (Real)Np;
        qe = c.electron_charge_sign * w;
        me = w;
        const Real qi = c.ion_charge_sign * w;
        std::normal_distribution<Real> normal(0.0, 1.0);
        for (int n = 0; n < Np; ++n) {
            const Real xc = 0.5 * c.Lx + c.cloud_sigma * normal(rng);
            const Real yc = 0.5 * c.Ly + c.cloud_sigma * normal(rng);
            const Real x = wrap_periodic(xc, c.Lx), y = wrap_periodic(yc, c.Ly);
            Real u = c.cloud_drift * c.c + c.cloud_thermal * normal(rng);
            Real t = c.cloud_drift * c.c + c.cloud_thermal * normal(rng);
            // cap at 0.2 c (rare large draws)
            const Real sp = std::sqrt(u * u + t * t + 0.0 * 0.0), vmax = 0.2 * c.c;
            if (sp > vmax) { u *= (vmax / sp); t *= (vmax / sp); }
            xs.push_back(x); ys.push_back(y); vx.push_back(u); vy.push_back(t);
            p2_scatter(L.ion_rho, c, x, y, qi);
        }
        L.ions = Np;
    }
    const int64_t np = (int64_t)xs.size();
    L.np = np;
    L.x.assign(3 * (size_t)np, 0.0); L.v.assign(3 * (size_t)np, 0.0); L.P.assign(3 * (size_t)np, 0.0);
    L.q.assign((size_t)np, qe); L.m.assign((size_t)np, me);
    for (int64_t i = 0; i < np; ++i) {
        L.x[i] = xs[i]; L.x[np + i] = ys[i];
        L.v[i] = vx[i]; L.v[np + i] = vy[i];
        const Real g = p2_gamma(vx[i], vy[i], 0.0, c.c);  // P = m gamma v (A = 0)
        L.P[i] = (me * g) * vx[i] + qe * 0.0;
        L.P[np + i] = (me * g) * vy[i] + qe * 0.0;
        L.P[2 * np + i] = (me * g) * 0.0 + qe * 0.0;
    }
    return L;
}

int run_p2(const std::string& deck_path) {
    const P2Config c = read_p2(deck_path);
    upic_params p{};
    p.code = UPIC_CODE_P2;
    p.nx = c.nx; p.ny = c.ny; p.nz = 1; p.Lx = c.Lx; p.Ly = c.Ly; p.Lz = 1.0; p.dt = c.dt;
    p.spline_order = 2; p.orbit_quad_order = 1;
    p.c = c.c; p.eps0 = c.eps0; p.mu0 = c.mu0; p.sigma1 = 1.0; p.sigma2 = 1.0;
    p.bdf_order = c.method == "bdf2" ? 2 : 1;
    p.charge_conserving = c.charge_update == "conserving" ? 1 : 0;
    p.precision = UPIC_PREC_FP64; p.sort_interval = 1;
    P2Load L = p2_load(c);
    Run r;
    r.np = L.np;
    r.N = c.nx * c.ny;
    check(upic_ctx_create(&p, &r.ctx), nullptr, "context");
    check(upic_sim_create(r.ctx, r.np, &r.sim), r.ctx, "simulation");
    std::vector<Real> fz(3 * (size_t)r.N, 0.0);
    upic_host_state hs{};
    hs.np = r.np;
    hs.x = L.x.data(); hs.v = L.v.data(); hs.v_prev = L.v.data(); hs.P = L.P.data(); hs.vbar_last = L.v.data();
    hs.q = L.q.data(); hs.m = L.m.data();
    hs.phi = fz.data(); hs.psi = fz.data(); hs.rho = fz.data(); hs.A = fz.data(); hs.U = fz.data(); hs.J = fz.data();
    hs.fixed_ion_rho = L.ion_rho.data();
    check(upic_sim_upload(r.sim, &hs), r.ctx, "upload");
    check(upic_sim_init_fields(r.sim, c.n0), r.ctx, "initial fields");
    {  // write_config_echo (P2:964-984)
        std::ofstream out(c.output_prefix + "_config_echo.txt");
        out << std::setprecision(17) << "example = " << c.example << "\nmethod = " << c.method
            << "\ncharge_update = " << c.charge_update << "\nnx = " << c.nx << "\nny = " << c.ny << "\nLx = " << c.Lx
            << "\nLy = " << c.Ly << "\ndt = " << c.dt << "\nn_steps = " << c.n_steps << "\nc = " << c.c
            << "\neps0 = " << c.eps0 << "\nmu0 = " << c.mu0 << "\nn0 = " << c.n0 << "\nmobile_electrons = " << L.np
            << "\nstatic_ions = " << L.ions << "\ncell_area = " << (c.Lx / c.nx) * (c.Ly / c.ny)
            << "\noutput_prefix = " << c.output_prefix << "\n";
    }
    std::ofstream diag(c.output_prefix + "_diagnostics.csv");
    diag << std::setprecision(17);
    diag << "step,time,requested_method,active_method,charge_update,"
         << "gauge_rms,gauss_rms,continuity_rms,Bz_l2,B_l2,E_l2,rho_l2,J_l2,"
         << "field_energy,kinetic_energy,total_energy,max_speed,mean_gamma\n";
    auto row = [&](int step, Real t, const char* active) {
        upic_diagnostics d{};
        check(upic_sim_diagnostics(r.sim, c.perturbation_mode, &d), r.ctx, "diagnostics");
        diag << step << ',' << t << ',' << c.method << ',' << active << ',' << c.charge_update << ',' << d.gauge_rms
             << ',' << d.gauss_rms << ',' << d.continuity_rms << ',' << d.Bz_l2 << ',' << d.B_l2 << ',' << d.E_l2
             << ',' << d.rho_l2 << ',' << d.J_l2 << ',' << d.field_energy << ',' << d.kinetic_energy << ','
             << d.total_energy << ',' << d.max_speed << ',' << d.mean_gamma << "\n";
    };
    row(0, 0.0, "initial");  // BDF1 with duplicated histories at t = 0
    for (int step = 1; step <= c.n_steps; ++step) {
        upic_step_info info{};
        check(upic_step_p2(r.sim, &info), r.ctx, "step");
        if (step % std::max(1, c.diagnostic_stride) == 0 || step == c.n_steps)
            row(step, step * c.dt, c.method.c_str());
    }
    {  // write_particle_sample (P2:986-1001)
        std::vector<Real> x(3 * (size_t)r.np), v(3 * (size_t)r.np), q((size_t)r.np), m((size_t)r.np);
        upic_host_state ds{};
        ds.np = r.np;
        ds.x = x.data(); ds.v = v.data(); ds.q = q.data(); ds.m = m.data();
        check(upic_sim_download(r.sim, &ds), r.ctx, "download");
        std::ofstream out(c.output_prefix + "_particle_sample.csv");
        out << std::setprecision(17) << "id,x,y,x_centered,y_centered,vx,vy,vz,q,m,gamma\n";
        const int64_t count = std::min<int64_t>(c.particle_sample_count, r.np);
        for (int64_t i = 0; i < count; ++i) {
            const Real vx = v[i], vy = v[r.np + i], vz = v[2 * r.np + i];
            out << i << ',' << x[i] << ',' << x[r.np + i] << ',' << x[i] - 0.5 * c.Lx << ','
                << x[r.np + i] - 0.5 * c.Ly << ',' << vx << ',' << vy << ',' << vz << ',' << q[i] << ',' << m[i]
                << ',' << p2_gamma(vx, vy, vz, c.c) << "\n";
        }
    }
    std::cout << "wrote " << c.output_prefix << "_diagnostics.csv\n"
              << "wrote " << c.output_prefix << "_particle_sample.csv\n"
              << "wrote " << c.output_prefix << "_config_echo.txt\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string code = argc == 3 ? argv[1] : "";
    if (code != "p1" && code != "p2" && code != "p3") {
        std::cerr << "usage: " << argv[0] << " p1|p2|p3 input_deck.txt\n";
        return 2;
    }
    try {
        return code == "p1" ? run_p1(argv[2]) : (code == "p2" ? run_p2(argv[2]) : run_p3(argv[2]));
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
