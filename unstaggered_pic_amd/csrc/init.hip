// init.hip -- particle loading (host-side C++): seeded loads for the
// synthetic benchmark workloads (upic_load_particles) and the reference's P3
// initial conditions (upic_p3_initial_state).
#include "api_internal.hpp"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): one 64-bit state per stream.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in (0, 1): 53 random bits, offset by half an ulp
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    // standard normal by Box-Muller (both variates used)
    bool have = false;
    double spare = 0.0;
    double normal() {
        if (have) { have = false; return spare; }
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double t = 6.283185307179586476925286766559 * u2;
        spare = r * std::sin(t);
        have = true;
        return r * std::cos(t);
    }
};

// x in [0, L) with density proportional to 1 + alpha cos(k x): Newton on the
// CDF u L = x + alpha sin(k x) / k.
double inverse_cdf_cos(double u, double L, double alpha, double k) {
    const double target = u * L;
    double x = target;
    for (int it = 0; it < 60; ++it) {
        const double f = x + alpha * std::sin(k * x) / k - target;
        const double df = 1.0 + alpha * std::cos(k * x);
        const double dx = f / df;
        x -= dx;
        if (std::abs(dx) < 1.0e-15 * L) break;
    }
    return x;
}

}  // namespace

extern "C" upic_status upic_load_particles(const upic_params* p, const upic_load_spec* sp, upic_host_state* st) {
    if (!p || !sp || !st) return UPIC_EINVAL;
    try {
        const int64_t ncell = (int64_t)p->nx * p->ny * p->nz;
        const int64_t np = ncell * sp->ppc;
        if (sp->ppc < 1 || st->np != np || !st->x || !st->v || !st->q || !st->m) return UPIC_EINVAL;
        const double L[3] = {p->Lx, p->Ly, p->Lz};
        const double h[3] = {p->Lx / p->nx, p->Ly / p->ny, p->Lz / p->nz};
        const double origin = (p->code == 2) ? 0.0 : -0.5;  // P2 grids start at 0
        const double q = -sp->n0 * L[0] * L[1] * L[2] / (double)np;
        const double kappa = p->c;
        // one independent stream per cell keeps the load reproducible and parallel
#pragma omp parallel for schedule(static)
        for (int64_t cell = 0; cell < ncell; ++cell) {
            SplitMix rng{sp->seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(cell + 1))};
            const int ix = (int)(cell / ((int64_t)p->ny * p->nz));
            const int iy = (int)((cell / p->nz) % p->ny);
            const int iz = (int)(cell % p->nz);
            for (int k = 0; k < sp->ppc; ++k) {
                const int64_t i = cell * sp->ppc + k;
                double x[3], v[3];
                if (sp->kind == UPIC_LOAD_BEAMS) {
                    // lattice of ppc points per cell along x, jittered
                    x[0] = (ix + (k + 0.5) / sp->ppc + sp->noise * (rng.uniform() - 0.5)) * h[0];
                    x[1] = (iy + 0.5 + sp->noise * (rng.uniform() - 0.5)) * h[1];
                    x[2] = (iz + 0.5) * h[2];
                } else if (sp->kind == UPIC_LOAD_LANDAU) {
                    // the cell index only picks the stream; x follows the perturbed density
                    const double kx = 6.283185307179586476925286766559 * sp->mode / L[0];
                    x[0] = inverse_cdf_cos(rng.uniform(), L[0], sp->alpha, kx);
                    x[1] = rng.uniform() * L[1];
                    x[2] = (iz + 0.5) * h[2];
                    if (p->ny > 1) x[1] = (iy + rng.uniform()) * h[1];
                } else {
                    x[0] = (ix + rng.uniform()) * h[0];
                    x[1] = (iy + rng.uniform()) * h[1];
                    x[2] = (p->nz > 1) ? (iz + rng.uniform()) * h[2] : (iz + 0.5) * h[2];
                }
                // P3 is relativistic: thermal spreads are drawn in momentum
                // (u = gamma v) so that no speed reaches c
                if (sp->kind == UPIC_LOAD_BIMAXWELL || (sp->kind == UPIC_LOAD_LANDAU && p->code == 3)) {
                    const double u[3] = {sp->vth[0] * rng.normal(), sp->vth[1] * rng.normal(), sp->vth[2] * rng.normal()};
                    const double g = std::sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / (kappa * kappa));
                    for (int d = 0; d < 3; ++d) v[d] = u[d] / g;
                } else {
                    for (int d = 0; d < 3; ++d) v[d] = sp->vth[d] * rng.normal();
                    if (sp->kind == UPIC_LOAD_BEAMS) v[0] += (k % 2 == 0) ? sp->v0 : -sp->v0;
                }
                for (int d = 0; d < 3; ++d) {
                    double xd = x[d] + origin * L[d];
                    // into [-L/2, L/2) (P1/P3) or [0, L) (P2)
                    const double lo = origin * L[d];
                    if (xd < lo) xd += L[d];
                    if (xd >= lo + L[d]) xd -= L[d];
                    st->x[d * np + i] = xd;
                    st->v[d * np + i] = v[d];
                }
                st->q[i] = q;
                st->m[i] = std::abs(q);
            }
        }
        return UPIC_OK;
    } catch (const std::exception&) {
        return UPIC_EINVAL;
    }
}

// ---------------------------------------------------------------------------
// The reference's P3 initial conditions (P3:3689-4233).  Same formulas and
// operation order as the reference, so the particle part matches it to the
// last bit; the field tail runs on the device (upic_sim_init_fields_AU).
// ---------------------------------------------------------------------------
namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct InitError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

double wrap_1(double x, double L) {  // wrap1 (P3:704-708): into [-L/2, L/2)
    double y = std::fmod(x + 0.5 * L, L);
    if (y < 0.0) y += L;
    return y - 0.5 * L;
}

// cell_centered_grid(n, L, true) (P3:3599-3607)
std::vector<double> centred_points(int n, double L) {
    std::vector<double> x((size_t)n);
    const double dx = L / (double)n;
    for (int i = 0; i < n; ++i) x[(size_t)i] = ((double)i + 0.5) * dx - 0.5 * L;
    return x;
}

// Inverse standard-normal CDF, P. J. Acklam's rational approximation (the
// published coefficient table; normal_quantile P3:3623-3669 uses the same
// method), |relative error| < 1.15e-9.
double normal_quantile(double p) {
    if (!(p > 0.0 && p < 1.0)) throw InitError("normal_quantile requires 0<p<1");
    static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                               1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                               6.680131188771972e+01, -1.328068155288572e+01};
    static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                               -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                               3.754408661907416e+00};
    const double plow = 0.02425;
    if (p < plow || p > 1.0 - plow) {
        // tails: rational function of sqrt(-2 log(min(p, 1 - p))), odd in p - 1/2
        const double t = std::sqrt(-2.0 * std::log(p < plow ? p : 1.0 - p));
        const double x = (((((c[0] * t + c[1]) * t + c[2]) * t + c[3]) * t + c[4]) * t + c[5]) /
                         ((((d[0] * t + d[1]) * t + d[2]) * t + d[3]) * t + 1.0);
        return p < plow ? x : -x;
    }
    const double u = p - 0.5, r = u * u;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * u /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
}

// Fourier factor of the degree-r centred B-spline: sinc(k dx / 2)^(r+1)
// (bspline_mode_factor_1d, P3:1157-1162)
double spline_mode_factor(int degree, double k, double dx) {
    const double z = 0.5 * k * dx;
    const double s = std::abs(z) < 1.0e-12 ? 1.0 : std::sin(z) / z;
    return std::pow(s, (double)(degree + 1));
}

// Lorentz factor of a speed along x; |v| < kappa (gamma_from_v, P3:623-629)
double lorentz_gamma(double v, double kappa) {
    const double beta2 = (v * v) / (kappa * kappa);
    if (beta2 >= 1.0) throw InitError("relativistic initial velocity must satisfy |v| < kappa");
    return 1.0 / std::sqrt(std::max(1.0e-300, 1.0 - beta2));
}

// Cold symmetric two-stream rate with omega_p^2 |S(k)|^2: the unstable root
// of 1 = a/(w - k v0)^2 + a/(w + k v0)^2, a = omega_p^2 / (2 gamma0^3)
// (cold_two_stream_growth_rate, P3:4160-4195)
double two_stream_rate(const upic_params& p, const upic_p3_case& cs) {
    const double k = 2.0 * kPi * (double)std::max(1, cs.perturbation_mode) / p.Lx;
    const double g0 = lorentz_gamma(cs.v0, p.c);
    const double S = cs.two_stream_finite_grid_theory ? spline_mode_factor(p.spline_order, k, p.Lx / p.nx) : 1.0;
    const double wp2 = (p.sigma1 * cs.n0) * S * S;
    const double a = wp2 / (2.0 * g0 * g0 * g0);
    const double b = k * k * cs.v0 * cs.v0;
    const double y = b + a - std::sqrt(a * a + 4.0 * a * b);
    return y < 0.0 ? std::sqrt(-y) : 0.0;
}

// Cold filamentation (Weibel) rate for k = k_y: Gamma^2 is the positive root
// of G^2 + (kappa^2 k^2 + wp2/gamma0^3) G - k^2 wp2 v0^2 / gamma0 = 0
// (cold_weibel_growth_rate, P3:4197-4226)
double weibel_rate(const upic_params& p, const upic_p3_case& cs) {
    const double k = 2.0 * kPi * (double)std::max(1, cs.perturbation_mode) / p.Ly;
    const double g0 = lorentz_gamma(cs.v0, p.c);
    const double S = cs.weibel_finite_grid_theory ? spline_mode_factor(p.spline_order, k, p.Ly / p.ny) : 1.0;
    const double wp2 = (p.sigma1 * cs.n0) * S * S;
    const double A = p.c * p.c * k * k + wp2 / (g0 * g0 * g0);
    const double C = k * k * wp2 * cs.v0 * cs.v0 / g0;
    const double G2 = 0.5 * (std::sqrt(A * A + 4.0 * C) - A);
    return G2 > 0.0 ? std::sqrt(G2) : 0.0;
}

double speed(const double* v, int64_t np, int64_t i) {
    return std::sqrt(v[i] * v[i] + v[np + i] * v[np + i] + v[2 * np + i] * v[2 * np + i]);
}

// Two cold streams at +-v0 along x on a shared lattice of nx*ppc x-points
// per (y, z) cell line, seeded by the growing cold-fluid eigenmode (or the
// legacy common velocity seed) (P3:3689-3796).
void init_two_stream(const upic_params& p, const upic_p3_case& cs, upic_host_state* st) {
    const int nxs = p.nx * cs.particles_per_cell_pair;
    const int64_t nbase = (int64_t)nxs * p.ny * p.nz, np = 2 * nbase;
    if (!st->x) { st->np = np; return; }
    if (st->np != np) throw InitError("upic_p3_initial_state: st->np does not match the deck");
    const double L[3] = {p.Lx, p.Ly, p.Lz};
    const double kx = 2.0 * kPi * (double)cs.perturbation_mode / p.Lx;
    if (!(kx > 0.0)) throw InitError("two-stream initialization needs perturbation_mode>0");
    const double g0 = lorentz_gamma(cs.v0, p.c), g03 = g0 * g0 * g0;
    const double ax = cs.two_stream_density_alpha >= 0.0 ? cs.two_stream_density_alpha : 0.0;
    if (std::abs(ax) >= 0.25)
        throw InitError("two_stream_density_alpha is a linear-mode amplitude and should be < 0.25");
    const bool eig = cs.two_stream_eigenmode && cs.two_stream_density_alpha >= 0.0;
    const double S = cs.two_stream_finite_grid_theory ? spline_mode_factor(p.spline_order, kx, p.Lx / p.nx) : 1.0;
    const double G = two_stream_rate(p, cs);
    // grid charge amplitude ax n0 -> grid field E0 = sigma1 ax n0 / kx, felt
    // by the particles as S E0
    const double Eg = (eig && ax > 0.0) ? (ax * p.sigma1 * cs.n0 / kx) : 0.0;
    const double Ep = S * Eg;
    const double den = kx * kx * cs.v0 * cs.v0 + G * G;
    const double vc = den > 0.0 ? (Ep * kx * cs.v0 / (g03 * den)) : 0.0;
    const double vs = den > 0.0 ? (Ep * G / (g03 * den)) : 0.0;
    const double nc = den > 0.0 ? (-kx * Ep * (kx * kx * cs.v0 * cs.v0 - G * G) / (g03 * den * den)) : 0.0;
    const double ns = den > 0.0 ? (2.0 * kx * kx * cs.v0 * G * Ep / (g03 * den * den)) : 0.0;
    const auto xl = centred_points(nxs, p.Lx), yl = centred_points(p.ny, p.Ly), zl = centred_points(p.nz, p.Lz);
    const double qm = -cs.n0 * (p.Lx * p.Ly * p.Lz) / (double)np;
    std::vector<double> w((size_t)np);
    int64_t b = 0;
    for (double y : yl)
        for (double z : zl)
            for (double x : xl) {
                const double pos[3] = {x, y, z};
                const double ph = kx * x, cph = std::cos(ph), sph = std::sin(ph);
                for (int d = 0; d < 3; ++d) st->x[d * np + b] = st->x[d * np + nbase + b] = wrap_1(pos[d], L[d]);
                double up, um, wp = 1.0, wm = 1.0;
                if (eig) {
                    // omega = i Gamma: opposite cos(kx x) and common sin(kx x)
                    // velocity parts, quadrature-phase stream densities
                    up = cs.v0 + (vc * cph - vs * sph);
                    um = -cs.v0 + (-vc * cph - vs * sph);
                    wp = 1.0 + nc * cph + ns * sph;
                    wm = 1.0 + nc * cph - ns * sph;
                } else {
                    const double seed = cs.perturbation * std::sin(ph);
                    up = cs.v0 + seed;
                    um = -cs.v0 + seed;
                    if (cs.two_stream_density_alpha >= 0.0) wp = wm = 1.0 + ax * std::cos(ph);
                }
                st->v[b] = up; st->v[nbase + b] = um;
                st->v[np + b] = st->v[np + nbase + b] = 0.0;
                st->v[2 * np + b] = st->v[2 * np + nbase + b] = 0.0;
                w[(size_t)b] = wp; w[(size_t)(nbase + b)] = wm;
                ++b;
            }
    // common transverse current seed (not an eigenmode; P3:3770-3778)
    if (cs.tsi_transverse_seed != 0.0)
        for (int64_t i = 0; i < np; ++i)
            st->v[i] += cs.tsi_transverse_seed * std::cos(2.0 * kPi * cs.tsi_transverse_mode * st->x[np + i] / p.Ly);
    for (int64_t i = 0; i < np; ++i)
        if (speed(st->v, np, i) >= p.c)
            throw InitError("two-stream initialization generated |v| >= kappa; reduce v0 or seed amplitude");
    for (int64_t i = 0; i < np; ++i) {
        if (!(w[(size_t)i] > 0.0))
            throw InitError("two-stream density/eigenmode seed produced a nonpositive macro weight; reduce alpha");
        st->q[i] = qm * w[(size_t)i];
        st->m[i] = std::abs(st->q[i]);
    }
}

// Quiet start: nv inverse-CDF Maxwellian samples along x at every cell
// centre, charge modulated by 1 + alpha cos(k x) (P3:3853-3937)
void init_landau(const upic_params& p, const upic_p3_case& cs, upic_host_state* st) {
    const int nv = cs.particles_per_cell_pair;
    if (nv < 2) throw InitError("Landau examples need at least two velocity samples per cell");
    const double alpha = cs.landau_alpha >= 0.0 ? cs.landau_alpha : cs.perturbation;
    if (std::abs(alpha) >= 1.0)
        throw InitError("Landau density amplitude must satisfy |alpha| < 1 so all macro weights stay positive");
    const int64_t np = (int64_t)p.nx * p.ny * p.nz * nv;
    if (!st->x) { st->np = np; return; }
    if (st->np != np) throw InitError("upic_p3_initial_state: st->np does not match the deck");
    const double L[3] = {p.Lx, p.Ly, p.Lz};
    std::vector<double> vq((size_t)nv);
    for (int j = 0; j < nv; ++j) vq[(size_t)j] = cs.thermal_velocity * normal_quantile(((double)j + 0.5) / (double)nv);
    for (double u : vq) lorentz_gamma(u, p.c);  // momentum_from_v rejects |v| >= kappa
    const double kx = 2.0 * kPi * (double)cs.perturbation_mode / p.Lx;
    const double q0 = -cs.n0 * (p.Lx * p.Ly * p.Lz) / (double)np;
    const auto xl = centred_points(p.nx, p.Lx), yl = centred_points(p.ny, p.Ly), zl = centred_points(p.nz, p.Lz);
    int64_t i = 0;
    for (double y : yl)
        for (double z : zl)
            for (double x : xl) {
                const double w = 1.0 + alpha * std::cos(kx * x);
                const double pos[3] = {wrap_1(x, L[0]), wrap_1(y, L[1]), wrap_1(z, L[2])};
                for (int j = 0; j < nv; ++j, ++i) {
                    for (int d = 0; d < 3; ++d) st->x[d * np + i] = pos[d];
                    st->v[i] = vq[(size_t)j];
                    st->v[np + i] = st->v[2 * np + i] = 0.0;
                    st->q[i] = q0 * w;
                    st->m[i] = std::abs(st->q[i]);
                }
            }
}

// Cold relativistic filamentation: streams at +-v0 along x, k = k_y, the
// growing cold-fluid eigenmode in B_z (through A_x), U_x = Gamma A_x, the
// transverse velocities and the stream-opposite weights (P3:3961-4108)
void init_weibel(const upic_params& p, const upic_p3_case& cs, upic_host_state* st) {
    const int nxc = std::max(1, cs.particles_per_cell_pair);
    const int64_t nbase = (int64_t)p.nx * nxc * p.ny * p.nz, np = 2 * nbase;
    if (std::abs(cs.v0) >= p.c) throw InitError("Weibel initialization requires |v0| < kappa");
    const int mode = std::max(1, cs.perturbation_mode);
    const double ky = 2.0 * kPi * (double)mode / p.Ly;
    if (!(ky > 0.0)) throw InitError("Weibel initialization needs perturbation_mode>0");
    const double B0 = cs.weibel_B0 >= 0.0 ? cs.weibel_B0 : cs.perturbation;
    const double G = weibel_rate(p, cs);
    if (cs.weibel_eigenmode && !(G > 0.0))
        throw InitError("Weibel eigenmode requested, but the selected mode is not unstable");
    const double g0 = lorentz_gamma(cs.v0, p.c);
    const double S = cs.weibel_finite_grid_theory ? spline_mode_factor(p.spline_order, ky, p.Ly / p.ny) : 1.0;
    // amplitudes relative to the mesh B_z amplitude B0 (q/m = -1)
    const bool eig = cs.weibel_eigenmode && G > 0.0;
    const double dvx = eig ? (S * B0 / (ky * g0 * g0 * g0)) : 0.0;
    const double dvy = eig ? (S * cs.v0 * B0 / (g0 * G)) : 0.0;
    const double dn = eig ? (S * cs.v0 * ky * B0 / (g0 * G * G)) : 0.0;
    if (std::abs(dn) >= 0.25) throw InitError("Weibel eigenmode density perturbation is too large; reduce weibel_B0");
    if (!st->x) { st->np = np; return; }
    if (st->np != np) throw InitError("upic_p3_initial_state: st->np does not match the deck");
    const double L[3] = {p.Lx, p.Ly, p.Lz};
    const auto xl = centred_points(p.nx * nxc, p.Lx), yl = centred_points(p.ny, p.Ly), zl = centred_points(p.nz, p.Lz);
    const double qm = -cs.n0 * (p.Lx * p.Ly * p.Lz) / (double)np;
    std::vector<double> w((size_t)np);
    int64_t b = 0;
    for (double y : yl)
        for (double z : zl)
            for (double x : xl) {
                const double pos[3] = {x, y, z};
                const double ph = ky * y, cph = std::cos(ph), sph = std::sin(ph);
                for (int d = 0; d < 3; ++d) st->x[d * np + b] = st->x[d * np + nbase + b] = wrap_1(pos[d], L[d]);
                st->v[b] = cs.v0 + dvx * cph;
                st->v[nbase + b] = -cs.v0 + dvx * cph;
                st->v[np + b] = dvy * sph;
                st->v[np + nbase + b] = -dvy * sph;
                st->v[2 * np + b] = st->v[2 * np + nbase + b] = 0.0;
                w[(size_t)b] = 1.0 - dn * cph;
                w[(size_t)(nbase + b)] = 1.0 + dn * cph;
                ++b;
            }
    for (int64_t i = 0; i < np; ++i)
        if (speed(st->v, np, i) >= p.c)
            throw InitError("Weibel initialization generated |v| >= kappa; reduce v0 or weibel_B0");
    for (int64_t i = 0; i < np; ++i) {
        if (!(w[(size_t)i] > 0.0))
            throw InitError("Weibel eigenmode produced a nonpositive stream weight; reduce weibel_B0");
        st->q[i] = qm * w[(size_t)i];
        st->m[i] = std::abs(st->q[i]);
    }
    // A_x = (B0/k) cos(k y) at the nodes, U_x = Gamma A_x (P3:4081-4092)
    const int N = p.nx * p.ny * p.nz;
    if (st->A) std::fill(st->A, st->A + 3 * (size_t)N, 0.0);
    if (st->U) std::fill(st->U, st->U + 3 * (size_t)N, 0.0);
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j) {
            const double y = -0.5 * p.Ly + (double)j * (p.Ly / p.ny);
            const double Ax = (B0 / ky) * std::cos(ky * y);
            for (int k = 0; k < p.nz; ++k) {
                const size_t id = ((size_t)i * p.ny + j) * p.nz + k;
                if (st->A) st->A[id] = Ax;
                if (st->U) st->U[id] = cs.weibel_eigenmode ? (G * Ax) : 0.0;
            }
        }
}

}  // namespace

extern "C" void upic_p3_case_defaults(int32_t kind, upic_p3_case* cs) {
    if (!cs) return;
    std::memset(cs, 0, sizeof(*cs));
    cs->kind = kind;
    cs->particles_per_cell_pair = 4;
    cs->perturbation_mode = 1;
    cs->tsi_transverse_mode = 1;
    cs->two_stream_eigenmode = cs->two_stream_finite_grid_theory = 1;
    cs->weibel_eigenmode = cs->weibel_finite_grid_theory = 1;
    cs->n0 = 1.0; cs->v0 = 0.3; cs->perturbation = 0.02;
    cs->two_stream_density_alpha = cs->weibel_B0 = cs->landau_alpha = -1.0;
    cs->thermal_velocity = 1.0;
    cs->tsi_transverse_seed = 0.0;
}

extern "C" upic_status upic_p3_initial_state(const upic_params* p, const upic_p3_case* cs, upic_host_state* st,
                                             double* growth_rate) {
    if (!p || !cs || !st) return UPIC_EINVAL;
    try {
        if (p->code != 3) throw InitError("upic_p3_initial_state needs P3 parameters (code 3)");
        if (p->nx < 1 || p->ny < 1 || p->nz < 1 || !(p->Lx > 0.0 && p->Ly > 0.0 && p->Lz > 0.0))
            throw InitError("upic_p3_initial_state: bad grid");
        if (st->x && (!st->v || !st->q || !st->m)) throw InitError("upic_p3_initial_state: x, v, q, m are required");
        double rate = 0.0;
        if (cs->kind == UPIC_P3_TWO_STREAM) {
            init_two_stream(*p, *cs, st);
            rate = two_stream_rate(*p, *cs);
        } else if (cs->kind == UPIC_P3_WEIBEL) {
            init_weibel(*p, *cs, st);
            rate = weibel_rate(*p, *cs);
        } else if (cs->kind == UPIC_P3_LANDAU) {
            init_landau(*p, *cs, st);
        } else {
            throw InitError("unknown test_case/initial_condition kind " + std::to_string(cs->kind));
        }
        if (st->x && cs->kind != UPIC_P3_WEIBEL) {
            const size_t n3 = 3 * (size_t)p->nx * p->ny * p->nz;
            if (st->A) std::fill(st->A, st->A + n3, 0.0);
            if (st->U) std::fill(st->U, st->U + n3, 0.0);
        }
        if (growth_rate) *growth_rate = rate;
        return UPIC_OK;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return UPIC_EINVAL;
    }
}
