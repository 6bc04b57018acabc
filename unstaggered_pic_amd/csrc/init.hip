// init.hip -- seeded particle loading for the synthetic benchmark workloads
// (host-side C++; upic_load_particles in include/upic.h).
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
