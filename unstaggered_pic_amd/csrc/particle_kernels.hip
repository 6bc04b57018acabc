// particle_kernels.hip -- particle-side HIP kernels for gfx950.
//
// One work-item per particle (64-wide wavefronts, 256-wide workgroups);
// particle arrays are SoA FP64 (x[0..np) | y | z).  Every kernel is
// instantiated for spline degree R = 1..4 and for 2-D (nz = 1) or 3-D grids.
#include "kernels.hpp"
#include "particles.hpp"
#include "orbit2d.hpp"

namespace upic {

__device__ __forceinline__ void ld3(const double* __restrict__ a, int64_t np, int64_t p, double o[3]) {
    o[0] = a[p]; o[1] = a[np + p]; o[2] = a[2 * np + p];
}
__device__ __forceinline__ void st3(double* __restrict__ a, int64_t np, int64_t p, const double v[3]) {
    a[p] = v[0]; a[np + p] = v[1]; a[2 * np + p] = v[2];
}
__device__ __forceinline__ bool finite3(const double v[3]) {
    return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}
__device__ __forceinline__ double norm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// ---------------------------------------------------------------------------
// Group A kernels
// ---------------------------------------------------------------------------
template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_deposit_orbit(GridDev g, QuadDev q, OrbitOpt opt, const double* __restrict__ x0,
                                                       const double* __restrict__ dx, const double* __restrict__ vbar,
                                                       const double* __restrict__ qa, int64_t np, double* __restrict__ J,
                                                       unsigned long long* __restrict__ stats) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int nseg = 0;
    if (p < np) {
        double X[3], D[3], V[3];
        ld3(x0, np, p, X); ld3(dx, np, p, D); ld3(vbar, np, p, V);
        nseg = deposit_orbit_one<R, Z3>(g, q, opt, X, D, V, qa[p], J);
    }
    path_stats_add(nseg, stats);
}

template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_gather_orbit(GridDev g, QuadDev q, OrbitOpt opt, const double* __restrict__ x0,
                                                      const double* __restrict__ dx, const double* __restrict__ F,
                                                      int64_t np, double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double X[3], D[3], G[3], Dm[3][3];
    ld3(x0, np, p, X); ld3(dx, np, p, D);
    orbit_gather_grad<R, Z3, true, false>(g, q, opt, X, D, F, nullptr, nullptr, 0.0, 1.0, false, G, Dm);
    st3(out, np, p, G);
}

template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_grad_A(GridDev g, QuadDev q, OrbitOpt opt, const double* __restrict__ x0,
                                                const double* __restrict__ dx, const double* __restrict__ A0,
                                                const double* __restrict__ A1, double th0, double th1, int p1_theta,
                                                int64_t np, double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double X[3], D[3], G[3], Dm[3][3];
    ld3(x0, np, p, X); ld3(dx, np, p, D);
    orbit_gather_grad<R, Z3, false, true>(g, q, opt, X, D, nullptr, A0, A1, th0, th1, p1_theta != 0, G, Dm);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) out[(a * 3 + b) * np + p] = Dm[a][b];
}

template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_gather_point(GridDev g, const double* __restrict__ pos,
                                                      const double* __restrict__ F, int64_t np, double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double X[3], o[3];
    ld3(pos, np, p, X);
    gather3<R, Z3>(g, F, X, o);
    st3(out, np, p, o);
}

template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_deposit_scalar_point(GridDev g, const double* __restrict__ pos,
                                                              const double* __restrict__ val, int64_t np, double scale,
                                                              double* __restrict__ mesh) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double X[3];
    ld3(pos, np, p, X);
    PointSt<R, Z3> st;
    point_stencil<R, Z3>(g, X, st);
    const double v = val[p];
    for_each_node<R, Z3>(g, st, [&](int id, double w, const double*) { atomic_add(&mesh[id], v * w * scale); });
}

// ---------------------------------------------------------------------------
// P1: global Picard (P1:1455-1552)
// ---------------------------------------------------------------------------

// First deposit of a step: v_guess = 2 v^n - v^{n-1} (P1:1514), vbar, orbit
// current (P1:1467-1473).  Stores v_guess.
template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_p1_deposit_guess(GridDev g, QuadDev q, OrbitOpt opt, P1Args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double tile_mem[kTile * kTile * 3];
    DepTile T{};
    if constexpr (!Z3) T = tile_begin<R>(g, tile_mem, a.x, a.np, (int64_t)blockIdx.x * blockDim.x);
    int nseg = 0;
    if (p < a.np) {
        double X[3], V[3], Vp[3], G[3], Vb[3], D[3];
        ld3(a.x, a.np, p, X); ld3(a.v, a.np, p, V); ld3(a.v_prev, a.np, p, Vp);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            G[d] = 2.0 * V[d] - Vp[d];
            Vb[d] = 0.5 * (G[d] + V[d]);
            D[d] = a.dt * Vb[d];
        }
        st3(a.vg_out, a.np, p, G);
        if constexpr (Z3) nseg = deposit_orbit_one<R, Z3>(g, q, opt, X, D, Vb, a.q[p], a.J_out);
        else nseg = deposit_orbit_2d<R>(g, q, opt, X, D, Vb, a.q[p], T, a.J_out);
    }
    if constexpr (!Z3) tile_end(g, T, a.J_out);
    block_partials<0, 0>(nullptr, nullptr, nseg, a.part);
}

// Map evaluation tail (P1:1481-1499) for one iterate: orbit gather of
// grad phi_mid, orbit D_A, A_end at x_end, canonical push.  In iterate mode
// it also accumulates the velocity-error sums (P1:188-195) and deposits the
// current of the next map evaluation (vbar = (v_next + v^n)/2) so each Picard
// iteration is a single particle sweep.  In final mode it writes the state.
template <int R, bool Z3, bool FINAL>
__global__ __launch_bounds__(256) void k_p1_push(GridDev g, QuadDev q, OrbitOpt opt, P1Args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double tile_mem[kTile * kTile * 3];
    DepTile T{};
    if constexpr (!Z3 && !FINAL) T = tile_begin<R>(g, tile_mem, a.x, a.np, (int64_t)blockIdx.x * blockDim.x);
    double err[3] = {0.0, 0.0, 0.0};
    int nseg = 0;
    if (p < a.np) {
        double X[3], V[3], Gs[3], Pp[3];
        ld3(a.x, a.np, p, X); ld3(a.v, a.np, p, V); ld3(a.vg, a.np, p, Gs); ld3(a.P, a.np, p, Pp);
        const double qp = a.q[p], mp = a.m[p];
        double Vb[3], Dx[3], Xe[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Vb[d] = 0.5 * (Gs[d] + V[d]);
            Dx[d] = a.dt * Vb[d];
        }
        Xe[0] = wrap1(X[0] + Dx[0], g.L[0]);
        Xe[1] = wrap1(X[1] + Dx[1], g.L[1]);
        Xe[2] = wrap1(X[2] + Dx[2], g.L[2]);
        double gpb[3], DA[3][3];
        double Ae[3];
        if constexpr (Z3) {
            orbit_gather_grad<R, Z3, true, true>(g, q, opt, X, Dx, a.gpm, a.A_n, a.A_end, 0.0, 1.0, true, gpb, DA);
            gather3<R, Z3>(g, a.A_end, Xe, Ae);
        } else {
            orbit_gg_2d<R>(g, q, opt, X, Dx, a.NR, 0.0, 1.0, true, gpb, DA);
            gather_A_rec<R>(g, a.NR, 1, Xe, Ae);
        }
        double Pe[3], Vn[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double DATv = DA[0][b] * Vb[0] + DA[1][b] * Vb[1] + DA[2][b] * Vb[2];
            const double f = (-qp) * gpb[b] + qp * DATv;
            Pe[b] = Pp[b] + a.dt * f;
            Vn[b] = (Pe[b] - qp * Ae[b]) / mp;
        }
        if constexpr (FINAL) {
            st3(a.x_out, a.np, p, Xe);
            st3(a.v_out, a.np, p, Vn);
            st3(a.P_out, a.np, p, Pe);
            st3(a.vprev_out, a.np, p, V);
            st3(a.vbar_out, a.np, p, Vb);
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double df = Vn[d] - Gs[d];
                err[0] += df * df;
                err[1] += Vn[d] * Vn[d];
                err[2] += Gs[d] * Gs[d];
            }
            st3(a.vg_out, a.np, p, Vn);
            double Vb2[3], D2[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Vb2[d] = 0.5 * (Vn[d] + V[d]);
                D2[d] = a.dt * Vb2[d];
            }
            if constexpr (Z3) nseg = deposit_orbit_one<R, Z3>(g, q, opt, X, D2, Vb2, qp, a.J_out);
            else nseg = deposit_orbit_2d<R>(g, q, opt, X, D2, Vb2, qp, T, a.J_out);
        }
    }
    if constexpr (!Z3 && !FINAL) tile_end(g, T, a.J_out);
    if constexpr (!FINAL) block_partials<3, 0>(err, nullptr, nseg, a.part);
}

// ---------------------------------------------------------------------------
// P3: per-particle inner solve fused with the orbit deposit
// (evaluate_selective_current_map_at_layout P3:5235-5431 with
//  solve_local_particle_substep 4930-4983 and
//  evaluate_local_particle_substep 4843-4884)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gamma_p(const double p[3], double m, double kappa) {
    const double mk = m * kappa;
    return sqrt(1.0 + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / (mk * mk));
}

// higuera_cary_rotate_with_K (P3:4636-4655) with B_eff from K (4605-4616).
__device__ __forceinline__ void hc_rotate(const double pm[3], const double B[3], double q, double m, double kappa,
                                          double dt, double out[3]) {
    const double inv = 1.0 / (m * kappa);
    const double u[3] = {inv * pm[0], inv * pm[1], inv * pm[2]};
    const double gm = sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const double c = q * dt / (2.0 * m);
    const double tau[3] = {c * B[0], c * B[1], c * B[2]};
    const double tau2 = tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2];
    const double udt = u[0] * tau[0] + u[1] * tau[1] + u[2] * tau[2];
    const double sigma = gm * gm - tau2;
    const double ghc2 = 0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + udt * udt)));
    const double ghc = sqrt(fmax(ghc2, 1.0));
    const double ig = 1.0 / ghc;
    const double t[3] = {ig * tau[0], ig * tau[1], ig * tau[2]};
    const double sc = 2.0 / (1.0 + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
    const double sv[3] = {sc * t[0], sc * t[1], sc * t[2]};
    const double up[3] = {u[0] + (u[1] * t[2] - u[2] * t[1]), u[1] + (u[2] * t[0] - u[0] * t[2]),
                          u[2] + (u[0] * t[1] - u[1] * t[0])};
    const double upl[3] = {u[0] + (up[1] * sv[2] - up[2] * sv[1]), u[1] + (up[2] * sv[0] - up[0] * sv[2]),
                           u[2] + (up[0] * sv[1] - up[1] * sv[0])};
    const double mk = m * kappa;
    out[0] = mk * upl[0]; out[1] = mk * upl[1]; out[2] = mk * upl[2];
}

// gather_vector_time_interp_at (P3:2927-2933); theta 0 and 1 reduce to a
// single gather exactly as (1-0)a0 + 0 a1 and 0 a0 + 1 a1 do for finite data.
template <int R, bool Z3>
__device__ __forceinline__ void gather_time_interp(const GridDev& g, const double* A0, const double* A1, double theta,
                                                   const double x[3], double out[3]) {
    if (theta == 0.0) { gather3<R, Z3>(g, A0, x, out); return; }
    if (theta == 1.0) { gather3<R, Z3>(g, A1, x, out); return; }
    double a0[3], a1[3];
    gather3<R, Z3>(g, A0, x, a0);
    gather3<R, Z3>(g, A1, x, a1);
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = (1.0 - theta) * a0[d] + theta * a1[d];
}

// 2-D time-interpolated gather from the node records
template <int R>
__device__ __forceinline__ void gather_interp_rec(const GridDev& g, const double* NR, double theta, const double x[3],
                                                  double out[3]) {
    if (theta == 0.0) { gather_A_rec<R>(g, NR, 0, x, out); return; }
    if (theta == 1.0) { gather_A_rec<R>(g, NR, 1, x, out); return; }
    double a0[3], a1[3];
    gather_A_rec<R>(g, NR, 0, x, a0);
    gather_A_rec<R>(g, NR, 1, x, a1);
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = (1.0 - theta) * a0[d] + theta * a1[d];
}

struct LocalMap {
    double vbar[3], P_end[3], v_end[3];
};

// evaluate_local_particle_substep (P3:4843-4884).  A0p and p_start depend on
// the substep start only and are passed in.
template <int R, bool Z3>
__device__ __forceinline__ void p3_local_map(const GridDev& g, const QuadDev& qd, const OrbitOpt& opt, const P3Args& a,
                                             const double gs[3], const double x0[3], const double P0[3],
                                             const double A0p[3], const double pst[3], double q, double m, double h,
                                             double th0, double th1, LocalMap& out) {
    const double kappa = a.kappa;
    const double dxp[3] = {h * gs[0], h * gs[1], h * gs[2]};
    const double x1[3] = {wrap1(x0[0] + dxp[0], g.L[0]), wrap1(x0[1] + dxp[1], g.L[1]), wrap1(x0[2] + dxp[2], g.L[2])};
    double A1p[3];
    double gpb[3], D[3][3];
    if constexpr (Z3) {
        gather_time_interp<R, Z3>(g, a.A_n, a.A_end, th1, x1, A1p);
        orbit_gather_grad<R, Z3, true, true>(g, qd, opt, x0, dxp, a.gpm, a.A_n, a.A_end, th0, th1, false, gpb, D);
    } else {
        gather_interp_rec<R>(g, a.NR, th1, x1, A1p);
        orbit_gg_2d<R>(g, qd, opt, x0, dxp, a.NR, th0, th1, false, gpb, D);
    }
    const double Aep[3] = {0.5 * (A0p[0] + A1p[0]), 0.5 * (A0p[1] + A1p[1]), 0.5 * (A0p[2] + A1p[2])};
    double Dv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) Dv[i] = D[i][0] * gs[0] + D[i][1] * gs[1] + D[i][2] * gs[2];
    // K = D^T - D; B_eff = (K12, -K02, K01)
    const double B[3] = {D[2][1] - D[1][2], -(D[2][0] - D[0][2]), D[1][0] - D[0][1]};
    const double hq = 0.5 * q * h;
    double pm[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double t = P0[d];
        t = t - q * Aep[d];
        t = t - hq * gpb[d];
        t = t + hq * Dv[d];
        pm[d] = t;
    }
    double pp[3];
    hc_rotate(pm, B, q, m, kappa, h, pp);
    double pe[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double t = pp[d];
        t = t + q * Aep[d];
        t = t - hq * gpb[d];
        t = t + hq * Dv[d];
        out.P_end[d] = t;
        pe[d] = t - q * A1p[d];
    }
    const double ge = gamma_p(pe, m, kappa);
    const double ive = 1.0 / (m * ge);
#pragma unroll
    for (int d = 0; d < 3; ++d) out.v_end[d] = ive * pe[d];
    // kinetic_secant_velocity (P3:682-689)
    const double gmi = gamma_p(pm, m, kappa), gst = gamma_p(pst, m, kappa), gpl = gamma_p(pp, m, kappa);
    const double c1 = 1.0 / (m * (gmi + gst)), c2 = 1.0 / (m * (ge + gpl));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double em = c1 * (pm[d] + pst[d]);
        const double ep = c2 * (pe[d] + pp[d]);
        out.vbar[d] = 0.5 * (em + ep);
    }
}

// solve_local_particle_substep (P3:4930-4983): damped Picard on vbar from
// gs over the substep [th0, th1] of length h.  Returns the failure reason
// (0: converged; vacc and acc then hold the accepted iterate and its map).
template <int R, bool Z3>
__device__ __forceinline__ int p3_solve_substep(const GridDev& g, const QuadDev& qd, const OrbitOpt& opt,
                                                const P3Args& a, double gs[3], const double xpos[3],
                                                const double Pc[3], const double A0p[3], const double pst[3],
                                                double qp, double mp, double h, double th0, double th1,
                                                LocalMap& acc, double vacc[3], int& its, double& scaled) {
    its = 0;
    scaled = 0.0;
    int reason = 5;  // maximum_iterations
    for (int it = 0; it < a.max_iter; ++it) {
        LocalMap mpd;
        p3_local_map<R, Z3>(g, qd, opt, a, gs, xpos, Pc, A0p, pst, qp, mp, h, th0, th1, mpd);
        ++its;
        if (!finite3(mpd.vbar)) { reason = 1; break; }
        const double nm = norm3(mpd.vbar);
        if (nm >= a.kappa) { reason = 2; break; }
        const double res[3] = {mpd.vbar[0] - gs[0], mpd.vbar[1] - gs[1], mpd.vbar[2] - gs[2]};
        const double scale = 0.5 * (nm + norm3(gs));
        scaled = norm3(res) / (a.atol + a.rtol * scale);
        if (scaled <= 1.0) {
            vacc[0] = gs[0]; vacc[1] = gs[1]; vacc[2] = gs[2];
            acc = mpd;
            reason = 0;
            break;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) gs[d] = gs[d] + a.damping * res[d];
        if (!finite3(gs)) { reason = 3; break; }
        if (norm3(gs) >= a.kappa) { reason = 4; break; }
    }
    return reason;
}

// Two wavefronts per SIMD: the FP64 sweep then spills part of its state to
// scratch (which stays in cache) but hides latency better than one
// wavefront holding all of it in VGPRs and AGPRs (19.5 against 24.4 ms per
// 16.8M-particle cubic sweep).
// FIX: the deterministic deposit (fixed-point words in a.fix).
template <int R, bool Z3, bool FIX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p3_map(GridDev g, QuadDev qd,
                                                                                      OrbitOpt opt, P3Args a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double tile_mem[FIX ? kTileLds : kTile * kTile * 3];
    __shared__ double park[10 * 256];  // per-lane state parked across the inner solve (256 lanes)
    DepTile T{};
    if constexpr (!Z3) T = tile_begin<R, FIX>(g, tile_mem, a.x, a.np, (int64_t)blockIdx.x * blockDim.x, a.fix);
    // red: [0] sum of local iterations, [1] failed particles; redm: [0] max
    // local iterations, [1] max final scaled residual
    double rsum[2] = {0.0, 0.0}, rmax[2] = {0.0, 0.0};
    int nseg_total = 0;
    if (p < a.np) {
        const int nsub = a.layout ? (int)a.layout[p] : 1;
        const double h = a.dt / (double)nsub;
        double xpos[3], Pc[3], vc[3], vmac[3] = {0.0, 0.0, 0.0};
        ld3(a.x, a.np, p, xpos); ld3(a.P, a.np, p, Pc); ld3(a.v, a.np, p, vc);
        const double qp = a.q[p], mp = a.m[p];
        int reason = 0;
        for (int s = 0; s < nsub && reason == 0; ++s) {
            const double th0 = (double)s / (double)nsub, th1 = (double)(s + 1) / (double)nsub;
            double A0p[3];
            if constexpr (Z3) gather_time_interp<R, Z3>(g, a.A_n, a.A_end, th0, xpos, A0p);
            else gather_interp_rec<R>(g, a.NR, th0, xpos, A0p);
            const double pst[3] = {Pc[0] - qp * A0p[0], Pc[1] - qp * A0p[1], Pc[2] - qp * A0p[2]};
            double gs[3] = {vc[0], vc[1], vc[2]};
            if (a.guess && nsub == 1) ld3(a.guess, a.np, p, gs);
            LocalMap mp_acc;
            double vacc[3];
            int its = 0;
            double scaled = 0.0;
            // v, the substep-mean vbar and the sweep statistics wait in LDS
            // while the inner solve runs (the compiler barriers keep the
            // reloads from being forwarded): 20 VGPRs fewer live through the
            // orbit passes, 272 -> 208 bytes of spill per lane, 5% on the
            // cubic sweep
            double* pk = park + threadIdx.x;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                pk[(0 + d) * 256] = vc[d];
                pk[(3 + d) * 256] = vmac[d];
            }
            pk[6 * 256] = rsum[0];
            pk[7 * 256] = rsum[1];
            pk[8 * 256] = rmax[0];
            pk[9 * 256] = rmax[1];
            asm volatile("" ::: "memory");
            reason = p3_solve_substep<R, Z3>(g, qd, opt, a, gs, xpos, Pc, A0p, pst, qp, mp, h, th0, th1, mp_acc,
                                             vacc, its, scaled);
            asm volatile("" ::: "memory");
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                vc[d] = pk[(0 + d) * 256];
                vmac[d] = pk[(3 + d) * 256];
            }
            rsum[0] = pk[6 * 256];
            rsum[1] = pk[7 * 256];
            rmax[0] = pk[8 * 256];
            rmax[1] = pk[9 * 256];
            const bool conv = reason == 0;
            rsum[0] += (double)its;
            rmax[0] = fmax(rmax[0], (double)its);
            rmax[1] = fmax(rmax[1], scaled);
            if (!conv) break;
            if (a.write_out && a.sub.x0) {
                const int64_t j = a.sub.off[p] + s;
                st3(a.sub.x0, a.sub.n, j, xpos);
                st3(a.sub.vbar, a.sub.n, j, vacc);
            }
            const double dxp[3] = {h * vacc[0], h * vacc[1], h * vacc[2]};
            if (a.J_out) {
                if constexpr (Z3) nseg_total += deposit_orbit_one<R, Z3>(g, qd, opt, xpos, dxp, vacc, qp / (double)nsub, a.J_out, a.fix);
                else nseg_total += deposit_orbit_2d<R, FIX>(g, qd, opt, xpos, dxp, vacc, qp / (double)nsub, T, a.J_out);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xpos[d] = wrap1(xpos[d] + dxp[d], g.L[d]);
                Pc[d] = mp_acc.P_end[d];
                vc[d] = mp_acc.v_end[d];
                vmac[d] = vmac[d] + (1.0 / (double)nsub) * vacc[d];
            }
        }
        if (a.reason) a.reason[p] = (uint8_t)reason;
        if (reason != 0) rsum[1] = 1.0;
        if (reason == 0 && a.write_out) {
            st3(a.x_out, a.np, p, xpos);
            st3(a.P_out, a.np, p, Pc);
            st3(a.v_out, a.np, p, vc);
            st3(a.vbar_out, a.np, p, vmac);
        }
    }
    if constexpr (!Z3) {
        if (a.J_out) tile_end<FIX>(g, T, a.J_out);
    }
    block_partials<2, 2>(rsum, rmax, nseg_total, a.part);
}

// The inner solve alone for given particles and fields (one substep of
// length h over [th0, th1]): vbar, P_end, v_end of converged particles (zero
// for failed ones, whose reason is set), iterations.
template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_p3_local_solve(GridDev g, QuadDev qd, OrbitOpt opt, P3Args a, double h,
                                                        double th0, double th1, int32_t* __restrict__ iters) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    double xpos[3], Pc[3], gs[3];
    ld3(a.x, a.np, p, xpos); ld3(a.P, a.np, p, Pc); ld3(a.guess, a.np, p, gs);
    const double qp = a.q[p], mp = a.m[p];
    double A0p[3];
    if constexpr (Z3) gather_time_interp<R, Z3>(g, a.A_n, a.A_end, th0, xpos, A0p);
    else gather_interp_rec<R>(g, a.NR, th0, xpos, A0p);
    const double pst[3] = {Pc[0] - qp * A0p[0], Pc[1] - qp * A0p[1], Pc[2] - qp * A0p[2]};
    LocalMap acc;
    double vacc[3];
    int its = 0;
    double scaled = 0.0;
    const int reason = p3_solve_substep<R, Z3>(g, qd, opt, a, gs, xpos, Pc, A0p, pst, qp, mp, h, th0, th1, acc, vacc,
                                               its, scaled);
    const double zero[3] = {0.0, 0.0, 0.0};
    st3(a.vbar_out, a.np, p, reason == 0 ? vacc : zero);
    st3(a.P_out, a.np, p, reason == 0 ? acc.P_end : zero);
    st3(a.v_out, a.np, p, reason == 0 ? acc.v_end : zero);
    iters[p] = its;
    a.reason[p] = (uint8_t)reason;
}

// ---------------------------------------------------------------------------
// P2: explicit IAEM step (P2:1050-1092), quadratic nearest-node weights
// (P2:537-580), grid origin 0, positions in [0, L).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wrap_periodic(double x, double L) {
    x -= floor(x / L) * L;
    if (x >= L) x -= L;
    if (x < 0.0) x += L;
    return x;
}
struct W1 { int i[3]; double w[3]; };
__device__ __forceinline__ W1 quad_w(double x, double dx, int n) {
    const double s = x / dx;
    const int ic = (int)floor(s + 0.5);
    const double r = s - (double)ic;
    W1 o;
    o.i[0] = ((ic - 1) % n + n) % n; o.i[1] = (ic % n + n) % n; o.i[2] = ((ic + 1) % n + n) % n;
    o.w[0] = 0.5 * (0.5 - r) * (0.5 - r); o.w[1] = 0.75 - r * r; o.w[2] = 0.5 * (0.5 + r) * (0.5 + r);
    return o;
}

__global__ __launch_bounds__(256) void k_p2_advance_deposit(P2Args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const int64_t np = a.np;
    const double vx = a.v[p], vy = a.v[np + p], vz = a.v[2 * np + p];
    const double x = wrap_periodic(a.x[p] + a.dt * vx, a.Lx);
    const double y = wrap_periodic(a.x[np + p] + a.dt * vy, a.Ly);
    a.x[p] = x;
    a.x[np + p] = y;
    const W1 wx = quad_w(x, a.dx, a.nx), wy = quad_w(y, a.dy, a.ny);
    const double q = a.q[p];
    const double s = 1.0 / a.dA;
    const int N = a.nx * a.ny;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int id = wx.i[i] * a.ny + wy.i[j];
            const double w = wx.w[i] * wy.w[j] * s;
            atomic_add(&a.J[id], q * vx * w);
            atomic_add(&a.J[N + id], q * vy * w);
            atomic_add(&a.J[2 * N + id], q * vz * w);
            if (a.rho) atomic_add(&a.rho[id], q * w);
        }
}

// Push with 11 gathered scalar fields: A (3), grad phi (2), grad A (6):
// fields layout F[k*N + id], k = 0..10 as Ax,Ay,Az, phi_x,phi_y,
// dAx/dx,dAx/dy, dAy/dx,dAy/dy, dAz/dx,dAz/dy.
__global__ __launch_bounds__(256) void k_p2_push(P2Args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const int64_t np = a.np;
    const int N = a.nx * a.ny;
    const W1 wx = quad_w(a.x[p], a.dx, a.nx), wy = quad_w(a.x[np + p], a.dy, a.ny);
    double f[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) f[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int id = wx.i[i] * a.ny + wy.i[j];
            const double w = wx.w[i] * wy.w[j];
#pragma unroll
            for (int k = 0; k < 11; ++k) f[k] += a.F[k * N + id] * w;
        }
    double v[3], vp[3], P[3];
    ld3(a.v, np, p, v); ld3(a.v_prev, np, p, vp); ld3(a.P, np, p, P);
    const double q = a.q[p], m = a.m[p], c = a.c;
    const double vs[3] = {2.0 * v[0] - vp[0], 2.0 * v[1] - vp[1], 2.0 * v[2] - vp[2]};
    const double gav[3] = {f[5] * vs[0] + f[7] * vs[1] + f[9] * vs[2], f[6] * vs[0] + f[8] * vs[1] + f[10] * vs[2], 0.0};
    const double force[3] = {(-1.0) * f[3] + gav[0], (-1.0) * f[4] + gav[1], (-1.0) * 0.0 + gav[2]};
    double Pn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) Pn[d] = P[d] + (q * a.dt) * force[d];
    const double pm[3] = {Pn[0] - q * f[0], Pn[1] - q * f[1], Pn[2] - q * f[2]};
    const double den = sqrt(c * c * (pm[0] * pm[0] + pm[1] * pm[1] + pm[2] * pm[2]) + (m * c * c) * (m * c * c));
    double vn[3] = {0.0, 0.0, 0.0};
    if (den != 0.0) {
        const double k = c * c / den;
        vn[0] = k * pm[0]; vn[1] = k * pm[1]; vn[2] = k * pm[2];
    }
    st3(a.v_prev, np, p, v);
    st3(a.v, np, p, vn);
    st3(a.P, np, p, Pn);
}

// ---------------------------------------------------------------------------
// Host-side launch dispatch over (R, Z3)
// ---------------------------------------------------------------------------
#define UPIC_DISPATCH(R_, Z3_, BODY)                       \
    do {                                                   \
        switch ((R_) * 2 + ((Z3_) ? 1 : 0)) {              \
        case 2: { constexpr int R = 1; constexpr bool Z = false; BODY; } break; \
        case 3: { constexpr int R = 1; constexpr bool Z = true; BODY; } break;  \
        case 4: { constexpr int R = 2; constexpr bool Z = false; BODY; } break; \
        case 5: { constexpr int R = 2; constexpr bool Z = true; BODY; } break;  \
        case 6: { constexpr int R = 3; constexpr bool Z = false; BODY; } break; \
        case 7: { constexpr int R = 3; constexpr bool Z = true; BODY; } break;  \
        case 8: { constexpr int R = 4; constexpr bool Z = false; BODY; } break; \
        case 9: { constexpr int R = 4; constexpr bool Z = true; BODY; } break;  \
        default: break;                                    \
        }                                                  \
    } while (0)

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_deposit_orbit(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                          const double* dx, const double* vbar, const double* qa, int64_t np, double* J,
                          unsigned long long* stats) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_deposit_orbit<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, q, o, x0, dx, vbar, qa, np, J, stats));
}
void launch_gather_orbit(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                         const double* dx, const double* F, int64_t np, double* out) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_gather_orbit<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, q, o, x0, dx, F, np, out));
}
void launch_grad_A(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                   const double* dx, const double* A0, const double* A1, double th0, double th1, int p1_theta,
                   int64_t np, double* out) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_grad_A<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, q, o, x0, dx, A0, A1, th0, th1, p1_theta, np, out));
}
void launch_gather_point(hipStream_t s, const GridDev& g, const double* pos, const double* F, int64_t np, double* out) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_gather_point<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, pos, F, np, out));
}
void launch_deposit_scalar_point(hipStream_t s, const GridDev& g, const double* pos, const double* val, int64_t np,
                                 double scale, double* mesh) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_deposit_scalar_point<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, pos, val, np, scale, mesh));
}
void launch_p1_deposit_guess(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P1Args& a) {
    if (a.np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p1_deposit_guess<R, Z>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a));
    launch_finish_partials(s, a.part, nblk(a.np), 0, 0, nullptr, nullptr, a.stats);
}
void launch_p1_push(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P1Args& a, bool final_) {
    if (a.np <= 0) return;
    if (final_) {
        UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p1_push<R, Z, true>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a));
    } else {
        UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p1_push<R, Z, false>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a));
        launch_finish_partials(s, a.part, nblk(a.np), 3, 0, a.red, nullptr, a.stats);
    }
}
void launch_p3_map(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P3Args& args) {
    if (args.np <= 0) return;
    P3Args a = args;
    a.gate = current_gate();
    if (a.fix.hi)
        UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p3_map<R, Z, true>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a));
    else
        UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p3_map<R, Z, false>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a));
    launch_finish_partials(s, a.part, nblk(a.np), 2, 2, a.red, a.redmax, a.stats);
}
void launch_p3_local_solve(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P3Args& a,
                           double h, double th0, double th1, int32_t* iters) {
    if (a.np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_p3_local_solve<R, Z>), dim3(nblk(a.np)), dim3(256), 0, s, g, q, o, a, h, th0, th1, iters));
}

// Folding of the per-block partials of a sweep (slots 4, 5, 8 are maxima,
// the others sums): the 256 threads of a workgroup fold blocks [b0, b1) in a
// fixed order; the result is valid in thread 0.
__device__ __forceinline__ void fold_partials(const double* __restrict__ part, int64_t b0, int64_t b1,
                                              double t[kPartSlots]) {
    __shared__ double sh[kPartSlots][4];
    double v[kPartSlots];
#pragma unroll
    for (int k = 0; k < kPartSlots; ++k) v[k] = 0.0;
    for (int64_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        const double* p = part + b * kPartSlots;
#pragma unroll
        for (int k = 0; k < kPartSlots; ++k) {
            const bool m = (k == 4 || k == 5 || k == 8);
            v[k] = m ? fmax(v[k], p[k]) : v[k] + p[k];
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kPartSlots; ++k) {
        const bool m = (k == 4 || k == 5 || k == 8);
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_down(x, o, 64);
            x = m ? fmax(x, y) : x + y;
        }
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kPartSlots; ++k) {
            const bool m = (k == 4 || k == 5 || k == 8);
            t[k] = sh[k][0];
            for (int w = 1; w < 4; ++w) t[k] = m ? fmax(t[k], sh[k][w]) : t[k] + sh[k][w];
        }
    }
}

// First stage for many blocks: workgroup g folds blocks [g per, (g+1) per)
// into slot g of `out`.
__global__ __launch_bounds__(256) void k_fold_partials(const double* __restrict__ part, int64_t nb, int64_t per,
                                                       double* __restrict__ out, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t b0 = (int64_t)blockIdx.x * per;
    double t[kPartSlots];
    fold_partials(part, b0, b0 + per < nb ? b0 + per : nb, t);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kPartSlots; ++k) out[(int64_t)blockIdx.x * kPartSlots + k] = t[k];
}

// Last stage: one workgroup folds the (first-stage) partials and updates
// the accumulators.
__global__ __launch_bounds__(256) void k_finish_partials(const double* __restrict__ part, int64_t nb, int ns, int nm,
                                                         double* __restrict__ sum, double* __restrict__ mx,
                                                         unsigned long long* __restrict__ stats,
                                                         const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    double t[kPartSlots];
    fold_partials(part, 0, nb, t);
    if (threadIdx.x == 0) {
        for (int k = 0; k < ns && sum; ++k) sum[k] += t[k];
        for (int k = 0; k < nm && mx; ++k) mx[k] = fmax(mx[k], t[4 + k]);
        if (stats) {
            stats[0] += (unsigned long long)t[6];
            stats[1] += (unsigned long long)t[7];
            stats[2] = max(stats[2], (unsigned long long)t[8]);
        }
    }
}
// Partials of nblocks workgroups at part[0, nblocks); above kFoldMin blocks a
// first stage folds them into at most kFoldBlocks slots placed right after
// them (the buffer holds partial_blocks(np) slots, kFoldBlocks of them spare).
void launch_finish_partials(hipStream_t s, const double* part, int64_t nblocks, int ns, int nm, double* sum,
                            double* mx, unsigned long long* stats) {
    if (nblocks <= 0) return;
    constexpr int64_t kFoldMin = 2048;
    if (nblocks <= kFoldMin) {
        hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(256), 0, s, part, nblocks, ns, nm, sum, mx, stats,
                           current_gate());
        return;
    }
    const int64_t per = (nblocks + kFoldBlocks - 1) / kFoldBlocks;
    const int64_t g = (nblocks + per - 1) / per;
    double* out = const_cast<double*>(part) + nblocks * kPartSlots;
    hipLaunchKernelGGL(k_fold_partials, dim3((unsigned)g), dim3(256), 0, s, part, nblocks, per, out, current_gate());
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(256), 0, s, out, g, ns, nm, sum, mx, stats, current_gate());
}
// Step diagnostics (compute_step_diagnostics, P1:1384-1439 / P3:4501-4570)
__global__ __launch_bounds__(256) void k_step_diag_orbit(int64_t np, int midpoint, double dt, double Lx, double Ly,
                                                         double Lz, const double* __restrict__ x0,
                                                         const double* __restrict__ v0, const double* __restrict__ v1,
                                                         const double* __restrict__ vbl, double* __restrict__ vbar,
                                                         double* __restrict__ dx, double* __restrict__ x1) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double L[3] = {Lx, Ly, Lz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int64_t i = d * np + p;
        const double vb = midpoint ? 0.5 * (v1[i] + v0[i]) : vbl[i];
        const double dd = dt * vb;
        vbar[i] = vb;
        dx[i] = dd;
        x1[i] = wrap1(x0[i] + dd, L[d]);
    }
}

// R = (A^{n+1}(x1) - A^n(x0)) - dt (Ubar + D_A vbar) per particle
// (compute_RA_orbit_residual, P3:4242-4296) and the particle sums:
// part: sum |R|^2, sum |A_delta|^2, sum |rhs|^2, sum |R| and max |R|;
// part2: sum -q vbar.R (work defect), sum q vbar.Ebar (particle work / dt)
__global__ __launch_bounds__(256) void k_step_diag_particles(int64_t np, double dt, const double* __restrict__ q,
                                                             const double* __restrict__ vbar,
                                                             const double* __restrict__ Ubar,
                                                             const double* __restrict__ D,
                                                             const double* __restrict__ A0p,
                                                             const double* __restrict__ A1p,
                                                             const double* __restrict__ Ebar,
                                                             double* __restrict__ part, double* __restrict__ part2) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s4[4] = {0.0, 0.0, 0.0, 0.0}, mx[1] = {0.0}, w2[2] = {0.0, 0.0};
    if (p < np) {
        double vb[3], R[3], Ad[3], rhs[3], Eb[3];
        ld3(vbar, np, p, vb);
        ld3(Ebar, np, p, Eb);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Ad[a] = A1p[a * np + p] - A0p[a * np + p];
            const double DAv = D[(3 * a) * np + p] * vb[0] + D[(3 * a + 1) * np + p] * vb[1] +
                               D[(3 * a + 2) * np + p] * vb[2];
            rhs[a] = dt * (Ubar[a * np + p] + DAv);
            R[a] = Ad[a] - rhs[a];
        }
        const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double nr = sqrt(r2);
        s4[0] = r2;
        s4[1] = Ad[0] * Ad[0] + Ad[1] * Ad[1] + Ad[2] * Ad[2];
        s4[2] = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
        s4[3] = nr;
        mx[0] = nr;
        const double qp = q[p];
        w2[0] = -qp * (vb[0] * R[0] + vb[1] * R[1] + vb[2] * R[2]);
        w2[1] = qp * (vb[0] * Eb[0] + vb[1] * Eb[1] + vb[2] * Eb[2]);
    }
    block_partials<4, 1>(s4, mx, 0, part);
    block_partials<2, 0>(w2, nullptr, 0, part2);
}

// Segments of the knot-split orbit x0 -> x0 + dxp (the entries of
// path_breakpoints_for_particle minus one, P3:2555-2606).
__device__ __forceinline__ int orbit_segments(const GridDev& g, const OrbitOpt& opt, const double x0[3],
                                              const double dxp[3]) {
    if (!opt.split) return 1;
    BreakGen bg;
    breakgen_init(bg, g, x0, dxp, true);
    int n = 1;
    while (breakgen_next(bg) < 1.0) ++n;
    return n;
}

// compute_subcycled_step_diagnostics (P3:4363-4473): one work-item per
// particle walks its recorded substeps in order.  For substep s of nsub
// (h = dt / nsub, theta in [s, s + 1] / nsub): R = A(theta1, x1) -
// A(theta0, x0) - h (Ubar + D_A(theta0, theta1) vbar), the 1/nsub-weighted
// norms of R, A_delta and rhs, the work defect -q vbar.R, the particle work
// h q vbar.Ebar and the orbit's segments.
template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_step_diag_subcycled(GridDev g, QuadDev qd, OrbitOpt opt, int64_t np,
                                                             double dt, SubPaths sp, const double* __restrict__ qa,
                                                             const double* __restrict__ A0,
                                                             const double* __restrict__ A1,
                                                             const double* __restrict__ Umid,
                                                             const double* __restrict__ Emid,
                                                             double* __restrict__ part, double* __restrict__ part2) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s4[4] = {0.0, 0.0, 0.0, 0.0}, mx[1] = {0.0}, w3[3] = {0.0, 0.0, 0.0};
    double segs = 0.0, crossed = 0.0;
    if (p < np) {
        const int64_t b = sp.off[p], e = sp.off[p + 1];
        const int nsub = (int)(e - b);
        const double h = dt / (double)nsub, w = 1.0 / (double)nsub;
        const double qp = qa[p];
        for (int s = 0; s < nsub; ++s) {
            double x0[3], vb[3];
            ld3(sp.x0, sp.n, b + s, x0);
            ld3(sp.vbar, sp.n, b + s, vb);
            const double th0 = (double)s / (double)nsub, th1 = (double)(s + 1) / (double)nsub;
            const double dxp[3] = {h * vb[0], h * vb[1], h * vb[2]};
            const double x1[3] = {wrap1(x0[0] + dxp[0], g.L[0]), wrap1(x0[1] + dxp[1], g.L[1]),
                                  wrap1(x0[2] + dxp[2], g.L[2])};
            double A0p[3], A1p[3], G[3], DA[3][3], Ub[3], Eb[3], Dm[3][3];
            gather_time_interp<R, Z3>(g, A0, A1, th0, x0, A0p);
            gather_time_interp<R, Z3>(g, A0, A1, th1, x1, A1p);
            orbit_gather_grad<R, Z3, false, true>(g, qd, opt, x0, dxp, nullptr, A0, A1, th0, th1, false, G, DA);
            orbit_gather_grad<R, Z3, true, false>(g, qd, opt, x0, dxp, Umid, nullptr, nullptr, 0.0, 1.0, false, Ub, Dm);
            orbit_gather_grad<R, Z3, true, false>(g, qd, opt, x0, dxp, Emid, nullptr, nullptr, 0.0, 1.0, false, Eb, Dm);
            double Rv[3], Ad[3], rhs[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                Ad[a] = A1p[a] - A0p[a];
                rhs[a] = h * (Ub[a] + (DA[a][0] * vb[0] + DA[a][1] * vb[1] + DA[a][2] * vb[2]));
                Rv[a] = Ad[a] - rhs[a];
            }
            const double r2 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            const double nr = sqrt(r2);
            s4[0] += w * r2;
            s4[1] += w * (Ad[0] * Ad[0] + Ad[1] * Ad[1] + Ad[2] * Ad[2]);
            s4[2] += w * (rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2]);
            s4[3] += w * nr;
            mx[0] = fmax(mx[0], nr);
            w3[0] += -qp * (vb[0] * Rv[0] + vb[1] * Rv[1] + vb[2] * Rv[2]);
            w3[1] += h * qp * (vb[0] * Eb[0] + vb[1] * Eb[1] + vb[2] * Eb[2]);
            const int nseg = orbit_segments(g, opt, x0, dxp);
            segs += (double)nseg;
            w3[2] += (double)(nseg - 1);
            if (nseg > 1) crossed = 1.0;
        }
    }
    block_partials_stats<4, 1>(s4, mx, crossed, segs, p < np ? segs : 1.0, part);
    block_partials<3, 0>(w3, nullptr, 0, part2);
}

void launch_step_diag_subcycled(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, int64_t np,
                                double dt, const SubPaths& sp, const double* qa, const double* A0, const double* A1,
                                const double* Umid, const double* Emid, double* part, double* part2, double* sums,
                                double* max1, unsigned long long* stats) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_step_diag_subcycled<R, Z>), dim3(nblk(np)), dim3(256), 0, s,
                                                       g, q, o, np, dt, sp, qa, A0, A1, Umid, Emid, part, part2));
    launch_finish_partials(s, part, nblk(np), 4, 1, sums, max1, stats);
    launch_finish_partials(s, part2, nblk(np), 3, 0, sums + 4, nullptr, nullptr);
}

void launch_step_diag_orbit(hipStream_t s, int64_t np, int midpoint, double dt, const double L[3], const double* x0,
                            const double* v0, const double* v1, const double* vbar_last, double* vbar, double* dx,
                            double* x1) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_step_diag_orbit, dim3(nblk(np)), dim3(256), 0, s, np, midpoint, dt, L[0], L[1], L[2], x0, v0,
                       v1, vbar_last, vbar, dx, x1);
}
void launch_step_diag_particles(hipStream_t s, int64_t np, double dt, const double* q, const double* vbar,
                                const double* Ubar, const double* D, const double* A0p, const double* A1p,
                                const double* Ebar, double* part, double* part2, double* sums6, double* max1) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_step_diag_particles, dim3(nblk(np)), dim3(256), 0, s, np, dt, q, vbar, Ubar, D, A0p, A1p,
                       Ebar, part, part2);
    launch_finish_partials(s, part, nblk(np), 4, 1, sums6, max1, nullptr);
    launch_finish_partials(s, part2, nblk(np), 2, 0, sums6 + 4, nullptr, nullptr);
}

void launch_p2_advance_deposit(hipStream_t s, const P2Args& a) {
    if (a.np <= 0) return;
    hipLaunchKernelGGL(k_p2_advance_deposit, dim3(nblk(a.np)), dim3(256), 0, s, a);
}
void launch_p2_push(hipStream_t s, const P2Args& a) {
    if (a.np <= 0) return;
    hipLaunchKernelGGL(k_p2_push, dim3(nblk(a.np)), dim3(256), 0, s, a);
}

}  // namespace upic

namespace upic {
template <int R, bool Z3>
__global__ __launch_bounds__(256) void k_init_momenta(GridDev g, int64_t np, int code, double kappa, double dt,
                                                      const double* __restrict__ x, const double* __restrict__ v,
                                                      const double* __restrict__ q, const double* __restrict__ m,
                                                      const double* __restrict__ E, const double* __restrict__ B,
                                                      const double* __restrict__ A, double* __restrict__ P,
                                                      double* __restrict__ v_prev, double* __restrict__ vbar_last) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double X[3], V[3], Ep[3];
    ld3(x, np, p, X); ld3(v, np, p, V);
    gather3<R, Z3>(g, E, X, Ep);
    const double qp = q[p], mp = m[p];
    double Pn[3], Vp[3];
    if (code == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Pn[d] = mp * V[d];
            Vp[d] = V[d] - dt * ((qp / mp) * Ep[d]);
        }
    } else {
        double Bp[3], Ap[3];
        gather3<R, Z3>(g, B, X, Bp);
        gather3<R, Z3>(g, A, X, Ap);
        const double beta2 = (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) / (kappa * kappa);
        const double gam = 1.0 / sqrt(fmax(1.0e-300, 1.0 - beta2));
        double pn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            pn[d] = (mp * gam) * V[d];
            Pn[d] = pn[d] + qp * Ap[d];
        }
        const double vxB[3] = {V[1] * Bp[2] - V[2] * Bp[1], V[2] * Bp[0] - V[0] * Bp[2], V[0] * Bp[1] - V[1] * Bp[0]};
        double pp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) pp[d] = pn[d] - dt * (qp * (Ep[d] + vxB[d]));
        const double gp = gamma_p(pp, mp, kappa);
#pragma unroll
        for (int d = 0; d < 3; ++d) Vp[d] = (1.0 / (mp * gp)) * pp[d];
    }
    st3(P, np, p, Pn);
    st3(v_prev, np, p, Vp);
    st3(vbar_last, np, p, V);
}
void launch_init_momenta(hipStream_t s, const GridDev& g, int64_t np, int code, double kappa, double dt,
                         const double* x, const double* v, const double* q, const double* m, const double* E,
                         const double* B, const double* A, double* P, double* v_prev, double* vbar_last) {
    if (np <= 0) return;
    UPIC_DISPATCH(g.order, g.nz > 1, hipLaunchKernelGGL((k_init_momenta<R, Z>), dim3(nblk(np)), dim3(256), 0, s, g, np, code, kappa, dt, x, v, q, m, E, B, A, P, v_prev, vbar_last));
}
}  // namespace upic

namespace upic {
__global__ void k_fix_to_double(int64_t n, unsigned long long* __restrict__ hi, unsigned long long* __restrict__ lo,
                                double inv_scale, double* __restrict__ out, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long h = (long long)hi[i], l = (long long)lo[i];
    out[i] = ((double)h + (double)l * (1.0 / kFixLoScale)) * inv_scale;
    hi[i] = 0ull;
    lo[i] = 0ull;
}
void launch_fix_to_double(hipStream_t s, int64_t n, const FixAcc& fx, double* out) {
    if (n > 0)
        hipLaunchKernelGGL(k_fix_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, fx.hi, fx.lo,
                           1.0 / fx.scale, out, current_gate());
}
__global__ void k_pack_nodes(int N, const double* __restrict__ gpm, const double* __restrict__ A0,
                             const double* __restrict__ A1, double* __restrict__ NR, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    double2* o = reinterpret_cast<double2*>(NR + (size_t)8 * id);
    o[0] = make_double2(gpm[id], gpm[N + id]);
    o[1] = make_double2(A0[id], A0[N + id]);
    o[2] = make_double2(A0[2 * N + id], A1[id]);
    o[3] = make_double2(A1[N + id], A1[2 * N + id]);
}
void launch_pack_nodes(hipStream_t s, int N, const double* gpm, const double* A_n, const double* A_end, double* NR) {
    hipLaunchKernelGGL(k_pack_nodes, dim3((N + 255) / 256), dim3(256), 0, s, N, gpm, A_n, A_end, NR, current_gate());
}
}  // namespace upic
