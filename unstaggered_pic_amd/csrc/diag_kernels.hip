// diag_kernels.hip -- the fused 2-D pass of the per-step diagnostics
// (compute_step_diagnostics, Project1/ec_pic_serial_demo.cpp:1384-1439,
// Project3/ec_pic_nyquist_projected.cpp:4501-4570) for gfx950.
#include "kernels.hpp"
#include "particles.hpp"
#include "orbit2d.hpp"

namespace upic {

namespace {
__device__ __forceinline__ void ld3(const double* __restrict__ a, int64_t np, int64_t p, double o[3]) {
    o[0] = a[p]; o[1] = a[np + p]; o[2] = a[2 * np + p];
}
inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }
}  // namespace

// Node record of the fused 2-D step diagnostics: U_mid (3), E_mid (3), A^n
// (3), A^{n+1} (3) -- six 16-byte loads per node.
__global__ void k_pack_diag_nodes(int N, const double* __restrict__ U, const double* __restrict__ E,
                                  const double* __restrict__ A0, const double* __restrict__ A1,
                                  double* __restrict__ NR) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    double2* o = reinterpret_cast<double2*>(NR + (size_t)12 * id);
    o[0] = make_double2(U[id], U[N + id]);
    o[1] = make_double2(U[2 * N + id], E[id]);
    o[2] = make_double2(E[N + id], E[2 * N + id]);
    o[3] = make_double2(A0[id], A0[N + id]);
    o[4] = make_double2(A0[2 * N + id], A1[id]);
    o[5] = make_double2(A1[N + id], A1[2 * N + id]);
}

// point gather of the three record entries from offset `off` (A^n: 6,
// A^{n+1}: 9)
template <int R>
__device__ __forceinline__ void gather_rec12(const GridDev& g, const double* __restrict__ NR, int off, const double x[3],
                                             double out[3]) {
    AxisSt<R> sx, sy;
    axis_stencil<R>(wrap1(x[0], g.L[0]), g.nx, g.L[0], g.h[0], sx);
    axis_stencil<R>(wrap1(x[1], g.L[1]), g.ny, g.L[1], g.h[1], sy);
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap_idx(sx.base + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap_idx(sy.base + b, g.ny);
            const double* rp = NR + (size_t)12 * (ia * g.ny + ib) + off;
            const double w = sx.w[a] * sy.w[b];
            o0 += rp[0] * w;
            o1 += rp[1] * w;
            o2 += rp[2] * w;
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2;
}

// Orbit sums of one particle in the fused diagnostics pass: orbit averages
// of U_mid and E_mid, the x/y columns of D_A, and the deposit weights of the
// current stencil base (flushed into the tile when the base changes).
template <int R>
struct DiagAcc {
    double u0, u1, u2, e0, e1, e2;
    double d00, d01, d10, d11, d20, d21;
    double Wd[R + 1][R + 1];
    int dbx, dby;
    bool have;
};

template <int R>
__device__ __forceinline__ void diag_flush(const GridDev& g, const DepTile& T, double* __restrict__ Jbar,
                                           DiagAcc<R>& A, double j0, double j1, double j2) {
    if (!A.have) return;
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap_idx(A.dbx + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap_idx(A.dby + b, g.ny);
            const double w = A.Wd[a][b];
            tile_add(g, T, Jbar, ia, ib, w * j0, w * j1, w * j2);
        }
    }
    A.have = false;
}

template <int R>
__device__ __forceinline__ void diag_start(const GridDev& g, const DepTile& T, double* __restrict__ Jbar,
                                           DiagAcc<R>& A, int bx, int by, double j0, double j1, double j2) {
    if (A.have && bx == A.dbx && by == A.dby) return;
    diag_flush<R>(g, T, Jbar, A, j0, j1, j2);
    A.dbx = bx;
    A.dby = by;
#pragma unroll
    for (int a = 0; a <= R; ++a)
#pragma unroll
        for (int b = 0; b <= R; ++b) A.Wd[a][b] = 0.0;
    A.have = true;
}

// one node of one fold: value weight W, gradient weights G.0 ((1 - theta)
// part) and G.1 (theta part); record = U_mid | E_mid | A^n | A^{n+1}
template <int R>
__device__ __forceinline__ void diag_node(const double* __restrict__ rec, DiagAcc<R>& A, double& wd, double W,
                                          double Gx0, double Gx1, double Gy0, double Gy1) {
    const double2* rp = reinterpret_cast<const double2*>(rec);
    const double2 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3], r4 = rp[4], r5 = rp[5];
    A.u0 += W * r0.x; A.u1 += W * r0.y; A.u2 += W * r1.x;
    A.e0 += W * r1.y; A.e1 += W * r2.x; A.e2 += W * r2.y;
    // A^n = (r3.x, r3.y, r4.x), A^{n+1} = (r4.y, r5.x, r5.y)
    A.d00 += r3.x * Gx0 + r4.y * Gx1;
    A.d01 += r3.x * Gy0 + r4.y * Gy1;
    A.d10 += r3.y * Gx0 + r5.x * Gx1;
    A.d11 += r3.y * Gy0 + r5.x * Gy1;
    A.d20 += r4.x * Gx0 + r5.y * Gx1;
    A.d21 += r4.x * Gy0 + r5.y * Gy1;
    wd += W;
}

// compute_step_diagnostics (P1:1384-1439, P3:4501-4570) on a 2-D mesh in one
// orbit pass per particle: the orbit averages of U_mid and E_mid, the orbit
// gradient D_A over theta in [0, 1] (compute_RA_orbit_residual,
// P3:4242-4296), the point values A^n(x0) and A^{n+1}(x1), the orbit current
// deposit of q vbar into Jbar (through the workgroup LDS tile) and the
// particle sums of k_step_diag_particles.  The five orbit sums share their
// stencil weights, folded per Gauss-point pair as in orbit_gg_2d; the
// results are the separate kernels' sums reordered (roundoff).
// One wavefront per SIMD holding the unrolled node loop in VGPRs and AGPRs
// (10.6 ms per 16.8M-particle cubic pass) measured faster than two with
// spills (12.3 ms).
template <int R>
__global__ __launch_bounds__(256) void k_step_diag_2d(GridDev g, QuadDev qd, OrbitOpt opt, int64_t np, int midpoint,
                                                      double dt, const double* __restrict__ x0,
                                                      const double* __restrict__ v0, const double* __restrict__ v1,
                                                      const double* __restrict__ vbl, const double* __restrict__ qa,
                                                      const double* __restrict__ NR, double* __restrict__ Jbar,
                                                      double* __restrict__ part, double* __restrict__ part2) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double tile_mem[kTile * kTile * 3];
    const DepTile T = tile_begin<R>(g, tile_mem, x0, np, (int64_t)blockIdx.x * blockDim.x);
    double s4[4] = {0.0, 0.0, 0.0, 0.0}, mx[1] = {0.0}, w2[2] = {0.0, 0.0};
    int nseg = 0;
    if (p < np) {
        double X[3], vb[3];
        ld3(x0, np, p, X);
        if (midpoint) {
#pragma unroll
            for (int d = 0; d < 3; ++d) vb[d] = 0.5 * (v1[d * np + p] + v0[d * np + p]);
        } else {
            ld3(vbl, np, p, vb);
        }
        const double dxp[3] = {dt * vb[0], dt * vb[1], dt * vb[2]};
        const double X1[3] = {wrap1(X[0] + dxp[0], g.L[0]), wrap1(X[1] + dxp[1], g.L[1]), wrap1(X[2] + dxp[2], g.L[2])};
        double A0p[3], A1p[3];
        gather_rec12<R>(g, NR, 6, X, A0p);
        gather_rec12<R>(g, NR, 9, X1, A1p);
        const double qp = qa[p];
        const double qs = qp / g.dV;
        const double j0 = qs * vb[0], j1 = qs * vb[1], j2 = qs * vb[2];
        DiagAcc<R> A;
        A.u0 = A.u1 = A.u2 = A.e0 = A.e1 = A.e2 = 0.0;
        A.d00 = A.d01 = A.d10 = A.d11 = A.d20 = A.d21 = 0.0;
        A.dbx = A.dby = 0;
        A.have = false;
        BreakGen bg;
        breakgen_init(bg, g, X, dxp, opt.split != 0);
        double sa = 0.0;
        for (;;) {
            const double sb = opt.split ? breakgen_next(bg) : 1.0;
            const double len = sb - sa;
            if (len > 0.0) {
                for (int q0 = 0; q0 < qd.n; q0 += kChunk) {
                    double wx[kChunk][R + 1], dwx[kChunk][R + 1], wy[kChunk][R + 1], dwy[kChunk][R + 1], wq[kChunk],
                        th[kChunk];
                    int bx[kChunk], by[kChunk];
#pragma unroll
                    for (int k = 0; k < kChunk; ++k) {
                        const int iq = q0 + k < qd.n ? q0 + k : q0;
                        const double s = sa + len * qd.s[iq];
                        wq[k] = (q0 + k < qd.n) ? len * qd.w[iq] : 0.0;
                        double xq0 = X[0] + s * dxp[0], xq1 = X[1] + s * dxp[1];
                        if (opt.wrap_path) {
                            xq0 = wrap1(xq0, g.L[0]);
                            xq1 = wrap1(xq1, g.L[1]);
                        }
                        AxisSt<R> sx, sy;
                        axis_stencil<R>(wrap1(xq0, g.L[0]), g.nx, g.L[0], g.h[0], sx);
                        axis_stencil<R>(wrap1(xq1, g.L[1]), g.ny, g.L[1], g.h[1], sy);
                        bx[k] = sx.base;
                        by[k] = sy.base;
#pragma unroll
                        for (int a = 0; a <= R; ++a) {
                            wx[k][a] = sx.w[a]; dwx[k][a] = sx.dw[a];
                            wy[k][a] = sy.w[a]; dwy[k][a] = sy.dw[a];
                        }
                        th[k] = s;
                    }
                    bool same = true;
#pragma unroll
                    for (int k = 1; k < kChunk; ++k) same = same && bx[k] == bx[0] && by[k] == by[0];
                    if (same) {
                        diag_start<R>(g, T, Jbar, A, bx[0], by[0], j0, j1, j2);
#pragma unroll
                        for (int a = 0; a <= R; ++a) {
                            const int ia = wrap_idx(bx[0] + a, g.nx);
                            double u[kChunk], du[kChunk], tdu[kChunk], tu[kChunk];
#pragma unroll
                            for (int k = 0; k < kChunk; ++k) {
                                u[k] = wq[k] * wx[k][a];
                                du[k] = wq[k] * dwx[k][a];
                                tu[k] = th[k] * u[k];
                                tdu[k] = th[k] * du[k];
                            }
#pragma unroll
                            for (int b = 0; b <= R; ++b) {
                                const int ib = wrap_idx(by[0] + b, g.ny);
                                double W = 0.0, Gx0 = 0.0, Gx1 = 0.0, Gy0 = 0.0, Gy1 = 0.0;
#pragma unroll
                                for (int k = 0; k < kChunk; ++k) {
                                    W += u[k] * wy[k][b];
                                    Gx0 += du[k] * wy[k][b];
                                    Gx1 += tdu[k] * wy[k][b];
                                    Gy0 += u[k] * dwy[k][b];
                                    Gy1 += tu[k] * dwy[k][b];
                                }
                                diag_node<R>(NR + (size_t)12 * (ia * g.ny + ib), A, A.Wd[a][b], W, Gx0 - Gx1, Gx1,
                                             Gy0 - Gy1, Gy1);
                            }
                        }
                    } else {
                        // the stencil changes inside the fold (rare): point by point
#pragma unroll 1
                        for (int k = 0; k < kChunk && q0 + k < qd.n; ++k) {
                            const double s = sa + len * qd.s[q0 + k];
                            const double wk = len * qd.w[q0 + k];
                            double xq0 = X[0] + s * dxp[0], xq1 = X[1] + s * dxp[1];
                            if (opt.wrap_path) {
                                xq0 = wrap1(xq0, g.L[0]);
                                xq1 = wrap1(xq1, g.L[1]);
                            }
                            AxisSt<R> sx, sy;
                            axis_stencil<R>(wrap1(xq0, g.L[0]), g.nx, g.L[0], g.h[0], sx);
                            axis_stencil<R>(wrap1(xq1, g.L[1]), g.ny, g.L[1], g.h[1], sy);
                            diag_start<R>(g, T, Jbar, A, sx.base, sy.base, j0, j1, j2);
#pragma unroll
                            for (int a = 0; a <= R; ++a) {
                                const int ia = wrap_idx(sx.base + a, g.nx);
#pragma unroll
                                for (int b = 0; b <= R; ++b) {
                                    const int ib = wrap_idx(sy.base + b, g.ny);
                                    const double ws = wk * sy.w[b];
                                    const double W = sx.w[a] * ws;
                                    const double sx_ = sx.dw[a] * ws;
                                    const double sy_ = sx.w[a] * wk * sy.dw[b];
                                    const double Gx1 = s * sx_, Gy1 = s * sy_;
                                    diag_node<R>(NR + (size_t)12 * (ia * g.ny + ib), A, A.Wd[a][b], W, sx_ - Gx1, Gx1,
                                                 sy_ - Gy1, Gy1);
                                }
                            }
                        }
                    }
                }
            }
            ++nseg;
            sa = sb;
            if (sb >= 1.0) break;
        }
        diag_flush<R>(g, T, Jbar, A, j0, j1, j2);
        // R = (A^{n+1}(x1) - A^n(x0)) - dt (Ubar + D_A vbar); the z column of D_A is zero in 2-D
        const double Ub[3] = {A.u0, A.u1, A.u2}, Eb[3] = {A.e0, A.e1, A.e2};
        const double DAv[3] = {A.d00 * vb[0] + A.d01 * vb[1], A.d10 * vb[0] + A.d11 * vb[1],
                               A.d20 * vb[0] + A.d21 * vb[1]};
        double Rv[3], Ad[3], rhs[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Ad[a] = A1p[a] - A0p[a];
            rhs[a] = dt * (Ub[a] + DAv[a]);
            Rv[a] = Ad[a] - rhs[a];
        }
        const double r2 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        const double nr = sqrt(r2);
        s4[0] = r2;
        s4[1] = Ad[0] * Ad[0] + Ad[1] * Ad[1] + Ad[2] * Ad[2];
        s4[2] = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
        s4[3] = nr;
        mx[0] = nr;
        w2[0] = -qp * (vb[0] * Rv[0] + vb[1] * Rv[1] + vb[2] * Rv[2]);
        w2[1] = qp * (vb[0] * Eb[0] + vb[1] * Eb[1] + vb[2] * Eb[2]);
    }
    tile_end(g, T, Jbar);
    block_partials<4, 1>(s4, mx, nseg, part);
    block_partials<2, 0>(w2, nullptr, 0, part2);
}

void launch_step_diag_2d(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, int64_t np,
                         int midpoint, double dt, const double* x0, const double* v0, const double* v1,
                         const double* vbar_last, const double* qa, const double* Umid, const double* Emid,
                         const double* A0, const double* A1, double* NR12, double* Jbar, double* part, double* part2,
                         double* sums6, double* max1, unsigned long long* stats) {
    const int N = g.N;
    hipLaunchKernelGGL(k_pack_diag_nodes, dim3((N + 255) / 256), dim3(256), 0, s, N, Umid, Emid, A0, A1, NR12);
    if (np <= 0) return;
    switch (g.order) {
        case 1: hipLaunchKernelGGL(k_step_diag_2d<1>, dim3(nblk(np)), dim3(256), 0, s, g, q, o, np, midpoint, dt, x0, v0, v1, vbar_last, qa, NR12, Jbar, part, part2); break;
        case 2: hipLaunchKernelGGL(k_step_diag_2d<2>, dim3(nblk(np)), dim3(256), 0, s, g, q, o, np, midpoint, dt, x0, v0, v1, vbar_last, qa, NR12, Jbar, part, part2); break;
        case 3: hipLaunchKernelGGL(k_step_diag_2d<3>, dim3(nblk(np)), dim3(256), 0, s, g, q, o, np, midpoint, dt, x0, v0, v1, vbar_last, qa, NR12, Jbar, part, part2); break;
        case 4: hipLaunchKernelGGL(k_step_diag_2d<4>, dim3(nblk(np)), dim3(256), 0, s, g, q, o, np, midpoint, dt, x0, v0, v1, vbar_last, qa, NR12, Jbar, part, part2); break;
        default: break;
    }
    launch_finish_partials(s, part, nblk(np), 4, 1, sums6, max1, stats);
    launch_finish_partials(s, part2, nblk(np), 2, 0, sums6 + 4, nullptr, nullptr);
}

}  // namespace upic
