// orbit2d.hpp -- the 2-D (nz = 1) fast path of the orbit integrals.
//
// Mesh data a particle push reads at a node are packed into one 64-byte
// record (grad phi_mid x, y; A^n x, y, z; A^{n+1} x, y, z) so a node costs
// four 16-byte loads.  Gauss points are processed in chunks of kChunk: while
// the stencil base is the same for the whole chunk (always, between knots),
// the node loop is outermost and each record is loaded once per chunk, with
// the quadrature folded into five node weights
//   W = sum w S, Gx0 = sum w (1-theta) dS/dx, Gx1 = sum w theta dS/dx, (y alike)
// so that gather += W grad_phi and D[c][.] += A0_c G.0 + A1_c G.1.
// This is the same sum as the reference's point-by-point loops
// (path_average_gather_and_gradient, P3:3016-3071; orbit_discrete_gradient_A,
// P1:908-934) reordered; results agree to roundoff.
//
// The current deposit accumulates per segment in registers and flushes into a
// workgroup LDS tile (8 x 8 nodes x 3 components, FP64 LDS atomics) when the
// node falls inside it, else straight to HBM with FP64 atomics; the tile is
// added to the mesh once per workgroup.  With cell-sorted particles a
// workgroup's particles sit in one or two cells, so nearly every deposit
// lands in LDS.  The deterministic mode keeps the same path with exact
// fixed-point words (FixAcc) in place of the FP64 atomics.
#pragma once

#include "particles.hpp"

namespace upic {

constexpr int kTile = 8;
// Gauss points per chunk of the FP64 2-D orbit pass: their stencil weights
// are live together (2 points: 32 doubles at cubic order; 4 points measured
// slower in the two-wavefront sweep)
constexpr int kChunk = 2;

struct NodeRec {  // 8 doubles
    double gx, gy, a0x, a0y, a0z, a1x, a1y, a1z;
};

__device__ __forceinline__ void load_rec(const double* __restrict__ NR, int id, double r[8]) {
    const double2* p = reinterpret_cast<const double2*>(NR + (size_t)8 * id);
    const double2 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
    r[0] = v0.x; r[1] = v0.y; r[2] = v1.x; r[3] = v1.y; r[4] = v2.x; r[5] = v2.y; r[6] = v3.x; r[7] = v3.y;
}

// gather of A^{n+1} (which = 1) or A^n (which = 0) at a point from the records
template <int R>
__device__ __forceinline__ void gather_A_rec(const GridDev& g, const double* __restrict__ NR, int which, const double x[3],
                                             double out[3]) {
    AxisSt<R> sx, sy;
    axis_stencil<R>(wrap1(x[0], g.L[0]), g.nx, g.L[0], g.h[0], sx);
    axis_stencil<R>(wrap1(x[1], g.L[1]), g.ny, g.L[1], g.h[1], sy);
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;
    const int off = which ? 5 : 2;
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap_idx(sx.base + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap_idx(sy.base + b, g.ny);
            const double* rp = NR + (size_t)8 * (ia * g.ny + ib) + off;
            const double w = sx.w[a] * sy.w[b];
            o0 += rp[0] * w;
            o1 += rp[1] * w;
            o2 += rp[2] * w;
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2;
}

// Fused orbit gather of grad phi_mid and orbit gradient of A, 2-D.
// D[c][0..1] = d A_c / d(x, y); the z column is identically zero in 2-D.
template <int R>
__device__ __forceinline__ int orbit_gg_2d(const GridDev& g, const QuadDev& q, const OrbitOpt& opt, const double x0[3],
                                           const double dxp[3], const double* __restrict__ NR, double th0, double th1,
                                           bool p1_theta, double gat[3], double D[3][3]) {
    double g0 = 0.0, g1 = 0.0;
    double d00 = 0.0, d01 = 0.0, d10 = 0.0, d11 = 0.0, d20 = 0.0, d21 = 0.0;
    BreakGen bg;
    breakgen_init(bg, g, x0, dxp, opt.split != 0);
    double sa = 0.0;
    int nseg = 0;
    for (;;) {
        const double sb = opt.split ? breakgen_next(bg) : 1.0;
        const double len = sb - sa;
        if (len > 0.0) {
            for (int q0 = 0; q0 < q.n; q0 += kChunk) {
                double wx[kChunk][R + 1], dwx[kChunk][R + 1], wy[kChunk][R + 1], dwy[kChunk][R + 1], wq[kChunk],
                    th[kChunk];
                int bx[kChunk], by[kChunk];
#pragma unroll
                for (int k = 0; k < kChunk; ++k) {
                    const int iq = q0 + k < q.n ? q0 + k : q0;
                    const double s = sa + len * q.s[iq];
                    wq[k] = (q0 + k < q.n) ? len * q.w[iq] : 0.0;
                    double xq0 = x0[0] + s * dxp[0], xq1 = x0[1] + s * dxp[1];
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
                    th[k] = p1_theta ? s : th0 + (th1 - th0) * s;
                }
                bool same = true;
#pragma unroll
                for (int k = 1; k < kChunk; ++k) same = same && bx[k] == bx[0] && by[k] == by[0];
                if (same) {
#pragma unroll
                    for (int a = 0; a <= R; ++a) {
                        const int ia = wrap_idx(bx[0] + a, g.nx);
                        // x factors of this stencil row's node weights
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
                                Gx0 += du[k] * wy[k][b];   // w dS/dx
                                Gx1 += tdu[k] * wy[k][b];
                                Gy0 += u[k] * dwy[k][b];   // w dS/dy
                                Gy1 += tu[k] * dwy[k][b];
                            }
                            Gx0 -= Gx1;  // sum w (1 - theta) dS/dx
                            Gy0 -= Gy1;
                            double r[8];
                            load_rec(NR, ia * g.ny + ib, r);
                            g0 += W * r[0];
                            g1 += W * r[1];
                            d00 += r[2] * Gx0 + r[5] * Gx1;
                            d01 += r[2] * Gy0 + r[5] * Gy1;
                            d10 += r[3] * Gx0 + r[6] * Gx1;
                            d11 += r[3] * Gy0 + r[6] * Gy1;
                            d20 += r[4] * Gx0 + r[7] * Gx1;
                            d21 += r[4] * Gy0 + r[7] * Gy1;
                        }
                    }
                } else {
                    // the stencil changes inside the chunk (rare): point by
                    // point, recomputing each point's stencil in a rolled loop
#pragma unroll 1
                    for (int k = 0; k < kChunk && q0 + k < q.n; ++k) {
                        const double s = sa + len * q.s[q0 + k];
                        const double wk = len * q.w[q0 + k];
                        const double thk = p1_theta ? s : th0 + (th1 - th0) * s;
                        double xq0 = x0[0] + s * dxp[0], xq1 = x0[1] + s * dxp[1];
                        if (opt.wrap_path) {
                            xq0 = wrap1(xq0, g.L[0]);
                            xq1 = wrap1(xq1, g.L[1]);
                        }
                        AxisSt<R> sx, sy;
                        axis_stencil<R>(wrap1(xq0, g.L[0]), g.nx, g.L[0], g.h[0], sx);
                        axis_stencil<R>(wrap1(xq1, g.L[1]), g.ny, g.L[1], g.h[1], sy);
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
                                const double Gx1 = thk * sx_, Gx0 = sx_ - Gx1;
                                const double Gy1 = thk * sy_, Gy0 = sy_ - Gy1;
                                double r[8];
                                load_rec(NR, ia * g.ny + ib, r);
                                g0 += W * r[0];
                                g1 += W * r[1];
                                d00 += r[2] * Gx0 + r[5] * Gx1;
                                d01 += r[2] * Gy0 + r[5] * Gy1;
                                d10 += r[3] * Gx0 + r[6] * Gx1;
                                d11 += r[3] * Gy0 + r[6] * Gy1;
                                d20 += r[4] * Gx0 + r[7] * Gx1;
                                d21 += r[4] * Gy0 + r[7] * Gy1;
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
    gat[0] = g0; gat[1] = g1; gat[2] = 0.0;
    D[0][0] = d00; D[0][1] = d01; D[0][2] = 0.0;
    D[1][0] = d10; D[1][1] = d11; D[1][2] = 0.0;
    D[2][0] = d20; D[2][1] = d21; D[2][2] = 0.0;
    return nseg;
}

// Workgroup LDS tile for the current deposit (2-D).  In the deterministic
// mode (fx.hi set) the tile holds the two fixed-point words of FixAcc per
// entry and the deposit goes to fx instead of J.
struct DepTile {
    double* t;   // kTile*kTile*3 doubles in LDS (2 x that many words in the fixed-point mode)
    int ox, oy;  // wrapped mesh index of tile node (0,0)
    FixAcc fx;
};
// LDS doubles a deposit tile needs in either mode
constexpr int kTileLds = 2 * kTile * kTile * 3;

template <bool FIX = false>
__device__ __forceinline__ void tile_add(const GridDev& g, const DepTile& T, double* __restrict__ J, int ia, int ib,
                                         double j0, double j1, double j2) {
    int dx = ia - T.ox;
    if (dx < 0) dx += g.nx;
    int dy = ib - T.oy;
    if (dy < 0) dy += g.ny;
    const bool in = dx < kTile && dy < kTile;
    if constexpr (FIX) {
        if (in) {
            unsigned long long* c = reinterpret_cast<unsigned long long*>(T.t) + 3 * (dx * kTile + dy);
            constexpr int lo = kTile * kTile * 3;
            fix_add(c, c + lo, j0, T.fx.scale);
            fix_add(c + 1, c + 1 + lo, j1, T.fx.scale);
            fix_add(c + 2, c + 2 + lo, j2, T.fx.scale);
        } else {
            const int id = ia * g.ny + ib;
            fix_add(T.fx.hi + id, T.fx.lo + id, j0, T.fx.scale);
            fix_add(T.fx.hi + g.N + id, T.fx.lo + g.N + id, j1, T.fx.scale);
            fix_add(T.fx.hi + 2 * g.N + id, T.fx.lo + 2 * g.N + id, j2, T.fx.scale);
        }
        return;
    }
    if (in) {
        double* c = T.t + 3 * (dx * kTile + dy);
        atomic_add(c, j0);
        atomic_add(c + 1, j1);
        atomic_add(c + 2, j2);
    } else {
        const int id = ia * g.ny + ib;
        atomic_add(&J[id], j0);
        atomic_add(&J[g.N + id], j1);
        atomic_add(&J[2 * g.N + id], j2);
    }
}

// Orbit current deposit of one particle into the tile (2-D); FIX: the
// tile holds fixed-point words.
template <int R, bool FIX = false>
__device__ __forceinline__ int deposit_orbit_2d(const GridDev& g, const QuadDev& q, const OrbitOpt& opt,
                                                const double x0[3], const double dxp[3], const double vel[3],
                                                double charge, const DepTile& T, double* __restrict__ J) {
    double W[R + 1][R + 1];
    int bx = -1, by = -1;
    bool have = false;
    const double qs = charge / g.dV;
    auto flush = [&]() {
        if (!have) return;
#pragma unroll
        for (int a = 0; a <= R; ++a) {
            const int ia = wrap_idx(bx + a, g.nx);
#pragma unroll
            for (int b = 0; b <= R; ++b) {
                const int ib = wrap_idx(by + b, g.ny);
                const double w = W[a][b] * qs;
                tile_add<FIX>(g, T, J, ia, ib, w * vel[0], w * vel[1], w * vel[2]);
            }
        }
        have = false;
    };
    const int nseg = for_each_orbit_point(g, q, opt, x0, dxp, [&](double, double w, const double* xq) {
        AxisSt<R> sx, sy;
        axis_stencil<R>(wrap1(xq[0], g.L[0]), g.nx, g.L[0], g.h[0], sx);
        axis_stencil<R>(wrap1(xq[1], g.L[1]), g.ny, g.L[1], g.h[1], sy);
        if (!(have && sx.base == bx && sy.base == by)) {
            flush();
            bx = sx.base;
            by = sy.base;
#pragma unroll
            for (int a = 0; a <= R; ++a)
#pragma unroll
                for (int b = 0; b <= R; ++b) W[a][b] = 0.0;
            have = true;
        }
#pragma unroll
        for (int a = 0; a <= R; ++a) {
            const double wa = w * sx.w[a];
#pragma unroll
            for (int b = 0; b <= R; ++b) W[a][b] += wa * sy.w[b];
        }
    });
    flush();
    return nseg;
}

// Tile set-up: origin two nodes below the stencil base of the workgroup's
// first particle; zeroed; must be called by every thread of the workgroup.
// With fx.hi set the tile takes fixed-point words (lds holds kTileLds).
template <int R, bool FIX = false>
__device__ __forceinline__ DepTile tile_begin(const GridDev& g, double* lds, const double* __restrict__ x, int64_t np,
                                              int64_t p_first, FixAcc fx = FixAcc{}) {
    __shared__ int org[2];
    if (threadIdx.x == 0) {
        int ox = 0, oy = 0;
        if (p_first < np) {
            AxisSt<R> sx, sy;
            axis_stencil<R>(wrap1(x[p_first], g.L[0]), g.nx, g.L[0], g.h[0], sx);
            axis_stencil<R>(wrap1(x[np + p_first], g.L[1]), g.ny, g.L[1], g.h[1], sy);
            ox = sx.base - 2;
            oy = sy.base - 2;
            ox = ((ox % g.nx) + g.nx) % g.nx;
            oy = ((oy % g.ny) + g.ny) % g.ny;
        }
        org[0] = ox;
        org[1] = oy;
    }
    constexpr int nz = FIX ? kTileLds : kTile * kTile * 3;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) lds[i] = 0.0;
    __syncthreads();
    return DepTile{lds, org[0], org[1], fx};
}

template <bool FIX = false>
__device__ __forceinline__ void tile_end(const GridDev& g, const DepTile& T, double* __restrict__ J) {
    __syncthreads();
    if constexpr (FIX) {
        const unsigned long long* w = reinterpret_cast<const unsigned long long*>(T.t);
        constexpr int lo = kTile * kTile * 3;
        for (int i = threadIdx.x; i < lo; i += blockDim.x) {
            const unsigned long long h = w[i], l = w[lo + i];
            if (h == 0 && l == 0) continue;
            const int comp = i % 3, cell = i / 3;
            const int dx = cell / kTile, dy = cell % kTile;
            if (dx >= g.nx || dy >= g.ny) continue;
            const int ia = (T.ox + dx) % g.nx, ib = (T.oy + dy) % g.ny;
            const size_t id = (size_t)comp * g.N + ia * g.ny + ib;
            if (h) atomicAdd(T.fx.hi + id, h);
            if (l) atomicAdd(T.fx.lo + id, l);
        }
        return;
    }
    for (int i = threadIdx.x; i < kTile * kTile * 3; i += blockDim.x) {
        const double v = T.t[i];
        if (v != 0.0) {
            const int comp = i % 3, cell = i / 3;
            const int dx = cell / kTile, dy = cell % kTile;
            if (dx >= g.nx || dy >= g.ny) continue;
            const int ia = (T.ox + dx) % g.nx, ib = (T.oy + dy) % g.ny;
            atomic_add(&J[(size_t)comp * g.N + ia * g.ny + ib], v);
        }
    }
}

}  // namespace upic
