// mixed.hip -- the mixed-precision particle path of the P3 step (2-D, nz = 1).
//
// Particle state is held as an int32 cell index c = i*ny + j plus in-cell
// offsets t = (u - i, v - j, w) in cell units (u = (x + Lx/2)/hx; w is the z
// position in units of Lz, the single z cell), momenta P, velocities v and
// orbit-averaged velocities in FP32.  The orbit integrals, the GM-HC map and
// the deposit run in FP32 on cell-relative coordinates; node records are an
// FP32 copy of the FP64 mesh fields; the deposit accumulates in an FP32 LDS
// tile that is flushed into the FP64 mesh current with FP64 atomics.  The
// inner convergence test, every reduction and the field solve stay in FP64.
//
// The arithmetic follows the FP64 path (particle_kernels.hip, orbit2d.hpp):
//   evaluate_local_particle_substep   P3:4843-4884
//   solve_local_particle_substep      P3:4930-4983
//   path_average_gather_and_gradient  P3:3016-3071
//   path_average_deposit_current      P3:2784-2857
//   path_breakpoints_for_particle     P3:2555-2606
// with positions relative to the particle's start cell, so that FP32 keeps
// ~1e-7 of a cell of resolution anywhere on the grid.
#include <algorithm>

#include "kernels.hpp"
#include "mlp.hpp"
#include "particles.hpp"

namespace upic {

namespace {

__device__ __forceinline__ int wrapn(int i, int n) {
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    return i;
}
__device__ __forceinline__ int wrap1n(int i, int n) { return i >= n ? i - n : i; }

// Centred B-spline of degree R and its derivative (P3:1856-1947), FP32.
template <int R>
__device__ __forceinline__ void bsp_f(float r, float& w, float& dw) {
    if constexpr (R == 1) {
        if (r >= -1.0f && r < 0.0f) { w = 1.0f + r; dw = 1.0f; }
        else if (r >= 0.0f && r < 1.0f) { w = 1.0f - r; dw = -1.0f; }
        else { w = 0.0f; dw = 0.0f; }
    } else {
        const float a = fabsf(r);
        const float sg = (r < 0.0f) ? -1.0f : 1.0f;
        if constexpr (R == 2) {
            if (a < 0.5f) { w = 0.75f - a * a; dw = -2.0f * r; }
            else if (a < 1.5f) { const float z = 1.5f - a; w = 0.5f * z * z; dw = -sg * z; }
            else { w = 0.0f; dw = 0.0f; }
        } else if constexpr (R == 3) {
            if (a < 1.0f) { w = (2.0f / 3.0f) - a * a + 0.5f * a * a * a; dw = sg * (-2.0f * a + 1.5f * a * a); }
            else if (a < 2.0f) { const float z = 2.0f - a; w = z * z * z * (1.0f / 6.0f); dw = -0.5f * sg * z * z; }
            else { w = 0.0f; dw = 0.0f; }
        } else {
            if (a < 0.5f) {
                const float a2 = a * a;
                w = 115.0f / 192.0f - (5.0f / 8.0f) * a2 + 0.25f * a2 * a2;
                dw = sg * (-(5.0f / 4.0f) * a + a * a * a);
            } else if (a < 1.5f) {
                w = 55.0f / 96.0f + (5.0f / 24.0f) * a - 1.25f * a * a + (5.0f / 6.0f) * a * a * a -
                    (1.0f / 6.0f) * a * a * a * a;
                dw = sg * (5.0f / 24.0f - 2.5f * a + 2.5f * a * a - (2.0f / 3.0f) * a * a * a);
            } else if (a < 2.5f) {
                const float z = 2.5f - a;
                const float z2 = z * z;
                w = z2 * z2 * (1.0f / 24.0f);
                dw = -sg * z * z * z * (1.0f / 6.0f);
            } else { w = 0.0f; dw = 0.0f; }
        }
    }
}

// One axis of the stencil at cell-relative coordinate t: nodes at cell
// offsets fl + a, fl = floor(t - (R-1)/2) (axis_stencil, P3:1971-1986).
template <int R>
struct AxF {
    int fl;
    float w[R + 1], dw[R + 1];  // dw already divided by h
};
// One weight is the complement of the others and its derivative weight
// minus their sum: the FP32 closed forms round with a bias (2/3 rounds up,
// so the cubic weights sum to 1 + 4e-8 on average), and a partition of unity
// off by a bias scales every gather and deposit coherently, which shows up as
// an energy drift; the complement leaves only unbiased rounding.
//
// Quadratic and cubic take the weights as polynomials of the in-cell
// fraction u (the distance past the stencil's centre knot, in [0, 1)): the
// same B-spline pieces bsp_f selects by |r|, without evaluating and selecting
// every piece per node; the complement goes to a middle weight, the largest.
// Other degrees evaluate bsp_f per node, with the last weight the complement.
template <int R>
__device__ __forceinline__ void axis_f(float t, float ih, AxF<R>& s) {
    const float f = floorf(t - 0.5f * (float)(R - 1));
    s.fl = (int)f;
    if constexpr (R == 3) {
        // nodes at r = 1+u, u, u-1, u-2 from the particle
        const float u = (t - f) - 1.0f, v = 1.0f - u;
        const float u2 = u * u, v2 = v * v;
        const float w0 = v2 * v * (1.0f / 6.0f), w3 = u2 * u * (1.0f / 6.0f);
        const float w1 = (2.0f / 3.0f) - u2 + 0.5f * u2 * u;
        const float d0 = -0.5f * v2, d1 = u * (1.5f * u - 2.0f), d3 = 0.5f * u2;
        s.w[0] = w0; s.w[1] = w1; s.w[2] = 1.0f - (w0 + w1 + w3); s.w[3] = w3;
        s.dw[0] = d0 * ih; s.dw[1] = d1 * ih; s.dw[2] = -(d0 + d1 + d3) * ih; s.dw[3] = d3 * ih;
    } else if constexpr (R == 2) {
        // nodes at r = u+1/2, u-1/2, u-3/2
        const float u = (t - f) - 0.5f, v = 1.0f - u;
        const float w0 = 0.5f * v * v, w2 = 0.5f * u * u;
        s.w[0] = w0; s.w[1] = 1.0f - (w0 + w2); s.w[2] = w2;
        s.dw[0] = -v * ih; s.dw[1] = (v - u) * ih; s.dw[2] = u * ih;
    } else {
        const float r0 = t - f;
        float sw = 0.0f, sdw = 0.0f;
#pragma unroll
        for (int a = 0; a < R; ++a) {
            float w, dw;
            bsp_f<R>(r0 - (float)a, w, dw);
            s.w[a] = w;
            s.dw[a] = dw;
            sw += w;
            sdw += dw;
        }
        s.w[R] = 1.0f - sw;
        s.dw[R] = -sdw;
#pragma unroll
        for (int a = 0; a <= R; ++a) s.dw[a] *= ih;
    }
}

// Knot crossings of the two in-plane axes in increasing s (the merge of
// path_breakpoints_for_particle, P3:2555-2606).  The FP32 filter and
// de-duplication widths are 1e-6: a cut that close to another or to an end
// changes the orbit integral by far less than FP32 resolution.
struct CrossF {
    float a0, du;
    int m, m_end, step;
};
struct BreakF {
    CrossF ax[2];
    float last;
};
__device__ __forceinline__ void breakf_init(BreakF& bg, int order, float t0x, float t0y, float dux, float duy) {
    bg.last = 0.0f;
    const float off = 0.5f * (float)(order - 1);
    const float t0[2] = {t0x, t0y}, du[2] = {dux, duy};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        CrossF& c = bg.ax[d];
        c.a0 = t0[d] - off;
        c.du = du[d];
        c.m = 0; c.m_end = 0; c.step = 1;
        if (du[d] == 0.0f) continue;
        const float a1 = c.a0 + du[d];
        const float lo = fminf(c.a0, a1), hi = fmaxf(c.a0, a1);
        const int first = (int)floorf(lo) + 1, last = (int)ceilf(hi) - 1;
        if (last < first) continue;
        if (du[d] > 0.0f) { c.m = first; c.m_end = last + 1; c.step = 1; }
        else { c.m = last; c.m_end = first - 1; c.step = -1; }
    }
}
__device__ __forceinline__ float breakf_next(BreakF& bg) {
    const float tol = 1.0e-6f;
    for (;;) {
        const float s0 = (bg.ax[0].m == bg.ax[0].m_end) ? 2.0f : ((float)bg.ax[0].m - bg.ax[0].a0) / bg.ax[0].du;
        const float s1 = (bg.ax[1].m == bg.ax[1].m_end) ? 2.0f : ((float)bg.ax[1].m - bg.ax[1].a0) / bg.ax[1].du;
        const bool first = s0 <= s1;
        const float s = first ? s0 : s1;
        if (s > 1.5f) { bg.last = 1.0f; return 1.0f; }
        if (first) bg.ax[0].m += bg.ax[0].step;
        else bg.ax[1].m += bg.ax[1].step;
        if (!(s > tol && s < 1.0f - tol)) continue;
        if (fabsf(s - bg.last) > tol) { bg.last = s; return s; }
    }
}

// Scheduling fence between stencil rows: keeps the compiler from hoisting
// every node record load of the chunk at once (register pressure).
#ifndef UPIC_MIXED_SCHED
#define UPIC_MIXED_SCHED 1
#endif
#if UPIC_MIXED_SCHED
#define UPIC_MIXED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define UPIC_MIXED_FENCE() ((void)0)
#endif

// FP32 node record: (gx, gy, A0x, A0y) (A0z, A1x, A1y, A1z)
__device__ __forceinline__ void rec_f(const float4* __restrict__ NR, int id, float4& r0, float4& r1) {
    r0 = NR[2 * id];
    r1 = NR[2 * id + 1];
}

struct MGrid {
    int nx, ny, N, order;
    float ihx, ihy;  // 1 / h
    float hx, hy;
};

// A^n (which = 0) or A^{n+1} (which = 1) at cell-relative (tx, ty) of cell (ci, cj)
template <int R>
__device__ __forceinline__ void gatherA_f(const MGrid& g, const float4* __restrict__ NR, int ci, int cj, float tx,
                                          float ty, int which, float out[3]) {
    AxF<R> sx, sy;
    axis_f<R>(tx, g.ihx, sx);
    axis_f<R>(ty, g.ihy, sy);
    const int bx = wrapn(ci + sx.fl, g.nx), by = wrapn(cj + sy.fl, g.ny);
    float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap1n(bx + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap1n(by + b, g.ny);
            float4 r0, r1;
            rec_f(NR, ia * g.ny + ib, r0, r1);
            const float w = sx.w[a] * sy.w[b];
            if (which) {
                o0 += r1.y * w; o1 += r1.z * w; o2 += r1.w * w;
            } else {
                o0 += r0.z * w; o1 += r0.w * w; o2 += r1.x * w;
            }
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2;
}

template <int R>
__device__ __forceinline__ void gather_interp_f(const MGrid& g, const float4* NR, int ci, int cj, float tx, float ty,
                                                float theta, float out[3]) {
    if (theta == 0.0f) { gatherA_f<R>(g, NR, ci, cj, tx, ty, 0, out); return; }
    if (theta == 1.0f) { gatherA_f<R>(g, NR, ci, cj, tx, ty, 1, out); return; }
    float a0[3], a1[3];
    gatherA_f<R>(g, NR, ci, cj, tx, ty, 0, a0);
    gatherA_f<R>(g, NR, ci, cj, tx, ty, 1, a1);
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = (1.0f - theta) * a0[d] + theta * a1[d];
}

// Fused orbit gather of grad phi_mid and the orbit gradient of A over the
// path t(s) = t0 + s du (cell units), theta(s) = th0 + (th1 - th0) s.
// Gauss points in chunks of four share one node loop while their stencil
// base agrees (the quadrature folded into node weights, as in the FP64 path).
// FOLD (single substep, th1 = 1): the node loop of the last segment also
// gathers A^{n+1} at the end point t0 + du when the end point's stencil is
// the chunk's (A1, a1_ok set); otherwise the caller gathers it separately.
template <int R, bool FOLD = false>
__device__ __forceinline__ int orbit_gg_f(const MGrid& g, const QuadF& q, bool split, const float4* __restrict__ NR,
                                          int ci, int cj, float t0x, float t0y, float dux, float duy, float th0,
                                          float th1, float gat[2], float D[3][2], float* A1 = nullptr,
                                          bool* a1_ok = nullptr) {
    float g0 = 0.0f, g1 = 0.0f, d00 = 0.0f, d01 = 0.0f, d10 = 0.0f, d11 = 0.0f, d20 = 0.0f, d21 = 0.0f;
    float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
    bool got = false;
    AxF<R> ex, ey;
    if constexpr (FOLD) {
        axis_f<R>(t0x + dux, g.ihx, ex);
        axis_f<R>(t0y + duy, g.ihy, ey);
    }
    BreakF bg;
    breakf_init(bg, R, t0x, t0y, split ? dux : 0.0f, split ? duy : 0.0f);
    float sa = 0.0f;
    int nseg = 0;
    for (;;) {
        const float sb = breakf_next(bg);
        const float len = sb - sa;
        if (len > 0.0f) {
            for (int q0 = 0; q0 < q.n; q0 += 4) {
                float wx[4][R + 1], dwx[4][R + 1], wy[4][R + 1], dwy[4][R + 1], wq[4], th[4];
                int bx[4], by[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int iq = q0 + k < q.n ? q0 + k : q0;
                    const float s = sa + len * q.s[iq];
                    wq[k] = (q0 + k < q.n) ? len * q.w[iq] : 0.0f;
                    AxF<R> sx, sy;
                    axis_f<R>(t0x + s * dux, g.ihx, sx);
                    axis_f<R>(t0y + s * duy, g.ihy, sy);
                    bx[k] = sx.fl;
                    by[k] = sy.fl;
#pragma unroll
                    for (int a = 0; a <= R; ++a) {
                        wx[k][a] = sx.w[a]; dwx[k][a] = sx.dw[a];
                        wy[k][a] = sy.w[a]; dwy[k][a] = sy.dw[a];
                    }
                    th[k] = th0 + (th1 - th0) * s;
                }
                const bool same = bx[1] == bx[0] && bx[2] == bx[0] && bx[3] == bx[0] && by[1] == by[0] &&
                                  by[2] == by[0] && by[3] == by[0];
                if (same) {
                    const int ox = wrapn(ci + bx[0], g.nx), oy = wrapn(cj + by[0], g.ny);
                    bool take = false;
                    if constexpr (FOLD) {
                        take = sb >= 1.0f && q0 == 0 && ex.fl == bx[0] && ey.fl == by[0];
                        got = got || take;
                    }
#pragma unroll
                    for (int a = 0; a <= R; ++a) {
                        const int ia = wrap1n(ox + a, g.nx);
                        const float xe = FOLD && take ? ex.w[a] : 0.0f;
                        // x factors of the node weights of this stencil row:
                        // u = wq wx, du = wq dwx and their theta-weighted forms
                        float u[4], du[4], tu[4], tdu[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            u[k] = wq[k] * wx[k][a];
                            du[k] = wq[k] * dwx[k][a];
                            tu[k] = th[k] * u[k];
                            tdu[k] = th[k] * du[k];
                        }
#pragma unroll
                        for (int b = 0; b <= R; ++b) {
                            const int ib = wrap1n(oy + b, g.ny);
                            // W = sum_k u wy, S = sum_k (du wy, u dwy), G1 = theta-weighted S
                            float W = u[0] * wy[0][b], Sx = du[0] * wy[0][b], Gx1 = tdu[0] * wy[0][b];
                            float Sy = u[0] * dwy[0][b], Gy1 = tu[0] * dwy[0][b];
#pragma unroll
                            for (int k = 1; k < 4; ++k) {
                                W += u[k] * wy[k][b];
                                Sx += du[k] * wy[k][b];
                                Gx1 += tdu[k] * wy[k][b];
                                Sy += u[k] * dwy[k][b];
                                Gy1 += tu[k] * dwy[k][b];
                            }
                            const float Gx0 = Sx - Gx1, Gy0 = Sy - Gy1;
                            float4 r0, r1;
                            rec_f(NR, ia * g.ny + ib, r0, r1);
                            g0 += W * r0.x;
                            g1 += W * r0.y;
                            d00 += r0.z * Gx0 + r1.y * Gx1;
                            d01 += r0.z * Gy0 + r1.y * Gy1;
                            d10 += r0.w * Gx0 + r1.z * Gx1;
                            d11 += r0.w * Gy0 + r1.z * Gy1;
                            d20 += r1.x * Gx0 + r1.w * Gx1;
                            d21 += r1.x * Gy0 + r1.w * Gy1;
                            if constexpr (FOLD) {
                                const float we = xe * ey.w[b];
                                e0 += we * r1.y;
                                e1 += we * r1.z;
                                e2 += we * r1.w;
                            }
                        }
                        UPIC_MIXED_FENCE();
                    }
                } else {
                    // the stencil changes inside the chunk (rare): point by
                    // point, recomputing each point's stencil in a rolled loop
#pragma unroll 1
                    for (int k = 0; k < 4 && q0 + k < q.n; ++k) {
                        const float s = sa + len * q.s[q0 + k];
                        const float wk = len * q.w[q0 + k];
                        const float thk = th0 + (th1 - th0) * s;
                        AxF<R> sx, sy;
                        axis_f<R>(t0x + s * dux, g.ihx, sx);
                        axis_f<R>(t0y + s * duy, g.ihy, sy);
                        const int ox = wrapn(ci + sx.fl, g.nx), oy = wrapn(cj + sy.fl, g.ny);
#pragma unroll
                        for (int a = 0; a <= R; ++a) {
                            const int ia = wrap1n(ox + a, g.nx);
#pragma unroll
                            for (int b = 0; b <= R; ++b) {
                                const int ib = wrap1n(oy + b, g.ny);
                                const float W = wk * sx.w[a] * sy.w[b];
                                const float sx_ = wk * sx.dw[a] * sy.w[b];
                                const float sy_ = wk * sx.w[a] * sy.dw[b];
                                const float Gx1 = thk * sx_, Gx0 = sx_ - Gx1;
                                const float Gy1 = thk * sy_, Gy0 = sy_ - Gy1;
                                float4 r0, r1;
                                rec_f(NR, ia * g.ny + ib, r0, r1);
                                g0 += W * r0.x;
                                g1 += W * r0.y;
                                d00 += r0.z * Gx0 + r1.y * Gx1;
                                d01 += r0.z * Gy0 + r1.y * Gy1;
                                d10 += r0.w * Gx0 + r1.z * Gx1;
                                d11 += r0.w * Gy0 + r1.z * Gy1;
                                d20 += r1.x * Gx0 + r1.w * Gx1;
                                d21 += r1.x * Gy0 + r1.w * Gy1;
                            }
                        }
                    }
                }
            }
        }
        ++nseg;
        sa = sb;
        if (sb >= 1.0f) break;
    }
    gat[0] = g0; gat[1] = g1;
    if constexpr (FOLD) {
        A1[0] = e0; A1[1] = e1; A1[2] = e2;
        *a1_ok = got;
    }
    D[0][0] = d00; D[0][1] = d01;
    D[1][0] = d10; D[1][1] = d11;
    D[2][0] = d20; D[2][1] = d21;
    return nseg;
}

// FP32 LDS deposit tile of kTileX x kTileY nodes x 3 components; origin
// three cells below the workgroup's first particle cell.  Long in y: the
// particles of a workgroup are in cell order (cell = i ny + j), so a
// workgroup whose particles are spread over several cells (few particles per
// cell, or the group of knot-crossing orbits) covers a run of cells along y.
constexpr int kTileX = 8, kTileY = 32;
// The tile's shape (tx x ty nodes, tx ty = kTileX kTileY) is a launch
// parameter of the wave deposit, which the deposit autotuner varies; an
// empty tile (tx = 0) sends everything to the mesh.
struct DepTileF {
    float* t;
    int ox, oy;
    int tx, ty;
};

__device__ __forceinline__ void tilef_add(const MGrid& g, const DepTileF& T, double* __restrict__ J, int ia, int ib,
                                          float j0, float j1, float j2) {
    int dx = ia - T.ox;
    if (dx < 0) dx += g.nx;
    int dy = ib - T.oy;
    if (dy < 0) dy += g.ny;
    if (dx < T.tx && dy < T.ty) {
        float* c = T.t + 3 * (dx * T.ty + dy);
        atomicAdd(c, j0);
        atomicAdd(c + 1, j1);
        atomicAdd(c + 2, j2);
    } else {
        const int id = ia * g.ny + ib;
        unsafeAtomicAdd(&J[id], (double)j0);
        unsafeAtomicAdd(&J[g.N + id], (double)j1);
        unsafeAtomicAdd(&J[2 * g.N + id], (double)j2);
    }
}

// one component of tilef_add
__device__ __forceinline__ void tilef_add1(const MGrid& g, const DepTileF& T, double* __restrict__ J, int ia, int ib,
                                           int comp, float v) {
    int dx = ia - T.ox;
    if (dx < 0) dx += g.nx;
    int dy = ib - T.oy;
    if (dy < 0) dy += g.ny;
    if (dx < T.tx && dy < T.ty) atomicAdd(T.t + 3 * (dx * T.ty + dy) + comp, v);
    else unsafeAtomicAdd(&J[(size_t)comp * g.N + ia * g.ny + ib], (double)v);
}

// Orbit current deposit of one particle (weights per stencil window kept in
// registers, flushed when the window moves).
template <int R>
__device__ __forceinline__ int deposit_orbit_f(const MGrid& g, const QuadF& q, bool split, int ci, int cj, float t0x,
                                               float t0y, float dux, float duy, const float vel[3], float qs,
                                               const DepTileF& T, double* __restrict__ J) {
    float W[R + 1][R + 1];
    int bx = 0, by = 0;
    bool have = false;
    auto flush = [&]() {
        if (!have) return;
        const int ox = wrapn(ci + bx, g.nx), oy = wrapn(cj + by, g.ny);
#pragma unroll
        for (int a = 0; a <= R; ++a) {
            const int ia = wrap1n(ox + a, g.nx);
#pragma unroll
            for (int b = 0; b <= R; ++b) {
                const int ib = wrap1n(oy + b, g.ny);
                const float w = W[a][b] * qs;
                tilef_add(g, T, J, ia, ib, w * vel[0], w * vel[1], w * vel[2]);
            }
        }
        have = false;
    };
    BreakF bg;
    breakf_init(bg, R, t0x, t0y, split ? dux : 0.0f, split ? duy : 0.0f);
    float sa = 0.0f;
    int nseg = 0;
    for (;;) {
        const float sb = breakf_next(bg);
        const float len = sb - sa;
        if (len > 0.0f) {
            for (int iq = 0; iq < q.n; ++iq) {
                const float s = sa + len * q.s[iq];
                const float w = len * q.w[iq];
                AxF<R> sx, sy;
                axis_f<R>(t0x + s * dux, g.ihx, sx);
                axis_f<R>(t0y + s * duy, g.ihy, sy);
                if (!(have && sx.fl == bx && sy.fl == by)) {
                    flush();
                    bx = sx.fl;
                    by = sy.fl;
#pragma unroll
                    for (int a = 0; a <= R; ++a)
#pragma unroll
                        for (int b = 0; b <= R; ++b) W[a][b] = 0.0f;
                    have = true;
                }
#pragma unroll
                for (int a = 0; a <= R; ++a) {
                    const float wa = w * sx.w[a];
#pragma unroll
                    for (int b = 0; b <= R; ++b) W[a][b] += wa * sy.w[b];
                }
            }
        }
        ++nseg;
        sa = sb;
        if (sb >= 1.0f) break;
    }
    flush();
    return nseg;
}

// Tile set-up: origin tx/2 - 1 cells before the workgroup's first particle
// cell in x and 3 before it in y (the particles of a workgroup run forward
// along y); zeroed; must be called by every thread of the workgroup.
__device__ __forceinline__ DepTileF tilef_begin(const MGrid& g, float* lds, const int* __restrict__ cell, int64_t np,
                                                int64_t p_first, int tx = kTileX, int ty = kTileY) {
    __shared__ int org[2];
    if (threadIdx.x == 0) {
        int ox = 0, oy = 0;
        if (p_first < np) {
            const int c = cell[p_first];
            ox = wrapn(c / g.ny - (tx / 2 - 1), g.nx);
            oy = wrapn(c % g.ny - 3, g.ny);
        }
        org[0] = ox;
        org[1] = oy;
    }
    for (int i = threadIdx.x; i < kTileX * kTileY * 3; i += blockDim.x) lds[i] = 0.0f;
    __syncthreads();
    return DepTileF{lds, org[0], org[1], tx, ty};
}

__device__ __forceinline__ void tilef_end(const MGrid& g, const DepTileF& T, double* __restrict__ J) {
    __syncthreads();
    for (int i = threadIdx.x; i < T.tx * T.ty * 3; i += blockDim.x) {
        const float v = T.t[i];
        if (v != 0.0f) {
            const int comp = i % 3, cell = i / 3;
            const int dx = cell / T.ty, dy = cell % T.ty;
            if (dx >= g.nx || dy >= g.ny) continue;
            const int ia = (T.ox + dx) % g.nx, ib = (T.oy + dy) % g.ny;
            unsafeAtomicAdd(&J[(size_t)comp * g.N + ia * g.ny + ib], (double)v);
        }
    }
}

__device__ __forceinline__ float gamma_f(const float p[3], float mk) {
    const float imk = 1.0f / mk;
    const float u0 = p[0] * imk, u1 = p[1] * imk, u2 = p[2] * imk;
    return sqrtf(1.0f + (u0 * u0 + u1 * u1 + u2 * u2));
}

// higuera_cary_rotate_with_K (P3:4636-4655), FP32.  The rotation is linear
// and homogeneous in u = p/(m kappa), so it is applied to p directly; 1/(m
// kappa) enters only the dimensionless gamma terms.  Rotating u and scaling
// back by m kappa multiplies every momentum by m kappa * fl(1/(m kappa)),
// a factor 1 +- 6e-8 that is the same for all particles of one mass: a
// coherent shrink or growth of the kinetic energy every step.
__device__ __forceinline__ void hc_rotate_f(const float pm[3], const float B[3], float q, float m, float kappa, float dt,
                                            float out[3]) {
    const float imk = 1.0f / (m * kappa);
    const float u2 = (pm[0] * pm[0] + pm[1] * pm[1] + pm[2] * pm[2]) * (imk * imk);
    const float gm2 = 1.0f + u2;
    const float c = q * dt / (2.0f * m);
    const float tau[3] = {c * B[0], c * B[1], c * B[2]};
    const float tau2 = tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2];
    const float udt = (pm[0] * tau[0] + pm[1] * tau[1] + pm[2] * tau[2]) * imk;
    const float sigma = gm2 - tau2;
    const float ghc2 = 0.5f * (sigma + sqrtf(sigma * sigma + 4.0f * (tau2 + udt * udt)));
    const float ghc = sqrtf(fmaxf(ghc2, 1.0f));
    const float ig = 1.0f / ghc;
    const float t[3] = {ig * tau[0], ig * tau[1], ig * tau[2]};
    const float sc = 2.0f / (1.0f + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
    const float sv[3] = {sc * t[0], sc * t[1], sc * t[2]};
    const float up[3] = {pm[0] + (pm[1] * t[2] - pm[2] * t[1]), pm[1] + (pm[2] * t[0] - pm[0] * t[2]),
                         pm[2] + (pm[0] * t[1] - pm[1] * t[0])};
    out[0] = pm[0] + (up[1] * sv[2] - up[2] * sv[1]);
    out[1] = pm[1] + (up[2] * sv[0] - up[0] * sv[2]);
    out[2] = pm[2] + (up[0] * sv[1] - up[1] * sv[0]);
}

struct LocalMapF {
    float vbar[3], P_end[3], v_end[3];
};

// evaluate_local_particle_substep (P3:4843-4884) in cell-relative FP32.
template <int R, bool FOLD = false>
__device__ __forceinline__ void local_map_f(const MGrid& g, const QuadF& qd, bool split, const float4* NR,
                                            const float gs[3], int ci, int cj, float tx, float ty, const float P0[3],
                                            const float A0p[3], const float pst[3], float q, float m, float kappa,
                                            float h, float th0, float th1, LocalMapF& out) {
    const float dux = h * gs[0] * g.ihx, duy = h * gs[1] * g.ihy;
    float A1p[3];
    float gpb[2], D[3][2];
    if constexpr (FOLD) {
        // single substep: th1 = 1, A at the end point is A^{n+1}
        bool ok;
        orbit_gg_f<R, true>(g, qd, split, NR, ci, cj, tx, ty, dux, duy, th0, th1, gpb, D, A1p, &ok);
        if (!ok) gather_interp_f<R>(g, NR, ci, cj, tx + dux, ty + duy, th1, A1p);
    } else {
        gather_interp_f<R>(g, NR, ci, cj, tx + dux, ty + duy, th1, A1p);
        orbit_gg_f<R>(g, qd, split, NR, ci, cj, tx, ty, dux, duy, th0, th1, gpb, D);
    }
    const float Aep[3] = {0.5f * (A0p[0] + A1p[0]), 0.5f * (A0p[1] + A1p[1]), 0.5f * (A0p[2] + A1p[2])};
    float Dv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) Dv[i] = D[i][0] * gs[0] + D[i][1] * gs[1];
    // K = D^T - D with the z column of D zero: B_eff = (K12, -K02, K01)
    const float B[3] = {D[2][1], -D[2][0], D[1][0] - D[0][1]};
    const float hq = 0.5f * q * h;
    const float gp3[3] = {gpb[0], gpb[1], 0.0f};
    float pm[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) pm[d] = P0[d] - q * Aep[d] - hq * gp3[d] + hq * Dv[d];
    float pp[3];
    hc_rotate_f(pm, B, q, m, kappa, h, pp);
    float pe[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float t = pp[d] + q * Aep[d] - hq * gp3[d] + hq * Dv[d];
        out.P_end[d] = t;
        pe[d] = t - q * A1p[d];
    }
    const float mk = m * kappa;
    const float ge = gamma_f(pe, mk);
    const float ive = 1.0f / (m * ge);
#pragma unroll
    for (int d = 0; d < 3; ++d) out.v_end[d] = ive * pe[d];
    // kinetic_secant_velocity (P3:682-689)
    const float gmi = gamma_f(pm, mk), gst = gamma_f(pst, mk), gpl = gamma_f(pp, mk);
    const float c1 = 1.0f / (m * (gmi + gst)), c2 = 1.0f / (m * (ge + gpl));
#pragma unroll
    for (int d = 0; d < 3; ++d) out.vbar[d] = 0.5f * (c1 * (pm[d] + pst[d]) + c2 * (pe[d] + pp[d]));
}

__device__ __forceinline__ void ld3f(const float* __restrict__ a, int64_t np, int64_t p, float o[3]) {
    o[0] = a[p]; o[1] = a[np + p]; o[2] = a[2 * np + p];
}
__device__ __forceinline__ void st3f(float* __restrict__ a, int64_t np, int64_t p, const float v[3]) {
    a[p] = v[0]; a[np + p] = v[1]; a[2 * np + p] = v[2];
}
__device__ __forceinline__ double norm3f(const float v[3]) {
    const double a = v[0], b = v[1], c = v[2];
    return sqrt(a * a + b * b + c * c);
}
__device__ __forceinline__ bool finite3f(const float v[3]) {
    return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

// advance a cell-relative coordinate by du cells and renormalise into [0, 1)
__device__ __forceinline__ void move_axis(float& t, float du, int& c, int n) {
    const float u = t + du;
    const float f = floorf(u);
    float r = u - f;
    int k = (int)f;
    if (r >= 1.0f) { r = 0.0f; k += 1; }
    t = r;
    c = wrapn(c + k, n);
}

// Initial guess of the inner solve (no counterpart in the reference, which
// starts from v^n, P3:5318).  Gathered at the start point: E = -grad phi_mid
// - (A^{n+1} - A^n)/h, B = curl (A^n + A^{n+1})/2 and the convective
// derivative (v . grad) E; the physics predictor of the orbit-averaged
// velocity is half a relativistic Lorentz step,
//   delta = (h/2) q/(m gamma) (F - v (v.F)/kappa^2),  F = E + v x B,
// and the network (mlp.hpp) adds a learned correction from the 12 features
// (v, E, B, (v . grad) E).
// FULL = false computes only the predictor (E and B at the start point; f
// is left untouched), for sweeps that neither run the network nor collect
// training rows.
template <int R, bool FULL = true>
__device__ __forceinline__ void guess_features(const MGrid& g, const float4* __restrict__ NR, int ci, int cj, float tx,
                                               float ty, const float v[3], float q, float m, float kappa, float h,
                                               float f[12], float dphys[3], float A0[3]) {
    AxF<R> sx, sy;
    axis_f<R>(tx, g.ihx, sx);
    axis_f<R>(ty, g.ihy, sy);
    const int bx = wrapn(ci + sx.fl, g.nx), by = wrapn(cj + sy.fl, g.ny);
    float gp0 = 0.0f, gp1 = 0.0f, dg00 = 0.0f, dg01 = 0.0f, dg10 = 0.0f, dg11 = 0.0f;
    float dA[3] = {0.0f, 0.0f, 0.0f}, ddA[3][2] = {}, dAm[3][2] = {};
    float a00 = 0.0f, a01 = 0.0f, a02 = 0.0f;
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap1n(bx + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap1n(by + b, g.ny);
            float4 r0, r1;
            rec_f(NR, ia * g.ny + ib, r0, r1);
            const float W = sx.w[a] * sy.w[b], Wx = sx.dw[a] * sy.w[b], Wy = sx.w[a] * sy.dw[b];
            gp0 += W * r0.x;
            gp1 += W * r0.y;
            const float da[3] = {r1.y - r0.z, r1.z - r0.w, r1.w - r1.x};
            const float am[3] = {0.5f * (r0.z + r1.y), 0.5f * (r0.w + r1.z), 0.5f * (r1.x + r1.w)};
#pragma unroll
            for (int c = 0; c < 3; ++c) dA[c] += W * da[c];
            // A^n at the start point (the push's theta = 0 gather)
            a00 += W * r0.z;
            a01 += W * r0.w;
            a02 += W * r1.x;
            // B = curl A_mid needs d/dy of A_mid x, z and d/dx of A_mid y, z
            dAm[0][1] += Wy * am[0];
            dAm[1][0] += Wx * am[1];
            dAm[2][0] += Wx * am[2];
            dAm[2][1] += Wy * am[2];
            if constexpr (FULL) {
                dg00 += Wx * r0.x; dg01 += Wy * r0.x;
                dg10 += Wx * r0.y; dg11 += Wy * r0.y;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ddA[c][0] += Wx * da[c];
                    ddA[c][1] += Wy * da[c];
                }
            }
        }
    }
    A0[0] = a00; A0[1] = a01; A0[2] = a02;
    const float ih = 1.0f / h;
    const float E[3] = {-gp0 - dA[0] * ih, -gp1 - dA[1] * ih, -dA[2] * ih};
    const float Bm[3] = {dAm[2][1], -dAm[2][0], dAm[1][0] - dAm[0][1]};
    if constexpr (FULL) {
        const float vgE[3] = {-(v[0] * dg00 + v[1] * dg01) - (v[0] * ddA[0][0] + v[1] * ddA[0][1]) * ih,
                              -(v[0] * dg10 + v[1] * dg11) - (v[0] * ddA[1][0] + v[1] * ddA[1][1]) * ih,
                              -(v[0] * ddA[2][0] + v[1] * ddA[2][1]) * ih};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            f[c] = v[c];
            f[3 + c] = E[c];
            f[6 + c] = Bm[c];
            f[9 + c] = vgE[c];
        }
    }
    const float ik2 = 1.0f / (kappa * kappa);
    const float v2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * ik2;
    const float gam = 1.0f / sqrtf(fmaxf(1.0f - v2, 1.0e-12f));
    const float F[3] = {E[0] + (v[1] * Bm[2] - v[2] * Bm[1]), E[1] + (v[2] * Bm[0] - v[0] * Bm[2]),
                        E[2] + (v[0] * Bm[1] - v[1] * Bm[0])};
    const float vF = (v[0] * F[0] + v[1] * F[1] + v[2] * F[2]) * ik2;
    const float coef = 0.5f * h * q / (m * gam);
#pragma unroll
    for (int c = 0; c < 3; ++c) dphys[c] = coef * (F[c] - v[c] * vF);
}

// One particle of the sweep: the substep loop of
// evaluate_selective_current_map_at_layout (P3:5290-5431) in cell-relative
// FP32.  gin: the initial guess of the (single-substep) inner solve; it0: the
// Picard iterations already spent on it (a deferred retry continues at 1).
// With defer set, a particle that the first Picard iteration leaves
// unconverged is queued (index and updated guess) for k_p3_map_retry and
// nothing else is written for it here: a handful of such lanes would
// otherwise keep their whole wavefront iterating.
template <int R, bool COLLECT, bool FUSED, bool FOLD = false>
__device__ __forceinline__ void map_particle(const MGrid& g, const QuadF& qd, int split, const P3MArgs& a, int64_t p,
                                             const float gin[3], int it0, bool defer, const float feat[12],
                                             const float dph[3], bool has_a0, float a0x, float a0y, float a0z,
                                             const DepTileF& T, double rsum[2], double rmax[2], int& nseg_total) {
    // Only the fused sweep runs per-particle substep layouts; the others
    // (and the retry sweep) are launched for the uniform layout, so their
    // substep loop, start-point gather and substep-path recording compile
    // away (they had cost the later cubic sweep 244 VGPRs and 14%).
    const int nsub = FUSED && a.layout ? (int)a.layout[p] : 1;
    const float h = (float)(a.dt / (double)nsub);
    const int c0 = a.cell[p];
    int ci = c0 / g.ny, cj = c0 - (c0 / g.ny) * g.ny;
    float tx = a.off[p], ty = a.off[a.np + p], tz = a.off[2 * a.np + p];
    float Pc[3], vc[3];
    ld3f(a.P, a.np, p, Pc);
    ld3f(a.v, a.np, p, vc);
    float vmac[3] = {0.0f, 0.0f, 0.0f};
    const float qp = a.q[p], mp = a.m[p];
    const float kappa = (float)a.kappa;
    const float rtol = (float)a.rtol, atol = (float)a.atol, damp = (float)a.damping;
    const float inv_nsub = 1.0f / (float)nsub;
    int reason = 0;
    for (int s = 0; s < nsub && reason == 0; ++s) {
        const float th0 = (float)s * inv_nsub, th1 = (float)(s + 1) * inv_nsub;
        float A0p[3];
        // A^n at the start point, by value (kept in registers) when the
        // caller has it
        if (has_a0 && nsub == 1) {
            A0p[0] = a0x; A0p[1] = a0y; A0p[2] = a0z;
        } else {
            gather_interp_f<R>(g, a.NR, ci, cj, tx, ty, th0, A0p);
        }
        const float pst[3] = {Pc[0] - qp * A0p[0], Pc[1] - qp * A0p[1], Pc[2] - qp * A0p[2]};
        float gs[3] = {vc[0], vc[1], vc[2]};
        if (nsub == 1) { gs[0] = gin[0]; gs[1] = gin[1]; gs[2] = gin[2]; }
        LocalMapF acc;
        float vacc[3];
        bool conv = false;
        int its = nsub == 1 ? it0 : 0;
        float scaled = 0.0f;
        reason = 5;  // maximum_iterations
        for (int it = its; it < a.max_iter; ++it) {
            LocalMapF mpd;
            if (FOLD && nsub == 1)
                local_map_f<R, true>(g, qd, split != 0, a.NR, gs, ci, cj, tx, ty, Pc, A0p, pst, qp, mp, kappa, h, th0,
                                     th1, mpd);
            else
                local_map_f<R>(g, qd, split != 0, a.NR, gs, ci, cj, tx, ty, Pc, A0p, pst, qp, mp, kappa, h, th0, th1,
                               mpd);
            ++its;
            if (!finite3f(mpd.vbar)) { reason = 1; break; }
            // convergence test of solve_local_particle_substep (P3:4951-4968)
            const float nm = sqrtf(mpd.vbar[0] * mpd.vbar[0] + mpd.vbar[1] * mpd.vbar[1] + mpd.vbar[2] * mpd.vbar[2]);
            if (nm >= kappa) { reason = 2; break; }
            const float res[3] = {mpd.vbar[0] - gs[0], mpd.vbar[1] - gs[1], mpd.vbar[2] - gs[2]};
            const float ng = sqrtf(gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]);
            const float nr = sqrtf(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
            const float tol = atol + rtol * (0.5f * (nm + ng));
            scaled = nr / tol;
            if (nr <= tol) {
                vacc[0] = gs[0]; vacc[1] = gs[1]; vacc[2] = gs[2];
                acc = mpd;
                conv = true;
                reason = 0;
                break;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) gs[d] = gs[d] + damp * res[d];
            if (!finite3f(gs)) { reason = 3; break; }
            if (gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2] >= kappa * kappa) { reason = 4; break; }
            if (defer && nsub == 1) {
                // one queue reservation for the lanes of the wavefront deferring here
                const unsigned long long grp = __ballot(1);
                const int lane = threadIdx.x & 63, leader = __ffsll((long long)grp) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(a.retry_n, __popcll(grp));
                base = __shfl(base, leader, 64);
                const int k = base + __popcll(grp & ((1ull << lane) - 1ull));
                a.retry_idx[k] = (int)p;
                if (a.retry_old) st3f(a.retry_old, a.np, k, gin);
                st3f(a.retry_gs, a.np, k, gs);
                return;
            }
        }
        rsum[0] += (double)its;
        rmax[0] = fmax(rmax[0], (double)its);
        rmax[1] = fmax(rmax[1], (double)scaled);
        if (!conv) break;
        if constexpr (COLLECT) {
            if (a.write_out && nsub == 1 && a.train && (p % a.train_stride) == 0) {
                const int64_t row = a.train_offset + p / a.train_stride;
                if (row < a.train_cap) {
                    float* o = a.train + 15 * row;
#pragma unroll
                    for (int k = 0; k < 12; ++k) o[k] = feat[k];
#pragma unroll
                    for (int d = 0; d < 3; ++d) o[12 + d] = vacc[d] - vc[d] - dph[d];
                }
            }
        }
        if constexpr (FUSED) {
            if (a.write_out && a.sub.x0) {
                // the substep's start point and vbar, as FP64 positions (the
                // FP64 view's -L/2 + (cell + offset) h)
                const int64_t j = a.sub.off[p] + s, n = a.sub.n;
                a.sub.x0[j] = a.lo_x + ((double)ci + (double)tx) * a.hx;
                a.sub.x0[n + j] = a.lo_y + ((double)cj + (double)ty) * a.hy;
                a.sub.x0[2 * n + j] = a.lo_z + (double)tz * a.Lz;
#pragma unroll
                for (int d = 0; d < 3; ++d) a.sub.vbar[d * n + j] = (double)vacc[d];
            }
        }
        const float dux = h * vacc[0] * g.ihx, duy = h * vacc[1] * g.ihy;
        if constexpr (FUSED) {
            if (a.J_out)
                nseg_total += deposit_orbit_f<R>(g, qd, split != 0, ci, cj, tx, ty, dux, duy, vacc,
                                                 qp * inv_nsub * a.inv_dV, T, a.J_out);
        } else {
            st3f(a.vdep, a.np, p, vacc);
        }
        move_axis(tx, dux, ci, g.nx);
        move_axis(ty, duy, cj, g.ny);
        int cz = 0;
        move_axis(tz, h * vacc[2] * a.inv_Lz, cz, 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Pc[d] = acc.P_end[d];
            vc[d] = acc.v_end[d];
            vmac[d] += vacc[d] * inv_nsub;
        }
    }
    if (a.reason) a.reason[p] = (uint8_t)reason;
    if (reason != 0) rsum[1] += 1.0;
    if (reason == 0 && a.write_out) {
        a.cell_out[p] = ci * g.ny + cj;
        a.off_out[p] = tx;
        a.off_out[a.np + p] = ty;
        a.off_out[2 * a.np + p] = tz;
        st3f(a.P_out, a.np, p, Pc);
        st3f(a.v_out, a.np, p, vc);
        st3f(a.vbar_out, a.np, p, vmac);
    }
}

// GM: initial guess of the inner solve, 0 = v^n (the reference's) or
// a.guess when given, 1 = physics predictor, 2 = predictor + network,
// 3 = as 0 with A^n at the start point read from a.a0_in;
// COLLECT: write (features, converged correction) training rows in the
// write-out pass.
// FUSED: deposit inside the sweep (per-particle substep layouts); otherwise
// the converged vbar goes to a.vdep and the deposit kernel deposits it,
// which keeps the deposit's registers out of the Picard sweep.
template <int R, int GM, bool COLLECT, bool FUSED>
__device__ __forceinline__ void p3_map_mixed_body(const MGrid& g, const QuadF& qd, int split, const P3MArgs& a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float tile_mem[kTileX * kTileY * 3];
    extern __shared__ unsigned char mlp_lds[];
    DepTileF T{};
    if constexpr (FUSED) T = tilef_begin(g, tile_mem, a.cell, a.np, (int64_t)blockIdx.x * blockDim.x);
    // features and guess at the start point (all lanes: the network runs on
    // whole wavefronts)
    float feat[12] = {}, dph[3] = {0.0f, 0.0f, 0.0f}, lg[3] = {0.0f, 0.0f, 0.0f}, a0s[3] = {0.0f, 0.0f, 0.0f};
    constexpr bool PRED = GM == 1 || GM == 2;
    if constexpr (PRED || COLLECT) {
        if (p < a.np && (!FUSED || !a.layout)) {
            const int c0 = a.cell[p];
            float v0[3];
            ld3f(a.v, a.np, p, v0);
            guess_features<R, GM == 2 || COLLECT>(g, a.NR, c0 / g.ny, c0 - (c0 / g.ny) * g.ny, a.off[p],
                                                   a.off[a.np + p], v0, a.q[p], a.m[p], (float)a.kappa,
                                                   (float)a.dt, feat, dph, a0s);
            if (a.a0_out) st3f(a.a0_out, a.np, p, a0s);
        }
        if constexpr (GM == 2) mlp_wave_infer(a.mlp, feat, lg, mlp_lds + (threadIdx.x / 64) * kMlpLdsWave);
    }
    double rsum[2] = {0.0, 0.0}, rmax[2] = {0.0, 0.0};
    int nseg_total = 0;
    if constexpr (GM == 3) {
        if (p < a.np) ld3f(a.a0_in, a.np, p, a0s);  // from the step's first sweep
    }
    if (p < a.np) {
        float gin[3];
        ld3f(a.v, a.np, p, gin);
        if (a.guess) ld3f(a.guess, a.np, p, gin);
        if constexpr (PRED) {
#pragma unroll
            for (int d = 0; d < 3; ++d) gin[d] += dph[d] + (GM == 2 ? lg[d] : 0.0f);
        }
        // (the unfused sweeps never run a substep layout; they all fold the
        // end-point gather into the orbit's node loop: with the network in
        // the same sweep that also keeps it at 238 VGPRs without scratch,
        // against 256 and 48 bytes of scratch for the separate gather)
        const bool uniform = !FUSED || !a.layout;
        map_particle<R, COLLECT, FUSED, !FUSED>(g, qd, split, a, p, gin, 0,
                                                           a.retry_n != nullptr && uniform, feat, dph,
                                        (PRED || COLLECT || GM == 3) && uniform, a0s[0], a0s[1], a0s[2], T,
                                        rsum, rmax, nseg_total);
    }
    if constexpr (FUSED) {
        if (a.J_out) tilef_end(g, T, a.J_out);
    }
    block_partials<2, 2>(rsum, rmax, nseg_total, a.part);
}

// The sweeps are capped at 256 registers (VGPRs + AGPRs, two wavefronts
// per SIMD): left to itself the compiler can settle one register past it and
// halve their occupancy (1.39 against 2.50 ms per 16.8M-particle cubic
// sweep).
template <int R, int GM, bool COLLECT, bool FUSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p3_map_mixed(MGrid g, QuadF qd,
                                                                                            int split, P3MArgs a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    p3_map_mixed_body<R, GM, COLLECT, FUSED>(g, qd, split, a);
}

// The collection sweeps (training rows) hold the features as well; the same
// cap.
template <int R, int GM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p3_map_mixed_collect(
    MGrid g, QuadF qd, int split, P3MArgs a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    p3_map_mixed_body<R, GM, true, false>(g, qd, split, a);
}

// The deferred particles of a sweep (uniform layout), packed into full
// wavefronts: queue entries [0, *retry_n), grid-stride.  Picks the Picard
// iteration up at its second step from the queued guess.  Capped at two
// wavefronts per SIMD like the sweeps.
template <int R, bool COLLECT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p3_map_retry(MGrid g, QuadF qd,
                                                                                            int split, P3MArgs a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    const int n = *a.retry_n;
    double rsum[2] = {0.0, 0.0}, rmax[2] = {0.0, 0.0};
    int nseg_total = 0;
    const DepTileF T{};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int64_t p = a.retry_idx[k];
        float gin[3];
        ld3f(a.retry_gs, a.np, k, gin);
        float feat[12] = {}, dph[3] = {0.0f, 0.0f, 0.0f};
        if constexpr (COLLECT) {
            const int c0 = a.cell[p];
            float v0[3];
            ld3f(a.v, a.np, p, v0);
            float a0s[3];
            guess_features<R>(g, a.NR, c0 / g.ny, c0 - (c0 / g.ny) * g.ny, a.off[p], a.off[a.np + p], v0, a.q[p],
                              a.m[p], (float)a.kappa, (float)a.dt, feat, dph, a0s);
        }
        map_particle<R, COLLECT, false, true>(g, qd, split, a, p, gin, 1, false, feat, dph, false, 0.0f, 0.0f, 0.0f,
                                              T, rsum, rmax, nseg_total);
    }
    block_partials<2, 2>(rsum, rmax, nseg_total, a.part);
}

// Orbit current deposit of the converged single-substep sweep: particle p
// moves from its start point by dt vbar (vbar from the sweep), FP32 LDS tile
// flushed into the FP64 mesh current (path_average_deposit_current,
// P3:2784-2857, with the sweep's q/dV weight).  Particles that failed the
// inner solve deposit nothing (the step is then rejected anyway).
template <int R>
__global__ __launch_bounds__(256) void k_deposit_mixed(MGrid g, QuadF qd, int split, P3MArgs a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float tile_mem[kTileX * kTileY * 3];
    const DepTileF T = tilef_begin(g, tile_mem, a.cell, a.np, (int64_t)blockIdx.x * blockDim.x);
    int nseg = 0;
    if (p < a.np && a.reason[p] == 0) {
        const float h = (float)a.dt;
        const int c0 = a.cell[p];
        const int ci = c0 / g.ny, cj = c0 - (c0 / g.ny) * g.ny;
        float vb[3];
        ld3f(a.vdep, a.np, p, vb);
        const float dux = h * vb[0] * g.ihx, duy = h * vb[1] * g.ihy;
        nseg = deposit_orbit_f<R>(g, qd, split != 0, ci, cj, a.off[p], a.off[a.np + p], dux, duy, vb,
                                  a.q[p] * a.inv_dV, T, a.J_out);
    }
    tilef_end(g, T, a.J_out);
    block_partials<0, 0>(nullptr, nullptr, nseg, a.part);
}

// Incremental deposit of a sweep that started every particle from the
// previous sweep's vbar: a particle that converged at its first Picard
// iteration accepted that vbar unchanged, so its current is the one already
// in the kept deposit; only the deferred particles (queue entries [0,
// *retry_n)) changed, and each moves its current from the starting vbar
// (retry_old) to the converged one (vdep).  The partial slots of all
// workgroups are written (neutral ones past the queue).  sums: [0] segment count change,
// [1] crossing count change; maxs: [0] largest new segment count.
template <int R>
__global__ __launch_bounds__(256) void k_deposit_delta_mixed(MGrid g, QuadF qd, int split, P3MArgs a) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    const int n = *a.retry_n;
    __shared__ float tile_mem[kTileX * kTileY * 3];
    double sums[2] = {0.0, 0.0}, maxs[1] = {0.0};
    // A workgroup takes 128 queue entries at a time: the 32 entries of each
    // wavefront are spread over its 64 lanes, the lower half removing the
    // old orbit's current and the upper half adding the new one, so a lane
    // runs one orbit (the queue is short and scattered: parallelism, not
    // the tile, sets the speed).  Grid-stride; k0 is uniform in the
    // workgroup, so every thread takes part in each chunk's tile.
    const int lane = threadIdx.x & 63;
    const bool add = lane >= 32;
    for (int64_t k0 = (int64_t)blockIdx.x * 128; k0 < n; k0 += (int64_t)gridDim.x * 128) {
        const DepTileF T = tilef_begin(g, tile_mem, a.cell, a.np, (int64_t)a.retry_idx[k0]);
        const int64_t k = k0 + 32 * (threadIdx.x >> 6) + (lane & 31);
        if (k < n && a.reason[a.retry_idx[k]] == 0) {
            const int64_t p = a.retry_idx[k];
            const float h = (float)a.dt;
            const int c0 = a.cell[p];
            const int ci = c0 / g.ny, cj = c0 - (c0 / g.ny) * g.ny;
            const float qs = a.q[p] * a.inv_dV;
            float vb[3];
            if (add) ld3f(a.vdep, a.np, p, vb);
            else ld3f(a.retry_old, a.np, k, vb);
            const int sg = deposit_orbit_f<R>(g, qd, split != 0, ci, cj, a.off[p], a.off[a.np + p],
                                              h * vb[0] * g.ihx, h * vb[1] * g.ihy, vb, add ? qs : -qs, T, a.J_out);
            sums[0] += add ? (double)sg : -(double)sg;
            sums[1] += sg > 1 ? (add ? 1.0 : -1.0) : 0.0;
            if (add) maxs[0] = fmax(maxs[0], (double)sg);
        }
        tilef_end(g, T, a.J_out);
        __syncthreads();  // the flush has read the tile before the next chunk clears it
    }
    block_partials<2, 1>(sums, maxs, 0, a.part);
}

// Cell-mapped deposit (cell-sorted particles): a workgroup owns a block of
// 64 cells, bw rows by 64 / bw columns (bw = 1: 64 consecutive cells of one
// grid row; bw = 8: an 8 x 8 block); lane l of every wave takes one cell of
// the block, and the kCellWaves waves of the kCellSlices workgroups (grid y)
// of that block split each cell's particles.  The 64 lanes of an LDS atomic
// then address 64 different stencil windows instead of one (particles of one
// cell share their stencil), which removes the 64-way LDS bank serialisation
// of the particle-order deposit.  The LDS tile spans rows i0-3..i0+bw+4 and
// columns j0-3..j0+64/bw+4 of the block.  The square block flushes about
// half as many nodes per cell to the mesh as the row segment (each node of
// a block's tile is added to the mesh with one FP64 atomic per component,
// and neighbouring blocks repeat the nodes they share).
constexpr int kCellWaves = 4, kCellSlices = 8, kCTR = 8, kCTC = 72;
constexpr int kCTiles = kCTR * kCTC;  // LDS nodes, either shape (8 x 72 or 16 x 16)

template <int R>
__global__ __launch_bounds__(256) void k_deposit_cells_mixed(MGrid g, QuadF qd, int split, P3MArgs a,
                                                             const int* __restrict__ cell_end, int bw) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    __shared__ float tile[kCTiles * 3];
    const int bh = 64 / bw;  // block columns
    const int nbj = g.ny / bh;
    const int i0 = (blockIdx.x / nbj) * bw, j0 = (blockIdx.x - (blockIdx.x / nbj) * nbj) * bh;
    // tile rows and columns: the block, 3 before and 5 after
    const int tr = bw == 1 ? kCTR : bw + 8, tc = bw == 1 ? kCTC : bh + 8;
    const int ox = wrapn(i0 - 3, g.nx), oy = wrapn(j0 - 3, g.ny);
    for (int e = threadIdx.x; e < kCTiles * 3; e += blockDim.x) tile[e] = 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int ci = i0 + lane / bh, cj = j0 + lane - (lane / bh) * bh;
    const int c = ci * g.ny + cj;
    const int64_t kb = c == 0 ? 0 : cell_end[c - 1], ke = cell_end[c];
    double cross = 0.0, segs = 0.0, mseg = 0.0;
    const int stride = kCellWaves * gridDim.y;
    for (int64_t p = kb + sub + kCellWaves * blockIdx.y; p < ke; p += stride) {
        if (a.reason[p] != 0) continue;
        float vb[3];
        ld3f(a.vdep, a.np, p, vb);
        const float h = (float)a.dt;
        const float dux = h * vb[0] * g.ihx, duy = h * vb[1] * g.ihy;
        const float qs = a.q[p] * a.inv_dV;
        // the orbit deposit with the block's tile (window flush as in deposit_orbit_f)
        float W[R + 1][R + 1];
        int bx = 0, by = 0;
        bool have = false;
        auto flush = [&]() {
            if (!have) return;
            const int gx = wrapn(ci + bx, g.nx), gy = wrapn(cj + by, g.ny);
#pragma unroll
            for (int aa = 0; aa <= R; ++aa) {
                const int ia = wrap1n(gx + aa, g.nx);
                int dx = ia - ox;
                if (dx < 0) dx += g.nx;
#pragma unroll
                for (int bb = 0; bb <= R; ++bb) {
                    const int ib = wrap1n(gy + bb, g.ny);
                    int dy = ib - oy;
                    if (dy < 0) dy += g.ny;
                    const float w = W[aa][bb] * qs;
                    if (dx < tr && dy < tc) {
                        float* t = tile + 3 * (dx * tc + dy);
                        atomicAdd(t, w * vb[0]);
                        atomicAdd(t + 1, w * vb[1]);
                        atomicAdd(t + 2, w * vb[2]);
                    } else {
                        const int id = ia * g.ny + ib;
                        unsafeAtomicAdd(&a.J_out[id], (double)(w * vb[0]));
                        unsafeAtomicAdd(&a.J_out[g.N + id], (double)(w * vb[1]));
                        unsafeAtomicAdd(&a.J_out[2 * g.N + id], (double)(w * vb[2]));
                    }
                }
            }
            have = false;
        };
        BreakF bg;
        const float t0x = a.off[p], t0y = a.off[a.np + p];
        breakf_init(bg, R, t0x, t0y, split ? dux : 0.0f, split ? duy : 0.0f);
        float sa = 0.0f;
        int nseg = 0;
        for (;;) {
            const float sb = breakf_next(bg);
            const float len = sb - sa;
            if (len > 0.0f) {
                for (int iq = 0; iq < qd.n; ++iq) {
                    const float s_ = sa + len * qd.s[iq];
                    const float w = len * qd.w[iq];
                    AxF<R> sx, sy;
                    axis_f<R>(t0x + s_ * dux, g.ihx, sx);
                    axis_f<R>(t0y + s_ * duy, g.ihy, sy);
                    if (!(have && sx.fl == bx && sy.fl == by)) {
                        flush();
                        bx = sx.fl;
                        by = sy.fl;
#pragma unroll
                        for (int aa = 0; aa <= R; ++aa)
#pragma unroll
                            for (int bb = 0; bb <= R; ++bb) W[aa][bb] = 0.0f;
                        have = true;
                    }
#pragma unroll
                    for (int aa = 0; aa <= R; ++aa) {
                        const float wa = w * sx.w[aa];
#pragma unroll
                        for (int bb = 0; bb <= R; ++bb) W[aa][bb] += wa * sy.w[bb];
                    }
                }
            }
            ++nseg;
            sa = sb;
            if (sb >= 1.0f) break;
        }
        flush();
        cross += nseg > 1 ? 1.0 : 0.0;
        segs += (double)nseg;
        mseg = fmax(mseg, (double)nseg);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < tr * tc * 3; e += blockDim.x) {
        const float v = tile[e];
        if (v != 0.0f) {
            const int comp = e % 3, node = e / 3;
            const int dx = node / tc, dy = node - (node / tc) * tc;
            if (dx >= g.nx || dy >= g.ny) continue;
            const int ia = (ox + dx) % g.nx, ib = (oy + dy) % g.ny;
            unsafeAtomicAdd(&a.J_out[(size_t)comp * g.N + ia * g.ny + ib], (double)v);
        }
    }
    block_partials_stats<0, 0>(nullptr, nullptr, cross, segs, mseg,
                               a.part + (size_t)blockIdx.y * gridDim.x * kPartSlots);
}

// Particle-order deposit with a wavefront reduction.  With cell-sorted
// particles the 64 lanes of a wave almost always share one stencil window
// (same start cell, no knot crossing), so each lane writes its 16 node
// weights W and its scaled velocity qs v to LDS and lane j < 48 sums column
// (node j / 3, component j % 3) over the 64 lanes: one LDS atomic per node
// and component per wave instead of one per lane, the per-lane float
// atomics being the bottleneck of the plain tile deposit.  Lanes whose
// window differs from the wave's leading lane, and the earlier windows of
// orbits that cross a knot, deposit with per-lane atomics.
template <int R>
__global__ __launch_bounds__(256) void k_deposit_wave_mixed(MGrid g, QuadF qd, int split, P3MArgs a, int tx, int ty) {
    if (UPIC_GATE_CLOSED(a.gate)) return;
    constexpr int NW = (R + 1) * (R + 1);
    constexpr int ROW = NW + 4;  // W (NW) | qs v (3) | pad
    __shared__ float tile_mem[kTileX * kTileY * 3];
    __shared__ float red[4][64 * ROW];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const DepTileF T = tilef_begin(g, tile_mem, a.cell, a.np, (int64_t)blockIdx.x * blockDim.x, tx, ty);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float W[R + 1][R + 1];
#pragma unroll
    for (int i = 0; i <= R; ++i)
#pragma unroll
        for (int j = 0; j <= R; ++j) W[i][j] = 0.0f;
    float sv[3] = {0.0f, 0.0f, 0.0f};
    int key = -1, gx = 0, gy = 0, nseg = 0;
    const bool valid = p < a.np && a.reason[p] == 0;
    if (valid) {
        const float h = (float)a.dt;
        const int c0 = a.cell[p];
        const int ci = c0 / g.ny, cj = c0 - (c0 / g.ny) * g.ny;
        float vb[3];
        ld3f(a.vdep, a.np, p, vb);
        const float dux = h * vb[0] * g.ihx, duy = h * vb[1] * g.ihy;
        const float qs = a.q[p] * a.inv_dV;
        sv[0] = qs * vb[0]; sv[1] = qs * vb[1]; sv[2] = qs * vb[2];
        const float t0x = a.off[p], t0y = a.off[a.np + p];
        int bx = 0, by = 0;
        bool have = false;
        BreakF bg;
        breakf_init(bg, R, t0x, t0y, split ? dux : 0.0f, split ? duy : 0.0f);
        float sa = 0.0f;
        for (;;) {
            const float sb = breakf_next(bg);
            const float len = sb - sa;
            if (len > 0.0f) {
                for (int iq = 0; iq < qd.n; ++iq) {
                    const float s_ = sa + len * qd.s[iq];
                    const float w = len * qd.w[iq];
                    AxF<R> sx, sy;
                    axis_f<R>(t0x + s_ * dux, g.ihx, sx);
                    axis_f<R>(t0y + s_ * duy, g.ihy, sy);
                    if (have && (sx.fl != bx || sy.fl != by)) {
                        // the window moved: the finished window goes out with per-lane atomics
                        const int ox = wrapn(ci + bx, g.nx), oy = wrapn(cj + by, g.ny);
#pragma unroll
                        for (int aa = 0; aa <= R; ++aa) {
                            const int ia = wrap1n(ox + aa, g.nx);
#pragma unroll
                            for (int bb = 0; bb <= R; ++bb) {
                                const int ib = wrap1n(oy + bb, g.ny);
                                tilef_add(g, T, a.J_out, ia, ib, W[aa][bb] * sv[0], W[aa][bb] * sv[1], W[aa][bb] * sv[2]);
                                W[aa][bb] = 0.0f;
                            }
                        }
                    }
                    bx = sx.fl;
                    by = sy.fl;
                    have = true;
#pragma unroll
                    for (int aa = 0; aa <= R; ++aa) {
                        const float wa = w * sx.w[aa];
#pragma unroll
                        for (int bb = 0; bb <= R; ++bb) W[aa][bb] += wa * sy.w[bb];
                    }
                }
            }
            ++nseg;
            sa = sb;
            if (sb >= 1.0f) break;
        }
        gx = wrapn(ci + bx, g.nx);
        gy = wrapn(cj + by, g.ny);
        key = gx * g.ny + gy;
    }
    // Lanes whose final windows agree are reduced together: each round takes
    // the window of the first pending lane, sums W (masked to that window's
    // lanes) times q v / dV over the wavefront through LDS, one lane per
    // (node, component), and adds the sums to the tile.  A wavefront of
    // cell-ordered particles holds one or two windows, one of the group of
    // knot-crossing orbits several (one per cell it spans); lanes still
    // pending after kRounds rounds use per-lane atomics.  Eight rounds rather
    // than two: on cold counter-streaming beams (1024^2, 64 per cell) the
    // per-lane LDS float atomics of the crossing group cost 10.4 against 7.1
    // ms per deposit, and their order-dependent FP32 sums left the outer
    // solve 7.7 instead of 5 evaluations per step at the FP32 tolerance floor
    // (232 -> 148 ms per step); the Weibel and Landau decks are unchanged or
    // faster.
    constexpr int kRounds = 8;
    float* rw = red[wv] + lane * ROW;
    rw[NW] = sv[0];
    rw[NW + 1] = sv[1];
    rw[NW + 2] = sv[2];
    unsigned long long pending = __ballot(key >= 0);
    for (int round = 0; round < kRounds && pending; ++round) {
        const int lead = __ffsll((long long)pending) - 1;
        const int lkey = __shfl(key, lead, 64);
        const int lgx = __shfl(gx, lead, 64), lgy = __shfl(gy, lead, 64);
        const unsigned long long grp = __ballot(key == lkey) & pending;
        const bool in = (grp >> lane) & 1ull;
#pragma unroll
        for (int aa = 0; aa <= R; ++aa)
#pragma unroll
            for (int bb = 0; bb <= R; ++bb) rw[aa * (R + 1) + bb] = in ? W[aa][bb] : 0.0f;
        wave_sync_lds();
        // column (node, component) sums over the window's lanes (in lane
        // order; the other lanes hold zero weights, which add nothing), 3
        // (R+1)^2 columns
        for (int col = lane; col < 3 * NW; col += 64) {
            const int node = col / 3, comp = col - 3 * (col / 3);
            float acc = 0.0f;
            for (unsigned long long mk = grp; mk; mk &= mk - 1ull) {
                const int l = __ffsll((long long)mk) - 1;
                acc += red[wv][l * ROW + node] * red[wv][l * ROW + NW + comp];
            }
            if (acc != 0.0f) {
                const int ia = wrap1n(lgx + node / (R + 1), g.nx), ib = wrap1n(lgy + node % (R + 1), g.ny);
                tilef_add1(g, T, a.J_out, ia, ib, comp, acc);
            }
        }
        wave_sync_lds();
        pending &= ~grp;
    }
    if ((pending >> lane) & 1ull) {
#pragma unroll
        for (int aa = 0; aa <= R; ++aa) {
            const int ia = wrap1n(gx + aa, g.nx);
#pragma unroll
            for (int bb = 0; bb <= R; ++bb) {
                const int ib = wrap1n(gy + bb, g.ny);
                tilef_add(g, T, a.J_out, ia, ib, W[aa][bb] * sv[0], W[aa][bb] * sv[1], W[aa][bb] * sv[2]);
            }
        }
    }
    tilef_end(g, T, a.J_out);
    block_partials<0, 0>(nullptr, nullptr, nseg, a.part);
}

// FP32 node records from the FP64 fields
__global__ void k_pack_nodes_f32(int N, const double* __restrict__ gpm, const double* __restrict__ A0,
                                 const double* __restrict__ A1, float4* __restrict__ NR, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    NR[2 * id] = make_float4((float)gpm[id], (float)gpm[N + id], (float)A0[id], (float)A0[N + id]);
    NR[2 * id + 1] = make_float4((float)A0[2 * N + id], (float)A1[id], (float)A1[N + id], (float)A1[2 * N + id]);
}

// FP64 particle arrays -> mixed state.  q and m are rounded to FP32 and the
// FP64 copies replaced by the rounded values: every particle shares one
// charge and one mass, so a rounding left only on the FP32 side would be a
// coherent bias between the FP32 push and the FP64 field solve and
// diagnostics (an energy drift of order 1e-7 per step).
__global__ void k_mixed_compress(GridDev g, int64_t np, const double* __restrict__ x, const double* __restrict__ P,
                                 const double* __restrict__ v, const double* __restrict__ vp,
                                 const double* __restrict__ vb, double* __restrict__ q, double* __restrict__ m,
                                 MixedState s) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[2];
    const int n[2] = {g.nx, g.ny};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const double u = (wrap1(x[d * np + p], g.L[d]) + 0.5 * g.L[d]) / g.h[d];
        int i = (int)floor(u);
        float t = (float)(u - (double)i);
        if (t >= 1.0f) { t = 0.0f; i += 1; }
        if (i >= n[d]) i -= n[d];
        if (i < 0) i += n[d];
        c[d] = i;
        s.off[d * np + p] = t;
    }
    {
        float t = (float)((wrap1(x[2 * np + p], g.L[2]) + 0.5 * g.L[2]) / g.L[2]);
        if (t >= 1.0f) t = 0.0f;
        s.off[2 * np + p] = t;
    }
    s.cell[p] = c[0] * g.ny + c[1];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s.P[d * np + p] = (float)P[d * np + p];
        s.v[d * np + p] = (float)v[d * np + p];
        s.vp[d * np + p] = (float)vp[d * np + p];
        s.vb[d * np + p] = (float)vb[d * np + p];
    }
    const float qf = (float)q[p], mf = (float)m[p];
    s.q[p] = qf;
    s.m[p] = mf;
    q[p] = (double)qf;
    m[p] = (double)mf;
}

// mixed state -> FP64 particle arrays (x in [-L/2, L/2))
__global__ void k_mixed_expand(GridDev g, int64_t np, MixedState s, double* __restrict__ x, double* __restrict__ P,
                               double* __restrict__ v, double* __restrict__ vp, double* __restrict__ vb,
                               double* __restrict__ q, double* __restrict__ m) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int c = s.cell[p];
    const int ci = c / g.ny, cj = c - (c / g.ny) * g.ny;
    x[p] = -0.5 * g.L[0] + ((double)ci + (double)s.off[p]) * g.h[0];
    x[np + p] = -0.5 * g.L[1] + ((double)cj + (double)s.off[np + p]) * g.h[1];
    x[2 * np + p] = -0.5 * g.L[2] + (double)s.off[2 * np + p] * g.L[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        P[d * np + p] = (double)s.P[d * np + p];
        v[d * np + p] = (double)s.v[d * np + p];
        vp[d * np + p] = (double)s.vp[d * np + p];
        vb[d * np + p] = (double)s.vb[d * np + p];
    }
    // the FP64 charge and mass were rounded to FP32 when the state was
    // compressed, so the FP32 copies restore them exactly
    q[p] = (double)s.q[p];
    m[p] = (double)s.m[p];
}

// Counting-sort keys: the cell index, plus N when `split` and the orbit
// t -> t + dt vbar / h over the stored vbar (the last step's) crosses a knot
// of the degree-R splines (integer t - (R - 1)/2) along x or y.
__global__ void k_mixed_keys(MGrid g, int64_t np, const int* __restrict__ cell, const float* __restrict__ off,
                             const float* __restrict__ vb, float dt, int split, int* __restrict__ keys,
                             int* __restrict__ counts) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < np;
    int key = 0;
    if (active) {
        key = cell[p];
        if (split) {
            const float koff = 0.5f * (float)(g.order - 1);
            const float ax = off[p] - koff, ay = off[np + p] - koff;
            const float bx = ax + dt * vb[p] * g.ihx, by = ay + dt * vb[np + p] * g.ihy;
            if (floorf(ax) != floorf(bx) || floorf(ay) != floorf(by)) key += g.N;
        }
        keys[p] = key;
    }
    (void)wave_claim(counts, key, active);
}

// All per-particle arrays of the mixed state through one permutation:
// dst[i] = src[perm[i]] for the cell index, the 3-vectors (offsets, P, v,
// v_prev, vbar), q, m and the original particle index.
// HIST: also v_prev and vbar (else the caller permutes them later, or never:
// a committed step overwrites both)
template <bool HIST>
__global__ void k_mixed_permute(int64_t np, const int* __restrict__ perm, MixedState src, MixedState dst,
                                const int64_t* __restrict__ id_src, int64_t* __restrict__ id_dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t j = perm[i];
    dst.cell[i] = src.cell[j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dst.off[c * np + i] = src.off[c * np + j];
        dst.P[c * np + i] = src.P[c * np + j];
        dst.v[c * np + i] = src.v[c * np + j];
        if constexpr (HIST) {
            dst.vp[c * np + i] = src.vp[c * np + j];
            dst.vb[c * np + i] = src.vb[c * np + j];
        }
    }
    dst.q[i] = src.q[j];
    dst.m[i] = src.m[j];
    id_dst[i] = id_src[j];
}

inline unsigned mblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// v_prev and vbar of a permutation whose other fields have been moved
__global__ void k_mixed_permute_hist(int64_t np, const int* __restrict__ perm, const float* __restrict__ vp,
                                     const float* __restrict__ vb, float* __restrict__ vp_out,
                                     float* __restrict__ vb_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t j = perm[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vp_out[c * np + i] = vp[c * np + j];
        vb_out[c * np + i] = vb[c * np + j];
    }
}

__global__ void k_zero_int(int* __restrict__ x, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    if (threadIdx.x == 0) *x = 0;
}

// Sum over particles of |q| |v| (the gross current, times the cell volume
// and the number of nodes, of a uniform plasma).
__global__ __launch_bounds__(256) void k_mixed_gross(int64_t np, const float* __restrict__ q,
                                                     const float* __restrict__ v, double* __restrict__ part) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double sums[1] = {0.0};
    if (p < np) {
        const double a = v[p], b = v[np + p], c = v[2 * np + p];
        sums[0] = fabs((double)q[p]) * sqrt(a * a + b * b + c * c);
    }
    block_partials<1, 0>(sums, nullptr, 0, part);
}

}  // namespace

void launch_p3_map_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& args) {
    if (args.np <= 0) return;
    P3MArgs a = args;
    a.gate = current_gate();
    if (a.retry_n && !a.layout) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, a.retry_n, current_gate());
    MGrid mg;
    mg.nx = g.nx; mg.ny = g.ny; mg.N = g.N; mg.order = g.order;
    mg.hx = (float)g.h[0]; mg.hy = (float)g.h[1];
    mg.ihx = (float)(1.0 / g.h[0]); mg.ihy = (float)(1.0 / g.h[1]);
    int gm = (a.layout || (a.guess_mode == 2 && !a.mlp)) ? 0 : a.guess_mode;
    const bool col = a.train != nullptr && !a.layout;
    const size_t lds = gm == 2 ? 4 * (size_t)kMlpLdsWave : 0;
#define UPIC_MIXED_LAUNCH(R_)                                                                                       \
    do {                                                                                                            \
        const dim3 gr(mblk(a.np)), bl(256);                                                                         \
        if (a.layout) {                                                                                             \
            hipLaunchKernelGGL((k_p3_map_mixed<R_, 0, false, true>), gr, bl, 0, s, mg, q, split, a);               \
        } else if (col) {                                                                                           \
            if (gm == 2) hipLaunchKernelGGL((k_p3_map_mixed_collect<R_, 2>), gr, bl, lds, s, mg, q, split, a);     \
            else if (gm == 1) hipLaunchKernelGGL((k_p3_map_mixed_collect<R_, 1>), gr, bl, 0, s, mg, q, split, a);  \
            else hipLaunchKernelGGL((k_p3_map_mixed_collect<R_, 0>), gr, bl, 0, s, mg, q, split, a);               \
        } else {                                                                                                    \
            if (gm == 2) hipLaunchKernelGGL((k_p3_map_mixed<R_, 2, false, false>), gr, bl, lds, s, mg, q, split, a); \
            else if (gm == 1) hipLaunchKernelGGL((k_p3_map_mixed<R_, 1, false, false>), gr, bl, 0, s, mg, q, split, a); \
            else if (a.a0_in) hipLaunchKernelGGL((k_p3_map_mixed<R_, 3, false, false>), gr, bl, 0, s, mg, q, split, a); \
            else hipLaunchKernelGGL((k_p3_map_mixed<R_, 0, false, false>), gr, bl, 0, s, mg, q, split, a);         \
        }                                                                                                           \
    } while (0)
    switch (g.order) {
        case 1: UPIC_MIXED_LAUNCH(1); break;
        case 2: UPIC_MIXED_LAUNCH(2); break;
        case 3: UPIC_MIXED_LAUNCH(3); break;
        case 4: UPIC_MIXED_LAUNCH(4); break;
        default: break;
    }
#undef UPIC_MIXED_LAUNCH
    launch_finish_partials(s, a.part, mblk(a.np), 2, 2, a.red, a.redmax, a.layout ? a.stats : nullptr);
    if (a.retry_n && !a.layout) {
        // the deferred particles; their partials reuse the buffer after the
        // fold above
        const unsigned nb = std::min<unsigned>(mblk(a.np), 1024u);
#define UPIC_RETRY_LAUNCH(R_)                                                                            \
    do {                                                                                                 \
        if (col) hipLaunchKernelGGL((k_p3_map_retry<R_, true>), dim3(nb), dim3(256), 0, s, mg, q, split, a);  \
        else hipLaunchKernelGGL((k_p3_map_retry<R_, false>), dim3(nb), dim3(256), 0, s, mg, q, split, a);    \
    } while (0)
        switch (g.order) {
            case 1: UPIC_RETRY_LAUNCH(1); break;
            case 2: UPIC_RETRY_LAUNCH(2); break;
            case 3: UPIC_RETRY_LAUNCH(3); break;
            case 4: UPIC_RETRY_LAUNCH(4); break;
            default: break;
        }
#undef UPIC_RETRY_LAUNCH
        launch_finish_partials(s, a.part, nb, 2, 2, a.red, a.redmax, nullptr);
    }
}

void launch_mixed_gross(hipStream_t s, int64_t np, const MixedState& ms, double* part, double* out) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_mixed_gross, dim3(mblk(np)), dim3(256), 0, s, np, ms.q, ms.v, part);
    launch_finish_partials(s, part, mblk(np), 1, 0, out, nullptr, nullptr);
}

void launch_deposit_delta_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& args,
                                double* out) {
    if (args.np <= 0 || !args.retry_n || !args.retry_old) return;
    P3MArgs a = args;
    a.gate = current_gate();
    MGrid mg;
    mg.nx = g.nx; mg.ny = g.ny; mg.N = g.N; mg.order = g.order;
    mg.hx = (float)g.h[0]; mg.hy = (float)g.h[1];
    mg.ihx = (float)(1.0 / g.h[0]); mg.ihy = (float)(1.0 / g.h[1]);
    // at most kDeltaBlocks workgroups striding over the queue (its length is
    // on the device): a short queue costs a small grid and a small fold
    constexpr unsigned kDeltaBlocks = 1024;
    const unsigned nb = std::min(mblk(a.np), kDeltaBlocks);
    switch (g.order) {
        case 1: hipLaunchKernelGGL(k_deposit_delta_mixed<1>, dim3(nb), dim3(256), 0, s, mg, q, split, a); break;
        case 2: hipLaunchKernelGGL(k_deposit_delta_mixed<2>, dim3(nb), dim3(256), 0, s, mg, q, split, a); break;
        case 3: hipLaunchKernelGGL(k_deposit_delta_mixed<3>, dim3(nb), dim3(256), 0, s, mg, q, split, a); break;
        case 4: hipLaunchKernelGGL(k_deposit_delta_mixed<4>, dim3(nb), dim3(256), 0, s, mg, q, split, a); break;
        default: break;
    }
    launch_finish_partials(s, a.part, nb, 2, 1, out, out + 2, nullptr);
}

// 0: wavefront-reduced deposit on an 8 x 32 tile, 1 and 2: cell-mapped
// deposits (1 or kCellSlices workgroups per 64 cells), 3: per-lane atomics,
// 4 and 5: the wavefront-reduced deposit on 4 x 64 and 16 x 16 tiles
int deposit_variants() { return 8; }
bool deposit_variant_usable(const GridDev& g, int variant, bool cells_fresh) {
    if (variant == 1 || variant == 2) return cells_fresh && g.ny % 64 == 0 && g.nx >= kCTR;
    // 8 x 8 blocks; the 16 x 16 tile must not wrap onto itself
    if (variant == 6 || variant == 7) return cells_fresh && g.nx % 8 == 0 && g.ny % 8 == 0 && g.nx >= 16 && g.ny >= 16;
    return true;
}

void launch_deposit_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& args,
                          const int* cell_end, int variant) {
    if (args.np <= 0 || args.layout || !args.J_out) return;
    P3MArgs a = args;
    a.gate = current_gate();
    if (!deposit_variant_usable(g, variant, cell_end != nullptr)) variant = 0;
    MGrid mg;
    mg.nx = g.nx; mg.ny = g.ny; mg.N = g.N; mg.order = g.order;
    mg.hx = (float)g.h[0]; mg.hy = (float)g.h[1];
    mg.ihx = (float)(1.0 / g.h[0]); mg.ihy = (float)(1.0 / g.h[1]);
    if (variant == 0 || variant == 4 || variant == 5) {
        const dim3 gr(mblk(a.np)), bl(256);
        const int tx = variant == 4 ? 4 : variant == 5 ? 16 : kTileX;
        const int ty = kTileX * kTileY / tx;
        switch (g.order) {
            case 1: hipLaunchKernelGGL(k_deposit_wave_mixed<1>, gr, bl, 0, s, mg, q, split, a, tx, ty); break;
            case 2: hipLaunchKernelGGL(k_deposit_wave_mixed<2>, gr, bl, 0, s, mg, q, split, a, tx, ty); break;
            case 3: hipLaunchKernelGGL(k_deposit_wave_mixed<3>, gr, bl, 0, s, mg, q, split, a, tx, ty); break;
            case 4: hipLaunchKernelGGL(k_deposit_wave_mixed<4>, gr, bl, 0, s, mg, q, split, a, tx, ty); break;
            default: break;
        }
        launch_finish_partials(s, a.part, mblk(a.np), 0, 0, nullptr, nullptr, a.stats);
        return;
    }
    if (variant == 1 || variant == 2 || variant == 6 || variant == 7) {
        const int nb = g.N / 64;
        // the partial slots of nb * slices workgroups must fit the
        // per-particle-block buffer (sized for ceil(np / 256) blocks)
        const int want = (variant == 1 || variant == 6) ? 1 : kCellSlices;
        const int sl = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)mblk(a.np) / std::max(nb, 1)));
        const int bw = variant >= 6 ? 8 : 1;
        const dim3 gr(nb, sl), bl(64 * kCellWaves);
        switch (g.order) {
            case 1: hipLaunchKernelGGL(k_deposit_cells_mixed<1>, gr, bl, 0, s, mg, q, split, a, cell_end, bw); break;
            case 2: hipLaunchKernelGGL(k_deposit_cells_mixed<2>, gr, bl, 0, s, mg, q, split, a, cell_end, bw); break;
            case 3: hipLaunchKernelGGL(k_deposit_cells_mixed<3>, gr, bl, 0, s, mg, q, split, a, cell_end, bw); break;
            case 4: hipLaunchKernelGGL(k_deposit_cells_mixed<4>, gr, bl, 0, s, mg, q, split, a, cell_end, bw); break;
            default: break;
        }
        launch_finish_partials(s, a.part, (int64_t)nb * sl, 0, 0, nullptr, nullptr, a.stats);
        return;
    }
    const dim3 gr(mblk(a.np)), bl(256);
    switch (g.order) {
        case 1: hipLaunchKernelGGL(k_deposit_mixed<1>, gr, bl, 0, s, mg, q, split, a); break;
        case 2: hipLaunchKernelGGL(k_deposit_mixed<2>, gr, bl, 0, s, mg, q, split, a); break;
        case 3: hipLaunchKernelGGL(k_deposit_mixed<3>, gr, bl, 0, s, mg, q, split, a); break;
        case 4: hipLaunchKernelGGL(k_deposit_mixed<4>, gr, bl, 0, s, mg, q, split, a); break;
        default: break;
    }
    launch_finish_partials(s, a.part, mblk(a.np), 0, 0, nullptr, nullptr, a.stats);
}

void launch_pack_nodes_f32(hipStream_t s, int N, const double* gpm, const double* A_n, const double* A_end, float* NR) {
    hipLaunchKernelGGL(k_pack_nodes_f32, dim3(mblk(N)), dim3(256), 0, s, N, gpm, A_n, A_end,
                       reinterpret_cast<float4*>(NR), current_gate());
}

void launch_mixed_compress(hipStream_t s, const GridDev& g, int64_t np, const double* x, const double* P,
                           const double* v, const double* vp, const double* vb, double* q, double* m,
                           const MixedState& ms) {
    if (np > 0) hipLaunchKernelGGL(k_mixed_compress, dim3(mblk(np)), dim3(256), 0, s, g, np, x, P, v, vp, vb, q, m, ms);
}

void launch_mixed_expand(hipStream_t s, const GridDev& g, int64_t np, const MixedState& ms, double* x, double* P,
                         double* v, double* vp, double* vb, double* q, double* m) {
    if (np > 0) hipLaunchKernelGGL(k_mixed_expand, dim3(mblk(np)), dim3(256), 0, s, g, np, ms, x, P, v, vp, vb, q, m);
}

void launch_mixed_keys(hipStream_t s, const GridDev& g, int64_t np, const MixedState& ms, double dt, int split,
                       int* keys, int* counts) {
    if (np <= 0) return;
    MGrid mg;
    mg.nx = g.nx; mg.ny = g.ny; mg.N = g.N; mg.order = g.order;
    mg.hx = (float)g.h[0]; mg.hy = (float)g.h[1];
    mg.ihx = (float)(1.0 / g.h[0]); mg.ihy = (float)(1.0 / g.h[1]);
    hipLaunchKernelGGL(k_mixed_keys, dim3(mblk(np)), dim3(256), 0, s, mg, np, ms.cell, ms.off, ms.vb, (float)dt, split,
                       keys, counts);
}

void launch_mixed_permute(hipStream_t s, int64_t np, const int* perm, const MixedState& src, const MixedState& dst,
                          const int64_t* id_src, int64_t* id_dst, bool hist) {
    if (np <= 0) return;
    if (hist) hipLaunchKernelGGL(k_mixed_permute<true>, dim3(mblk(np)), dim3(256), 0, s, np, perm, src, dst, id_src, id_dst);
    else hipLaunchKernelGGL(k_mixed_permute<false>, dim3(mblk(np)), dim3(256), 0, s, np, perm, src, dst, id_src, id_dst);
}
void launch_mixed_permute_hist(hipStream_t s, int64_t np, const int* perm, const float* vp, const float* vb,
                               float* vp_out, float* vb_out) {
    if (np > 0)
        hipLaunchKernelGGL(k_mixed_permute_hist, dim3(mblk(np)), dim3(256), 0, s, np, perm, vp, vb, vp_out, vb_out);
}

}  // namespace upic
