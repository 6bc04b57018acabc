// particles.hpp -- per-particle orbit integration on the GPU: point stencils,
// point gather/deposit, and the orbit traversal shared by the deposit,
// gather, orbit-discrete-gradient and fused push kernels.
#pragma once

#include "device_common.hpp"

namespace upic {

// Launch-time options common to the particle kernels.
struct OrbitOpt {
    int split;      // split_orbit_at_knots
    int wrap_path;  // P1 wraps each path point before the stencil (P1:778-781)
};

template <int R, bool Z3>
struct PointSt {
    AxisSt<R> s[Z3 ? 3 : 2];
};

// Stencil at a point; the position is wrapped exactly as wrap_pos
// (P1:166-168, P3:723-725) before axis_stencil.
template <int R, bool Z3>
__device__ __forceinline__ void point_stencil(const GridDev& g, const double x[3], PointSt<R, Z3>& st) {
    axis_stencil<R>(wrap1(x[0], g.L[0]), g.nx, g.L[0], g.h[0], st.s[0]);
    axis_stencil<R>(wrap1(x[1], g.L[1]), g.ny, g.L[1], g.h[1], st.s[1]);
    if constexpr (Z3) axis_stencil<R>(wrap1(x[2], g.L[2]), g.nz, g.L[2], g.h[2], st.s[2]);
}

// Calls f(id, wxyz, dS[3]) for each of the (R+1)^d stencil nodes.  For 2-D
// grids (nz = 1) the z factor is the partition of unity (weights summing to
// one, derivative zero), which the reference's (R+1)^3 loop reproduces to
// roundoff.
template <int R, bool Z3, class F>
__device__ __forceinline__ void for_each_node(const GridDev& g, const PointSt<R, Z3>& st, F&& f) {
#pragma unroll
    for (int a = 0; a <= R; ++a) {
        const int ia = wrap_idx(st.s[0].base + a, g.nx);
#pragma unroll
        for (int b = 0; b <= R; ++b) {
            const int ib = wrap_idx(st.s[1].base + b, g.ny);
            const double wx = st.s[0].w[a], wy = st.s[1].w[b];
            const double dwx = st.s[0].dw[a], dwy = st.s[1].dw[b];
            if constexpr (Z3) {
#pragma unroll
                for (int c = 0; c <= R; ++c) {
                    const int ic = wrap_idx(st.s[2].base + c, g.nz);
                    const double wz = st.s[2].w[c], dwz = st.s[2].dw[c];
                    const double dS[3] = {dwx * wy * wz, wx * dwy * wz, wx * wy * dwz};
                    f((ia * g.ny + ib) * g.nz + ic, wx * wy * wz, dS);
                }
            } else {
                const double dS[3] = {dwx * wy, wx * dwy, 0.0};
                f(ia * g.ny + ib, wx * wy, dS);
            }
        }
    }
}

// gather_vector_at (P1:666-672, P3:2267-2286)
template <int R, bool Z3>
__device__ __forceinline__ void gather3(const GridDev& g, const double* __restrict__ F, const double x[3], double out[3]) {
    PointSt<R, Z3> st;
    point_stencil<R, Z3>(g, x, st);
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;
    const int N = g.N;
    for_each_node<R, Z3>(g, st, [&](int id, double w, const double*) {
        o0 += F[id] * w;
        o1 += F[N + id] * w;
        o2 += F[2 * N + id] * w;
    });
    out[0] = o0; out[1] = o1; out[2] = o2;
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// Exact fixed-point accumulation of a current (the deterministic deposit,
// which the reference obtains from per-thread meshes summed in a fixed order,
// P3:2194-2216): a value x enters as two integers, hi = floor(x 2^s) and lo =
// the remaining fraction times 2^40, each added with a 64-bit integer atomic.
// Integer sums do not depend on the order of the additions, so the result is
// bitwise reproducible.  With 2^s chosen so that every node sum stays below
// 2^61 / 2^s in magnitude, neither word overflows (lo < 2^40 per term), and
// the part of each value below 2^-(s+40) that is dropped is far below FP64
// resolution of the node sums (FixAcc, device_common.hpp).
__device__ __forceinline__ void fix_add(unsigned long long* hi, unsigned long long* lo, double x, double scale) {
    const double y = x * scale;  // exact: a power-of-two scaling
    const double f = floor(y);
    const unsigned long long h = (unsigned long long)(long long)f;
    const unsigned long long l = (unsigned long long)(long long)((y - f) * kFixLoScale);  // y - f exact, in [0, 1)
    if (h) atomicAdd(hi, h);
    if (l) atomicAdd(lo, l);
}

// Orbit traversal: for each knot-split segment (P1:858-906, P3:2689-2857)
// and each Gauss point, calls f(s, w, xq) with xq the path point (wrapped
// for P1).  Returns the number of segments (br.size() - 1).
template <class F>
__device__ __forceinline__ int for_each_orbit_point(const GridDev& g, const QuadDev& q, const OrbitOpt& opt,
                                                    const double x0[3], const double dxp[3], F&& f) {
    BreakGen bg;
    breakgen_init(bg, g, x0, dxp, opt.split != 0);
    double a = 0.0;
    int nseg = 0;
    for (;;) {
        const double b = opt.split ? breakgen_next(bg) : 1.0;
        const double len = b - a;
        if (len > 0.0) {
            for (int iq = 0; iq < q.n; ++iq) {
                const double s = a + len * q.s[iq];
                const double w = len * q.w[iq];
                double xq[3] = {x0[0] + s * dxp[0], x0[1] + s * dxp[1], x0[2] + s * dxp[2]};
                if (opt.wrap_path) {
                    xq[0] = wrap1(xq[0], g.L[0]);
                    xq[1] = wrap1(xq[1], g.L[1]);
                    xq[2] = wrap1(xq[2], g.L[2]);
                }
                f(s, w, xq);
            }
        }
        ++nseg;
        a = b;
        if (b >= 1.0) break;
    }
    return nseg;
}

// Orbit-averaged current deposit of one particle (path_average_deposit_current
// P1:883-906, P3:2784-2857): J[d] += w q vel[d] S / dV at every Gauss point.
// Node weights of a Gauss point are accumulated per segment in registers while
// the stencil base is unchanged and flushed with one atomic per node and
// component; a point whose stencil differs (a knot met by roundoff, or
// unsplit orbits) is flushed on its own.
template <int R, bool Z3>
__device__ __forceinline__ int deposit_orbit_one(const GridDev& g, const QuadDev& q, const OrbitOpt& opt,
                                                 const double x0[3], const double dxp[3], const double vel[3],
                                                 double charge, double* __restrict__ J, FixAcc fx = FixAcc{}) {
    constexpr int NZ = Z3 ? R + 1 : 1;
    double W[R + 1][R + 1][NZ];
    int base[3] = {-1, -1, -1};
    bool have = false;
    const double qs = charge / g.dV;
    const int N = g.N;
    auto flush = [&]() {
        if (!have) return;
#pragma unroll
        for (int a = 0; a <= R; ++a) {
            const int ia = wrap_idx(base[0] + a, g.nx);
#pragma unroll
            for (int b = 0; b <= R; ++b) {
                const int ib = wrap_idx(base[1] + b, g.ny);
#pragma unroll
                for (int c = 0; c < NZ; ++c) {
                    const int ic = Z3 ? wrap_idx(base[2] + c, g.nz) : 0;
                    const int id = (ia * g.ny + ib) * g.nz + ic;
                    const double w = W[a][b][c] * qs;
                    if (fx.hi) {
                        fix_add(fx.hi + id, fx.lo + id, w * vel[0], fx.scale);
                        fix_add(fx.hi + N + id, fx.lo + N + id, w * vel[1], fx.scale);
                        fix_add(fx.hi + 2 * N + id, fx.lo + 2 * N + id, w * vel[2], fx.scale);
                        continue;
                    }
                    atomic_add(&J[id], w * vel[0]);
                    atomic_add(&J[N + id], w * vel[1]);
                    atomic_add(&J[2 * N + id], w * vel[2]);
                }
            }
        }
        have = false;
    };
    const int nseg = for_each_orbit_point(g, q, opt, x0, dxp, [&](double, double w, const double* xq) {
        PointSt<R, Z3> st;
        point_stencil<R, Z3>(g, xq, st);
        const bool same = have && st.s[0].base == base[0] && st.s[1].base == base[1] && (!Z3 || st.s[Z3 ? 2 : 0].base == base[2]);
        if (!same) {
            flush();
            base[0] = st.s[0].base;
            base[1] = st.s[1].base;
            if constexpr (Z3) base[2] = st.s[2].base;
#pragma unroll
            for (int a = 0; a <= R; ++a)
#pragma unroll
                for (int b = 0; b <= R; ++b)
#pragma unroll
                    for (int c = 0; c < NZ; ++c) W[a][b][c] = 0.0;
            have = true;
        }
#pragma unroll
        for (int a = 0; a <= R; ++a)
#pragma unroll
            for (int b = 0; b <= R; ++b) {
                const double wxy = w * st.s[0].w[a] * st.s[1].w[b];
                if constexpr (Z3) {
#pragma unroll
                    for (int c = 0; c <= R; ++c) W[a][b][c] += wxy * st.s[2].w[c];
                } else {
                    W[a][b][0] += wxy;
                }
            }
    });
    flush();
    return nseg;
}

// Fused orbit average of a vector field and orbit-discrete gradient of A
// (path_average_gather_and_gradient P3:3016-3071; orbit_discrete_gradient_A
// P1:908-934 / _interval P3:2958-2984).  theta(s) = th0 + (th1 - th0) s;
// p1_theta uses theta = s exactly as P1 does.  Either output may be skipped.
template <int R, bool Z3, bool WANT_GATHER, bool WANT_GRAD>
__device__ __forceinline__ int orbit_gather_grad(const GridDev& g, const QuadDev& q, const OrbitOpt& opt,
                                                 const double x0[3], const double dxp[3],
                                                 const double* __restrict__ F, const double* __restrict__ A0,
                                                 const double* __restrict__ A1, double th0, double th1, bool p1_theta,
                                                 double gat[3], double D[3][3]) {
    const int N = g.N;
    if constexpr (WANT_GATHER) { gat[0] = gat[1] = gat[2] = 0.0; }
    if constexpr (WANT_GRAD) {
#pragma unroll
        for (int i = 0; i < 3; ++i) D[i][0] = D[i][1] = D[i][2] = 0.0;
    }
    return for_each_orbit_point(g, q, opt, x0, dxp, [&](double s, double w, const double* xq) {
        PointSt<R, Z3> st;
        point_stencil<R, Z3>(g, xq, st);
        const double theta = p1_theta ? s : th0 + (th1 - th0) * s;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        double G[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        for_each_node<R, Z3>(g, st, [&](int id, double wn, const double* dS) {
            if constexpr (WANT_GATHER) {
                v0 += F[id] * wn;
                v1 += F[N + id] * wn;
                v2 += F[2 * N + id] * wn;
            }
            if constexpr (WANT_GRAD) {
#pragma unroll
                for (int comp = 0; comp < 3; ++comp) {
                    const double Av = (1.0 - theta) * A0[comp * N + id] + theta * A1[comp * N + id];
                    G[comp][0] += Av * dS[0];
                    G[comp][1] += Av * dS[1];
                    if constexpr (Z3) G[comp][2] += Av * dS[2];
                }
            }
        });
        if constexpr (WANT_GATHER) { gat[0] += w * v0; gat[1] += w * v1; gat[2] += w * v2; }
        if constexpr (WANT_GRAD) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) D[a][b] += w * G[a][b];
        }
    });
}

// Wavefront + block reduction of a few doubles into global accumulators.
template <int K>
__device__ __forceinline__ void block_reduce_add(double v[K], double* __restrict__ out) {
    __shared__ double red[K][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int i = 0; i < nw; ++i) s += red[k][i];
            atomic_add(&out[k], s);
        }
    }
}

template <int K>
__device__ __forceinline__ void block_reduce_max(double v[K], double* __restrict__ out) {
    __shared__ double redm[K][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
        if (lane == 0) redm[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = redm[k][0];
            for (int i = 1; i < nw; ++i) s = fmax(s, redm[k][i]);
            // non-negative doubles order like their bit patterns
            atomicMax(reinterpret_cast<unsigned long long*>(&out[k]), (unsigned long long)__double_as_longlong(s));
        }
    }
}

// Per-block partial results of a particle sweep.  Tens of thousands of
// blocks adding into the same few global words serialise in L2 (about a
// million same-address atomics per 16.8M-particle sweep), so each block
// writes its partials to its own slots and launch_finish_partials combines
// them.  Slot layout per block (kPartSlots doubles): [0, 4) sums, [4, 6)
// maxima, [6] particles with more than one path segment, [7] total
// segments, [8] maximum segments.
template <int NS, int NM>
__device__ __forceinline__ void block_partials_stats(const double* sums, const double* maxs, double cross,
                                                     double segs, double maxseg, double* __restrict__ part);

template <int NS, int NM>
__device__ __forceinline__ void block_partials(const double* sums, const double* maxs, int nseg,
                                               double* __restrict__ part) {
    block_partials_stats<NS, NM>(sums, maxs, nseg > 1 ? 1.0 : 0.0, (double)nseg, (double)nseg, part);
}

// as block_partials, with the thread's path statistics given directly
// (particles with more than one segment, total segments, maximum segments)
template <int NS, int NM>
__device__ __forceinline__ void block_partials_stats(const double* sums, const double* maxs, double cross,
                                                     double segs, double maxseg, double* __restrict__ part) {
    static_assert(NS <= 4 && NM <= 2, "partial slots");
    __shared__ double sh[kPartSlots][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double v[kPartSlots];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = k < NS ? sums[k] : 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) v[4 + k] = k < NM ? maxs[k] : 0.0;
    v[6] = cross;
    v[7] = segs;
    v[8] = maxseg;
#pragma unroll
    for (int k = 0; k < kPartSlots; ++k) {
        double x = v[k];
        const bool mx = (k == 4 || k == 5 || k == 8);
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_down(x, o, 64);
            x = mx ? fmax(x, y) : x + y;
        }
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < kPartSlots) {
        const int k = threadIdx.x;
        const bool mx = (k == 4 || k == 5 || k == 8);
        const int nw = (blockDim.x + 63) >> 6;
        double x = sh[k][0];
        for (int i = 1; i < nw; ++i) x = mx ? fmax(x, sh[k][i]) : x + sh[k][i];
        part[(size_t)blockIdx.x * kPartSlots + k] = x;
    }
}

// Path statistics accumulator: [0] crossing particles, [1] total segments,
// [2] max segments (PathStats, P1:771-776).
__device__ __forceinline__ void path_stats_add(int nseg, unsigned long long* __restrict__ stats) {
    if (!stats) return;
    const int lane = threadIdx.x & 63;
    unsigned long long cross = nseg > 1 ? 1ull : 0ull, segs = (unsigned long long)nseg;
    int mx = nseg;
    for (int o = 32; o > 0; o >>= 1) {
        cross += __shfl_down(cross, o, 64);
        segs += __shfl_down(segs, o, 64);
        mx = max(mx, __shfl_down(mx, o, 64));
    }
    if (lane == 0) {
        atomicAdd(&stats[0], cross);
        atomicAdd(&stats[1], segs);
        atomicMax(&stats[2], (unsigned long long)mx);
    }
}

}  // namespace upic
