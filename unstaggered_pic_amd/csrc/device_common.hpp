// device_common.hpp -- device-side numerics shared by the particle kernels:
// grid conventions, wrapping, centred B-splines, per-axis stencils, and the
// on-the-fly orbit breakpoint generator.
//
// Conventions follow the three reference codes (paths relative to the
// reference root):
//   P1 Project1/ec_pic_serial_demo.cpp, P2 Project2/spectral_pic_part3.cpp,
//   P3 Project3/ec_pic_nyquist_projected.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace upic {

constexpr int kMaxQuad = 32;

// slots of the per-block partials of a particle sweep (particles.hpp)
constexpr int kPartSlots = 9;

// Claims slots of per-key counters for the active lanes of a wavefront:
// lanes with equal keys are grouped by ballot, one atomicAdd per distinct
// key reserves the group's slots, and each lane receives old value + its
// rank among the group's lanes (lane order).  Particles stored in cell order
// put one or two keys in a wavefront, so the counters of a cell see a few
// atomics instead of one per particle.  Every lane of the wavefront must
// call it (inactive lanes pass active = false and get 0).
__device__ __forceinline__ int wave_claim(int* __restrict__ ctr, int key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    int slot = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lk = __shfl(key, leader, 64);
        const unsigned long long grp = __ballot(active && key == lk) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lk], __popcll(grp));
        base = __shfl(base, leader, 64);
        if ((grp >> lane) & 1ull) slot = base + __popcll(grp & ((1ull << lane) - 1ull));
        pending &= ~grp;
    }
    return slot;
}

// Target of the deterministic (fixed-point) current deposit: two 64-bit
// integer words per mesh entry and the scale 2^s (fix_add, particles.hpp).
struct FixAcc {
    unsigned long long* hi;
    unsigned long long* lo;
    double scale;
};
constexpr double kFixLoScale = 1099511627776.0;  // 2^40

// A gated kernel returns at once when its gate (device int) reads 0.
#define UPIC_GATE_CLOSED(g) ((g) != nullptr && *(g) == 0)

// Grid description passed by value to every kernel.
struct GridDev {
    int nx, ny, nz, N;
    double L[3];
    double h[3];       // cell widths L/n
    double half_L[3];  // 0.5 L
    double dV;
    int code;          // 1, 2, 3 (upic_code)
    int order;         // spline degree r
};

struct QuadDev {
    int n;
    double s[kMaxQuad];
    double w[kMaxQuad];
};

// The far branch of wrap1 (|x| beyond 3L/2): the software fmod, kept out of
// line.  Inlined at every wrap of an unrolled orbit pass it made the FP64
// cubic sweep a 19K-instruction kernel spilling 468 bytes per lane (15K and
// 272 bytes with the call, 7% faster); the call costs nothing on the path
// every in-range particle takes.
static __device__ __attribute__((noinline)) double wrap_far(double y, double L) {
    y = fmod(y, L);
    return y < 0.0 ? y + L : y;
}

// wrap1 (P1:160-164, P3:704-708): into [-L/2, L/2).  The reference uses
// fmod(x + L/2, L); for x + L/2 in [-L, 2L) fmod returns y, y - L or (after
// its y += L fix-up) y + L, which the branch below reproduces bit for bit
// without the slow software fmod.  Farther values take fmod itself.
__device__ __forceinline__ double wrap1(double x, double L) {
    double y = x + 0.5 * L;
    if (y >= L) {
        y = (y < 2.0 * L) ? y - L : wrap_far(y, L);
    } else if (y < 0.0) {
        y = (y >= -L) ? y + L : wrap_far(y, L);
    }
    return y - 0.5 * L;
}

// Centred B-spline of degree R and its derivative in closed form
// (P3:1856-1947).  P1 evaluates the same piecewise polynomials by the
// Cox-de Boor recursion (P1:552-572); the two agree to roundoff.
template <int R>
__device__ __forceinline__ void bspline_wd(double r, double& w, double& dw) {
    if constexpr (R == 1) {
        if (r >= -1.0 && r < 0.0) { w = 1.0 + r; dw = 1.0; }
        else if (r >= 0.0 && r < 1.0) { w = 1.0 - r; dw = -1.0; }
        else { w = 0.0; dw = 0.0; }
    } else {
        const double a = fabs(r);
        const double sign = (r < 0.0) ? -1.0 : 1.0;
        if constexpr (R == 2) {
            if (a < 0.5) { w = 0.75 - a * a; dw = -2.0 * r; }
            else if (a < 1.5) { const double z = 1.5 - a; w = 0.5 * z * z; dw = -sign * z; }
            else { w = 0.0; dw = 0.0; }
        } else if constexpr (R == 3) {
            if (a < 1.0) { w = (2.0 / 3.0) - a * a + 0.5 * a * a * a; dw = sign * (-2.0 * a + 1.5 * a * a); }
            else if (a < 2.0) { const double z = 2.0 - a; w = z * z * z / 6.0; dw = -0.5 * sign * z * z; }
            else { w = 0.0; dw = 0.0; }
        } else {
            if (a < 0.5) {
                const double a2 = a * a;
                w = 115.0 / 192.0 - (5.0 / 8.0) * a2 + 0.25 * a2 * a2;
                dw = sign * (-(5.0 / 4.0) * a + a * a * a);
            } else if (a < 1.5) {
                w = 55.0 / 96.0 + (5.0 / 24.0) * a - 1.25 * a * a + (5.0 / 6.0) * a * a * a - (1.0 / 6.0) * a * a * a * a;
                dw = sign * (5.0 / 24.0 - 2.5 * a + 2.5 * a * a - (2.0 / 3.0) * a * a * a);
            } else if (a < 2.5) {
                const double z = 2.5 - a;
                const double z2 = z * z;
                w = z2 * z2 / 24.0;
                dw = -sign * z * z * z / 6.0;
            } else { w = 0.0; dw = 0.0; }
        }
    }
}

// One axis of the tensor-product stencil (axis_stencil, P1:580-595,
// P3:1971-1986): nodes at -L/2 + i h, i_left = floor(u - (r-1)/2).
template <int R>
struct AxisSt {
    int base;          // wrapped index of node a = 0
    double w[R + 1];
    double dw[R + 1];  // d/dx (already divided by h)
};

//
// Quadratic and cubic evaluate, for each node, the one piece its distance
// r falls in (node a's |r| lies in a fixed piece for every particle) with
// the same operations bspline_wd applies to that piece, so the weights are
// the same bits without its per-node compare-and-select of every piece
// (the one exception: a particle exactly on a knot, where one cubic weight
// comes from the adjacent piece, equal to within one rounding).
template <int R>
__device__ __forceinline__ void axis_stencil(double x, int n, double L, double h, AxisSt<R>& s) {
    const double u = (x + 0.5 * L) / h;
    const int i_left = (int)floor(u - 0.5 * (double)(R - 1));
    int b = i_left % n;
    if (b < 0) b += n;
    s.base = b;
    if constexpr (R == 3) {
        // r = 1+f, f, f-1, f-2 with f in [0, 1): pieces 2, 1, 1, 2
        const double a0 = u - (double)i_left, a1 = u - (double)(i_left + 1);
        const double a2 = fabs(u - (double)(i_left + 2)), a3 = fabs(u - (double)(i_left + 3));
        const double z0 = 2.0 - a0, z3 = 2.0 - a3;
        s.w[0] = z0 * z0 * z0 / 6.0;
        s.dw[0] = -0.5 * z0 * z0 / h;
        s.w[1] = (2.0 / 3.0) - a1 * a1 + 0.5 * a1 * a1 * a1;
        s.dw[1] = (-2.0 * a1 + 1.5 * a1 * a1) / h;
        s.w[2] = (2.0 / 3.0) - a2 * a2 + 0.5 * a2 * a2 * a2;
        s.dw[2] = -(-2.0 * a2 + 1.5 * a2 * a2) / h;
        s.w[3] = z3 * z3 * z3 / 6.0;
        s.dw[3] = 0.5 * z3 * z3 / h;
    } else if constexpr (R == 2) {
        // r = 1/2+f, f-1/2, f-3/2 with f in [0, 1): pieces 2, 1, 2
        const double r0 = u - (double)i_left, r1 = u - (double)(i_left + 1);
        const double a2 = fabs(u - (double)(i_left + 2));
        const double z0 = 1.5 - r0, z2 = 1.5 - a2;
        s.w[0] = 0.5 * z0 * z0;
        s.dw[0] = -z0 / h;
        s.w[1] = 0.75 - fabs(r1) * fabs(r1);
        s.dw[1] = -2.0 * r1 / h;
        s.w[2] = 0.5 * z2 * z2;
        s.dw[2] = z2 / h;
    } else {
#pragma unroll
        for (int a = 0; a <= R; ++a) {
            const double r = u - (double)(i_left + a);
            double w, dw;
            bspline_wd<R>(r, w, dw);
            s.w[a] = w;
            s.dw[a] = dw / h;
        }
    }
}

// Index of node base+a wrapped into [0, n).
__device__ __forceinline__ int wrap_idx(int i, int n) { return (i >= n) ? i - n : i; }

// ---------------------------------------------------------------------------
// Orbit breakpoints (path_breakpoints_for_particle, P1:782-833,
// P3:2555-2606) generated in increasing order without storage: each axis'
// knot crossings are monotone in s, so a three-way merge yields the sorted
// list.  The tolerance filter (1e-13) and the 1e-12 de-duplication against
// the last kept breakpoint are the reference's.
// ---------------------------------------------------------------------------
struct AxisCross {
    double a0, du;
    long long m, m_end;  // next knot index and one-past-last
    int step;            // +1 or -1
    bool active;
};

struct BreakGen {
    AxisCross ax[3];
    double last;   // last kept breakpoint
    bool done;

    __device__ __forceinline__ double peek_axis(int d) const {
        const AxisCross& c = ax[d];
        if (!c.active || c.m == c.m_end) return 2.0;
        return ((double)c.m - c.a0) / c.du;
    }
};

// x0 is the start point (unwrapped is fine; wrapped internally as in the
// reference), dxp the path displacement.  ndim = 3 always: axes with n = 1
// still produce crossings exactly as in the reference, which only changes
// how the path is cut, not the integrand.
__device__ __forceinline__ void breakgen_init(BreakGen& bg, const GridDev& g, const double x0[3], const double dxp[3],
                                              bool split) {
    bg.last = 0.0;
    bg.done = false;
    const double offset = 0.5 * (double)(g.order - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        AxisCross& c = bg.ax[d];
        c.active = false;
        c.m = 0; c.m_end = 0; c.step = 1; c.a0 = 0.0; c.du = 1.0;
        if (!split) continue;
        const double Ld = g.L[d], hd = g.h[d];
        const double u0 = (wrap1(x0[d], Ld) + 0.5 * Ld) / hd;
        const double du = dxp[d] / hd;
        if (fabs(du) < 1.0e-300) continue;
        const double a0 = u0 - offset;
        const double a1 = a0 + du;
        const double lo = fmin(a0, a1), hi = fmax(a0, a1);
        const long long first = (long long)floor(lo) + 1;
        const long long last = (long long)ceil(hi) - 1;
        if (last < first) continue;
        c.active = true;
        c.a0 = a0;
        c.du = du;
        if (du > 0.0) { c.m = first; c.m_end = last + 1; c.step = 1; }
        else { c.m = last; c.m_end = first - 1; c.step = -1; }
    }
}

// Returns the next breakpoint in (last, 1]; 1.0 terminates.
__device__ __forceinline__ double breakgen_next(BreakGen& bg) {
    const double tol = 1.0e-13;
    for (;;) {
        const double s0 = bg.peek_axis(0), s1 = bg.peek_axis(1), s2 = bg.peek_axis(2);
        int d = 0;
        double s = s0;
        if (s1 < s) { s = s1; d = 1; }
        if (s2 < s) { s = s2; d = 2; }
        if (s > 1.5) { bg.done = true; bg.last = 1.0; return 1.0; }
        // advance the chosen axis with constant indices only: an index known
        // only at run time would keep the whole generator in scratch memory
        if (d == 0) bg.ax[0].m += bg.ax[0].step;
        else if (d == 1) bg.ax[1].m += bg.ax[1].step;
        else bg.ax[2].m += bg.ax[2].step;
        if (!(s > tol && s < 1.0 - tol)) continue;
        if (fabs(s - bg.last) > 1.0e-12) { bg.last = s; return s; }
    }
}

}  // namespace upic
