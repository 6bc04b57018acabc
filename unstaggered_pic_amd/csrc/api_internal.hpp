// api_internal.hpp -- host-side internals shared by the C-ABI translation
// units: error type, device allocation, the context object.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/upic.h"
#include "kernels.hpp"
#include "particles.hpp"

using namespace upic;

// message of the last failed call that has no context (upic_last_error(NULL))
extern thread_local std::string g_create_err;

namespace upic_host {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw Error(UPIC_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
inline T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
    if (e != hipSuccess) throw Error(UPIC_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return p;
}
template <class T>
inline void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Gauss-Legendre on [0,1] (gauss_legendre_01, P1:739-763), evaluated in long
// double so the nodes are correctly rounded.
inline void gauss_legendre_01(int n, double* s, double* w) {
    const long double PI = 3.141592653589793238462643383279502884L;
    const int m = (n + 1) / 2;
    for (int i = 0; i < m; ++i) {
        long double x = cosl(PI * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const long double pk = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = pk;
            }
            if (n == 1) { p1 = x; p0 = 1.0L; }
            dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
            const long double dx = -p1 / dp;
            x += dx;
            if (fabsl(dx) < 1.0e-19L) break;
        }
        {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const long double pk = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = pk;
            }
            dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
        }
        const long double wl = 2.0L / ((1.0L - x * x) * dp * dp);
        s[i] = (double)(0.5L * (1.0L - x));
        w[i] = (double)(0.5L * wl);
        s[n - 1 - i] = (double)(0.5L * (1.0L + x));
        w[n - 1 - i] = (double)(0.5L * wl);
    }
}

}  // namespace upic_host
using namespace upic_host;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
// FP32 copy of a Gauss-Legendre rule on [0, 1] that keeps the rule's
// symmetry exactly: the upper nodes are rounded and the lower ones are their
// exact mirrors 1 - s (Sterbenz), weights are mirrored, and the largest
// weight absorbs the rounding so that the weights sum to 1 in FP32.  Plain
// rounding leaves the first moment sum w s off 1/2 by ~1e-8, a systematic
// shift of every orbit average that shows up as an energy drift.
inline QuadF fp32_rule(const QuadDev& q) {
    QuadF r{};
    r.n = q.n <= 8 ? q.n : 0;
    const int n = r.n;
    for (int i = n / 2; i < n; ++i) {  // s >= 1/2
        r.s[i] = (float)q.s[i];
        r.w[i] = (float)q.w[i];
    }
    for (int i = 0; i < n / 2; ++i) {
        r.s[i] = 1.0f - r.s[n - 1 - i];
        r.w[i] = r.w[n - 1 - i];
    }
    if (n % 2 == 1) r.s[n / 2] = 0.5f;
    // fix the total in FP32: adjust the central weight(s)
    const int mid = n / 2;
    for (int pass = 0; pass < 4 && n > 0; ++pass) {
        float tot = 0.0f;
        for (int i = 0; i < n; ++i) tot += r.w[i];
        if (tot == 1.0f) break;
        if (n % 2 == 1) {
            r.w[mid] += 1.0f - tot;
        } else {
            const float d = 0.5f * (1.0f - tot);
            r.w[mid] += d;
            r.w[mid - 1] = r.w[mid];
        }
    }
    return r;
}

// Slots of a context's reduction area (dred, mirrored in hred): [0, 8) the
// particle sweeps' sums and maxima, [8, 14) residual norms, 14 the absolute
// outer tolerance of the step, [16, 64) Gram entries and other sums,
// [64, 74) the device outer-solve control (Anderson status, nonfinite flag,
// coefficients), [80, 86) the norms of the step's initial predictor.
constexpr int kRedSlots = 128;
constexpr int kRedAtol = 14;
constexpr int kRedGram = 16;
constexpr int kRedCtl = 64;
constexpr int kRedPredictor = 80;

struct upic_ctx {
    upic_params p{};
    GridDev g{};
    QuadDev q{};
    QuadF qf{};          // FP32 copy of the rule (mixed precision)
    OrbitOpt opt{};
    SpecArgs sp{};
    FFTPlan plan{};
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
    double* dred = nullptr;               // small device reduction area (kRedSlots doubles)
    unsigned long long* dstats = nullptr; // 8 counters
    double* hred = nullptr;               // pinned mirror
    unsigned long long* hstats = nullptr;
    // group-A workspace, grown on demand
    double* work = nullptr;
    size_t work_n = 0;
    // per-kernel-family timing with HIP events on `stream`: each timed launch
    // is bracketed by two events from a pool; the pairs are read at the
    // stream's next synchronisation (no wait per launch)
    struct KTimer { int64_t launches = 0; double ms = 0.0; };
    struct PendingTimer { const char* name; hipEvent_t e0, e1; };
    bool timing = false;
    std::map<std::string, KTimer> timers;
    std::vector<PendingTimer> pending_timers;
    std::vector<hipEvent_t> event_pool, all_events;
    // replay the field solve of a map evaluation, or a whole P3 outer solve,
    // from a captured hipGraph; graph_slots > 0 fixes the number of map
    // evaluations a step graph unrolls (0: adapted to the steps)
    bool graphs = false;
    int graph_slots = 0;
    // deterministic mode (upic_ctx_set_deterministic): fixed-point deposit in
    // the FP64 P3 sweep, within-cell particle order by storage order in the
    // FP64 cell sort, one-workgroup reductions
    bool det = false;

    int N() const { return g.N; }
    double* workspace(size_t n) {
        if (n > work_n) {
            dfree(work);
            work = dalloc<double>(n);
            work_n = n;
        }
        return work;
    }
    void sync() {
        HIPCHK(hipStreamSynchronize(stream));
        resolve_timers();
    }
    hipEvent_t take_event() {
        if (event_pool.empty()) {
            hipEvent_t e = nullptr;
            HIPCHK(hipEventCreate(&e));
            all_events.push_back(e);
            return e;
        }
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    // Reads the event pairs of completed timed launches (the stream is idle).
    void resolve_timers() {
        for (const PendingTimer& t : pending_timers) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, t.e0, t.e1));
            auto& k = timers[t.name];
            k.launches += 1;
            k.ms += ms;
            event_pool.push_back(t.e0);
            event_pool.push_back(t.e1);
        }
        pending_timers.clear();
    }
    double c2() const { return p.c * p.c; }
    double phi_src() const { return p.code == 3 ? p.sigma1 : 1.0 / p.eps0; }
    double A_src() const { return p.code == 3 ? p.sigma2 : p.mu0; }
};

namespace upic_host {

// The particle kernels wrap a stencil's node indices by one period, which
// needs at least as many nodes as the spline degree on every axis they span
// (field-only operations have no such limit).
inline void require_stencil_grid(const upic_params& p) {
    const int order = p.code == 2 ? 2 : p.spline_order;
    if (p.nx < order || p.ny < order || (p.nz > 1 && p.nz < order))
        throw Error(UPIC_EINVAL, "particle operations need at least spline_order grid nodes on every axis (nz may be 1)");
}

inline void validate(const upic_params& p) {
    if (p.code < 1 || p.code > 3) throw Error(UPIC_EINVAL, "code must be 1, 2 or 3");
    if (!is_pow2(p.nx) || !is_pow2(p.ny) || !is_pow2(p.nz))
        throw Error(UPIC_EINVAL, "the in-file radix-2 FFT requires nx, ny, nz to be powers of two");
    if (p.nx > 2048 || p.ny > 2048 || p.nz > 2048) throw Error(UPIC_EINVAL, "grid lines longer than 2048 are not supported");
    if (p.code == 2 && p.nz != 1) throw Error(UPIC_EINVAL, "P2 is a 2-D code: nz must be 1");
    if (p.code != 2 && (p.spline_order < 1 || p.spline_order > 4)) throw Error(UPIC_EINVAL, "spline_order must be 1, 2, 3, or 4");
    if (p.code != 2 && (p.orbit_quad_order < 1 || p.orbit_quad_order > kMaxQuad))
        throw Error(UPIC_EINVAL, "orbit_quad_order must be in [1, 32]");
    if (!(p.dt > 0.0)) throw Error(UPIC_EINVAL, "time stepping requires dt>0");
    if (!(p.Lx > 0.0) || !(p.Ly > 0.0) || !(p.Lz > 0.0)) throw Error(UPIC_EINVAL, "domain lengths must be positive");
    if (p.code == 3) {  // P3:1349-1433
        if (!(p.c > 0.0) || !(p.sigma1 > 0.0) || !(p.sigma2 > 0.0))
            throw Error(UPIC_EINVAL, "nondimensional coefficients require kappa>0, sigma1>0, sigma2>0");
        if (!(p.nonlinear_rtol > 0.0) || !(p.nonlinear_atol > 0.0) || p.nonlinear_max_iter <= 0)
            throw Error(UPIC_EINVAL, "nonlinear controls require nonlinear_rtol>0, nonlinear_atol>0, and nonlinear_max_iter>0");
        if (p.anderson_memory < 1 || p.anderson_start < 0)
            throw Error(UPIC_EINVAL, "Anderson controls require anderson_memory>=1 and anderson_start>=0");
        if (p.anderson_memory > 8) throw Error(UPIC_EINVAL, "anderson_memory above 8 is not supported");
        if (!(p.anderson_damping > 0.0) || !(p.anderson_damping <= 1.0))
            throw Error(UPIC_EINVAL, "anderson_damping must satisfy 0 < damping <= 1");
        if (!(p.anderson_regularization >= 0.0) || !(p.anderson_growth_limit > 1.0))
            throw Error(UPIC_EINVAL, "Anderson controls require regularization>=0 and growth_limit>1");
        if (!(p.local_rtol > 0.0) || !(p.local_atol > 0.0) || p.local_max_iter < 1 || !(p.local_damping > 0.0) ||
            !(p.local_damping <= 1.0))
            throw Error(UPIC_EINVAL, "local particle solve requires positive rtol/atol/max_iter and 0<damping<=1");
        const int ms = p.selective_particle_max_substeps;
        if (ms < 1 || ms > 64 || (ms & (ms - 1)) != 0)
            throw Error(UPIC_EINVAL, "selective particle fallback requires a power-of-two max_substeps in [1,64]");
        const double compat = p.c * p.c * p.sigma2;
        if (std::abs(p.sigma1 - compat) / std::max(1.0, std::abs(p.sigma1)) > 1.0e-10)
            throw Error(UPIC_EINVAL, "Maxwell-compatible nondimensionalization requires sigma1 = kappa^2*sigma2");
    } else {
        if (!(p.c > 0.0) || !(p.eps0 > 0.0) || !(p.mu0 > 0.0)) throw Error(UPIC_EINVAL, "c, eps0 and mu0 must be positive");
    }
    if (p.code == 1 && p.picard_max_iter < 1) throw Error(UPIC_EINVAL, "picard_max_iter must be positive");
    if (p.code == 2 && p.bdf_order != 1 && p.bdf_order != 2) throw Error(UPIC_EINVAL, "method must be bdf1 or bdf2");
    if (p.precision != UPIC_PREC_FP64 && p.precision != UPIC_PREC_MIXED) throw Error(UPIC_EINVAL, "unknown precision");
    if (p.learned_guess < 0 || p.learned_guess > 2) throw Error(UPIC_EINVAL, "learned_guess must be 0, 1 or 2");
    if (p.learned_guess > 0 && p.precision != UPIC_PREC_MIXED)
        throw Error(UPIC_EINVAL, "the initial-guess modes are implemented for the mixed-precision P3 step");
    if (p.precision == UPIC_PREC_MIXED) {
        if (p.code != 3 || p.nz != 1)
            throw Error(UPIC_EINVAL, "mixed precision is implemented for the P3 step on 2-D grids (nz = 1)");
        if (p.orbit_quad_order > 8) throw Error(UPIC_EINVAL, "mixed precision supports orbit_quad_order <= 8");
        if (p.nx * p.ny >= (1 << 30)) throw Error(UPIC_EINVAL, "mixed precision needs nx*ny < 2^30");
    }
}

template <class F>
int guarded(upic_ctx* c, F&& f) {
    // the context's reduction mode holds for the launches of this call
    // (on this host thread)
    const FixedOrderScope fo(c && c->det);
    try {
        f();
        if (c) c->err.clear();
        return UPIC_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return UPIC_EINVAL;
    }
}

// Runs `launch` between two events on the context stream when timing is on
// and accumulates the elapsed time under `name`.
// `name` must have static storage (a string literal).
template <class F>
void timed(upic_ctx* c, const char* name, F&& launch) {
    if (!c->timing) { launch(); return; }
    if (c->pending_timers.size() >= 4096) c->sync();
    hipEvent_t e0 = c->take_event(), e1 = c->take_event();
    HIPCHK(hipEventRecord(e0, c->stream));
    launch();
    HIPCHK(hipEventRecord(e1, c->stream));
    c->pending_timers.push_back({name, e0, e1});
}

// real fields -> spectra; spectra -> real (destroys the spectra)
inline void to_spectra(upic_ctx* c, const double* re, double* zh, int batch) {
    launch_fft(c->stream, c->plan, zh, batch, false, re, nullptr);
}
inline void from_spectra(upic_ctx* c, double* zh, double* re, int batch) {
    launch_fft(c->stream, c->plan, zh, batch, true, nullptr, re);
}

}  // namespace upic_host
