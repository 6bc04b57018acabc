// api.hip -- C ABI (include/upic.h): context and the kernel-level (group A)
// entry points.  The per-step drivers are in sim.hip.
#include <mutex>

#include "api_internal.hpp"

// message of the last failed call that has no context (upic_last_error(NULL))
thread_local std::string g_create_err;

// Device scratch of upic_local_substep_noalloc: the 2-D packed node records
// (8 doubles per mesh node) of at most kNoAllocMaxNodes nodes, 16 MiB per
// GPU, held from module load to process end.  Host code never passes the
// array's name to a kernel: its device address is resolved once per device
// (resolve_noalloc_scratch, called by upic_ctx_create with the device
// current) and kept in g_scratch_dev.
constexpr size_t kNoAllocMaxNodes = (size_t)1 << 18;  // 512 x 512
static __device__ __align__(256) double scratch[8 * kNoAllocMaxNodes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxScratchDevices] = {};

void resolve_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        throw Error(UPIC_EDEVICE, "device ordinal beyond the scratch table (64 devices)");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(UPIC_EDEVICE, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_dev[device] = static_cast<double*>(ptr);
}

double* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxScratchDevices) ? g_scratch_dev[device] : nullptr;
}
}  // namespace

extern "C" {

upic_status upic_ctx_create(const upic_params* params, upic_ctx** out) {
    if (!params || !out) return UPIC_EINVAL;
    *out = nullptr;
    upic_ctx* c = new upic_ctx();
    const int rc = guarded(c, [&] {
        validate(*params);
        c->p = *params;
        const upic_params& p = c->p;
        HIPCHK(hipGetDevice(&c->device));
        resolve_noalloc_scratch(c->device);
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        GridDev& g = c->g;
        g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
        g.N = p.nx * p.ny * p.nz;
        g.L[0] = p.Lx; g.L[1] = p.Ly; g.L[2] = p.Lz;
        g.h[0] = p.Lx / p.nx; g.h[1] = p.Ly / p.ny; g.h[2] = p.Lz / p.nz;
        for (int d = 0; d < 3; ++d) g.half_L[d] = 0.5 * g.L[d];
        g.dV = g.h[0] * g.h[1] * g.h[2];
        g.code = p.code;
        g.order = p.code == 2 ? 2 : p.spline_order;
        c->q.n = p.code == 2 ? 1 : p.orbit_quad_order;
        gauss_legendre_01(c->q.n, c->q.s, c->q.w);
        c->qf = fp32_rule(c->q);
        c->opt.split = p.split_orbit_at_knots ? 1 : 0;
        c->opt.wrap_path = p.code == 1 ? 1 : 0;
        SpecArgs& sp = c->sp;
        sp.nx = p.nx; sp.ny = p.ny; sp.nz = p.nz; sp.N = g.N; sp.code = p.code;
        sp.projection = (p.code == 3 && p.nyquist_projection) ? 1 : 0;
        sp.L[0] = p.Lx; sp.L[1] = p.Ly; sp.L[2] = p.Lz;
        const int n3[3] = {p.nx, p.ny, p.nz};
        for (int d = 0; d < 3; ++d) {
            c->plan.n[d] = n3[d];
            c->plan.tw[d] = nullptr;
            const int n = n3[d];
            if (n <= 1) continue;
            std::vector<double> tw(n);  // n/2 complex twiddles exp(-2 pi i k / n)
            const long double PI = 3.141592653589793238462643383279502884L;
            for (int k = 0; k < n / 2; ++k) {
                const long double ang = 2.0L * PI * (long double)k / (long double)n;
                tw[2 * k] = (double)cosl(ang);
                tw[2 * k + 1] = (double)(-sinl(ang));
            }
            c->plan.tw[d] = dalloc<double>(n);
            HIPCHK(hipMemcpy(c->plan.tw[d], tw.data(), n * sizeof(double), hipMemcpyHostToDevice));
        }
        // the 8 path-statistic counters follow the reduction slots, so that
        // one copy brings both to the host
        c->dred = dalloc<double>(kRedSlots + 8);
        c->dstats = reinterpret_cast<unsigned long long*>(c->dred + kRedSlots);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hred), (kRedSlots + 8) * sizeof(double), hipHostMallocDefault));
        c->hstats = reinterpret_cast<unsigned long long*>(c->hred + kRedSlots);
    });
    if (rc != UPIC_OK) {
        g_create_err = c->err;
        upic_ctx_destroy(c);
        return (upic_status)rc;
    }
    *out = c;
    return UPIC_OK;
}

upic_status upic_ctx_destroy(upic_ctx* c) {
    if (!c) return UPIC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int d = 0; d < 3; ++d) dfree(c->plan.tw[d]);
    dfree(c->dred);  // dstats lies in the same allocation, hstats in hred's
    dfree(c->work);
    if (c->hred) (void)hipHostFree(c->hred);
    for (hipEvent_t e : c->all_events) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return UPIC_OK;
}

upic_status upic_ctx_sync(upic_ctx* c) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] { c->sync(); });
}

upic_status upic_ctx_set_timing(upic_ctx* c, int enable) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        c->sync();
        c->timing = enable != 0;
        c->timers.clear();
    });
}

upic_status upic_ctx_set_graphs(upic_ctx* c, int enable) {
    if (!c) return UPIC_EINVAL;
    if (enable < 0) return UPIC_EINVAL;
    c->graphs = enable != 0;
    c->graph_slots = enable > 1 ? enable : 0;
    return UPIC_OK;
}

upic_status upic_ctx_set_deterministic(upic_ctx* c, int enable) {
    if (!c) return UPIC_EINVAL;
    c->det = enable != 0;
    return UPIC_OK;
}

upic_status upic_ctx_timing(upic_ctx* c, const char* name, int64_t* launches, double* total_ms) {
    if (!c || !name || !launches || !total_ms) return UPIC_EINVAL;
    const int rc = guarded(c, [&] { c->sync(); });
    if (rc != UPIC_OK) return (upic_status)rc;
    auto it = c->timers.find(name);
    *launches = it == c->timers.end() ? 0 : it->second.launches;
    *total_ms = it == c->timers.end() ? 0.0 : it->second.ms;
    return UPIC_OK;
}

const char* upic_last_error(const upic_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

upic_status upic_quadrature(const upic_ctx* c, double* s, double* w) {
    if (!c || !s || !w) return UPIC_EINVAL;
    std::memcpy(s, c->q.s, c->q.n * sizeof(double));
    std::memcpy(w, c->q.w, c->q.n * sizeof(double));
    return UPIC_OK;
}

// ---- device memory plumbing ---------------------------------------------
upic_status upic_device_count(int* count) {
    if (!count) return UPIC_EINVAL;
    *count = 0;
    if (hipGetDeviceCount(count) != hipSuccess) { *count = 0; (void)hipGetLastError(); }
    return UPIC_OK;
}
upic_status upic_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return UPIC_EINVAL;
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) { (void)hipGetLastError(); return UPIC_ENOMEM; }
    return UPIC_OK;
}
upic_status upic_device_free(void* ptr) {
    if (ptr && hipFree(ptr) != hipSuccess) return UPIC_EDEVICE;
    return UPIC_OK;
}
upic_status upic_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? UPIC_OK : UPIC_EDEVICE;
}
upic_status upic_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? UPIC_OK : UPIC_EDEVICE;
}
upic_status upic_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? UPIC_OK : UPIC_EDEVICE;
}

// ---- Group A --------------------------------------------------------------
static void read_stats(upic_ctx* c, upic_path_stats* st, int64_t np) {
    HIPCHK(hipMemcpyAsync(c->hstats, c->dstats, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    st->n_particles = np;
    st->crossing_particles = (int64_t)c->hstats[0];
    st->total_segments = (int64_t)c->hstats[1];
    st->max_segments = std::max<int32_t>(1, (int32_t)c->hstats[2]);
    st->pad = 0;
}

upic_status upic_deposit_current_orbit(upic_ctx* c, const double* x0, const double* dx, const double* vbar,
                                       const double* q, int64_t np, double* J, upic_path_stats* stats) {
    if (!c || np < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        launch_zero(c->stream, 3 * (int64_t)c->N(), J);
        HIPCHK(hipMemsetAsync(c->dstats, 0, 3 * sizeof(unsigned long long), c->stream));
        launch_deposit_orbit(c->stream, c->g, c->q, c->opt, x0, dx, vbar, q, np, J, c->dstats);
        launch_project_nyquist(c->stream, c->sp, J, 3);
        if (stats) read_stats(c, stats, np);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_gather_orbit(upic_ctx* c, const double* x0, const double* dx, const double* f3, int64_t np, double* out) {
    if (!c || np < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        launch_gather_orbit(c->stream, c->g, c->q, c->opt, x0, dx, f3, np, out);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_orbit_grad_A(upic_ctx* c, const double* x0, const double* dx, const double* A_n, const double* A_np1,
                              double theta0, double theta1, int64_t np, double* D) {
    if (!c || np < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        const int p1_theta = (c->p.code == 1 && theta0 == 0.0 && theta1 == 1.0) ? 1 : 0;
        launch_grad_A(c->stream, c->g, c->q, c->opt, x0, dx, A_n, A_np1, theta0, theta1, p1_theta, np, D);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_bdf_wave(upic_ctx* c, const double* src, const double* hist, int order, double* out) {
    if (!c || !src || !hist || !out || (order != 1 && order != 2)) return UPIC_EINVAL;
    if (c->p.code != 2) {
        c->err = "upic_bdf_wave needs a P2 context";
        return UPIC_EINVAL;
    }
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        const std::vector<double> a = order == 2 ? std::vector<double>{1.5, -2.0, 0.5} : std::vector<double>{1.0, -1.0};
        std::vector<double> b(2 * a.size() - 1, 0.0);
        for (size_t i = 0; i < a.size(); ++i)
            for (size_t j = 0; j < a.size(); ++j) b[i + j] += a[i] * a[j];
        const double scale = 1.0 / (c->p.c * c->p.c * c->p.dt * c->p.dt);
        // by linearity: one transform of the combined real right-hand side
        double* w = c->workspace(3 * (size_t)N);
        LinComb lc{};
        lc.k = 0;
        lc.c[lc.k] = 1.0; lc.f[lc.k++] = src;
        for (size_t l = 1; l < b.size(); ++l) { lc.c[lc.k] = -scale * b[l]; lc.f[lc.k++] = hist + (l - 1) * (size_t)N; }
        launch_lincomb(c->stream, N, lc, w);
        double* zh = w + N;
        to_spectra(c, w, zh, 1);
        launch_spectral_div_shift(c->stream, c->sp, scale * b[0], zh, 1);
        from_spectra(c, zh, out, 1);
        HIPCHK(hipGetLastError());
    });
}

// Body of upic_local_substep(_noalloc).  node_scratch: device memory for the
// 2-D node records (8 N doubles); NULL = the context workspace (grown with
// hipMalloc on demand).
static upic_status local_substep(upic_ctx* c, const char* name, double* node_scratch, const double* guess,
                                 const double* x0, const double* P0, const double* q, const double* m, double h,
                                 double theta0, double theta1, const double* A_n, const double* A_end,
                                 const double* gpm, int64_t np, double* vbar, double* P_end, double* v_end,
                                 int32_t* iterations, uint8_t* reasons) {
    if (!c || np < 0 || !guess || !x0 || !P0 || !q || !m || !A_n || !A_end || !gpm || !vbar || !P_end || !v_end ||
        !iterations || !reasons)
        return UPIC_EINVAL;
    if (c->p.code != 3) {
        c->err = std::string(name) + " needs a P3 context";
        return UPIC_EINVAL;
    }
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        const upic_params& p = c->p;
        P3Args a{};
        a.np = np; a.dt = h; a.kappa = p.c; a.rtol = p.local_rtol; a.atol = p.local_atol;
        a.damping = p.local_damping; a.max_iter = p.local_max_iter;
        a.x = x0; a.P = P0; a.q = q; a.m = m; a.guess = guess;
        a.A_n = A_n; a.A_end = A_end; a.gpm = gpm;
        if (c->g.nz == 1) {  // 2-D: the node records the sweeps read
            double* nr = node_scratch ? node_scratch : c->workspace(8 * (size_t)c->N());
            launch_pack_nodes(c->stream, c->N(), gpm, A_n, A_end, nr);
            a.NR = nr;
        }
        a.vbar_out = vbar; a.P_out = P_end; a.v_out = v_end; a.reason = reasons;
        launch_p3_local_solve(c->stream, c->g, c->q, c->opt, a, h, theta0, theta1, iterations);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_local_substep(upic_ctx* c, const double* guess, const double* x0, const double* P0,
                               const double* q, const double* m, double h, double theta0, double theta1,
                               const double* A_n, const double* A_end, const double* gpm, int64_t np,
                               double* vbar, double* P_end, double* v_end, int32_t* iterations, uint8_t* reasons) {
    return local_substep(c, "upic_local_substep", nullptr, guess, x0, P0, q, m, h, theta0, theta1, A_n, A_end, gpm,
                         np, vbar, P_end, v_end, iterations, reasons);
}

upic_status upic_local_substep_noalloc(upic_ctx* c, const double* guess, const double* x0, const double* P0,
                                       const double* q, const double* m, double h, double theta0, double theta1,
                                       const double* A_n, const double* A_end, const double* gpm, int64_t np,
                                       double* vbar, double* P_end, double* v_end, int32_t* iterations,
                                       uint8_t* reasons) {
    if (!c) return UPIC_EINVAL;
    double* nr = nullptr;
    if (c->g.nz == 1) {
        if ((size_t)c->N() > kNoAllocMaxNodes) {
            c->err = "upic_local_substep_noalloc: a 2-D mesh of " + std::to_string(c->N()) +
                     " nodes exceeds the static scratch of " + std::to_string(kNoAllocMaxNodes) + " nodes";
            return UPIC_EINVAL;
        }
        nr = noalloc_scratch(c->device);
        if (!nr) {
            c->err = "upic_local_substep_noalloc: no scratch address for device " + std::to_string(c->device);
            return UPIC_EDEVICE;
        }
    }
    return local_substep(c, "upic_local_substep_noalloc", nr, guess, x0, P0, q, m, h, theta0, theta1, A_n, A_end,
                         gpm, np, vbar, P_end, v_end, iterations, reasons);
}

upic_status upic_gather_point(upic_ctx* c, const double* pos, const double* f3, int64_t np, double* out) {
    if (!c || np < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        launch_gather_point(c->stream, c->g, pos, f3, np, out);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_deposit_scalar_point(upic_ctx* c, const double* pos, const double* values, int64_t np,
                                      int divide_by_dV, double* mesh) {
    if (!c || np < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        require_stencil_grid(c->p);
        launch_zero(c->stream, c->N(), mesh);
        launch_deposit_scalar_point(c->stream, c->g, pos, values, np, divide_by_dV ? 1.0 / c->g.dV : 1.0, mesh);
        launch_project_nyquist(c->stream, c->sp, mesh, 1);
        HIPCHK(hipGetLastError());
    });
}

upic_status upic_fft(upic_ctx* c, double* data, int inverse) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        launch_fft(c->stream, c->plan, data, 1, inverse != 0);
        HIPCHK(hipGetLastError());
    });
}



upic_status upic_spectral_gradient(upic_ctx* c, const double* f, double* out3) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        double* w = c->workspace(8 * (size_t)N);
        to_spectra(c, f, w, 1);
        launch_spectral_grad(c->stream, c->sp, w, w + 2 * N);
        from_spectra(c, w + 2 * N, out3, 3);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_spectral_divergence(upic_ctx* c, const double* V3, double* out) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        double* w = c->workspace(8 * (size_t)N);
        to_spectra(c, V3, w, 3);
        launch_spectral_div(c->stream, c->sp, w, w + 6 * N);
        from_spectra(c, w + 6 * N, out, 1);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_spectral_curl(upic_ctx* c, const double* A3, double* out3) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        double* w = c->workspace(12 * (size_t)N);
        to_spectra(c, A3, w, 3);
        launch_spectral_curl(c->stream, c->sp, w, w + 6 * N);
        from_spectra(c, w + 6 * N, out3, 3);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_continuity(upic_ctx* c, const double* rho_n, const double* J3, double dt, double* rho_out) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        double* w = c->workspace(10 * (size_t)N);
        to_spectra(c, J3, w, 3);
        to_spectra(c, rho_n, w + 6 * N, 1);
        launch_spectral_continuity(c->stream, c->sp, dt, w + 6 * N, w, w + 8 * N);
        from_spectra(c, w + 8 * N, rho_out, 1);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_cn_potentials(upic_ctx* c, const double* phi, const double* psi, const double* A3, const double* U3,
                               const double* rho_mid, const double* J3, double dt, double* phi_out, double* psi_out,
                               double* A_out3, double* U_out3) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        const size_t Z = 2 * (size_t)N;
        double* w = c->workspace(20 * Z);
        double *ph = w, *sh = w + Z, *Ah = w + 2 * Z, *Uh = w + 5 * Z, *rh = w + 8 * Z, *Jh = w + 9 * Z;
        double *po = w + 12 * Z, *so = w + 13 * Z, *Ao = w + 14 * Z, *Uo = w + 17 * Z;
        to_spectra(c, phi, ph, 1);
        to_spectra(c, psi, sh, 1);
        to_spectra(c, A3, Ah, 3);
        to_spectra(c, U3, Uh, 3);
        to_spectra(c, rho_mid, rh, 1);
        to_spectra(c, J3, Jh, 3);
        launch_spectral_cn(c->stream, c->sp, dt, c->c2(), c->phi_src(), ph, sh, rh, po, so);
        launch_zero_mean_mode(c->stream, po, 2, N);  // subtract_mean of phi, psi (P1:983-984)
        for (int d = 0; d < 3; ++d)
            launch_spectral_cn(c->stream, c->sp, dt, c->c2(), c->A_src(), Ah + d * Z, Uh + d * Z, Jh + d * Z, Ao + d * Z, Uo + d * Z);
        from_spectra(c, po, phi_out, 1);
        from_spectra(c, so, psi_out, 1);
        from_spectra(c, Ao, A_out3, 3);
        from_spectra(c, Uo, U_out3, 3);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_project_nyquist(upic_ctx* c, double* f, int ncomp) {
    if (!c || ncomp < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        launch_project_nyquist(c->stream, c->sp, f, ncomp);
        HIPCHK(hipGetLastError());
    });
}
upic_status upic_poisson(upic_ctx* c, const double* rho, double* phi) {
    if (!c) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        const int N = c->N();
        double* w = c->workspace(4 * (size_t)N);
        to_spectra(c, rho, w, 1);
        launch_spectral_poisson(c->stream, c->sp, c->phi_src(), w, w + 2 * N);
        from_spectra(c, w + 2 * N, phi, 1);
        HIPCHK(hipGetLastError());
    });
}

}  // extern "C"
