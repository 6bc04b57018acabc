// field_kernels.hip -- mesh-side HIP kernels: LDS-staged radix-2 FFT,
// k-space field solves, projector, reductions, and the particle cell sort.
#include "kernels.hpp"
#include "particles.hpp"
#include "dense_solve.hpp"

namespace upic {

// ---------------------------------------------------------------------------
// FFT: one workgroup transforms 2^llpb lines of length n = 2^logn staged in
// LDS (bit-reversed load, logn radix-2 Danielson-Lanczos stages, store), the
// algorithm of fft1d (P1:424-454) with tabulated twiddles.  Lines along a
// strided axis (stride 2^lstride) are loaded line-fastest so neighbouring
// lanes touch neighbouring addresses.  Every size is a power of two (the
// grid's sides are), so the index arithmetic is shifts and masks.  TWL: the
// n/2 twiddles are staged in LDS with the lines; each stage then waits on
// LDS only, not on a global load (the stages are a chain of barriers, and
// the 256^2 transforms are latency-bound).  re_in (first pass of a forward
// transform): the lines are read from real fields, imaginary parts zero;
// re_out (last pass of an inverse transform): only the real parts are
// stored, into real fields; either saves the separate conversion pass.
// ---------------------------------------------------------------------------
template <bool TWL>
__global__ __launch_bounds__(256) void k_fft_axis(double* __restrict__ data, int N, int logn, int lstride, int lines,
                                                  int llpb, const double* __restrict__ tw, int inverse,
                                                  const double* __restrict__ re_in, double* __restrict__ re_out,
                                                  const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    extern __shared__ double lds[];  // lpb * n complex (+ n/2 twiddles)
    double2* a = reinterpret_cast<double2*>(lds);
    const int n = 1 << logn, lpb = 1 << llpb;
    const int total = lpb << logn;
    const double2* twg = reinterpret_cast<const double2*>(tw);
    double2* twl = a + total;
    if constexpr (TWL)
        for (int e = threadIdx.x; e < (n >> 1); e += blockDim.x) twl[e] = twg[e];
    const int field = blockIdx.y;
    double2* base_ptr = reinterpret_cast<double2*>(data) + (int64_t)field * N;
    const int line0 = blockIdx.x << llpb;
    const int smask = (1 << lstride) - 1;
    // load with bit reversal
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int lb, t;
        if (lstride == 0) { lb = e >> logn; t = e & (n - 1); }
        else { t = e >> llpb; lb = e & (lpb - 1); }
        const int l = line0 + lb;
        if (l < lines) {
            const int base = ((l >> lstride) << (logn + lstride)) + (l & smask);
            const int r = (int)(__brev((unsigned)t) >> (32 - logn));
            a[(lb << logn) + r] = re_in ? make_double2(re_in[(int64_t)field * N + base + (t << lstride)], 0.0)
                                        : base_ptr[base + (t << lstride)];
        }
    }
    __syncthreads();
    const int nbf = total >> 1;
    for (int s = 1; s <= logn; ++s) {
        const int hs = s - 1;  // log2 of the half length
        for (int bi = threadIdx.x; bi < nbf; bi += blockDim.x) {
            const int lb = bi >> (logn - 1), j = bi & ((n >> 1) - 1);
            const int grp = j >> hs, pos = j & ((1 << hs) - 1);
            const int i0 = (lb << logn) + (grp << s) + pos, i1 = i0 + (1 << hs);
            const double2 w = TWL ? twl[pos << (logn - s)] : twg[pos << (logn - s)];
            const double wr = w.x, wi = inverse ? -w.y : w.y;
            const double2 u = a[i0], x = a[i1];
            const double2 v = make_double2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
            a[i0] = make_double2(u.x + v.x, u.y + v.y);
            a[i1] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
    const double sc = inverse ? 1.0 / (double)n : 1.0;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int lb, t;
        if (lstride == 0) { lb = e >> logn; t = e & (n - 1); }
        else { t = e >> llpb; lb = e & (lpb - 1); }
        const int l = line0 + lb;
        if (l < lines) {
            const int base = ((l >> lstride) << (logn + lstride)) + (l & smask);
            double2 v = a[(lb << logn) + t];
            if (inverse) { v.x *= sc; v.y *= sc; }
            if (re_out) re_out[(int64_t)field * N + base + (t << lstride)] = v.x;
            else base_ptr[base + (t << lstride)] = v;
        }
    }
}

void launch_fft(hipStream_t s, const FFTPlan& plan, double* data, int batch, bool inverse, const double* re_in,
                double* re_out) {
    const int nx = plan.n[0], ny = plan.n[1], nz = plan.n[2];
    const int N = nx * ny * nz;
    const int strides[3] = {ny * nz, nz, 1};
    auto log2i = [](int v) { int k = 0; while ((1 << k) < v) ++k; return k; };
    int last = -1;
    for (int ax = 0; ax < 3; ++ax)
        if (plan.n[ax] > 1) { last = ax; break; }
    if (last < 0) {  // a single point: the transform is the identity
        if (re_in) launch_real_to_complex(s, re_in, data, (int64_t)batch * N);
        if (re_out) launch_complex_real(s, data, re_out, (int64_t)batch * N);
        return;
    }
    bool first = true;
    // z, y, x order (P1:456-485)
    for (int ax = 2; ax >= 0; --ax) {
        const int n = plan.n[ax];
        if (n <= 1) continue;
        const double* rin = first ? re_in : nullptr;
        double* rout = ax == last ? re_out : nullptr;
        first = false;
        const int logn = log2i(n);
        const int lines = N / n;
        const int lpb = n >= 512 ? 1 : 512 / n;
        const int blocks = (lines + lpb - 1) / lpb;
        const size_t lds = (size_t)lpb * n * sizeof(double2);
        const size_t lds_tw = lds + (size_t)(n / 2) * sizeof(double2);
        if (lds_tw <= 64 * 1024)
            hipLaunchKernelGGL(k_fft_axis<true>, dim3(blocks, batch), dim3(256), lds_tw, s, data, N, logn,
                               log2i(strides[ax]), lines, log2i(lpb), plan.tw[ax], inverse ? 1 : 0, rin, rout,
                               current_gate());
        else
            hipLaunchKernelGGL(k_fft_axis<false>, dim3(blocks, batch), dim3(256), lds, s, data, N, logn,
                               log2i(strides[ax]), lines, log2i(lpb), plan.tw[ax], inverse ? 1 : 0, rin, rout,
                               current_gate());
    }
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
__global__ void k_r2c(const double* __restrict__ re, double* __restrict__ z, int64_t n, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { z[2 * i] = re[i]; z[2 * i + 1] = 0.0; }
}
__global__ void k_c2r(const double* __restrict__ z, double* __restrict__ re, int64_t n, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) re[i] = z[2 * i];
}
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (x ? a * x[i] : 0.0) + b * y[i];
}
__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}
__global__ void k_fill(int64_t n, double v, double* __restrict__ y, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}
static inline unsigned gblk(int64_t n) { return (unsigned)((n + 255) / 256); }
void launch_real_to_complex(hipStream_t s, const double* re, double* z, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_r2c, dim3(gblk(n)), dim3(256), 0, s, re, z, n, current_gate());
}
void launch_complex_real(hipStream_t s, const double* z, double* re, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_c2r, dim3(gblk(n)), dim3(256), 0, s, z, re, n, current_gate());
}
void launch_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, double* y) {
    if (n > 0) hipLaunchKernelGGL(k_axpby, dim3(gblk(n)), dim3(256), 0, s, n, a, x, b, y);
}
void launch_copy(hipStream_t s, int64_t n, const double* x, double* y) {
    if (n > 0) hipLaunchKernelGGL(k_copy, dim3(gblk(n)), dim3(256), 0, s, n, x, y, current_gate());
}
void launch_fill(hipStream_t s, int64_t n, double v, double* y) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(gblk(n)), dim3(256), 0, s, n, v, y, current_gate());
}
void launch_zero(hipStream_t s, int64_t n, double* y) { launch_fill(s, n, 0.0, y); }

// ---------------------------------------------------------------------------
// Spectral symbols (fft_mode_number P1:386-392, P2:320-330, P3:1485-1504)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mode_number(int i, int n, int code) {
    if (code == 2) {
        if (n > 1 && i == n / 2) return 0;
        return (i < n / 2) ? i : i - n;
    }
    if (n == 1) return 0;
    if (code == 3 && (n % 2 == 0) && i == n / 2) return 0;
    return (i <= n / 2) ? i : i - n;
}
__device__ __forceinline__ bool even_nyq(int i, int n) { return n > 1 && (n % 2 == 0) && i == n / 2; }

// k[d] is the first-derivative symbol, k2 the Laplacian symbol.  P1 keeps
// the Nyquist wave number +n/2 in k2 (P1:386-392); its first derivatives
// are always followed by a transform back to a real grid function
// (ifft_real, P1:494-499), which removes the imaginary Nyquist component --
// the same as a zero derivative symbol there.  Using the zero symbol keeps
// chained k-space operations (divergence -> continuity -> wave solve)
// identical to the reference's real-space round trips.
struct KVec {
    double k[3], k2;
    bool excluded;
};
__device__ __forceinline__ KVec kvec(const SpecArgs& sp, int id) {
    const int i = id / (sp.ny * sp.nz), j = (id / sp.nz) % sp.ny, kk = id % sp.nz;
    const double twopi = 6.283185307179586476925286766559;
    KVec o;
    const int mi = mode_number(i, sp.nx, sp.code), mj = mode_number(j, sp.ny, sp.code);
    const int mk = (sp.code == 2) ? 0 : mode_number(kk, sp.nz, sp.code);
    const double ki = twopi * (double)mi / sp.L[0];
    const double kj = twopi * (double)mj / sp.L[1];
    const double kz = (sp.code == 2) ? 0.0 : twopi * (double)mk / sp.L[2];
    o.k2 = ki * ki + kj * kj + kz * kz;
    o.k[0] = even_nyq(i, sp.nx) ? 0.0 : ki;
    o.k[1] = even_nyq(j, sp.ny) ? 0.0 : kj;
    o.k[2] = even_nyq(kk, sp.nz) ? 0.0 : kz;
    o.excluded = sp.projection && (even_nyq(i, sp.nx) || even_nyq(j, sp.ny) || even_nyq(kk, sp.nz));
    return o;
}
__device__ __forceinline__ double2 ld2(const double* p, int id) { return reinterpret_cast<const double2*>(p)[id]; }
__device__ __forceinline__ void st2(double* p, int id, double2 v) { reinterpret_cast<double2*>(p)[id] = v; }
// i * k * z
__device__ __forceinline__ double2 ik(double k, double2 z) { return make_double2(-k * z.y, k * z.x); }

__global__ void k_spectral_solve(SolveArgs a, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = a.sp.N;
    if (id >= N) return;
    const KVec kv = kvec(a.sp, id);
    const double2 J0 = ld2(a.Jh, id), J1 = ld2(a.Jh + 2 * N, id), J2 = ld2(a.Jh + 4 * N, id);
    // i k . J
    const double2 divJ = make_double2(-(kv.k[0] * J0.y + kv.k[1] * J1.y + kv.k[2] * J2.y),
                                      kv.k[0] * J0.x + kv.k[1] * J1.x + kv.k[2] * J2.x);
    const double2 rn = ld2(a.rho_n_h, id);
    double2 re = make_double2(rn.x - a.dt * divJ.x, rn.y - a.dt * divJ.y);
    if (id == 0 || kv.excluded) re = make_double2(0.0, 0.0);
    st2(a.rho_end_h, id, re);
    const double2 rm = make_double2(0.5 * (re.x + rn.x), 0.5 * (re.y + rn.y));
    const double theta = 0.25 * a.c2 * a.dt * a.dt * kv.k2;
    const double den = 1.0 + theta;
    const double dt = a.dt;
    // scalar potential (cn_fft_wave_step P1:940-957, P3:3131-3153), mean removed
    {
        const double sc = a.c2 * a.phi_src;
        const double2 u = ld2(a.phi_n_h, id), v = ld2(a.psi_n_h, id);
        double2 uo = make_double2(((1.0 - theta) * u.x + dt * v.x + 0.5 * sc * dt * dt * rm.x) / den,
                                  ((1.0 - theta) * u.y + dt * v.y + 0.5 * sc * dt * dt * rm.y) / den);
        double2 vo = make_double2(((1.0 - theta) * v.x - a.c2 * dt * kv.k2 * u.x + sc * dt * rm.x) / den,
                                  ((1.0 - theta) * v.y - a.c2 * dt * kv.k2 * u.y + sc * dt * rm.y) / den);
        if (id == 0 || kv.excluded) { uo = make_double2(0.0, 0.0); vo = uo; }
        st2(a.phi_end_h, id, uo);
        st2(a.psi_end_h, id, vo);
        const double2 ps = make_double2(0.5 * (u.x + uo.x), 0.5 * (u.y + uo.y));
#pragma unroll
        for (int d = 0; d < 3; ++d) st2(a.gpm_h + 2 * d * N, id, ik(kv.k[d], ps));
    }
    const double sc = a.c2 * a.A_src;
    const double2 Jd[3] = {J0, J1, J2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double2 u = ld2(a.A_n_h + 2 * d * N, id), v = ld2(a.U_n_h + 2 * d * N, id);
        const double2 S = Jd[d];
        double2 uo = make_double2(((1.0 - theta) * u.x + dt * v.x + 0.5 * sc * dt * dt * S.x) / den,
                                  ((1.0 - theta) * u.y + dt * v.y + 0.5 * sc * dt * dt * S.y) / den);
        double2 vo = make_double2(((1.0 - theta) * v.x - a.c2 * dt * kv.k2 * u.x + sc * dt * S.x) / den,
                                  ((1.0 - theta) * v.y - a.c2 * dt * kv.k2 * u.y + sc * dt * S.y) / den);
        if (kv.excluded) { uo = make_double2(0.0, 0.0); vo = uo; }
        st2(a.A_end_h + 2 * d * N, id, uo);
        st2(a.U_end_h + 2 * d * N, id, vo);
    }
}
void launch_spectral_solve(hipStream_t s, const SolveArgs& a) {
    hipLaunchKernelGGL(k_spectral_solve, dim3(gblk(a.sp.N)), dim3(256), 0, s, a, current_gate());
}

__global__ void k_spectral_grad(SpecArgs sp, const double* __restrict__ fh, double* __restrict__ o) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    const double2 f = ld2(fh, id);
#pragma unroll
    for (int d = 0; d < 3; ++d) st2(o + 2 * d * sp.N, id, ik(kv.k[d], f));
}
__global__ void k_spectral_div(SpecArgs sp, const double* __restrict__ V, double* __restrict__ o) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    const int N = sp.N;
    const double2 a = ld2(V, id), b = ld2(V + 2 * N, id), c = ld2(V + 4 * N, id);
    st2(o, id, make_double2(-(kv.k[0] * a.y + kv.k[1] * b.y + kv.k[2] * c.y), kv.k[0] * a.x + kv.k[1] * b.x + kv.k[2] * c.x));
}
__global__ void k_spectral_curl(SpecArgs sp, const double* __restrict__ A, double* __restrict__ o) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    const int N = sp.N;
    const double2 ax = ld2(A, id), ay = ld2(A + 2 * N, id), az = ld2(A + 4 * N, id);
    // i (k x A)
    const double2 c0 = make_double2(kv.k[1] * az.x - kv.k[2] * ay.x, kv.k[1] * az.y - kv.k[2] * ay.y);
    const double2 c1 = make_double2(kv.k[2] * ax.x - kv.k[0] * az.x, kv.k[2] * ax.y - kv.k[0] * az.y);
    const double2 c2 = make_double2(kv.k[0] * ay.x - kv.k[1] * ax.x, kv.k[0] * ay.y - kv.k[1] * ax.y);
    st2(o, id, make_double2(-c0.y, c0.x));
    st2(o + 2 * N, id, make_double2(-c1.y, c1.x));
    st2(o + 4 * N, id, make_double2(-c2.y, c2.x));
}
__global__ void k_spectral_poisson(SpecArgs sp, double src, const double* __restrict__ rh, double* __restrict__ ph) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    double2 o = make_double2(0.0, 0.0);
    if (!kv.excluded && kv.k2 > 0.0 && id != 0) {
        const double2 r = ld2(rh, id);
        o = make_double2(src * r.x / kv.k2, src * r.y / kv.k2);
    }
    st2(ph, id, o);
}
__global__ void k_spectral_cn(SpecArgs sp, double dt, double c2, double src, const double* __restrict__ uh,
                              const double* __restrict__ vh, const double* __restrict__ Sh, double* __restrict__ uo,
                              double* __restrict__ vo) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    const double theta = 0.25 * c2 * dt * dt * kv.k2, den = 1.0 + theta, sc = c2 * src;
    const double2 u = ld2(uh, id), v = ld2(vh, id), S = ld2(Sh, id);
    double2 a = make_double2(((1.0 - theta) * u.x + dt * v.x + 0.5 * sc * dt * dt * S.x) / den,
                             ((1.0 - theta) * u.y + dt * v.y + 0.5 * sc * dt * dt * S.y) / den);
    double2 b = make_double2(((1.0 - theta) * v.x - c2 * dt * kv.k2 * u.x + sc * dt * S.x) / den,
                             ((1.0 - theta) * v.y - c2 * dt * kv.k2 * u.y + sc * dt * S.y) / den);
    if (kv.excluded) { a = make_double2(0.0, 0.0); b = a; }
    st2(uo, id, a);
    st2(vo, id, b);
}
__global__ void k_zero_mode0(double* zh, int N) {
    reinterpret_cast<double2*>(zh)[(int64_t)blockIdx.x * N] = make_double2(0.0, 0.0);
}
__global__ void k_spectral_continuity(SpecArgs sp, double dt, const double* __restrict__ rn, const double* __restrict__ J,
                                      double* __restrict__ ro) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    const int N = sp.N;
    const double2 a = ld2(J, id), b = ld2(J + 2 * N, id), c = ld2(J + 4 * N, id), r = ld2(rn, id);
    const double2 dv = make_double2(-(kv.k[0] * a.y + kv.k[1] * b.y + kv.k[2] * c.y), kv.k[0] * a.x + kv.k[1] * b.x + kv.k[2] * c.x);
    double2 o = make_double2(r.x - dt * dv.x, r.y - dt * dv.y);
    if (id == 0 || kv.excluded) o = make_double2(0.0, 0.0);
    st2(ro, id, o);
}
void launch_spectral_grad(hipStream_t s, const SpecArgs& sp, const double* fh, double* outh3) {
    hipLaunchKernelGGL(k_spectral_grad, dim3(gblk(sp.N)), dim3(256), 0, s, sp, fh, outh3);
}
void launch_spectral_div(hipStream_t s, const SpecArgs& sp, const double* Vh3, double* outh) {
    hipLaunchKernelGGL(k_spectral_div, dim3(gblk(sp.N)), dim3(256), 0, s, sp, Vh3, outh);
}
void launch_spectral_curl(hipStream_t s, const SpecArgs& sp, const double* Ah3, double* outh3) {
    hipLaunchKernelGGL(k_spectral_curl, dim3(gblk(sp.N)), dim3(256), 0, s, sp, Ah3, outh3);
}
void launch_spectral_poisson(hipStream_t s, const SpecArgs& sp, double src, const double* rhoh, double* phih) {
    hipLaunchKernelGGL(k_spectral_poisson, dim3(gblk(sp.N)), dim3(256), 0, s, sp, src, rhoh, phih);
}
void launch_spectral_cn(hipStream_t s, const SpecArgs& sp, double dt, double c2, double src, const double* uh,
                        const double* vh, const double* Sh, double* uoh, double* voh) {
    hipLaunchKernelGGL(k_spectral_cn, dim3(gblk(sp.N)), dim3(256), 0, s, sp, dt, c2, src, uh, vh, Sh, uoh, voh);
}
void launch_zero_mean_mode(hipStream_t s, double* zh, int batch, int N) {
    hipLaunchKernelGGL(k_zero_mode0, dim3(batch), dim3(1), 0, s, zh, N);
}
void launch_spectral_continuity(hipStream_t s, const SpecArgs& sp, double dt, const double* rhonh, const double* Jh3,
                                double* rhoh) {
    hipLaunchKernelGGL(k_spectral_continuity, dim3(gblk(sp.N)), dim3(256), 0, s, sp, dt, rhonh, Jh3, rhoh);
}

// project_nyquist (P3:2044-2061): per line, subtract the alternating mean.
// A workgroup takes kProjLines lines with kProjThreads threads per line;
// each thread sums every kProjThreads-th element of its line and the line's
// partial sums are added in a fixed order.  ROWS (stride 1): a line's
// threads are adjacent and read adjacent elements; otherwise adjacent
// threads take adjacent lines, which are adjacent in memory.  Four lines of
// 64 threads: a 256^2 field is 64 workgroups per component, each thread
// four elements of its line (16 lines of 16 threads had made 16 workgroups
// of sixteen serial loads per thread, 8.8 us per launch).
constexpr int kProjLines = 4, kProjThreads = 64;
template <bool ROWS>
// src: the field read (the same as f for the in-place pass); each element
// is read and written by the same thread.
__global__ __launch_bounds__(256) void k_project_axis(double* f, const double* src, int N, int n, int stride,
                                                      int lines, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    __shared__ double part[kProjLines][kProjThreads + 1];
    const int t = threadIdx.x;
    const int li = ROWS ? t / kProjThreads : t % kProjLines;
    const int ch = ROWS ? t % kProjThreads : t / kProjLines;
    const int l = blockIdx.x * kProjLines + li;
    double* F = f + (int64_t)blockIdx.y * N;
    const double* S = src + (int64_t)blockIdx.y * N;
    double sum = 0.0;
    int base = 0;
    if (l < lines) {
        base = (l / stride) * (n * stride) + (l % stride);
        for (int j = ch; j < n; j += kProjThreads) sum += (j & 1 ? -1.0 : 1.0) * S[base + j * stride];
    }
    part[li][ch] = sum;
    __syncthreads();
    double tot = 0.0;
    for (int k = 0; k < kProjThreads; ++k) tot += part[li][k];
    const double mean = tot / n;
    if (l < lines)
        for (int j = ch; j < n; j += kProjThreads) F[base + j * stride] = S[base + j * stride] - (j & 1 ? -mean : mean);
}
void launch_project_nyquist_from(hipStream_t s, const SpecArgs& sp, const double* src, double* f, int batch) {
    const int dims[3] = {sp.nx, sp.ny, sp.nz};
    const int strides[3] = {sp.ny * sp.nz, sp.nz, 1};
    const double* from = src;
    if (sp.projection) {
        for (int ax = 0; ax < 3; ++ax) {
            const int n = dims[ax];
            if (n <= 1 || n % 2) continue;
            const int lines = sp.N / n;
            const dim3 gr((lines + kProjLines - 1) / kProjLines, batch), bl(kProjLines * kProjThreads);
            if (strides[ax] == 1)
                hipLaunchKernelGGL(k_project_axis<true>, gr, bl, 0, s, f, from, sp.N, n, strides[ax], lines,
                                   current_gate());
            else
                hipLaunchKernelGGL(k_project_axis<false>, gr, bl, 0, s, f, from, sp.N, n, strides[ax], lines,
                                   current_gate());
            from = f;
        }
    }
    if (from != f) launch_copy(s, (int64_t)batch * sp.N, src, f);
}
void launch_project_nyquist(hipStream_t s, const SpecArgs& sp, double* f, int batch) {
    launch_project_nyquist_from(s, sp, f, f, batch);
}

__global__ void k_zero2(double* __restrict__ a, int na, double* __restrict__ b, int nb, const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0.0;
    else if (i < na + nb) b[i - na] = 0.0;
}
void launch_zero2(hipStream_t s, double* a, int na, double* b, int nb) {
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((na + nb + 255) / 256)), dim3(256), 0, s, a, na, b, nb, current_gate());
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__global__ void k_sum_fields(const double* __restrict__ f, int N, double* __restrict__ out) {
    double v[1] = {0.0};
    const double* F = f + (int64_t)blockIdx.y * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) v[0] += F[i];
    block_reduce_add<1>(v, out + blockIdx.y);
}
__global__ void k_sub_mean(double* __restrict__ f, int N, const double* __restrict__ sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) f[(int64_t)blockIdx.y * N + i] -= sums[blockIdx.y] / (double)N;
}
void launch_subtract_mean(hipStream_t s, double* f, int batch, int N, double* scratch) {
    launch_zero(s, batch, scratch);
    const unsigned nb = (unsigned)std::min<int64_t>(gblk(N), 1024);
    hipLaunchKernelGGL(k_sum_fields, dim3(nb, batch), dim3(256), 0, s, f, N, scratch);
    hipLaunchKernelGGL(k_sub_mean, dim3(gblk(N), batch), dim3(256), 0, s, f, N, scratch);
}

namespace {
thread_local bool t_fixed_order = false;
thread_local const int* t_gate = nullptr;
// launch shape of a grid-stride reduction over n items with at most cap
// workgroups of 256
struct RedShape { dim3 grid, block; };
RedShape red_shape(int64_t n, int64_t cap) {
    if (t_fixed_order) return {dim3(1), dim3(1024)};
    return {dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(gblk(n), cap))), dim3(256)};
}
}  // namespace
bool fixed_order_reductions() { return t_fixed_order; }
const int* current_gate() { return t_gate; }
GateScope::GateScope(const int* g) : prev(t_gate) { t_gate = g; }
GateScope::~GateScope() { t_gate = prev; }
FixedOrderScope::FixedOrderScope(bool on) : prev(t_fixed_order) { t_fixed_order = on; }
FixedOrderScope::~FixedOrderScope() { t_fixed_order = prev; }

__global__ void k_diff_norms(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = a[i], y = b[i], d = x - y;
        v[0] += d * d; v[1] += x * x; v[2] += y * y;
    }
    block_reduce_add<3>(v, out);
}
void launch_diff_norms(hipStream_t s, int64_t n, const double* a, const double* b, double* out3) {
    const RedShape r = red_shape(n, 2048);
    hipLaunchKernelGGL(k_diff_norms, r.grid, r.block, 0, s, n, a, b, out3);
}

struct DotPairs { const double* x[16]; const double* y[16]; };
__global__ void k_dot_many(int64_t n, int np_, DotPairs dp, double* __restrict__ out) {
    const int k = blockIdx.y;
    double v[1] = {0.0};
    const double* x = dp.x[k];
    const double* y = dp.y[k];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[0] += x[i] * y[i];
    block_reduce_add<1>(v, out + k);
}
void launch_dot_many(hipStream_t s, int64_t n, int npairs, const double* const* xs, const double* const* ys, double* out) {
    for (int base = 0; base < npairs; base += 16) {
        DotPairs dp;
        const int cnt = std::min(16, npairs - base);
        for (int k = 0; k < cnt; ++k) { dp.x[k] = xs[base + k]; dp.y[k] = ys[base + k]; }
        const RedShape r = red_shape(n, 512);
        hipLaunchKernelGGL(k_dot_many, dim3(r.grid.x, cnt), r.block, 0, s, n, cnt, dp, out + base);
    }
}

__global__ void k_residual_norms(int N, const double* __restrict__ map, const double* __restrict__ guess, double floor_,
                                 double rtol, double atol_v, const double* __restrict__ atol_dev, double* __restrict__ out,
                                 const int* __restrict__ gate) {
    if (UPIC_GATE_CLOSED(gate)) return;
    const double atol = atol_dev ? *atol_dev : atol_v;
    double sv[3] = {0.0, 0.0, 0.0}, mv[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double m0 = map[i], m1 = map[N + i], m2 = map[2 * N + i];
        const double g0 = guess[i], g1 = guess[N + i], g2 = guess[2 * N + i];
        const double r0 = m0 - g0, r1 = m1 - g1, r2 = m2 - g2;
        const double rr = r0 * r0 + r1 * r1 + r2 * r2;
        const double mm = m0 * m0 + m1 * m1 + m2 * m2, gg = g0 * g0 + g1 * g1 + g2 * g2;
        sv[0] += rr; sv[1] += mm; sv[2] += gg;
        const double rn = sqrt(rr);
        const double vs = 0.5 * (sqrt(mm) + sqrt(gg));
        mv[0] = fmax(mv[0], rn / fmax(floor_, vs));
        mv[1] = fmax(mv[1], rn / (atol + rtol * vs));
        if (!(isfinite(m0) && isfinite(m1) && isfinite(m2))) mv[2] = 1.0;
    }
    block_reduce_add<3>(sv, out);
    block_reduce_max<3>(mv, out + 3);
}
void launch_residual_norms(hipStream_t s, int N, const double* map, const double* guess, double floor_, double rtol,
                           double atol, double* out, const double* atol_dev) {
    const RedShape r = red_shape(N, 1024);
    hipLaunchKernelGGL(k_residual_norms, r.grid, r.block, 0, s, N, map, guess, floor_, rtol, atol, atol_dev, out,
                       current_gate());
}

__global__ void k_outer_atol(const double* __restrict__ sums, int N, double dV, double atol, double rtol,
                             double floor_gross, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const double gross = sums[1] / ((double)N * dV);
    out[0] = fmax(fmax(atol, rtol * sqrt(sums[0] / (double)N)), floor_gross * gross);
}
void launch_outer_atol(hipStream_t s, const double* sums, int N, double dV, double atol, double rtol,
                       double floor_gross, double* out) {
    hipLaunchKernelGGL(k_outer_atol, dim3(1), dim3(64), 0, s, sums, N, dV, atol, rtol, floor_gross, out);
}

// The Gram solve of anderson_type2_candidate (P3:5763-5823) in one lane
// (anderson_gamma, dense_solve.hpp).
__global__ void k_anderson_gamma(int m, const double* __restrict__ gram, double reg, double* __restrict__ ctl) {
    if (threadIdx.x != 0) return;
    anderson_gamma<8>(m, gram, reg, ctl);  // anderson_memory <= 8
}
void launch_anderson_gamma(hipStream_t s, int m, const double* gram, double reg, double* ctl) {
    hipLaunchKernelGGL(k_anderson_gamma, dim3(1), dim3(64), 0, s, m, gram, reg, ctl);
}

// The accumulation is that of k_lincomb (s += coef f, in term order), so a
// combination with the same coefficients gives the same bits.
__global__ void k_lincomb_ctl(int64_t n, LinCombCtl lc, double* __restrict__ ctl, int check_finite,
                              double* __restrict__ out) {
    if (ctl[0] == 0.0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < lc.k; ++k) {
        const double ck = lc.g[k] >= 0 ? (-ctl[lc.g[k]]) * lc.c[k] : lc.c[k];
        s += ck * lc.f[k][i];
    }
    out[i] = s;
    if (check_finite && !isfinite(s)) ctl[1] = 1.0;
}
void launch_lincomb_ctl(hipStream_t s, int64_t n, const LinCombCtl& lc, double* ctl, int check_finite, double* out) {
    if (n > 0) hipLaunchKernelGGL(k_lincomb_ctl, dim3(gblk(n)), dim3(256), 0, s, n, lc, ctl, check_finite, out);
}

__global__ void k_trial_select(int64_t n, const double* __restrict__ guess, const double* __restrict__ resid,
                               double* __restrict__ cand, double one, double damping, double alpha,
                               int use_anderson, const double* __restrict__ ctl, double* __restrict__ trial) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool take = use_anderson && ctl[0] != 0.0 && ctl[1] == 0.0;
    const double g = guess[i];
    double full;
    if (take) {
        full = cand[i];
    } else {
        double s = 0.0;
        s += one * g;
        s += damping * resid[i];
        full = s;
        cand[i] = full;
    }
    double t = 0.0;
    t += one * g;
    t += alpha * full;
    t += (-alpha) * g;
    trial[i] = t;
}
void launch_trial_select(hipStream_t s, int64_t n, const double* guess, const double* resid, double* cand,
                         double damping, double alpha, int use_anderson, const double* ctl, double* trial) {
    if (n > 0)
        hipLaunchKernelGGL(k_trial_select, dim3(gblk(n)), dim3(256), 0, s, n, guess, resid, cand, 1.0, damping, alpha,
                           use_anderson, ctl, trial);
}

// Sums of |g - xl|^2 and |g - xl + r - rl|^2: the outer iterate's change and
// the map's change between two iterations (estimated_map_contraction,
// P3:5983-5996).
__global__ void k_contraction_sums(int64_t n, const double* __restrict__ g, const double* __restrict__ xl,
                                   const double* __restrict__ r, const double* __restrict__ rl,
                                   double* __restrict__ out) {
    double v[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double dx = g[i] - xl[i];
        const double dg = dx + (r[i] - rl[i]);
        v[0] += dx * dx;
        v[1] += dg * dg;
    }
    block_reduce_add<2>(v, out);
}
void launch_contraction_sums(hipStream_t s, int64_t n, const double* g, const double* xl, const double* r,
                             const double* rl, double* out) {
    const RedShape sh = red_shape(n, 1024);
    hipLaunchKernelGGL(k_contraction_sums, sh.grid, sh.block, 0, s, n, g, xl, r, rl, out);
}

__global__ void k_sum_squares(int64_t n, const double* __restrict__ f, double* __restrict__ out) {
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[0] += f[i] * f[i];
    block_reduce_add<1>(v, out);
}
__global__ void k_sum(int64_t n, const double* __restrict__ f, double* __restrict__ out) {
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[0] += f[i];
    block_reduce_add<1>(v, out);
}
__global__ void k_max_abs(int64_t n, const double* __restrict__ f, double* __restrict__ out) {
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[0] = fmax(v[0], fabs(f[i]));
    block_reduce_max<1>(v, out);
}
void launch_sum_squares(hipStream_t s, int64_t n, const double* f, double* out) {
    const RedShape r = red_shape(n, 1024);
    hipLaunchKernelGGL(k_sum_squares, r.grid, r.block, 0, s, n, f, out);
}
void launch_sum(hipStream_t s, int64_t n, const double* f, double* out) {
    const RedShape r = red_shape(n, 1024);
    hipLaunchKernelGGL(k_sum, r.grid, r.block, 0, s, n, f, out);
}
void launch_max_abs(hipStream_t s, int64_t n, const double* f, double* out) {
    hipLaunchKernelGGL(k_max_abs, dim3((unsigned)std::min<int64_t>(gblk(n), 1024)), dim3(256), 0, s, n, f, out);
}

// kinetic energy: P1 0.5 m |v|^2 (P1:1031-1035); P3 m c^2 (gamma - 1)
// (P3:3457-3467); P2 also max speed and sum of gamma (P2:946-956).
__global__ void k_kinetic(int64_t np, const double* __restrict__ v, const double* __restrict__ m, double c, int rel,
                          double* __restrict__ out) {
    double s[2] = {0.0, 0.0}, mx[1] = {0.0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const double a = v[p], b = v[np + p], cc = v[2 * np + p];
        const double v2 = a * a + b * b + cc * cc;
        if (rel == 0) {
            s[0] += 0.5 * m[p] * v2;
        } else {
            const double beta2 = v2 / (c * c);
            const double gam = (rel == 2) ? 1.0 / sqrt(fmax(1.0e-14, 1.0 - beta2)) : 1.0 / sqrt(fmax(1.0e-300, 1.0 - beta2));
            s[0] += m[p] * c * c * (gam - 1.0);
            s[1] += gam;
        }
        mx[0] = fmax(mx[0], sqrt(v2));
    }
    block_reduce_add<2>(s, out);
    block_reduce_max<1>(mx, out + 2);
}
void launch_kinetic(hipStream_t s, int64_t np, const double* v, const double* m, double c, int relativistic, double* out3) {
    const RedShape r = red_shape(np, 2048);
    hipLaunchKernelGGL(k_kinetic, r.grid, r.block, 0, s, np, v, m, c, relativistic, out3);
}

// ---------------------------------------------------------------------------
// Cell sort (counting sort by the cell containing x)
// ---------------------------------------------------------------------------
__global__ void k_cell_keys(GridDev g, int64_t np, const double* __restrict__ x, int* __restrict__ keys, int* __restrict__ counts) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < np;
    int key = 0;
    if (active) {
        int c[3];
        const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xw = wrap1(x[d * np + p], g.L[d]);
            int i = (int)floor((xw + 0.5 * g.L[d]) / g.h[d]);
            i = i < 0 ? 0 : (i >= n[d] ? n[d] - 1 : i);
            c[d] = i;
        }
        key = (c[0] * g.ny + c[1]) * g.nz + c[2];
        keys[p] = key;
    }
    (void)wave_claim(counts, key, active);
}
// Exclusive scan of the cell counts in three passes: workgroups of 1024
// threads scan chunks of kScanChunk counts (4 per thread) and record chunk
// totals, one workgroup scans the totals, and the totals are added back.
constexpr int kScanChunk = 4096;
__device__ __forceinline__ int block_exclusive_scan_1024(int v, int* sh, int& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int x = (t >= o) ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[1023];
    const int r = sh[t] - v;
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(1024) void k_scan_local(int ncell, const int* __restrict__ counts,
                                                     int* __restrict__ offsets, int* __restrict__ totals) {
    __shared__ int sh[1024];
    const int base = blockIdx.x * kScanChunk + 4 * threadIdx.x;
    int c[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = base + k < ncell ? counts[base + k] : 0;
        sum += c[k];
    }
    int total;
    int run = block_exclusive_scan_1024(sum, sh, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < ncell) offsets[base + k] = run;
        run += c[k];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}
__global__ __launch_bounds__(1024) void k_scan_totals(int nb, int* __restrict__ totals) {
    __shared__ int sh[1024];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int i = b0 + (int)threadIdx.x;
        const int v = i < nb ? totals[i] : 0;
        int total;
        const int r = block_exclusive_scan_1024(v, sh, total);
        if (i < nb) totals[i] = carry + r;
        carry += total;
    }
}
__global__ void k_scan_add(int ncell, int* __restrict__ offsets, const int* __restrict__ totals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) offsets[i] += totals[i / kScanChunk];
}
__global__ void k_scatter_perm(int64_t np, const int* __restrict__ keys, int* __restrict__ cursor, int* __restrict__ perm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < np;
    const int key = active ? keys[p] : 0;
    const int pos = wave_claim(cursor, key, active);
    if (active) perm[pos] = (int)p;
}
// One workgroup per cell: each entry's rank is the number of entries of the
// cell with a smaller old index (entries are distinct), read from LDS for
// cells of up to kStableLds entries and from HBM beyond.
constexpr int kStableLds = 4096;
__global__ __launch_bounds__(256) void k_stable_cells(int ncell, const int* __restrict__ cell_end,
                                                      const int* __restrict__ perm, int* __restrict__ out) {
    __shared__ int v[kStableLds];
    const int c = blockIdx.x;
    const int b = c ? cell_end[c - 1] : 0, e = cell_end[c], n = e - b;
    const bool lds = n <= kStableLds;
    if (lds)
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = perm[b + i];
    __syncthreads();
    const int* src = lds ? v : perm + b;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int x = src[i];
        int r = 0;
        for (int j = 0; j < n; ++j) r += src[j] < x ? 1 : 0;
        out[b + r] = x;
    }
}
void launch_stable_cells(hipStream_t s, int ncell, const int* cell_end, const int* perm, int* out) {
    if (ncell > 0) hipLaunchKernelGGL(k_stable_cells, dim3(ncell), dim3(256), 0, s, ncell, cell_end, perm, out);
}
__global__ void k_permute(int64_t np, int ncomp, const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t j = perm[i];
    for (int c = 0; c < ncomp; ++c) dst[c * np + i] = src[c * np + j];
}
__global__ void k_permute_i64(int64_t np, const int64_t* __restrict__ src, int64_t* __restrict__ dst, const int* __restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[i] = src[perm[i]];
}
void launch_cell_keys(hipStream_t s, const GridDev& g, int64_t np, const double* x, int* keys, int* counts) {
    if (np > 0) hipLaunchKernelGGL(k_cell_keys, dim3(gblk(np)), dim3(256), 0, s, g, np, x, keys, counts);
}
int scan_scratch_ints(int ncell) { return (ncell + kScanChunk - 1) / kScanChunk + 1; }
void launch_scan_counts(hipStream_t s, int ncell, int* counts, int* offsets, int* scratch) {
    if (ncell <= 0) return;
    const int nb = (ncell + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(1024), 0, s, ncell, counts, offsets, scratch);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, s, nb, scratch);
    hipLaunchKernelGGL(k_scan_add, dim3((ncell + 255) / 256), dim3(256), 0, s, ncell, offsets, scratch);
}
void launch_scatter_perm(hipStream_t s, int64_t np, const int* keys, int* cursor, int* perm) {
    if (np > 0) hipLaunchKernelGGL(k_scatter_perm, dim3(gblk(np)), dim3(256), 0, s, np, keys, cursor, perm);
}
void launch_permute(hipStream_t s, int64_t np, int ncomp, const double* src, double* dst, const int* perm) {
    if (np > 0) hipLaunchKernelGGL(k_permute, dim3(gblk(np)), dim3(256), 0, s, np, ncomp, src, dst, perm);
}
void launch_permute_i64(hipStream_t s, int64_t np, const int64_t* src, int64_t* dst, const int* perm) {
    if (np > 0) hipLaunchKernelGGL(k_permute_i64, dim3(gblk(np)), dim3(256), 0, s, np, src, dst, perm);
}

}  // namespace upic

namespace upic {
__global__ void k_lincomb(int64_t n, LinComb lc, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < lc.k; ++k) s += lc.c[k] * lc.f[k][i];
    out[i] = s;
}
void launch_lincomb(hipStream_t s, int64_t n, const LinComb& lc, double* out) {
    if (n > 0) hipLaunchKernelGGL(k_lincomb, dim3(gblk(n)), dim3(256), 0, s, n, lc, out);
}
__global__ void k_spectral_div_shift(SpecArgs sp, double a, double* zh) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= sp.N) return;
    const KVec kv = kvec(sp, id);
    double2* z = reinterpret_cast<double2*>(zh) + (int64_t)blockIdx.y * sp.N;
    const double den = a + kv.k2;
    double2 v = z[id];
    v.x /= den; v.y /= den;
    z[id] = v;
}
void launch_spectral_div_shift(hipStream_t s, const SpecArgs& sp, double a, double* zh, int batch) {
    hipLaunchKernelGGL(k_spectral_div_shift, dim3(gblk(sp.N), batch), dim3(256), 0, s, sp, a, zh);
}
}  // namespace upic
