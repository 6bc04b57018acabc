// oracle/upic_oracle.cpp -- CPU restatement of the reference FP64 hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU path is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it; the product library (unstaggered_pic_amd) never links,
// loads or calls it.
//
// It restates, loop for loop, the serial FP64 algorithms of the three
// reference codes (cited file:line, paths relative to the reference root):
//   P1 = Project1/ec_pic_serial_demo.cpp
//   P2 = Project2/spectral_pic_part3.cpp
//   P3 = Project3/ec_pic_nyquist_projected.cpp
// in their general 3-D form (nz >= 1), with the reference's summation orders,
// FFT twiddle recurrence and long-double accumulations, so that its results
// agree with the reference to roundoff.  Each part is pinned by the checks in
// tests/test_oracle.py (golden vectors, closed forms, brute force).
//
// Layout at the C boundary: the upic.h conventions (particle arrays 3*np
// component-major, vector fields 3*N component-major).

#include "../include/upic.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using Real = double;
using Complex = std::complex<Real>;
using Field = std::vector<Real>;
using CField = std::vector<Complex>;
using Vec3 = std::array<Real, 3>;
using Mat3 = std::array<std::array<Real, 3>, 3>;
using VecField = std::array<Field, 3>;

static constexpr Real PI = 3.141592653589793238462643383279502884;
static thread_local std::string g_err;

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------------------
// Small vector helpers (P1:136-227, P3:471-831)
// ---------------------------------------------------------------------------
static Vec3 add3(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
static Vec3 sub3(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static Vec3 mul3(Real c, const Vec3& a) { return {c * a[0], c * a[1], c * a[2]}; }
static Real dot3(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static Real norm3(const Vec3& a) { return std::sqrt(dot3(a, a)); }
static Vec3 cross3(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
static Vec3 matvec3(const Mat3& M, const Vec3& v) {
    return {M[0][0] * v[0] + M[0][1] * v[1] + M[0][2] * v[2],
            M[1][0] * v[0] + M[1][1] * v[1] + M[1][2] * v[2],
            M[2][0] * v[0] + M[2][1] * v[1] + M[2][2] * v[2]};
}
static Mat3 zero_mat3() { Mat3 M{}; for (auto& r : M) r = {0.0, 0.0, 0.0}; return M; }
static VecField zero_vecfield(std::size_t n) { return {Field(n, 0.0), Field(n, 0.0), Field(n, 0.0)}; }
static bool finite_vec3(const Vec3& v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// wrap1: P1:160-164, P3:704-708 -- fmod into [-L/2, L/2)
static Real wrap1(Real x, Real L) {
    Real y = std::fmod(x + 0.5 * L, L);
    if (y < 0.0) y += L;
    return y - 0.5 * L;
}
// wrap_periodic: P2:161-167 -- floor into [0, L)
static Real wrap_periodic(Real x, Real L) {
    x -= std::floor(x / L) * L;
    if (x >= L) x -= L;
    if (x < 0.0) x += L;
    return x;
}

// rms helpers with long-double accumulation (P1:166-213, P3:739-785)
static Real rms_field(const Field& f) {
    long double s = 0.0L;
    for (Real x : f) s += static_cast<long double>(x) * static_cast<long double>(x);
    return std::sqrt(static_cast<Real>(s / std::max<std::size_t>(1, f.size())));
}
static Real rms_vec_particles(const std::vector<Vec3>& a) {
    long double s = 0.0L;
    std::size_t count = 0;
    for (const auto& v : a)
        for (int d = 0; d < 3; ++d) { s += static_cast<long double>(v[d]) * static_cast<long double>(v[d]); ++count; }
    return std::sqrt(static_cast<Real>(s / std::max<std::size_t>(1, count)));
}
static Real relative_rms_error_vec_particles(const std::vector<Vec3>& a, const std::vector<Vec3>& b) {
    std::vector<Vec3> diff(a.size());
    for (std::size_t p = 0; p < a.size(); ++p) diff[p] = sub3(a[p], b[p]);
    const Real denom = 0.5 * (rms_vec_particles(a) + rms_vec_particles(b)) + 1.0e-300;
    return rms_vec_particles(diff) / denom;
}
static Real relative_rms_error_vecfield(const VecField& a, const VecField& b) {
    long double sd = 0.0L, sa = 0.0L, sb = 0.0L;
    std::size_t count = 0;
    for (int m = 0; m < 3; ++m)
        for (std::size_t i = 0; i < a[m].size(); ++i) {
            const long double da = static_cast<long double>(a[m][i] - b[m][i]);
            sd += da * da;
            sa += static_cast<long double>(a[m][i]) * static_cast<long double>(a[m][i]);
            sb += static_cast<long double>(b[m][i]) * static_cast<long double>(b[m][i]);
            ++count;
        }
    const Real rd = std::sqrt(static_cast<Real>(sd / std::max<std::size_t>(1, count)));
    const Real ra = std::sqrt(static_cast<Real>(sa / std::max<std::size_t>(1, count)));
    const Real rb = std::sqrt(static_cast<Real>(sb / std::max<std::size_t>(1, count)));
    return rd / (0.5 * (ra + rb) + 1.0e-300);
}
static Real mean_field(const Field& f) {
    long double s = 0.0L;
    for (Real x : f) s += x;
    return static_cast<Real>(s / std::max<std::size_t>(1, f.size()));
}
static void subtract_mean(Field& f) { const Real m = mean_field(f); for (Real& x : f) x -= m; }

// ---------------------------------------------------------------------------
// Grid (P1:373-418, P2:298-357, P3:1441-1546)
// ---------------------------------------------------------------------------
struct Grid {
    int code = 1;
    int nx = 0, ny = 0, nz = 0, N = 0;
    Vec3 L{}, dx{};
    Real dV = 0.0;
    Field kx, ky, kz, k2;
    bool nyquist_projection = false;
    int index(int i, int j, int k) const { return (i * ny + j) * nz + k; }
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
static bool is_even_nyquist_index(int i, int n) { return n > 1 && (n % 2 == 0) && i == n / 2; }

static int fft_mode_number(int i, int n, int code) {
    if (code == 1) {  // P1:386-392 keeps the Nyquist mode as +n/2
        if (n == 1) return 0;
        return (i <= n / 2) ? i : i - n;
    }
    if (code == 3) {  // P3:1485-1504 zeroes the even-grid Nyquist mode
        if (n == 1) return 0;
        if (is_even_nyquist_index(i, n)) return 0;
        return (i <= n / 2) ? i : i - n;
    }
    // P2:320-330
    if (n > 1 && i == n / 2) return 0;
    return (i < n / 2) ? i : i - n;
}

static Grid make_grid(const upic_params& p) {
    if (!is_pow2(p.nx) || !is_pow2(p.ny) || !is_pow2(p.nz))
        throw Failure(UPIC_EINVAL, "the in-file radix-2 FFT requires nx, ny, nz to be powers of two");
    Grid g;
    g.code = p.code;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.N = g.nx * g.ny * g.nz;
    g.L = {p.Lx, p.Ly, p.Lz};
    g.dx = {p.Lx / p.nx, p.Ly / p.ny, p.Lz / p.nz};
    g.dV = g.dx[0] * g.dx[1] * g.dx[2];
    g.nyquist_projection = (p.code == 3) && p.nyquist_projection;
    g.kx.assign(g.N, 0.0); g.ky.assign(g.N, 0.0); g.kz.assign(g.N, 0.0); g.k2.assign(g.N, 0.0);
    for (int i = 0; i < g.nx; ++i) {
        const Real kxi = 2.0 * PI * static_cast<Real>(fft_mode_number(i, g.nx, g.code)) / g.L[0];
        for (int j = 0; j < g.ny; ++j) {
            const Real kyj = 2.0 * PI * static_cast<Real>(fft_mode_number(j, g.ny, g.code)) / g.L[1];
            for (int k = 0; k < g.nz; ++k) {
                // P2 is a 2-D code with no z axis at all (P2:298-357)
                const Real kzk = (g.code == 2) ? 0.0 : 2.0 * PI * static_cast<Real>(fft_mode_number(k, g.nz, g.code)) / g.L[2];
                const int id = g.index(i, j, k);
                g.kx[id] = kxi; g.ky[id] = kyj; g.kz[id] = kzk;
                g.k2[id] = kxi * kxi + kyj * kyj + kzk * kzk;
            }
        }
    }
    return g;
}

// ---------------------------------------------------------------------------
// Radix-2 FFT (P1:424-499, P2:372-434, P3:1566-1700): bit reversal, then
// Danielson-Lanczos stages with the twiddle advanced by repeated multiply.
// ---------------------------------------------------------------------------
static void fft1d(std::vector<Complex>& a, bool inverse) {
    const int n = static_cast<int>(a.size());
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const Real angle = 2.0 * PI / static_cast<Real>(len) * (inverse ? 1.0 : -1.0);
        const Complex wlen(std::cos(angle), std::sin(angle));
        for (int i = 0; i < n; i += len) {
            Complex w(1.0, 0.0);
            for (int j = 0; j < len / 2; ++j) {
                const Complex u = a[i + j];
                const Complex v = a[i + j + len / 2] * w;
                a[i + j] = u + v;
                a[i + j + len / 2] = u - v;
                w *= wlen;
            }
        }
    }
    if (inverse) {
        const Real inv_n = 1.0 / static_cast<Real>(n);
        for (auto& z : a) z *= inv_n;
    }
}

// z, y, x passes (P1:456-485).  P2's fft2d (y then x, 406-424) is the same
// sequence with nz = 1 (the length-1 z pass is the identity).
static void fft3d(CField& data, const Grid& g, bool inverse) {
    std::vector<Complex> tmp(g.nz);
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j) {
            for (int k = 0; k < g.nz; ++k) tmp[k] = data[g.index(i, j, k)];
            fft1d(tmp, inverse);
            for (int k = 0; k < g.nz; ++k) data[g.index(i, j, k)] = tmp[k];
        }
    tmp.resize(g.ny);
    for (int i = 0; i < g.nx; ++i)
        for (int k = 0; k < g.nz; ++k) {
            for (int j = 0; j < g.ny; ++j) tmp[j] = data[g.index(i, j, k)];
            fft1d(tmp, inverse);
            for (int j = 0; j < g.ny; ++j) data[g.index(i, j, k)] = tmp[j];
        }
    tmp.resize(g.nx);
    for (int j = 0; j < g.ny; ++j)
        for (int k = 0; k < g.nz; ++k) {
            for (int i = 0; i < g.nx; ++i) tmp[i] = data[g.index(i, j, k)];
            fft1d(tmp, inverse);
            for (int i = 0; i < g.nx; ++i) data[g.index(i, j, k)] = tmp[i];
        }
}
static CField fft_real(const Field& f, const Grid& g) {
    CField out(g.N);
    for (int id = 0; id < g.N; ++id) out[id] = Complex(f[id], 0.0);
    fft3d(out, g, false);
    return out;
}
static Field ifft_real(CField fhat, const Grid& g) {
    fft3d(fhat, g, true);
    Field f(g.N);
    for (int id = 0; id < g.N; ++id) f[id] = fhat[id].real();
    return f;
}

// Nyquist projector (P3:2044-2098)
static void project_nyquist(Field& f, const Grid& g) {
    if (!g.nyquist_projection) return;
    const int dims[3] = {g.nx, g.ny, g.nz};
    const int strides[3] = {g.ny * g.nz, g.nz, 1};
    for (int axis = 0; axis < 3; ++axis) {
        const int n = dims[axis], stride = strides[axis];
        if (n <= 1 || n % 2) continue;
        const int lines = g.N / n;
        for (int line = 0; line < lines; ++line) {
            const int base = (line / stride) * (n * stride) + line % stride;
            long double sum = 0;
            for (int j = 0; j < n; ++j) sum += (j % 2 ? -1.0L : 1.0L) * f[base + j * stride];
            const Real mean = static_cast<Real>(sum / n);
            for (int j = 0; j < n; ++j) f[base + j * stride] -= (j % 2 ? -mean : mean);
        }
    }
}
static void project_nyquist(VecField& f, const Grid& g) { for (auto& c : f) project_nyquist(c, g); }
static bool excluded_nyquist_mode(int id, const Grid& g) {
    if (!g.nyquist_projection) return false;
    const int i = id / (g.ny * g.nz), j = (id / g.nz) % g.ny, k = id % g.nz;
    return is_even_nyquist_index(i, g.nx) || is_even_nyquist_index(j, g.ny) || is_even_nyquist_index(k, g.nz);
}

// Spectral operators (P1:505-546, P3:1721-1795)
static VecField spectral_gradient_scalar(const Field& f, const Grid& g) {
    const Complex I(0.0, 1.0);
    const CField fhat = fft_real(f, g);
    VecField out = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) {
        CField dhat(g.N);
        for (int id = 0; id < g.N; ++id) {
            const Real kk = (d == 0) ? g.kx[id] : (d == 1) ? g.ky[id] : g.kz[id];
            dhat[id] = I * kk * fhat[id];
        }
        out[d] = ifft_real(std::move(dhat), g);
    }
    return out;
}
static Field spectral_divergence_vector(const VecField& V, const Grid& g) {
    const Complex I(0.0, 1.0);
    const CField hx = fft_real(V[0], g), hy = fft_real(V[1], g), hz = fft_real(V[2], g);
    CField divhat(g.N);
    for (int id = 0; id < g.N; ++id) divhat[id] = I * (g.kx[id] * hx[id] + g.ky[id] * hy[id] + g.kz[id] * hz[id]);
    return ifft_real(std::move(divhat), g);
}
static VecField spectral_curl(const VecField& A, const Grid& g) {
    const Complex I(0.0, 1.0);
    const CField ax = fft_real(A[0], g), ay = fft_real(A[1], g), az = fft_real(A[2], g);
    std::array<CField, 3> ch{CField(g.N), CField(g.N), CField(g.N)};
    for (int id = 0; id < g.N; ++id) {
        ch[0][id] = I * (g.ky[id] * az[id] - g.kz[id] * ay[id]);
        ch[1][id] = I * (g.kz[id] * ax[id] - g.kx[id] * az[id]);
        ch[2][id] = I * (g.kx[id] * ay[id] - g.ky[id] * ax[id]);
    }
    VecField curl = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) curl[d] = ifft_real(std::move(ch[d]), g);
    return curl;
}

// ---------------------------------------------------------------------------
// B-splines (P1:552-595 recursion; P3:1856-1986 closed forms)
// ---------------------------------------------------------------------------
static Real cardinal_bspline(int order, Real t) {
    if (order == 1) return (t >= 0.0 && t < 1.0) ? 1.0 : 0.0;
    return (t / static_cast<Real>(order - 1)) * cardinal_bspline(order - 1, t)
         + ((static_cast<Real>(order) - t) / static_cast<Real>(order - 1)) * cardinal_bspline(order - 1, t - 1.0);
}
static Real cardinal_bspline_derivative(int order, Real t) {
    if (order <= 1) return 0.0;
    return cardinal_bspline(order - 1, t) - cardinal_bspline(order - 1, t - 1.0);
}
static Real closed_bspline(int degree, Real r) {
    if (degree == 1) {
        if (r >= -1.0 && r < 0.0) return 1.0 + r;
        if (r >= 0.0 && r < 1.0) return 1.0 - r;
        return 0.0;
    }
    const Real a = std::abs(r);
    if (degree == 2) {
        if (a < 0.5) return 0.75 - a * a;
        if (a < 1.5) { const Real z = 1.5 - a; return 0.5 * z * z; }
        return 0.0;
    }
    if (degree == 3) {
        if (a < 1.0) return (2.0 / 3.0) - a * a + 0.5 * a * a * a;
        if (a < 2.0) { const Real z = 2.0 - a; return z * z * z / 6.0; }
        return 0.0;
    }
    if (a < 0.5) { const Real a2 = a * a; return 115.0 / 192.0 - (5.0 / 8.0) * a2 + 0.25 * a2 * a2; }
    if (a < 1.5) return 55.0 / 96.0 + (5.0 / 24.0) * a - 1.25 * a * a + (5.0 / 6.0) * a * a * a - (1.0 / 6.0) * a * a * a * a;
    if (a < 2.5) { const Real z = 2.5 - a; const Real z2 = z * z; return z2 * z2 / 24.0; }
    return 0.0;
}
static Real closed_bspline_derivative(int degree, Real r) {
    if (degree == 1) {
        if (r >= -1.0 && r < 0.0) return 1.0;
        if (r >= 0.0 && r < 1.0) return -1.0;
        return 0.0;
    }
    const Real a = std::abs(r);
    const Real sign = (r < 0.0) ? -1.0 : 1.0;
    if (degree == 2) {
        if (a < 0.5) return -2.0 * r;
        if (a < 1.5) return -sign * (1.5 - a);
        return 0.0;
    }
    if (degree == 3) {
        if (a < 1.0) return sign * (-2.0 * a + 1.5 * a * a);
        if (a < 2.0) { const Real z = 2.0 - a; return -0.5 * sign * z * z; }
        return 0.0;
    }
    if (a < 0.5) return sign * (-(5.0 / 4.0) * a + a * a * a);
    if (a < 1.5) return sign * (5.0 / 24.0 - 2.5 * a + 2.5 * a * a - (2.0 / 3.0) * a * a * a);
    if (a < 2.5) { const Real z = 2.5 - a; return -sign * z * z * z / 6.0; }
    return 0.0;
}
static Real centered_bspline(int code, int degree, Real r) {
    if (degree < 1 || degree > 4) throw Failure(UPIC_EINVAL, "spline_order must be 1, 2, 3, or 4");
    if (code == 1) return cardinal_bspline(degree + 1, r + 0.5 * static_cast<Real>(degree + 1));
    return closed_bspline(degree, r);
}
static Real centered_bspline_derivative(int code, int degree, Real r) {
    if (degree < 1 || degree > 4) throw Failure(UPIC_EINVAL, "spline_order must be 1, 2, 3, or 4");
    if (code == 1) return cardinal_bspline_derivative(degree + 1, r + 0.5 * static_cast<Real>(degree + 1));
    return closed_bspline_derivative(degree, r);
}

struct AxisStencil {
    std::array<int, 5> idx{};
    std::array<Real, 5> w{};
    std::array<Real, 5> dw_dx{};
};

// axis_stencil (P1:580-595, P3:1971-1986): nodes at x = -L/2 + i dx.
static AxisStencil axis_stencil(int code, Real x, int n, Real L, int order) {
    AxisStencil s;
    const Real dx = L / static_cast<Real>(n);
    const Real u = (x + 0.5 * L) / dx;
    const int i_left = static_cast<int>(std::floor(u - 0.5 * static_cast<Real>(order - 1)));
    for (int a = 0; a <= order; ++a) {
        const int raw = i_left + a;
        int ii = raw % n;
        if (ii < 0) ii += n;
        const Real r = u - static_cast<Real>(raw);
        s.idx[a] = ii;
        s.w[a] = centered_bspline(code, order, r);
        s.dw_dx[a] = centered_bspline_derivative(code, order, r) / dx;
    }
    return s;
}

static Vec3 wrap_pos(const Vec3& x, const Grid& g) {
    return {wrap1(x[0], g.L[0]), wrap1(x[1], g.L[1]), wrap1(x[2], g.L[2])};
}

// ---------------------------------------------------------------------------
// Point deposit / gather (P1:601-710, P3:2009-2407)
// ---------------------------------------------------------------------------
struct Stencil3 { AxisStencil s[3]; };
static Stencil3 stencil_at(const Vec3& pos, const Grid& g, int order) {
    const Vec3 x = wrap_pos(pos, g);
    return {{axis_stencil(g.code, x[0], g.nx, g.L[0], order), axis_stencil(g.code, x[1], g.ny, g.L[1], order),
             axis_stencil(g.code, x[2], g.nz, g.L[2], order)}};
}
static void deposit_scalar_add(Field& mesh, const Vec3& pos, Real value, const Grid& g, int order, bool divide_by_dV) {
    const Stencil3 st = stencil_at(pos, g, order);
    const Real scale = divide_by_dV ? (1.0 / g.dV) : 1.0;
    for (int a = 0; a <= order; ++a)
        for (int b = 0; b <= order; ++b)
            for (int c = 0; c <= order; ++c) {
                const int id = g.index(st.s[0].idx[a], st.s[1].idx[b], st.s[2].idx[c]);
                mesh[id] += value * st.s[0].w[a] * st.s[1].w[b] * st.s[2].w[c] * scale;
            }
}
static Field deposit_scalar_to_mesh(const std::vector<Vec3>& pos, const std::vector<Real>& values, const Grid& g,
                                    int order, bool divide_by_dV) {
    Field mesh(g.N, 0.0);
    for (std::size_t p = 0; p < pos.size(); ++p) deposit_scalar_add(mesh, pos[p], values[p], g, order, divide_by_dV);
    project_nyquist(mesh, g);
    return mesh;
}
static void deposit_current_add(VecField& J, const Vec3& pos, const Vec3& vel, Real charge, const Grid& g, int order,
                                Real prefactor) {
    const Stencil3 st = stencil_at(pos, g, order);
    for (int a = 0; a <= order; ++a)
        for (int b = 0; b <= order; ++b)
            for (int c = 0; c <= order; ++c) {
                const int id = g.index(st.s[0].idx[a], st.s[1].idx[b], st.s[2].idx[c]);
                const Real w = prefactor * charge * st.s[0].w[a] * st.s[1].w[b] * st.s[2].w[c] / g.dV;
                for (int d = 0; d < 3; ++d) J[d][id] += w * vel[d];
            }
}
static VecField deposit_current_to_mesh(const std::vector<Vec3>& pos, const std::vector<Vec3>& vel,
                                        const std::vector<Real>& q, const Grid& g, int order) {
    VecField J = zero_vecfield(g.N);
    for (std::size_t p = 0; p < pos.size(); ++p) deposit_current_add(J, pos[p], vel[p], q[p], g, order, 1.0);
    project_nyquist(J, g);
    return J;
}
static Vec3 gather_vector_at(const VecField& mesh, const Vec3& pos, const Grid& g, int order) {
    const Stencil3 st = stencil_at(pos, g, order);
    Vec3 out{0.0, 0.0, 0.0};
    if (g.code == 1) {  // P1:649-672: one scalar gather per component
        for (int comp = 0; comp < 3; ++comp) {
            Real o = 0.0;
            for (int a = 0; a <= order; ++a)
                for (int b = 0; b <= order; ++b)
                    for (int c = 0; c <= order; ++c) {
                        const int id = g.index(st.s[0].idx[a], st.s[1].idx[b], st.s[2].idx[c]);
                        o += mesh[comp][id] * st.s[0].w[a] * st.s[1].w[b] * st.s[2].w[c];
                    }
            out[comp] = o;
        }
        return out;
    }
    // P3:2267-2286: shared xy weight product
    for (int a = 0; a <= order; ++a)
        for (int b = 0; b <= order; ++b) {
            const Real wxy = st.s[0].w[a] * st.s[1].w[b];
            for (int c = 0; c <= order; ++c) {
                const int id = g.index(st.s[0].idx[a], st.s[1].idx[b], st.s[2].idx[c]);
                const Real w = wxy * st.s[2].w[c];
                out[0] += mesh[0][id] * w;
                out[1] += mesh[1][id] * w;
                out[2] += mesh[2][id] * w;
            }
        }
    return out;
}
static Mat3 gather_shape_gradient_vector_interp_at(const VecField& A0, const VecField& A1, Real s_interp,
                                                   const Vec3& pos, const Grid& g, int order) {
    const Stencil3 st = stencil_at(pos, g, order);
    Mat3 out = zero_mat3();
    for (int a = 0; a <= order; ++a)
        for (int b = 0; b <= order; ++b)
            for (int c = 0; c <= order; ++c) {
                const int id = g.index(st.s[0].idx[a], st.s[1].idx[b], st.s[2].idx[c]);
                const Real wx = st.s[0].w[a], wy = st.s[1].w[b], wz = st.s[2].w[c];
                const Real dwx = st.s[0].dw_dx[a], dwy = st.s[1].dw_dx[b], dwz = st.s[2].dw_dx[c];
                const Real dSdx[3] = {dwx * wy * wz, wx * dwy * wz, wx * wy * dwz};
                for (int comp = 0; comp < 3; ++comp) {
                    const Real Aval = (1.0 - s_interp) * A0[comp][id] + s_interp * A1[comp][id];
                    for (int d = 0; d < 3; ++d) out[comp][d] += Aval * dSdx[d];
                }
            }
    return out;
}

// ---------------------------------------------------------------------------
// Quadrature (P1:719-763, P3:2434-2503)
// ---------------------------------------------------------------------------
struct Quadrature { std::vector<Real> s, w; };
static std::pair<Real, Real> legendre_value_and_derivative(int n, Real x) {
    Real p0 = 1.0, p1 = x;
    if (n == 0) return {p0, 0.0};
    if (n == 1) return {p1, 1.0};
    for (int k = 2; k <= n; ++k) {
        const Real pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / static_cast<Real>(k);
        p0 = p1; p1 = pk;
    }
    return {p1, static_cast<Real>(n) * (x * p1 - p0) / (x * x - 1.0)};
}
static Quadrature gauss_legendre_01(int n) {
    Quadrature q;
    q.s.assign(n, 0.0); q.w.assign(n, 0.0);
    const int m = (n + 1) / 2;
    for (int i = 0; i < m; ++i) {
        Real x = std::cos(PI * (static_cast<Real>(i) + 0.75) / (static_cast<Real>(n) + 0.5));
        for (int it = 0; it < 80; ++it) {
            const auto pd = legendre_value_and_derivative(n, x);
            const Real dx = -pd.first / pd.second;
            x += dx;
            if (std::abs(dx) < 1.0e-15) break;
        }
        const auto pd = legendre_value_and_derivative(n, x);
        const Real w = 2.0 / ((1.0 - x * x) * pd.second * pd.second);
        q.s[i] = 0.5 * (1.0 - x); q.w[i] = 0.5 * w;
        q.s[n - 1 - i] = 0.5 * (1.0 + x); q.w[n - 1 - i] = 0.5 * w;
    }
    return q;
}

// ---------------------------------------------------------------------------
// Orbit integration (P1:771-934, P3:2509-2984)
// ---------------------------------------------------------------------------
struct PathStats { std::size_t n_particles = 0, crossing_particles = 0, total_segments = 0; int max_segments = 1; };

// P1 wraps the path point (P1:778-781); P3 keeps it unwrapped (P3:2531-2536).
static Vec3 path_position(const Vec3& x0, const Vec3& dxp, Real s, const Grid& g) {
    const Vec3 x = add3(x0, mul3(s, dxp));
    return (g.code == 1) ? wrap_pos(x, g) : x;
}

static std::vector<Real> path_breakpoints_for_particle(const Vec3& x0, const Vec3& dx_path, const Grid& g, int order,
                                                       bool split_at_knots) {
    std::vector<Real> b;
    b.reserve(16);
    b.push_back(0.0);
    b.push_back(1.0);
    if (!split_at_knots) return b;
    const Real tol = 1.0e-13;
    for (int d = 0; d < 3; ++d) {
        const Real Ld = g.L[d], hd = g.dx[d];
        const Real u0 = (wrap1(x0[d], Ld) + 0.5 * Ld) / hd;
        const Real du = dx_path[d] / hd;
        if (std::abs(du) < 1.0e-300) continue;
        const Real offset = 0.5 * static_cast<Real>(order - 1);
        const Real a0 = u0 - offset, a1 = a0 + du;
        const Real lo = std::min(a0, a1), hi = std::max(a0, a1);
        const long long first = static_cast<long long>(std::floor(lo)) + 1;
        const long long last = static_cast<long long>(std::ceil(hi)) - 1;
        for (long long m = first; m <= last; ++m) {
            const Real s = (static_cast<Real>(m) - a0) / du;
            if (s > tol && s < 1.0 - tol) b.push_back(s);
        }
    }
    std::sort(b.begin(), b.end());
    std::vector<Real> out;
    out.reserve(b.size());
    for (Real x : b)
        if (out.empty() || std::abs(x - out.back()) > 1.0e-12) out.push_back(x);
    if (out.front() != 0.0) out.insert(out.begin(), 0.0);
    if (out.back() != 1.0) out.push_back(1.0);
    return out;
}
static void update_path_stats(PathStats& st, const std::vector<Real>& br) {
    const int nseg = std::max<int>(1, static_cast<int>(br.size()) - 1);
    ++st.n_particles;
    st.total_segments += static_cast<std::size_t>(nseg);
    st.max_segments = std::max(st.max_segments, nseg);
    if (nseg > 1) ++st.crossing_particles;
}
static PathStats compute_path_stats(const std::vector<Vec3>& x0, const std::vector<Vec3>& dxp, const Grid& g,
                                    int order, bool split) {
    PathStats st;
    for (std::size_t p = 0; p < x0.size(); ++p) update_path_stats(st, path_breakpoints_for_particle(x0[p], dxp[p], g, order, split));
    return st;
}

static Vec3 path_average_gather_vector_one(const Vec3& x0, const Vec3& dxp, const VecField& mesh, const Grid& g,
                                           int order, const Quadrature& quad, bool split) {
    Vec3 out{0.0, 0.0, 0.0};
    const auto br = path_breakpoints_for_particle(x0, dxp, g, order, split);
    for (std::size_t seg = 0; seg + 1 < br.size(); ++seg) {
        const Real a = br[seg], len = br[seg + 1] - br[seg];
        if (len <= 0.0) continue;
        for (std::size_t iq = 0; iq < quad.s.size(); ++iq) {
            const Real s = a + len * quad.s[iq];
            const Real w = len * quad.w[iq];
            const Vec3 val = gather_vector_at(mesh, path_position(x0, dxp, s, g), g, order);
            for (int d = 0; d < 3; ++d) out[d] += w * val[d];
        }
    }
    return out;
}
static std::vector<Vec3> path_average_gather_vector(const std::vector<Vec3>& x0, const std::vector<Vec3>& dxp,
                                                    const VecField& mesh, const Grid& g, int order,
                                                    const Quadrature& quad, bool split) {
    std::vector<Vec3> out(x0.size());
    for (std::size_t p = 0; p < x0.size(); ++p) out[p] = path_average_gather_vector_one(x0[p], dxp[p], mesh, g, order, quad, split);
    return out;
}
static VecField path_average_deposit_current(const std::vector<Vec3>& x0, const std::vector<Vec3>& dxp,
                                             const std::vector<Vec3>& vbar, const std::vector<Real>& q,
                                             const Grid& g, int order, const Quadrature& quad, bool split,
                                             PathStats* stats) {
    VecField J = zero_vecfield(g.N);
    if (stats) *stats = PathStats{};
    for (std::size_t p = 0; p < x0.size(); ++p) {
        const auto br = path_breakpoints_for_particle(x0[p], dxp[p], g, order, split);
        if (stats) update_path_stats(*stats, br);
        for (std::size_t seg = 0; seg + 1 < br.size(); ++seg) {
            const Real a = br[seg], len = br[seg + 1] - br[seg];
            if (len <= 0.0) continue;
            for (std::size_t iq = 0; iq < quad.s.size(); ++iq) {
                const Real s = a + len * quad.s[iq];
                const Real w = len * quad.w[iq];
                deposit_current_add(J, path_position(x0[p], dxp[p], s, g), vbar[p], q[p], g, order, w);
            }
        }
    }
    if (stats) stats->max_segments = std::max(1, stats->max_segments);
    project_nyquist(J, g);
    return J;
}
// orbit_discrete_gradient_A (P1:908-934) and _interval (P3:2958-2984):
// theta(s) = theta0 + (theta1 - theta0) s.
static Mat3 orbit_discrete_gradient_A_interval(const Vec3& x0, const Vec3& dxp, const VecField& A_n,
                                               const VecField& A_np1, Real theta0, Real theta1, const Grid& g,
                                               int order, const Quadrature& quad, bool split) {
    Mat3 D = zero_mat3();
    const auto br = path_breakpoints_for_particle(x0, dxp, g, order, split);
    for (std::size_t seg = 0; seg + 1 < br.size(); ++seg) {
        const Real a = br[seg], len = br[seg + 1] - br[seg];
        if (len <= 0.0) continue;
        for (std::size_t iq = 0; iq < quad.s.size(); ++iq) {
            const Real s = a + len * quad.s[iq];
            const Real w = len * quad.w[iq];
            // P1 interpolates with s itself; theta0=0, theta1=1 reproduces it
            // bit for bit since 0 + 1*s == s.
            const Real theta = (theta0 == 0.0 && theta1 == 1.0) ? s : theta0 + (theta1 - theta0) * s;
            const Mat3 G = gather_shape_gradient_vector_interp_at(A_n, A_np1, theta, path_position(x0, dxp, s, g), g, order);
            for (int a0 = 0; a0 < 3; ++a0)
                for (int b0 = 0; b0 < 3; ++b0) D[a0][b0] += w * G[a0][b0];
        }
    }
    return D;
}

// ---------------------------------------------------------------------------
// Field solve (P1:940-1001, P3:3077-3296)
// ---------------------------------------------------------------------------
struct WaveSpectra { CField phi, psi; std::array<CField, 3> A, U; };

static std::pair<Field, Field> cn_wave_from_spectra(const CField& uhat, const CField& vhat, const Field& S_mid,
                                                    Real dt, const Grid& g, Real c, Real source_scale) {
    const CField Shat = fft_real(S_mid, g);
    CField unew(g.N), vnew(g.N);
    const Real c2 = c * c;
    const Real source_coeff = c2 * source_scale;
    for (int id = 0; id < g.N; ++id) {
        if (excluded_nyquist_mode(id, g)) { unew[id] = vnew[id] = Complex(0.0, 0.0); continue; }
        const Real theta = 0.25 * c2 * dt * dt * g.k2[id];
        const Real denom = 1.0 + theta;
        unew[id] = ((1.0 - theta) * uhat[id] + dt * vhat[id] + 0.5 * source_coeff * dt * dt * Shat[id]) / denom;
        vnew[id] = ((1.0 - theta) * vhat[id] - c2 * dt * g.k2[id] * uhat[id] + source_coeff * dt * Shat[id]) / denom;
    }
    return {ifft_real(std::move(unew), g), ifft_real(std::move(vnew), g)};
}
static Field solve_periodic_poisson_for_phi(const Field& rho, const Grid& g, Real source /*1/eps0 or sigma1*/,
                                            bool p1_form, Real eps0) {
    CField rhohat = fft_real(rho, g);
    CField phihat(g.N, Complex(0.0, 0.0));
    for (int id = 0; id < g.N; ++id) {
        if (p1_form) {
            if (g.k2[id] > 0.0) phihat[id] = rhohat[id] / (eps0 * g.k2[id]);       // P1:963-965
        } else {
            if (!excluded_nyquist_mode(id, g) && g.k2[id] > 0.0) phihat[id] = source * rhohat[id] / g.k2[id];  // P3:3201
        }
    }
    Field phi = ifft_real(std::move(phihat), g);
    subtract_mean(phi);
    return phi;
}

struct State {
    std::vector<Vec3> x, v, v_prev, vbar_last, P;
    std::vector<Real> q, m;
    Field phi, psi, rho;
    VecField A, U, J;
};

struct Coef { Real c, phi_src, A_src; };  // P1: c, 1/eps0, mu0; P3: kappa, sigma1, sigma2
static Coef coefs(const upic_params& p) {
    if (p.code == 3) return {p.c, p.sigma1, p.sigma2};
    return {p.c, 1.0 / p.eps0, p.mu0};
}

static void cn_potential_update(const WaveSpectra& sp, const Field& rho_mid, const VecField& J_mid, Real dt,
                                const Grid& g, const Coef& cf, Field& phi_new, Field& psi_new, VecField& A_new,
                                VecField& U_new) {
    auto pp = cn_wave_from_spectra(sp.phi, sp.psi, rho_mid, dt, g, cf.c, cf.phi_src);
    phi_new = std::move(pp.first);
    psi_new = std::move(pp.second);
    subtract_mean(phi_new);
    subtract_mean(psi_new);
    A_new = zero_vecfield(g.N);
    U_new = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) {
        auto au = cn_wave_from_spectra(sp.A[d], sp.U[d], J_mid[d], dt, g, cf.c, cf.A_src);
        A_new[d] = std::move(au.first);
        U_new[d] = std::move(au.second);
    }
}
static WaveSpectra make_spectra(const Field& phi, const Field& psi, const VecField& A, const VecField& U, const Grid& g) {
    WaveSpectra s;
    s.phi = fft_real(phi, g); s.psi = fft_real(psi, g);
    for (int d = 0; d < 3; ++d) { s.A[d] = fft_real(A[d], g); s.U[d] = fft_real(U[d], g); }
    return s;
}
static Field continuity_charge_update(const Field& rho_n, const VecField& J_mid, Real dt, const Grid& g) {
    Field divJ = spectral_divergence_vector(J_mid, g);
    Field rho(g.N);
    for (int id = 0; id < g.N; ++id) rho[id] = rho_n[id] - dt * divJ[id];
    subtract_mean(rho);
    project_nyquist(rho, g);  // no-op unless P3 projection is on (P3:3294)
    return rho;
}

// ---------------------------------------------------------------------------
// Relativistic kinematics and HC rotation (P3:584-689, 4585-4655)
// ---------------------------------------------------------------------------
static Real gamma_from_p(const Vec3& p, Real m, Real kappa) {
    const Real mk = m * kappa;
    if (!(mk > 0.0)) throw Failure(UPIC_EINVAL, "gamma_from_p requires positive m and kappa");
    return std::sqrt(1.0 + dot3(p, p) / (mk * mk));
}
static Vec3 velocity_from_p(const Vec3& p, Real m, Real kappa) { return mul3(1.0 / (m * gamma_from_p(p, m, kappa)), p); }
static Real gamma_from_v(const Vec3& v, Real kappa) {
    const Real beta2 = dot3(v, v) / (kappa * kappa);
    if (beta2 >= 1.0) throw Failure(UPIC_EINVAL, "relativistic initial velocity must satisfy |v| < kappa");
    return 1.0 / std::sqrt(std::max<Real>(1.0e-300, 1.0 - beta2));
}
static Vec3 momentum_from_v(const Vec3& v, Real m, Real kappa) { return mul3(m * gamma_from_v(v, kappa), v); }
static Vec3 kinetic_secant_velocity(const Vec3& a, const Vec3& b, Real m, Real kappa) {
    const Real ga = gamma_from_p(a, m, kappa), gb = gamma_from_p(b, m, kappa);
    return mul3(1.0 / (m * (ga + gb)), add3(a, b));
}
static Mat3 skew_part_from_D(const Mat3& D) {
    Mat3 K = zero_mat3();
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) K[a][b] = D[b][a] - D[a][b];
    return K;
}
static Vec3 beff_from_skew_K(const Mat3& K) { return {K[1][2], -K[0][2], K[0][1]}; }
static Vec3 higuera_cary_rotate_with_K(const Vec3& p_minus, const Mat3& K, Real q, Real m, Real kappa, Real dt) {
    const Vec3 B = beff_from_skew_K(K);
    Vec3 u = mul3(1.0 / (m * kappa), p_minus);
    const Real gamma_minus = std::sqrt(1.0 + dot3(u, u));
    const Vec3 tau = mul3(q * dt / (2.0 * m), B);
    const Real tau2 = dot3(tau, tau);
    const Real udottau = dot3(u, tau);
    const Real sigma = gamma_minus * gamma_minus - tau2;
    const Real gamma_hc2 = 0.5 * (sigma + std::sqrt(sigma * sigma + 4.0 * (tau2 + udottau * udottau)));
    const Real gamma_hc = std::sqrt(std::max<Real>(gamma_hc2, 1.0));
    const Vec3 t = mul3(1.0 / gamma_hc, tau);
    const Vec3 svec = mul3(2.0 / (1.0 + dot3(t, t)), t);
    const Vec3 uprime = add3(u, cross3(u, t));
    const Vec3 uplus = add3(u, cross3(uprime, svec));
    return mul3(m * kappa, uplus);
}

// ---------------------------------------------------------------------------
// P1 Picard map and step (P1:1441-1552)
// ---------------------------------------------------------------------------
struct P1Trial {
    std::vector<Vec3> vbar, dx_path, x_end, P_end, v_next;
    VecField Jbar;
    Field rho_end, phi_end, psi_end;
    VecField A_end, U_end;
    PathStats stats;
};

struct Ctx {
    upic_params p;
    Grid g;
    Quadrature quad;
    Coef cf;
    explicit Ctx(const upic_params& pp) : p(pp), g(make_grid(pp)), quad(gauss_legendre_01(pp.orbit_quad_order)), cf(coefs(pp)) {}
    int order() const { return p.spline_order; }
    bool split() const { return p.split_orbit_at_knots != 0; }
};

static P1Trial p1_map(const Ctx& C, const std::vector<Vec3>& v_guess, const State& st, const VecField& grad_phi_n,
                      const WaveSpectra& sp) {
    const Grid& g = C.g;
    const std::size_t np = st.x.size();
    const Real dt = C.p.dt;
    P1Trial tr;
    tr.vbar.assign(np, {0, 0, 0}); tr.dx_path.assign(np, {0, 0, 0}); tr.x_end.assign(np, {0, 0, 0});
    tr.P_end.assign(np, {0, 0, 0}); tr.v_next.assign(np, {0, 0, 0});
    for (std::size_t p = 0; p < np; ++p) {
        tr.vbar[p] = mul3(0.5, add3(v_guess[p], st.v[p]));
        tr.dx_path[p] = mul3(dt, tr.vbar[p]);
        tr.x_end[p] = wrap_pos(add3(st.x[p], tr.dx_path[p]), g);
    }
    tr.Jbar = path_average_deposit_current(st.x, tr.dx_path, tr.vbar, st.q, g, C.order(), C.quad, C.split(), &tr.stats);
    tr.rho_end = continuity_charge_update(st.rho, tr.Jbar, dt, g);
    Field rho_mid(g.N);
    for (int id = 0; id < g.N; ++id) rho_mid[id] = 0.5 * (tr.rho_end[id] + st.rho[id]);
    cn_potential_update(sp, rho_mid, tr.Jbar, dt, g, C.cf, tr.phi_end, tr.psi_end, tr.A_end, tr.U_end);
    VecField grad_phi_end = spectral_gradient_scalar(tr.phi_end, g);
    VecField grad_phi_mid = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d)
        for (int id = 0; id < g.N; ++id) grad_phi_mid[d][id] = 0.5 * (grad_phi_n[d][id] + grad_phi_end[d][id]);
    const auto gpb = path_average_gather_vector(st.x, tr.dx_path, grad_phi_mid, g, C.order(), C.quad, C.split());
    for (std::size_t p = 0; p < np; ++p) {
        const Mat3 DA = orbit_discrete_gradient_A_interval(st.x[p], tr.dx_path[p], st.A, tr.A_end, 0.0, 1.0, g,
                                                           C.order(), C.quad, C.split());
        const Vec3 A_end_p = gather_vector_at(tr.A_end, tr.x_end[p], g, C.order());
        Vec3 DATv{0.0, 0.0, 0.0};
        for (int b = 0; b < 3; ++b) DATv[b] = DA[0][b] * tr.vbar[p][0] + DA[1][b] * tr.vbar[p][1] + DA[2][b] * tr.vbar[p][2];
        Vec3 forceP{0.0, 0.0, 0.0};
        for (int d = 0; d < 3; ++d) forceP[d] = (-st.q[p]) * gpb[p][d] + st.q[p] * DATv[d];
        tr.P_end[p] = add3(st.P[p], mul3(dt, forceP));
        for (int d = 0; d < 3; ++d) tr.v_next[p][d] = (tr.P_end[p][d] - st.q[p] * A_end_p[d]) / st.m[p];
    }
    return tr;
}

static State p1_step(const Ctx& C, const State& st, upic_step_info* info) {
    const Grid& g = C.g;
    const std::size_t np = st.x.size();
    VecField grad_phi_n = spectral_gradient_scalar(st.phi, g);
    const WaveSpectra sp = make_spectra(st.phi, st.psi, st.A, st.U, g);
    std::vector<Vec3> v_guess(np);
    for (std::size_t p = 0; p < np; ++p) v_guess[p] = sub3(mul3(2.0, st.v[p]), st.v_prev[p]);
    VecField J_prev = zero_vecfield(g.N);
    bool have_J_prev = false;
    P1Trial tr;
    int iters = 0;
    Real cur_err = std::numeric_limits<Real>::infinity(), vel_err = cur_err;
    for (int it = 0; it < C.p.picard_max_iter; ++it) {
        tr = p1_map(C, v_guess, st, grad_phi_n, sp);
        ++iters;
        vel_err = relative_rms_error_vec_particles(tr.v_next, v_guess);
        cur_err = have_J_prev ? relative_rms_error_vecfield(tr.Jbar, J_prev) : std::numeric_limits<Real>::infinity();
        if (it > 0 && std::max(cur_err, vel_err) < C.p.picard_tol) break;
        J_prev = tr.Jbar;
        have_J_prev = true;
        v_guess = tr.v_next;
    }
    if (C.p.final_consistency_sweep) tr = p1_map(C, tr.v_next, st, grad_phi_n, sp);
    State out;
    out.x = tr.x_end; out.v = tr.v_next; out.v_prev = st.v; out.P = tr.P_end;
    out.q = st.q; out.m = st.m;
    out.phi = tr.phi_end; out.psi = tr.psi_end; out.A = tr.A_end; out.U = tr.U_end;
    out.rho = tr.rho_end; out.J = tr.Jbar;
    out.vbar_last = tr.vbar;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = iters;
        info->converged = (std::max(cur_err, vel_err) < C.p.picard_tol) ? 1 : 0;
        info->current_error = cur_err;
        info->velocity_error = vel_err;
        info->path.n_particles = static_cast<int64_t>(tr.stats.n_particles);
        info->path.crossing_particles = static_cast<int64_t>(tr.stats.crossing_particles);
        info->path.total_segments = static_cast<int64_t>(tr.stats.total_segments);
        info->path.max_segments = tr.stats.max_segments;
    }
    return out;
}

// ---------------------------------------------------------------------------
// P3 nested solve (P3:4843-4983, 5235-5431, 5576-5850, 5878-6215)
// ---------------------------------------------------------------------------
struct LocalStep { Vec3 vbar_map{0, 0, 0}, P_end{0, 0, 0}, p_end{0, 0, 0}, v_end{0, 0, 0}; };
struct LocalReport {
    int iterations = 0;
    int reason = UPIC_FAIL_MAXIMUM_ITERATIONS;
    Real first_scaled = std::numeric_limits<Real>::infinity();
    Real final_scaled = std::numeric_limits<Real>::infinity();
    Real final_abs = std::numeric_limits<Real>::infinity();
};

static Vec3 gather_time_interp(const VecField& A_n, const VecField& A_np1, Real theta, const Vec3& pos, const Grid& g,
                               int order) {
    const Vec3 a0 = gather_vector_at(A_n, pos, g, order);
    const Vec3 a1 = gather_vector_at(A_np1, pos, g, order);
    return add3(mul3(1.0 - theta, a0), mul3(theta, a1));
}

static LocalStep p3_local_map(const Ctx& C, const Vec3& g_vbar, const Vec3& x0, const Vec3& P0, Real q, Real m, Real h,
                              Real theta0, Real theta1, const VecField& A_n, const VecField& A_end,
                              const VecField& grad_phi_mid) {
    const Grid& g = C.g;
    const Real kappa = C.p.c;
    LocalStep out;
    const Vec3 dxp = mul3(h, g_vbar);
    const Vec3 x1 = wrap_pos(add3(x0, dxp), g);
    const Vec3 A0p = gather_time_interp(A_n, A_end, theta0, x0, g, C.order());
    const Vec3 A1p = gather_time_interp(A_n, A_end, theta1, x1, g, C.order());
    const Vec3 Aep = mul3(0.5, add3(A0p, A1p));
    const Vec3 p_start = sub3(P0, mul3(q, A0p));
    const Vec3 gpb = path_average_gather_vector_one(x0, dxp, grad_phi_mid, g, C.order(), C.quad, C.split());
    const Mat3 DA = orbit_discrete_gradient_A_interval(x0, dxp, A_n, A_end, theta0, theta1, g, C.order(), C.quad, C.split());
    const Vec3 Dv = matvec3(DA, g_vbar);
    const Mat3 K = skew_part_from_D(DA);
    Vec3 p_minus = P0;
    p_minus = sub3(p_minus, mul3(q, Aep));
    p_minus = sub3(p_minus, mul3(0.5 * q * h, gpb));
    p_minus = add3(p_minus, mul3(0.5 * q * h, Dv));
    const Vec3 p_plus = higuera_cary_rotate_with_K(p_minus, K, q, m, kappa, h);
    out.P_end = p_plus;
    out.P_end = add3(out.P_end, mul3(q, Aep));
    out.P_end = sub3(out.P_end, mul3(0.5 * q * h, gpb));
    out.P_end = add3(out.P_end, mul3(0.5 * q * h, Dv));
    out.p_end = sub3(out.P_end, mul3(q, A1p));
    out.v_end = velocity_from_p(out.p_end, m, kappa);
    const Vec3 vEm = kinetic_secant_velocity(p_minus, p_start, m, kappa);
    const Vec3 vEp = kinetic_secant_velocity(out.p_end, p_plus, m, kappa);
    out.vbar_map = mul3(0.5, add3(vEm, vEp));
    return out;
}

static bool p3_local_solve(const Ctx& C, const Vec3& initial_guess, const Vec3& x0, const Vec3& P0, Real q, Real m,
                           Real h, Real theta0, Real theta1, const VecField& A_n, const VecField& A_end,
                           const VecField& grad_phi_mid, Vec3& accepted_vbar, LocalStep& accepted, LocalReport& rep) {
    Vec3 guess = initial_guess;
    rep = LocalReport{};
    const Real kappa = C.p.c;
    for (int it = 0; it < C.p.local_max_iter; ++it) {
        const LocalStep mapped = p3_local_map(C, guess, x0, P0, q, m, h, theta0, theta1, A_n, A_end, grad_phi_mid);
        ++rep.iterations;
        if (!finite_vec3(mapped.vbar_map)) { rep.reason = UPIC_FAIL_NONFINITE_MAP; return false; }
        if (norm3(mapped.vbar_map) >= kappa) { rep.reason = UPIC_FAIL_SUPERLUMINAL_MAP; return false; }
        const Vec3 residual = sub3(mapped.vbar_map, guess);
        const Real scale = 0.5 * (norm3(mapped.vbar_map) + norm3(guess));
        rep.final_abs = norm3(residual);
        rep.final_scaled = rep.final_abs / (C.p.local_atol + C.p.local_rtol * scale);
        if (rep.iterations == 1) rep.first_scaled = rep.final_scaled;
        if (rep.final_scaled <= 1.0) {
            accepted_vbar = guess;
            accepted = mapped;
            rep.reason = UPIC_FAIL_NONE;
            return true;
        }
        guess = add3(guess, mul3(C.p.local_damping, residual));
        if (!finite_vec3(guess)) { rep.reason = UPIC_FAIL_NONFINITE_ITERATE; return false; }
        if (norm3(guess) >= kappa) { rep.reason = UPIC_FAIL_SUPERLUMINAL_ITERATE; return false; }
    }
    rep.reason = UPIC_FAIL_MAXIMUM_ITERATIONS;
    return false;
}

struct P3Trial {
    std::vector<Vec3> current_map;  // node vectors
    std::vector<Vec3> sub_x0, sub_vbar;
    std::vector<Vec3> x_end, P_end, v_next, vbar_macro;
    VecField J_map;
    Field rho_end, phi_end, psi_end;
    VecField A_end, U_end;
    PathStats stats;
    Real local_scaled_max = 0.0;
    int local_iter_max = 0;
    Real local_iter_mean = 0.0;
    long long local_iter_total = 0;
    std::vector<std::size_t> failed;
    std::vector<int> reasons;
};

static VecField node_to_vecfield(const std::vector<Vec3>& v, const Grid& g) {
    VecField out = zero_vecfield(g.N);
    for (int id = 0; id < g.N; ++id) for (int d = 0; d < 3; ++d) out[d][id] = v[id][d];
    return out;
}
static std::vector<Vec3> vecfield_to_node(const VecField& f, const Grid& g) {
    std::vector<Vec3> out(g.N);
    for (int id = 0; id < g.N; ++id) for (int d = 0; d < 3; ++d) out[id][d] = f[d][id];
    return out;
}
static bool finite_nodes(const std::vector<Vec3>& v) { for (auto& x : v) if (!finite_vec3(x)) return false; return true; }

// evaluate_selective_current_map_at_layout (P3:5235-5431)
static P3Trial p3_current_map(const Ctx& C, const std::vector<Vec3>& guess, const State& st, const VecField& grad_phi_n,
                              const WaveSpectra& sp, const std::vector<int>& layout) {
    const Grid& g = C.g;
    const std::size_t np = st.x.size();
    const Real dt = C.p.dt;
    P3Trial tr;
    std::vector<std::size_t> offset(np + 1, 0);
    for (std::size_t p = 0; p < np; ++p) offset[p + 1] = offset[p] + static_cast<std::size_t>(layout[p]);
    const std::size_t total = offset[np];
    VecField J_field = node_to_vecfield(guess, g);
    project_nyquist(J_field, g);
    tr.rho_end = continuity_charge_update(st.rho, J_field, dt, g);
    Field rho_mid(g.N);
    for (int id = 0; id < g.N; ++id) rho_mid[id] = 0.5 * (st.rho[id] + tr.rho_end[id]);
    cn_potential_update(sp, rho_mid, J_field, dt, g, C.cf, tr.phi_end, tr.psi_end, tr.A_end, tr.U_end);
    const VecField gpe = spectral_gradient_scalar(tr.phi_end, g);
    VecField gpm = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) for (int id = 0; id < g.N; ++id) gpm[d][id] = 0.5 * (grad_phi_n[d][id] + gpe[d][id]);

    tr.sub_x0.assign(total, {0, 0, 0}); tr.sub_vbar.assign(total, {0, 0, 0});
    std::vector<Vec3> sub_dx(total, {0, 0, 0});
    std::vector<Real> qw(total, 0.0);
    tr.x_end.assign(np, {0, 0, 0}); tr.P_end.assign(np, {0, 0, 0}); tr.v_next.assign(np, {0, 0, 0});
    tr.vbar_macro.assign(np, {0, 0, 0});
    tr.reasons.assign(np, UPIC_FAIL_NONE);
    long long iter_total = 0;
    for (std::size_t p = 0; p < np; ++p) {
        const std::size_t nsub = static_cast<std::size_t>(layout[p]);
        const Real h = dt / static_cast<Real>(nsub);
        Vec3 xpos = st.x[p], Pcur = st.P[p], vcur = st.v[p], vmacro{0, 0, 0};
        bool ok = true;
        for (std::size_t s = 0; s < nsub; ++s) {
            const std::size_t j = offset[p] + s;
            const Real th0 = static_cast<Real>(s) / static_cast<Real>(nsub);
            const Real th1 = static_cast<Real>(s + 1) / static_cast<Real>(nsub);
            Vec3 vbar{0, 0, 0};
            LocalStep mapped;
            LocalReport rep;
            const bool conv = p3_local_solve(C, vcur, xpos, Pcur, st.q[p], st.m[p], h, th0, th1, st.A, tr.A_end, gpm,
                                             vbar, mapped, rep);
            iter_total += rep.iterations;
            tr.local_iter_max = std::max(tr.local_iter_max, rep.iterations);
            tr.local_scaled_max = std::max(tr.local_scaled_max, rep.final_scaled);
            if (!conv) { ok = false; tr.reasons[p] = rep.reason; break; }
            tr.sub_x0[j] = xpos;
            tr.sub_vbar[j] = vbar;
            sub_dx[j] = mul3(h, vbar);
            qw[j] = st.q[p] / static_cast<Real>(nsub);
            xpos = wrap_pos(add3(xpos, sub_dx[j]), g);
            Pcur = mapped.P_end;
            vcur = mapped.v_end;
            vmacro = add3(vmacro, mul3(1.0 / static_cast<Real>(nsub), vbar));
        }
        if (!ok) { tr.failed.push_back(p); continue; }
        tr.x_end[p] = xpos; tr.P_end[p] = Pcur; tr.v_next[p] = vcur; tr.vbar_macro[p] = vmacro;
    }
    tr.local_iter_total = iter_total;
    tr.local_iter_mean = total ? static_cast<Real>(iter_total) / static_cast<Real>(total) : 0.0;
    if (!tr.failed.empty()) return tr;
    tr.J_map = path_average_deposit_current(tr.sub_x0, sub_dx, tr.sub_vbar, qw, g, C.order(), C.quad, C.split(), &tr.stats);
    tr.current_map = vecfield_to_node(tr.J_map, g);
    return tr;
}

struct Norms { Real rms_relative, max_particle_relative, rms_absolute, rms_scaled, max_particle_scaled; };
static Norms residual_norms(const std::vector<Vec3>& map, const std::vector<Vec3>& guess, Real kappa, Real rtol, Real atol) {
    std::vector<Vec3> r(guess.size());
    Norms n{};
    Real mpr = 0.0, mps = 0.0;
    const Real floor_ = 64.0 * std::numeric_limits<Real>::epsilon() * kappa;
    for (std::size_t p = 0; p < guess.size(); ++p) {
        r[p] = sub3(map[p], guess[p]);
        const Real rn = norm3(r[p]);
        const Real vs = 0.5 * (norm3(map[p]) + norm3(guess[p]));
        mpr = std::max(mpr, rn / std::max(floor_, vs));
        mps = std::max(mps, rn / (atol + rtol * vs));
    }
    n.max_particle_relative = mpr;
    n.max_particle_scaled = mps;
    n.rms_absolute = rms_vec_particles(r);
    const Real vrms = 0.5 * (rms_vec_particles(map) + rms_vec_particles(guess));
    n.rms_relative = n.rms_absolute / (vrms + floor_);
    n.rms_scaled = n.rms_absolute / (atol + rtol * vrms);
    return n;
}
static long double particle_vector_dot(const std::vector<Vec3>& a, const std::vector<Vec3>& b) {
    long double s = 0.0L;
    for (std::size_t p = 0; p < a.size(); ++p)
        for (int d = 0; d < 3; ++d) s += static_cast<long double>(a[p][d]) * static_cast<long double>(b[p][d]);
    return s;
}
static bool solve_small_dense(std::vector<std::vector<Real>> A, std::vector<Real> b, std::vector<Real>& x) {
    const int n = static_cast<int>(b.size());
    Real scale = 0.0;
    for (const auto& row : A) for (Real v : row) scale = std::max(scale, std::abs(v));
    if (!(scale > 0.0) || !std::isfinite(scale)) return false;
    Real largest = 0.0, smallest = std::numeric_limits<Real>::infinity();
    for (int col = 0; col < n; ++col) {
        int pr = col;
        for (int row = col + 1; row < n; ++row) if (std::abs(A[row][col]) > std::abs(A[pr][col])) pr = row;
        const Real piv = std::abs(A[pr][col]);
        if (!std::isfinite(piv) || piv <= 100.0 * std::numeric_limits<Real>::epsilon() * scale) return false;
        largest = std::max(largest, piv);
        smallest = std::min(smallest, piv);
        if (pr != col) { std::swap(A[pr], A[col]); std::swap(b[pr], b[col]); }
        for (int row = col + 1; row < n; ++row) {
            const Real f = A[row][col] / A[col][col];
            for (int j = col; j < n; ++j) A[row][j] -= f * A[col][j];
            b[row] -= f * b[col];
        }
    }
    if (smallest <= 1.0e-12 * largest) return false;
    x.assign(n, 0.0);
    for (int row = n - 1; row >= 0; --row) {
        Real rhs = b[row];
        for (int j = row + 1; j < n; ++j) rhs -= A[row][j] * x[j];
        x[row] = rhs / A[row][row];
        if (!std::isfinite(x[row])) return false;
    }
    return true;
}
static bool anderson_candidate(const std::vector<std::vector<Vec3>>& xh, const std::vector<std::vector<Vec3>>& rh,
                               int memory, Real damping, Real reg, std::vector<Vec3>& cand) {
    const int available = static_cast<int>(xh.size()) - 1;
    const int m = std::min(memory, available);
    if (m <= 0) return false;
    const int first = static_cast<int>(xh.size()) - m - 1;
    const std::size_t n = xh.back().size();
    std::vector<std::vector<Vec3>> dX(m, std::vector<Vec3>(n)), dR(m, std::vector<Vec3>(n));
    for (int j = 0; j < m; ++j)
        for (std::size_t p = 0; p < n; ++p) {
            dX[j][p] = sub3(xh[first + j + 1][p], xh[first + j][p]);
            dR[j][p] = sub3(rh[first + j + 1][p], rh[first + j][p]);
        }
    std::vector<std::vector<Real>> gram(m, std::vector<Real>(m, 0.0));
    std::vector<Real> rhs(m, 0.0);
    Real trace = 0.0;
    for (int i = 0; i < m; ++i) {
        rhs[i] = static_cast<Real>(particle_vector_dot(dR[i], rh.back()));
        for (int j = 0; j <= i; ++j) {
            const Real v = static_cast<Real>(particle_vector_dot(dR[i], dR[j]));
            gram[i][j] = v; gram[j][i] = v;
        }
        trace += gram[i][i];
    }
    if (!(trace > 0.0) || !std::isfinite(trace)) return false;
    std::vector<Real> unreg;
    if (!solve_small_dense(gram, rhs, unreg)) return false;
    const Real rs = trace / static_cast<Real>(m);
    for (int i = 0; i < m; ++i) gram[i][i] += reg * rs;
    std::vector<Real> gamma;
    if (!solve_small_dense(gram, rhs, gamma)) return false;
    cand = xh.back();
    for (std::size_t p = 0; p < n; ++p) {
        cand[p] = add3(cand[p], mul3(damping, rh.back()[p]));
        for (int j = 0; j < m; ++j) cand[p] = sub3(cand[p], mul3(gamma[j], add3(dX[j][p], mul3(damping, dR[j][p]))));
    }
    return true;
}

struct RefineNeeded : Failure {
    std::vector<std::size_t> failed;
    RefineNeeded(const std::string& m, std::vector<std::size_t> f) : Failure(UPIC_EREFINE, m), failed(std::move(f)) {}
};

// nonlinear_step_uniform_subcycled with a selective layout (P3:5878-6215)
static State p3_step_layout(const Ctx& C, const State& st, const std::vector<int>& layout, upic_step_info* info) {
    const Grid& g = C.g;
    const upic_params& P = C.p;
    const VecField grad_phi_n = spectral_gradient_scalar(st.phi, g);
    const WaveSpectra sp = make_spectra(st.phi, st.psi, st.A, st.U, g);
    std::vector<Vec3> guess = vecfield_to_node(st.J, g);
    if (!finite_nodes(guess)) throw Failure(UPIC_ENONLINEAR, "nonfinite initial macro-current predictor");
    const bool picard = P.nonlinear_solver_picard != 0;
    std::vector<std::vector<Vec3>> xh, rh;
    Real prev_rms = std::numeric_limits<Real>::infinity();
    bool converged = false, have_prefetched = false;
    int anderson_restarts = 0, anderson_steps = 0, picard_steps = 0, backtracked = 0, maps = 0, iters = 0;
    Real min_alpha = 1.0;
    long long local_total = 0;
    P3Trial tr;
    auto eval = [&](const std::vector<Vec3>& gs) {
        ++maps;
        P3Trial t = p3_current_map(C, gs, st, grad_phi_n, sp, layout);
        local_total += t.local_iter_total;
        if (!t.failed.empty()) throw RefineNeeded("selective local particle solve failed", t.failed);
        return t;
    };
    for (int it = 0; it < P.nonlinear_max_iter; ++it) {
        if (have_prefetched) have_prefetched = false; else tr = eval(guess);
        if (!finite_nodes(tr.current_map)) throw Failure(UPIC_ENONLINEAR, "nonfinite macro current");
        std::vector<Vec3> residual(guess.size());
        for (std::size_t i = 0; i < guess.size(); ++i) residual[i] = sub3(tr.current_map[i], guess[i]);
        const Norms nm = residual_norms(tr.current_map, guess, P.c, P.nonlinear_rtol, P.nonlinear_atol);
        iters = it + 1;
        if (nm.rms_scaled <= 1.0 && nm.max_particle_scaled <= 1.0) { converged = true; break; }
        if (!picard && std::isfinite(prev_rms) && nm.rms_relative > P.anderson_growth_limit * prev_rms) {
            xh.clear(); rh.clear(); ++anderson_restarts;
        }
        xh.push_back(guess); rh.push_back(residual);
        const std::size_t keep = static_cast<std::size_t>(P.anderson_memory + 1);
        if (xh.size() > keep) { xh.erase(xh.begin()); rh.erase(rh.begin()); }
        std::vector<Vec3> cand;
        bool used_anderson = false;
        if (!picard && it >= P.anderson_start && xh.size() >= 2) {
            used_anderson = anderson_candidate(xh, rh, P.anderson_memory, P.anderson_damping, P.anderson_regularization, cand);
            if (!used_anderson || !finite_nodes(cand)) {
                xh.assign(1, guess); rh.assign(1, residual); ++anderson_restarts; used_anderson = false;
            }
        }
        auto backtrack = [&](const std::vector<Vec3>& full, std::vector<Vec3>& acc, P3Trial& acc_tr, Real& acc_alpha) {
            std::vector<Vec3> tg(guess.size());
            for (int bt = 0; bt <= 6; ++bt) {
                const Real alpha = std::ldexp(1.0, -bt);
                for (std::size_t i = 0; i < guess.size(); ++i) tg[i] = add3(guess[i], mul3(alpha, sub3(full[i], guess[i])));
                if (!finite_nodes(tg)) continue;
                P3Trial tm = eval(tg);  // RefineNeeded propagates (P3:6064-6065)
                const Norms tn = residual_norms(tm.current_map, tg, P.c, P.nonlinear_rtol, P.nonlinear_atol);
                const bool dec = tn.rms_scaled <= (1.0 - 1.0e-4 * alpha) * nm.rms_scaled;
                const bool safe = tn.max_particle_scaled <= P.anderson_growth_limit * nm.max_particle_scaled;
                if (dec && safe) { acc = tg; acc_tr = std::move(tm); acc_alpha = alpha; return true; }
            }
            return false;
        };
        bool accepted = false, as_anderson = false;
        Real alpha = 0.0;
        P3Trial acc_tr;
        if (used_anderson) {
            accepted = backtrack(cand, cand, acc_tr, alpha);
            as_anderson = accepted;
            if (!accepted) { xh.assign(1, guess); rh.assign(1, residual); ++anderson_restarts; }
        }
        if (!accepted) {
            const Real damping = picard ? 1.0 : P.anderson_damping;
            std::vector<Vec3> pc(guess.size());
            for (std::size_t i = 0; i < guess.size(); ++i) pc[i] = add3(guess[i], mul3(damping, residual[i]));
            accepted = backtrack(pc, cand, acc_tr, alpha);
            as_anderson = false;
        }
        if (!accepted) throw Failure(UPIC_ENONLINEAR, "outer current Anderson/Picard backtracking failed at the prescribed macro dt");
        if (as_anderson) ++anderson_steps; else ++picard_steps;
        if (alpha < 1.0) { ++backtracked; min_alpha = std::min(min_alpha, alpha); }
        prev_rms = nm.rms_relative;
        guess = std::move(cand);
        tr = std::move(acc_tr);
        have_prefetched = true;
    }
    if (!converged) throw Failure(UPIC_ENONLINEAR, "outer macro-current solve failed");
    tr = eval(guess);
    const Norms fn = residual_norms(tr.current_map, guess, P.c, P.nonlinear_rtol, P.nonlinear_atol);
    const bool ok = finite_nodes(tr.current_map) && fn.rms_scaled <= 1.0 && fn.max_particle_scaled <= 1.0;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = iters;
        info->converged = ok ? 1 : 0;
        info->anderson_restarts = anderson_restarts;
        info->anderson_steps = anderson_steps;
        info->picard_fallback_steps = picard_steps;
        info->backtracked_steps = backtracked;
        info->map_evaluations = maps;
        info->minimum_update_alpha = min_alpha;
        info->final_rms_residual = fn.rms_relative;
        info->final_max_particle_residual = fn.max_particle_relative;
        info->final_absolute_rms_residual = fn.rms_absolute;
        info->local_iterations_max = tr.local_iter_max;
        info->local_iterations_mean = tr.local_iter_mean;
        info->local_scaled_residual_max = tr.local_scaled_max;
        info->local_iterations_total = local_total;
        info->path.n_particles = static_cast<int64_t>(tr.stats.n_particles);
        info->path.crossing_particles = static_cast<int64_t>(tr.stats.crossing_particles);
        info->path.total_segments = static_cast<int64_t>(tr.stats.total_segments);
        info->path.max_segments = tr.stats.max_segments;
    }
    if (!ok) throw Failure(UPIC_ENONLINEAR, "final outer macro-current consistency evaluation failed");
    State out;
    out.x = tr.x_end; out.v = tr.v_next; out.v_prev = st.v; out.vbar_last = tr.vbar_macro; out.P = tr.P_end;
    out.q = st.q; out.m = st.m;
    out.phi = tr.phi_end; out.psi = tr.psi_end; out.A = tr.A_end; out.U = tr.U_end; out.rho = tr.rho_end;
    out.J = tr.J_map;
    return out;
}

// nonlinear_step_orbit (P3:6477-6526) + selective fallback (6328-6454)
static State p3_step(const Ctx& C, const State& st, upic_step_info* info) {
    std::vector<int> layout(st.x.size(), 1);
    try {
        return p3_step_layout(C, st, layout, info);
    } catch (const RefineNeeded& e) {
        if (!C.p.selective_particle_fallback)
            throw Failure(UPIC_ENONLINEAR, std::string("inner particle solve failed over full dt; ") + e.what());
        for (;;) {
            try {
                return p3_step_layout(C, st, layout, info);
            } catch (const RefineNeeded& e2) {
                for (std::size_t p : e2.failed)
                    if (layout[p] >= C.p.selective_particle_max_substeps)
                        throw Failure(UPIC_ENONLINEAR, "selective particle fallback exhausted");
                for (std::size_t p : e2.failed) layout[p] = std::min(C.p.selective_particle_max_substeps, 2 * layout[p]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// P2 explicit step (P2:537-612, 739-880, 1050-1189)
// ---------------------------------------------------------------------------
struct W1 { int i[3]; Real w[3]; };
static W1 quadratic_weights(Real x, Real dx, int n) {
    const Real s = x / dx;
    const int ic = static_cast<int>(std::floor(s + 0.5));
    const Real r = s - static_cast<Real>(ic);
    W1 o;
    o.i[0] = (ic - 1 + n) % n; o.i[1] = (ic + n) % n; o.i[2] = (ic + 1 + n) % n;
    o.w[0] = 0.5 * (0.5 - r) * (0.5 - r); o.w[1] = 0.75 - r * r; o.w[2] = 0.5 * (0.5 + r) * (0.5 + r);
    return o;
}
static int p2_index(const Grid& g, int i, int j) { i = (i % g.nx + g.nx) % g.nx; j = (j % g.ny + g.ny) % g.ny; return i * g.ny + j; }
static void p2_scatter(Field& f, const Grid& g, Real x, Real y, Real value) {
    const W1 wx = quadratic_weights(x, g.dx[0], g.nx), wy = quadratic_weights(y, g.dx[1], g.ny);
    const Real dA = g.dx[0] * g.dx[1];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) f[p2_index(g, wx.i[a], wy.i[b])] += value * wx.w[a] * wy.w[b] / dA;
}
static Real p2_gather(const Field& f, const Grid& g, Real x, Real y) {
    const W1 wx = quadratic_weights(x, g.dx[0], g.nx), wy = quadratic_weights(y, g.dx[1], g.ny);
    Real v = 0.0;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) v += f[p2_index(g, wx.i[a], wy.i[b])] * wx.w[a] * wy.w[b];
    return v;
}
static Field p2_deriv(const Field& f, const Grid& g, int dir) {
    const Complex I(0.0, 1.0);
    CField h = fft_real(f, g);
    for (int id = 0; id < g.N; ++id) h[id] *= I * ((dir == 0) ? g.kx[id] : g.ky[id]);
    return ifft_real(std::move(h), g);
}
static Field p2_div(const VecField& V, const Grid& g) {
    const Complex I(0.0, 1.0);
    CField hx = fft_real(V[0], g), hy = fft_real(V[1], g);
    CField d(g.N);
    for (int id = 0; id < g.N; ++id) d[id] = I * (g.kx[id] * hx[id] + g.ky[id] * hy[id]);
    return ifft_real(std::move(d), g);
}
static std::vector<Real> bdf_a(int order) { return order == 2 ? std::vector<Real>{1.5, -2.0, 0.5} : std::vector<Real>{1.0, -1.0}; }
static std::vector<Real> convolve(const std::vector<Real>& a) {
    std::vector<Real> b(2 * a.size() - 1, 0.0);
    for (std::size_t i = 0; i < a.size(); ++i) for (std::size_t j = 0; j < a.size(); ++j) b[i + j] += a[i] * a[j];
    return b;
}
struct P2State {
    std::vector<Real> x, y;         // positions in [0,L)
    std::vector<Vec3> P, v, v_prev;
    std::vector<Real> q, m;
    std::deque<Field> phi, rho;     // newest first, 5 deep
    std::array<std::deque<Field>, 3> A;
    VecField J;
    Field fixed_ion_rho;
};
static Field p2_bdf_wave(const Field& src, const std::deque<Field>& hist, const std::vector<Real>& a, Real c, Real dt,
                         const Grid& g) {
    const std::vector<Real> b = convolve(a);
    const Real scale = 1.0 / (c * c * dt * dt);
    CField rhs = fft_real(src, g);
    for (std::size_t l = 1; l < b.size(); ++l) {
        CField uh = fft_real(hist[l - 1], g);
        for (int id = 0; id < g.N; ++id) rhs[id] -= scale * b[l] * uh[id];
    }
    for (int id = 0; id < g.N; ++id) rhs[id] /= (scale * b[0] + g.k2[id]);
    return ifft_real(std::move(rhs), g);
}
static Vec3 p2_velocity_from_canonical(const Vec3& P, Real q, Real m, const Vec3& A, Real c) {
    const Vec3 p = sub3(P, mul3(q, A));
    const Real denom = std::sqrt(c * c * dot3(p, p) + (m * c * c) * (m * c * c));
    if (denom == 0.0) return {0, 0, 0};
    return mul3(c * c / denom, p);
}
static void p2_step(const Ctx& C, P2State& s) {
    const Grid& g = C.g;
    const upic_params& P = C.p;
    const Real dt = P.dt;
    const std::vector<Real> a = bdf_a(P.bdf_order);
    const std::size_t np = s.x.size();
    for (std::size_t p = 0; p < np; ++p) {
        s.x[p] = wrap_periodic(s.x[p] + dt * s.v[p][0], g.L[0]);
        s.y[p] = wrap_periodic(s.y[p] + dt * s.v[p][1], g.L[1]);
    }
    VecField Jn = zero_vecfield(g.N);
    for (std::size_t p = 0; p < np; ++p) {
        p2_scatter(Jn[0], g, s.x[p], s.y[p], s.q[p] * s.v[p][0]);
        p2_scatter(Jn[1], g, s.x[p], s.y[p], s.q[p] * s.v[p][1]);
        p2_scatter(Jn[2], g, s.x[p], s.y[p], s.q[p] * s.v[p][2]);
    }
    Field rho_new(g.N, 0.0);
    if (P.charge_conserving) {
        const Field divJ = p2_div(Jn, g);
        for (int id = 0; id < g.N; ++id) {
            Real hs = 0.0;
            for (std::size_t l = 1; l < a.size(); ++l) hs += a[l] * s.rho[l - 1][id];
            rho_new[id] = -(dt * divJ[id] + hs) / a[0];
        }
    } else {
        rho_new = s.fixed_ion_rho;
        for (std::size_t p = 0; p < np; ++p) p2_scatter(rho_new, g, s.x[p], s.y[p], s.q[p]);
    }
    Field phis(g.N);
    for (int id = 0; id < g.N; ++id) phis[id] = rho_new[id] / P.eps0;
    Field phi_new = p2_bdf_wave(phis, s.phi, a, P.c, dt, g);
    VecField A_new = zero_vecfield(g.N);
    for (int comp = 0; comp < 3; ++comp) {
        Field src(g.N);
        for (int id = 0; id < g.N; ++id) src[id] = P.mu0 * Jn[comp][id];
        A_new[comp] = p2_bdf_wave(src, s.A[comp], a, P.c, dt, g);
    }
    VecField gphi = zero_vecfield(g.N);
    gphi[0] = p2_deriv(phi_new, g, 0); gphi[1] = p2_deriv(phi_new, g, 1);
    Field gA[3][2];
    for (int comp = 0; comp < 3; ++comp) { gA[comp][0] = p2_deriv(A_new[comp], g, 0); gA[comp][1] = p2_deriv(A_new[comp], g, 1); }
    for (std::size_t p = 0; p < np; ++p) {
        const Real x = s.x[p], y = s.y[p];
        const Vec3 Ap{p2_gather(A_new[0], g, x, y), p2_gather(A_new[1], g, x, y), p2_gather(A_new[2], g, x, y)};
        const Vec3 gp{p2_gather(gphi[0], g, x, y), p2_gather(gphi[1], g, x, y), p2_gather(gphi[2], g, x, y)};
        const Vec3 vs = sub3(mul3(2.0, s.v[p]), s.v_prev[p]);
        const Real axx = p2_gather(gA[0][0], g, x, y), axy = p2_gather(gA[0][1], g, x, y);
        const Real ayx = p2_gather(gA[1][0], g, x, y), ayy = p2_gather(gA[1][1], g, x, y);
        const Real azx = p2_gather(gA[2][0], g, x, y), azy = p2_gather(gA[2][1], g, x, y);
        const Vec3 gav{axx * vs[0] + ayx * vs[1] + azx * vs[2], axy * vs[0] + ayy * vs[1] + azy * vs[2], 0.0};
        const Vec3 force = add3(mul3(-1.0, gp), gav);
        const Vec3 Pn = add3(s.P[p], mul3(s.q[p] * dt, force));
        const Vec3 vn = p2_velocity_from_canonical(Pn, s.q[p], s.m[p], Ap, P.c);
        s.v_prev[p] = s.v[p]; s.v[p] = vn; s.P[p] = Pn;
    }
    s.phi.push_front(std::move(phi_new)); while (s.phi.size() > 5) s.phi.pop_back();
    for (int comp = 0; comp < 3; ++comp) { s.A[comp].push_front(std::move(A_new[comp])); while (s.A[comp].size() > 5) s.A[comp].pop_back(); }
    s.rho.push_front(std::move(rho_new)); while (s.rho.size() > 5) s.rho.pop_back();
    s.J = std::move(Jn);
}

// ---------------------------------------------------------------------------
// Diagnostics (P1:1011-1087, P3:3338-3574)
// ---------------------------------------------------------------------------
static VecField compute_E(const State& st, const Grid& g) {
    VecField gp = spectral_gradient_scalar(st.phi, g);
    VecField E = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) for (int id = 0; id < g.N; ++id) E[d][id] = -gp[d][id] - st.U[d][id];
    return E;
}
static Real compute_field_energy(const State& st, const Grid& g, const upic_params& p) {
    const VecField E = compute_E(st, g);
    const VecField B = spectral_curl(st.A, g);
    long double se = 0.0L, sb = 0.0L;
    for (int d = 0; d < 3; ++d)
        for (int id = 0; id < g.N; ++id) {
            se += static_cast<long double>(E[d][id]) * static_cast<long double>(E[d][id]);
            sb += static_cast<long double>(B[d][id]) * static_cast<long double>(B[d][id]);
        }
    if (p.code == 3) return static_cast<Real>((0.5 / p.sigma1 * se + 0.5 / p.sigma2 * sb) * g.dV);
    return static_cast<Real>((0.5 * p.eps0 * se + 0.5 / p.mu0 * sb) * g.dV);
}
static Real compute_kinetic_energy(const State& st, const upic_params& p) {
    long double s = 0.0L;
    if (p.code == 3) {
        for (std::size_t i = 0; i < st.v.size(); ++i) {
            const Real gamma = gamma_from_v(st.v[i], p.c);
            s += static_cast<long double>(st.m[i]) * p.c * p.c * (gamma - 1.0);
        }
        return static_cast<Real>(s);
    }
    for (std::size_t i = 0; i < st.v.size(); ++i) s += 0.5L * st.m[i] * dot3(st.v[i], st.v[i]);
    return static_cast<Real>(s);
}
static upic_diagnostics compute_diagnostics(const State& st, const Grid& g, const upic_params& p, int mode) {
    upic_diagnostics d;
    std::memset(&d, 0, sizeof(d));
    const VecField E = compute_E(st, g);
    const Field divE = spectral_divergence_vector(E, g);
    const Real s1 = (p.code == 3) ? p.sigma1 : 1.0 / p.eps0;
    Field gauss(g.N);
    for (int id = 0; id < g.N; ++id)
        gauss[id] = (p.code == 3) ? divE[id] - p.sigma1 * st.rho[id] : divE[id] - st.rho[id] / p.eps0;
    d.gauss_rms = rms_field(gauss);
    for (Real x : gauss) d.gauss_max = std::max(d.gauss_max, std::abs(x));
    d.divE_rms = rms_field(divE);
    d.rho_rms = std::abs(s1) * rms_field(st.rho);
    const Field divA = spectral_divergence_vector(st.A, g);
    Field gauge(g.N);
    for (int id = 0; id < g.N; ++id) gauge[id] = st.psi[id] / (p.c * p.c) + divA[id];
    d.gauge_rms = rms_field(gauge);
    for (Real x : gauge) d.gauge_max = std::max(d.gauge_max, std::abs(x));
    d.psi_over_c2_rms = rms_field(st.psi) / (p.c * p.c);
    d.divA_rms = rms_field(divA);
    d.field_energy = compute_field_energy(st, g, p);
    d.kinetic_energy = compute_kinetic_energy(st, p);
    d.total_energy = d.field_energy + d.kinetic_energy;
    const CField rhohat = fft_real(st.rho, g);
    const CField Exhat = fft_real(E[0], g);
    const int imode = ((mode % g.nx) + g.nx) % g.nx;
    const int jmode = ((mode % g.ny) + g.ny) % g.ny;
    const int idx = g.index(imode, 0, 0), idy = g.index(0, jmode, 0);
    const Real invN = 1.0 / static_cast<Real>(g.N);
    d.rho_mode = std::abs(rhohat[idx]) * invN;
    d.E_mode_re = Exhat[idx].real() * invN;
    d.E_mode_im = Exhat[idx].imag() * invN;
    d.E_mode_abs = std::abs(Exhat[idx]) * invN;
    if (p.code == 3) {
        const VecField B = spectral_curl(st.A, g);
        const CField Bzhat = fft_real(B[2], g);
        d.Bz_y_mode_re = Bzhat[idy].real() * invN;
        d.Bz_y_mode_im = Bzhat[idy].imag() * invN;
        d.Bz_y_mode_abs = std::abs(Bzhat[idy]) * invN;
    }
    d.mean_rho = mean_field(st.rho);
    return d;
}

// compute_step_diagnostics (P1:1384-1439); out[0..18] in the CSV column order
// of P1:1819 (after step,time).
static void p1_step_diagnostics(const Ctx& C, const State& s0, const State& s1, double* out) {
    const Grid& g = C.g;
    const Real dt = C.p.dt;
    const std::size_t np = s0.x.size();
    std::vector<Vec3> vbar(np), dxp(np), x1(np);
    for (std::size_t p = 0; p < np; ++p) {
        vbar[p] = mul3(0.5, add3(s1.v[p], s0.v[p]));
        dxp[p] = mul3(dt, vbar[p]);
        x1[p] = wrap_pos(add3(s0.x[p], dxp[p]), g);
    }
    const PathStats ps = compute_path_stats(s0.x, dxp, g, C.order(), C.split());
    VecField U_mid = zero_vecfield(g.N);
    for (int d = 0; d < 3; ++d) for (int id = 0; id < g.N; ++id) U_mid[d][id] = 0.5 * s0.U[d][id] + 0.5 * s1.U[d][id];
    const auto Ubar = path_average_gather_vector(s0.x, dxp, U_mid, g, C.order(), C.quad, C.split());
    std::vector<Vec3> R(np), Ad(np), rhs(np);
    for (std::size_t p = 0; p < np; ++p) {
        const Mat3 DA = orbit_discrete_gradient_A_interval(s0.x[p], dxp[p], s0.A, s1.A, 0.0, 1.0, g, C.order(), C.quad, C.split());
        const Vec3 A0p = gather_vector_at(s0.A, s0.x[p], g, C.order());
        const Vec3 A1p = gather_vector_at(s1.A, x1[p], g, C.order());
        Ad[p] = sub3(A1p, A0p);
        Vec3 DAv{0, 0, 0};
        for (int a = 0; a < 3; ++a) DAv[a] = DA[a][0] * vbar[p][0] + DA[a][1] * vbar[p][1] + DA[a][2] * vbar[p][2];
        rhs[p] = mul3(dt, add3(Ubar[p], DAv));
        R[p] = sub3(Ad[p], rhs[p]);
    }
    const Real abs_rms = rms_vec_particles(R);
    const Real rel_rms = abs_rms / (0.5 * (rms_vec_particles(Ad) + rms_vec_particles(rhs)) + 1.0e-300);
    long double norm_sum = 0.0L, work_defect = 0.0L;
    Real maxn = 0.0;
    for (std::size_t p = 0; p < np; ++p) {
        const Real nr = norm3(R[p]);
        maxn = std::max(maxn, nr);
        norm_sum += nr;
        work_defect += -static_cast<long double>(s0.q[p]) * dot3(vbar[p], R[p]);
    }
    VecField gp0 = spectral_gradient_scalar(s0.phi, g), gp1 = spectral_gradient_scalar(s1.phi, g);
    VecField E_mid = zero_vecfield(g.N);
    for (int m = 0; m < 3; ++m) for (int id = 0; id < g.N; ++id) E_mid[m][id] = -0.5 * (gp0[m][id] + gp1[m][id]) - U_mid[m][id];
    const auto Ebar = path_average_gather_vector(s0.x, dxp, E_mid, g, C.order(), C.quad, C.split());
    const VecField Jbar = path_average_deposit_current(s0.x, dxp, vbar, s0.q, g, C.order(), C.quad, C.split(), nullptr);
    long double pw = 0.0L, gw = 0.0L;
    for (std::size_t p = 0; p < np; ++p) pw += static_cast<long double>(s0.q[p]) * dot3(vbar[p], Ebar[p]);
    for (int m = 0; m < 3; ++m) for (int id = 0; id < g.N; ++id) gw += static_cast<long double>(Jbar[m][id]) * E_mid[m][id];
    const Real particle_work = static_cast<Real>(dt * pw);
    const Real grid_work = static_cast<Real>(dt * gw * g.dV);
    const Real dK = compute_kinetic_energy(s1, C.p) - compute_kinetic_energy(s0, C.p);
    const Real dW = compute_field_energy(s1, g, C.p) - compute_field_energy(s0, g, C.p);
    out[0] = abs_rms; out[1] = rel_rms; out[2] = maxn;
    out[3] = static_cast<Real>(norm_sum / std::max<std::size_t>(1, np));
    out[4] = static_cast<Real>(work_defect);
    out[5] = dK; out[6] = dW; out[7] = dK + dW;
    out[8] = particle_work; out[9] = grid_work;
    out[10] = particle_work - grid_work; out[11] = dK - particle_work; out[12] = dW + grid_work;
    out[13] = static_cast<Real>(ps.crossing_particles);
    out[14] = static_cast<Real>(ps.total_segments >= ps.n_particles ? ps.total_segments - ps.n_particles : 0);
    out[15] = ps.n_particles ? static_cast<Real>(ps.total_segments) / static_cast<Real>(ps.n_particles) : 1.0;
    out[16] = static_cast<Real>(ps.max_segments);
}

// ---------------------------------------------------------------------------
// Initialisers used to reproduce the reference's stored runs
// ---------------------------------------------------------------------------
static std::vector<Real> cell_centered_grid(int n, Real L, bool centered) {
    std::vector<Real> x(n);
    const Real dx = L / static_cast<Real>(n);
    for (int i = 0; i < n; ++i) { x[i] = (static_cast<Real>(i) + 0.5) * dx; if (centered) x[i] -= 0.5 * L; }
    return x;
}
// Acklam's inverse normal CDF, coefficients as published by P. J. Acklam and
// used by P1:1103-1149.
static Real normal_quantile(Real p) {
    static const Real a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const Real b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
    static const Real c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const Real d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00, 3.754408661907416e+00};
    const Real plow = 0.02425, phigh = 1.0 - plow;
    Real q, r;
    if (p < plow) {
        q = std::sqrt(-2.0 * std::log(p));
        return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    if (p <= phigh) {
        q = p - 0.5; r = q * q;
        return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
               (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
    }
    q = std::sqrt(-2.0 * std::log(1.0 - p));
    return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
}

// Common tail of the P1 initialisers (P1:1196-1211, 1287-1303).
static void p1_init_tail(const Ctx& C, State& st, Real n0) {
    const Grid& g = C.g;
    const std::size_t np = st.x.size();
    Field rho_e = deposit_scalar_to_mesh(st.x, st.q, g, C.order(), true);
    st.rho.assign(g.N, 0.0);
    for (int id = 0; id < g.N; ++id) st.rho[id] = rho_e[id] + n0;
    subtract_mean(st.rho);
    st.J = deposit_current_to_mesh(st.x, st.v, st.q, g, C.order());
    st.phi = solve_periodic_poisson_for_phi(st.rho, g, 1.0 / C.p.eps0, true, C.p.eps0);
    st.psi.assign(g.N, 0.0);
    st.A = zero_vecfield(g.N);
    st.U = zero_vecfield(g.N);
    st.P.assign(np, {0, 0, 0});
    for (std::size_t p = 0; p < np; ++p) st.P[p] = mul3(st.m[p], st.v[p]);
    const VecField E0 = compute_E(st, g);
    st.v_prev.assign(np, {0, 0, 0});
    for (std::size_t p = 0; p < np; ++p) {
        const Vec3 Ep = gather_vector_at(E0, st.x[p], g, C.order());
        st.v_prev[p] = sub3(st.v[p], mul3(C.p.dt, mul3(st.q[p] / st.m[p], Ep)));
    }
    st.vbar_last = st.v;
}

// Tail of the P3 initialisers for given particles (P3:3920-3937): rho, J,
// phi, P = m gamma v + q A (A = 0), projection (3398-3405) and the
// relativistic history velocity (3423-3441).
static void p3_init_tail(const Ctx& C, State& st, Real n0) {
    const Grid& g = C.g;
    const std::size_t np = st.x.size();
    const Real kappa = C.p.c;
    Field rho_e = deposit_scalar_to_mesh(st.x, st.q, g, C.order(), true);
    st.rho.assign(g.N, 0.0);
    for (int id = 0; id < g.N; ++id) st.rho[id] = rho_e[id] + n0;
    subtract_mean(st.rho);
    st.J = deposit_current_to_mesh(st.x, st.v, st.q, g, C.order());
    st.phi = solve_periodic_poisson_for_phi(st.rho, g, C.p.sigma1, false, 1.0);
    st.psi.assign(g.N, 0.0);
    st.A = zero_vecfield(g.N);
    st.U = zero_vecfield(g.N);
    st.P.assign(np, {0, 0, 0});
    for (std::size_t p = 0; p < np; ++p) st.P[p] = momentum_from_v(st.v[p], st.m[p], kappa);
    if (g.nyquist_projection) {
        project_nyquist(st.rho, g); project_nyquist(st.phi, g); project_nyquist(st.psi, g);
        project_nyquist(st.A, g); project_nyquist(st.U, g); project_nyquist(st.J, g);
        for (std::size_t p = 0; p < np; ++p)
            st.P[p] = add3(momentum_from_v(st.v[p], st.m[p], kappa), mul3(st.q[p], gather_vector_at(st.A, st.x[p], g, C.order())));
    }
    const VecField E0 = compute_E(st, g);
    const VecField B0 = spectral_curl(st.A, g);
    st.vbar_last = st.v;
    st.v_prev.assign(np, {0, 0, 0});
    for (std::size_t p = 0; p < np; ++p) {
        const Vec3 Ep = gather_vector_at(E0, st.x[p], g, C.order());
        const Vec3 Bp = gather_vector_at(B0, st.x[p], g, C.order());
        const Vec3 pn = momentum_from_v(st.v[p], st.m[p], kappa);
        const Vec3 lor = mul3(st.q[p], add3(Ep, cross3(st.v[p], Bp)));
        st.v_prev[p] = velocity_from_p(sub3(pn, mul3(C.p.dt, lor)), st.m[p], kappa);
    }
}

// initialize_landau_damping_1d3d (P1:1214-1304)
static State p1_init_landau(const Ctx& C, int nv, Real alpha, Real vth, int mode, Real n0) {
    const Grid& g = C.g;
    State st;
    const auto xl = cell_centered_grid(g.nx, g.L[0], true), yl = cell_centered_grid(g.ny, g.L[1], true),
               zl = cell_centered_grid(g.nz, g.L[2], true);
    std::vector<Real> vq(nv);
    for (int j = 0; j < nv; ++j) vq[j] = vth * normal_quantile((static_cast<Real>(j) + 0.5) / static_cast<Real>(nv));
    const std::size_t np = static_cast<std::size_t>(g.nx) * g.ny * g.nz * nv;
    st.x.assign(np, {0, 0, 0}); st.v.assign(np, {0, 0, 0}); st.q.assign(np, 0.0); st.m.assign(np, 0.0);
    const Real kx = 2.0 * PI * static_cast<Real>(mode) / g.L[0];
    const Real q0 = -n0 * (g.L[0] * g.L[1] * g.L[2]) / static_cast<Real>(np);
    std::size_t pi = 0;
    for (Real y : yl) for (Real z : zl) for (Real x : xl) {
        const Real dw = 1.0 + alpha * std::cos(kx * x);
        for (int jv = 0; jv < nv; ++jv) {
            st.x[pi] = wrap_pos(Vec3{x, y, z}, g);
            st.v[pi] = {vq[jv], 0.0, 0.0};
            st.q[pi] = q0 * dw;
            st.m[pi] = std::abs(st.q[pi]);
            ++pi;
        }
    }
    p1_init_tail(C, st, n0);
    return st;
}
// initialize_two_stream_3d_linear_x (P1:1151-1212)
static State p1_init_two_stream(const Ctx& C, int ppc, Real v0, Real pert, int mode, Real n0) {
    const Grid& g = C.g;
    State st;
    const int nxl = g.nx * ppc;
    const auto xl = cell_centered_grid(nxl, g.L[0], true), yl = cell_centered_grid(g.ny, g.L[1], true),
               zl = cell_centered_grid(g.nz, g.L[2], true);
    std::vector<Vec3> base;
    for (Real y : yl) for (Real z : zl) for (Real x : xl) base.push_back({x, y, z});
    const std::size_t nb = base.size(), np = 2 * nb;
    st.x.assign(np, {0, 0, 0}); st.v.assign(np, {0, 0, 0}); st.q.assign(np, 0.0); st.m.assign(np, 0.0);
    const Real kx = 2.0 * PI * static_cast<Real>(mode) / g.L[0];
    for (std::size_t p = 0; p < nb; ++p) {
        const Real sv = pert * std::sin(kx * base[p][0]);
        st.x[p] = wrap_pos(base[p], g);
        st.x[nb + p] = wrap_pos(base[p], g);
        st.v[p] = {+v0 + sv, 0.0, 0.0};
        st.v[nb + p] = {-v0 + sv, 0.0, 0.0};
    }
    const Real qm = -n0 * (g.L[0] * g.L[1] * g.L[2]) / static_cast<Real>(np);
    for (std::size_t p = 0; p < np; ++p) { st.q[p] = qm; st.m[p] = std::abs(qm); }
    p1_init_tail(C, st, n0);
    return st;
}

// ---------------------------------------------------------------------------
// SoA <-> State conversion at the C boundary
// ---------------------------------------------------------------------------
static std::vector<Vec3> load3(const double* a, int64_t np) {
    std::vector<Vec3> v(static_cast<std::size_t>(np));
    for (int64_t p = 0; p < np; ++p) v[p] = {a[p], a[np + p], a[2 * np + p]};
    return v;
}
static void store3(const std::vector<Vec3>& v, double* a) {
    const int64_t np = static_cast<int64_t>(v.size());
    for (int64_t p = 0; p < np; ++p) { a[p] = v[p][0]; a[np + p] = v[p][1]; a[2 * np + p] = v[p][2]; }
}
static VecField loadF3(const double* a, int N) {
    VecField f = zero_vecfield(N);
    for (int d = 0; d < 3; ++d) std::copy(a + d * N, a + (d + 1) * N, f[d].begin());
    return f;
}
static void storeF3(const VecField& f, double* a) {
    const std::size_t N = f[0].size();
    for (int d = 0; d < 3; ++d) std::copy(f[d].begin(), f[d].end(), a + d * N);
}
static State load_state(const upic_host_state* s, int N) {
    State st;
    st.x = load3(s->x, s->np); st.v = load3(s->v, s->np);
    st.v_prev = s->v_prev ? load3(s->v_prev, s->np) : st.v;
    st.vbar_last = s->vbar_last ? load3(s->vbar_last, s->np) : st.v;
    st.P = load3(s->P, s->np);
    st.q.assign(s->q, s->q + s->np); st.m.assign(s->m, s->m + s->np);
    st.phi.assign(s->phi, s->phi + N); st.psi.assign(s->psi, s->psi + N); st.rho.assign(s->rho, s->rho + N);
    st.A = loadF3(s->A, N); st.U = loadF3(s->U, N); st.J = loadF3(s->J, N);
    return st;
}
static void store_state(const State& st, upic_host_state* s) {
    store3(st.x, s->x); store3(st.v, s->v);
    if (s->v_prev) store3(st.v_prev, s->v_prev);
    if (s->vbar_last && !st.vbar_last.empty()) store3(st.vbar_last, s->vbar_last);
    store3(st.P, s->P);
    std::copy(st.q.begin(), st.q.end(), s->q); std::copy(st.m.begin(), st.m.end(), s->m);
    std::copy(st.phi.begin(), st.phi.end(), s->phi); std::copy(st.psi.begin(), st.psi.end(), s->psi);
    std::copy(st.rho.begin(), st.rho.end(), s->rho);
    storeF3(st.A, s->A); storeF3(st.U, s->U); storeF3(st.J, s->J);
}

template <class F>
static int guarded(F&& f) {
    try { f(); return UPIC_OK; }
    catch (const Failure& e) { g_err = e.what(); return e.code; }
    catch (const std::exception& e) { g_err = e.what(); return UPIC_EINVAL; }
}

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_quadrature(int n, double* s, double* w) {
    return guarded([&] { const Quadrature q = gauss_legendre_01(n); std::copy(q.s.begin(), q.s.end(), s); std::copy(q.w.begin(), q.w.end(), w); });
}
int orc_bspline(int code, int degree, int n, const double* r, double* w, double* dw) {
    return guarded([&] { for (int i = 0; i < n; ++i) { w[i] = centered_bspline(code, degree, r[i]); dw[i] = centered_bspline_derivative(code, degree, r[i]); } });
}
int orc_breakpoints(const upic_params* p, const double* x0, const double* dx, int64_t np, int32_t* nseg, double* br, int max_br) {
    return guarded([&] {
        const Ctx C(*p);
        const auto X = load3(x0, np), D = load3(dx, np);
        for (int64_t i = 0; i < np; ++i) {
            const auto b = path_breakpoints_for_particle(X[i], D[i], C.g, C.order(), C.split());
            nseg[i] = static_cast<int32_t>(b.size()) - 1;
            for (int k = 0; k < max_br; ++k) br[i * max_br + k] = (k < static_cast<int>(b.size())) ? b[k] : -1.0;
        }
    });
}
int orc_fft(const upic_params* p, double* data, int inverse) {
    return guarded([&] {
        const Grid g = make_grid(*p);
        CField c(g.N);
        for (int i = 0; i < g.N; ++i) c[i] = Complex(data[2 * i], data[2 * i + 1]);
        fft3d(c, g, inverse != 0);
        for (int i = 0; i < g.N; ++i) { data[2 * i] = c[i].real(); data[2 * i + 1] = c[i].imag(); }
    });
}
int orc_deposit_current_orbit(const upic_params* p, const double* x0, const double* dx, const double* vbar,
                              const double* q, int64_t np, double* J, upic_path_stats* stats) {
    return guarded([&] {
        const Ctx C(*p);
        PathStats ps;
        const VecField Jf = path_average_deposit_current(load3(x0, np), load3(dx, np), load3(vbar, np),
                                                         std::vector<Real>(q, q + np), C.g, C.order(), C.quad, C.split(), &ps);
        storeF3(Jf, J);
        if (stats) {
            stats->n_particles = static_cast<int64_t>(ps.n_particles);
            stats->crossing_particles = static_cast<int64_t>(ps.crossing_particles);
            stats->total_segments = static_cast<int64_t>(ps.total_segments);
            stats->max_segments = ps.max_segments;
        }
    });
}
int orc_gather_orbit(const upic_params* p, const double* x0, const double* dx, const double* field3, int64_t np, double* out) {
    return guarded([&] {
        const Ctx C(*p);
        store3(path_average_gather_vector(load3(x0, np), load3(dx, np), loadF3(field3, C.g.N), C.g, C.order(), C.quad, C.split()), out);
    });
}
int orc_orbit_grad_A(const upic_params* p, const double* x0, const double* dx, const double* An, const double* Anp1,
                     double theta0, double theta1, int64_t np, double* D) {
    return guarded([&] {
        const Ctx C(*p);
        const auto X = load3(x0, np), Dx = load3(dx, np);
        const VecField A0 = loadF3(An, C.g.N), A1 = loadF3(Anp1, C.g.N);
        for (int64_t i = 0; i < np; ++i) {
            const Mat3 M = orbit_discrete_gradient_A_interval(X[i], Dx[i], A0, A1, theta0, theta1, C.g, C.order(), C.quad, C.split());
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) D[(a * 3 + b) * np + i] = M[a][b];
        }
    });
}
int orc_gather_point(const upic_params* p, const double* pos, const double* field3, int64_t np, double* out) {
    return guarded([&] {
        const Ctx C(*p);
        const auto X = load3(pos, np);
        const VecField F = loadF3(field3, C.g.N);
        std::vector<Vec3> o(np);
        for (int64_t i = 0; i < np; ++i) o[i] = gather_vector_at(F, X[i], C.g, C.order());
        store3(o, out);
    });
}
int orc_deposit_scalar_point(const upic_params* p, const double* pos, const double* values, int64_t np, int divide, double* mesh) {
    return guarded([&] {
        const Ctx C(*p);
        const Field m = deposit_scalar_to_mesh(load3(pos, np), std::vector<Real>(values, values + np), C.g, C.order(), divide != 0);
        std::copy(m.begin(), m.end(), mesh);
    });
}
int orc_spectral_gradient(const upic_params* p, const double* f, double* out3) {
    return guarded([&] { const Grid g = make_grid(*p); storeF3(spectral_gradient_scalar(Field(f, f + g.N), g), out3); });
}
int orc_spectral_divergence(const upic_params* p, const double* V3, double* out) {
    return guarded([&] { const Grid g = make_grid(*p); const Field d = spectral_divergence_vector(loadF3(V3, g.N), g); std::copy(d.begin(), d.end(), out); });
}
int orc_spectral_curl(const upic_params* p, const double* A3, double* out3) {
    return guarded([&] { const Grid g = make_grid(*p); storeF3(spectral_curl(loadF3(A3, g.N), g), out3); });
}
int orc_continuity(const upic_params* p, const double* rho_n, const double* J3, double dt, double* out) {
    return guarded([&] {
        const Grid g = make_grid(*p);
        const Field r = continuity_charge_update(Field(rho_n, rho_n + g.N), loadF3(J3, g.N), dt, g);
        std::copy(r.begin(), r.end(), out);
    });
}
int orc_cn_potentials(const upic_params* p, const double* phi, const double* psi, const double* A3, const double* U3,
                      const double* rho_mid, const double* J3, double dt, double* phi_o, double* psi_o, double* A_o,
                      double* U_o) {
    return guarded([&] {
        const Grid g = make_grid(*p);
        const WaveSpectra sp = make_spectra(Field(phi, phi + g.N), Field(psi, psi + g.N), loadF3(A3, g.N), loadF3(U3, g.N), g);
        Field pn, sn;
        VecField An, Un;
        cn_potential_update(sp, Field(rho_mid, rho_mid + g.N), loadF3(J3, g.N), dt, g, coefs(*p), pn, sn, An, Un);
        std::copy(pn.begin(), pn.end(), phi_o); std::copy(sn.begin(), sn.end(), psi_o);
        storeF3(An, A_o); storeF3(Un, U_o);
    });
}
int orc_project_nyquist(const upic_params* p, double* f, int ncomp) {
    return guarded([&] {
        const Grid g = make_grid(*p);
        for (int c = 0; c < ncomp; ++c) { Field x(f + c * g.N, f + (c + 1) * g.N); project_nyquist(x, g); std::copy(x.begin(), x.end(), f + c * g.N); }
    });
}
// solve_bdf_wave_scalar (P2:829-848); hist holds u^n, u^{n-1}, ... (N each)
int orc_bdf_wave(const upic_params* p, const double* src, const double* hist, int order, double* out) {
    return guarded([&] {
        const Ctx C(*p);
        const int N = C.g.N;
        const std::vector<Real> a = order == 2 ? std::vector<Real>{1.5, -2.0, 0.5} : std::vector<Real>{1.0, -1.0};
        std::deque<Field> h;
        for (std::size_t l = 0; l + 1 < 2 * a.size() - 1; ++l) h.push_back(Field(hist + l * N, hist + (l + 1) * N));
        const Field u = p2_bdf_wave(Field(src, src + N), h, a, C.p.c, C.p.dt, C.g);
        std::copy(u.begin(), u.end(), out);
    });
}
int orc_poisson(const upic_params* p, const double* rho, double* phi) {
    return guarded([&] {
        const Grid g = make_grid(*p);
        const Field r = (p->code == 3) ? solve_periodic_poisson_for_phi(Field(rho, rho + g.N), g, p->sigma1, false, 1.0)
                                       : solve_periodic_poisson_for_phi(Field(rho, rho + g.N), g, 1.0 / p->eps0, true, p->eps0);
        std::copy(r.begin(), r.end(), phi);
    });
}
// One P3 inner solve per particle (solve_local_particle_substep, P3:4930-4983),
// nsub = 1.  Outputs vbar (accepted guess), P_end, v_end (3*np), iterations,
// reasons (np).
int orc_p3_local_solve(const upic_params* p, const double* guess, const double* x0, const double* P0, const double* q,
                       const double* m, double h, double theta0, double theta1, const double* A_n, const double* A_end,
                       const double* gpm, int64_t np, double* vbar, double* P_end, double* v_end, int32_t* iters,
                       int32_t* reasons) {
    return guarded([&] {
        const Ctx C(*p);
        const auto G = load3(guess, np), X = load3(x0, np), P0v = load3(P0, np);
        const VecField An = loadF3(A_n, C.g.N), Ae = loadF3(A_end, C.g.N), Gp = loadF3(gpm, C.g.N);
        std::vector<Vec3> vb(np, {0, 0, 0}), Pe(np, {0, 0, 0}), ve(np, {0, 0, 0});
        for (int64_t i = 0; i < np; ++i) {
            LocalStep ls;
            LocalReport rep;
            p3_local_solve(C, G[i], X[i], P0v[i], q[i], m[i], h, theta0, theta1, An, Ae, Gp, vb[i], ls, rep);
            Pe[i] = ls.P_end; ve[i] = ls.v_end;
            iters[i] = rep.iterations; reasons[i] = rep.reason;
        }
        store3(vb, vbar); store3(Pe, P_end); store3(ve, v_end);
    });
}
int orc_step(const upic_params* p, const upic_host_state* in, upic_host_state* out, upic_step_info* info) {
    return guarded([&] {
        const Ctx C(*p);
        const State st = load_state(in, C.g.N);
        State o;
        if (p->code == 1) o = p1_step(C, st, info);
        else if (p->code == 3) o = p3_step(C, st, info);
        else throw Failure(UPIC_EINVAL, "orc_step: use orc_p2_run for P2");
        store_state(o, out);
    });
}
// Runs n_steps of the P1 or P3 driver, writing per-step diagnostics
// (rows of 19 doubles: time, total/field/kinetic energy, gauss_rms, gauge_rms,
// iterations, current_error, velocity_error, E_mode_abs, Bz_y_mode_abs,
// crossing_particles, avg segments, max segments, rho_mode, mean_rho,
// gauss_max, local_iterations_mean, map_evaluations).  Row 0 is the initial state.
int orc_run(const upic_params* p, upic_host_state* st_io, int n_steps, int mode, double* rows) {
    return guarded([&] {
        const Ctx C(*p);
        State st = load_state(st_io, C.g.N);
        auto row = [&](int n, const State& s, const upic_step_info* inf) {
            const upic_diagnostics d = compute_diagnostics(s, C.g, *p, mode);
            double* r = rows + static_cast<std::size_t>(n) * 19;
            r[0] = n * p->dt; r[1] = d.total_energy; r[2] = d.field_energy; r[3] = d.kinetic_energy;
            r[4] = d.gauss_rms; r[5] = d.gauge_rms;
            r[6] = inf ? inf->iterations : 0; r[7] = inf ? inf->current_error : NAN; r[8] = inf ? inf->velocity_error : NAN;
            r[9] = d.E_mode_abs; r[10] = d.Bz_y_mode_abs;
            r[11] = inf ? static_cast<double>(inf->path.crossing_particles) : 0.0;
            r[12] = (inf && inf->path.n_particles) ? static_cast<double>(inf->path.total_segments) / inf->path.n_particles : 0.0;
            r[13] = inf ? inf->path.max_segments : 1;
            r[14] = d.rho_mode; r[15] = d.mean_rho; r[16] = d.gauss_max;
            r[17] = inf ? inf->local_iterations_mean : 0.0; r[18] = inf ? inf->map_evaluations : 0;
        };
        row(0, st, nullptr);
        for (int n = 0; n < n_steps; ++n) {
            upic_step_info inf;
            st = (p->code == 1) ? p1_step(C, st, &inf) : p3_step(C, st, &inf);
            row(n + 1, st, &inf);
        }
        store_state(st, st_io);
    });
}
// Initial mesh state, momenta and history velocity from given x, v, q, m
// (tails of the P1 / P3 initialisers).
int orc_init_fields(const upic_params* p, upic_host_state* st_io, double n0) {
    return guarded([&] {
        const Ctx C(*p);
        State st;
        st.x = load3(st_io->x, st_io->np); st.v = load3(st_io->v, st_io->np);
        st.q.assign(st_io->q, st_io->q + st_io->np); st.m.assign(st_io->m, st_io->m + st_io->np);
        if (p->code == 1) p1_init_tail(C, st, n0); else p3_init_tail(C, st, n0);
        store_state(st, st_io);
    });
}
int orc_diagnostics(const upic_params* p, const upic_host_state* s, int mode, upic_diagnostics* out) {
    return guarded([&] { const Ctx C(*p); *out = compute_diagnostics(load_state(s, C.g.N), C.g, *p, mode); });
}
int orc_p1_step_diagnostics(const upic_params* p, const upic_host_state* s0, const upic_host_state* s1, double* out17) {
    return guarded([&] { const Ctx C(*p); p1_step_diagnostics(C, load_state(s0, C.g.N), load_state(s1, C.g.N), out17); });
}
// P1 initialisers.  kind 0 = Landau (P1:1214-1304), 1 = two-stream (P1:1151-1212).
// Call with st->x == NULL to get np in st->np.
int orc_p1_init(const upic_params* p, int kind, int ppc, double alpha, double vth, double v0, double pert, int mode,
                double n0, upic_host_state* st) {
    return guarded([&] {
        const Ctx C(*p);
        const std::size_t npx = (kind == 0) ? static_cast<std::size_t>(C.g.N) * ppc : 2ull * C.g.N * ppc;
        if (!st->x) { st->np = static_cast<int64_t>(npx); return; }
        const State s = (kind == 0) ? p1_init_landau(C, ppc, alpha, vth, mode, n0) : p1_init_two_stream(C, ppc, v0, pert, mode, n0);
        store_state(s, st);
    });
}
// P2 run (P2:1094-1196) from a given initial particle set.  Positions are in
// [0,L) (P2 convention) in st->x (x, y, unused z).  Histories are initialised
// from the initial fields as P2:1126-1131.  rows: per step (n_steps+1 rows of
// 8): gauge_rms, gauss_rms, continuity_rms, Bz_l2, field_energy,
// kinetic_energy, total_energy, max_speed.
int orc_p2_run(const upic_params* p, upic_host_state* st, const double* fixed_ion_rho, int n_steps, double* rows) {
    return guarded([&] {
        const Ctx C(*p);
        const Grid& g = C.g;
        P2State s;
        const int64_t np = st->np;
        const auto X = load3(st->x, np);
        s.x.resize(np); s.y.resize(np);
        for (int64_t i = 0; i < np; ++i) { s.x[i] = X[i][0]; s.y[i] = X[i][1]; }
        s.v = load3(st->v, np); s.v_prev = load3(st->v_prev, np); s.P = load3(st->P, np);
        s.q.assign(st->q, st->q + np); s.m.assign(st->m, st->m + np);
        s.fixed_ion_rho.assign(fixed_ion_rho, fixed_ion_rho + g.N);
        // initial rho = ions + electrons (naive), J, phi Poisson (P2:1121-1124, 812-827)
        Field rho0 = s.fixed_ion_rho;
        for (int64_t i = 0; i < np; ++i) p2_scatter(rho0, g, s.x[i], s.y[i], s.q[i]);
        VecField J0 = zero_vecfield(g.N);
        for (int64_t i = 0; i < np; ++i) {
            p2_scatter(J0[0], g, s.x[i], s.y[i], s.q[i] * s.v[i][0]);
            p2_scatter(J0[1], g, s.x[i], s.y[i], s.q[i] * s.v[i][1]);
            p2_scatter(J0[2], g, s.x[i], s.y[i], s.q[i] * s.v[i][2]);
        }
        CField rh = fft_real(rho0, g);
        CField ph(g.N, Complex(0.0, 0.0));
        for (int id = 0; id < g.N; ++id) if (g.k2[id] > 0.0) ph[id] = rh[id] / (p->eps0 * g.k2[id]);
        const Field phi0 = ifft_real(std::move(ph), g);
        s.phi.assign(5, phi0); s.rho.assign(5, rho0);
        for (int c = 0; c < 3; ++c) s.A[c].assign(5, Field(g.N, 0.0));
        s.J = J0;
        auto diag = [&](int n, const Field& phi_new, const VecField& A_new, const Field& rho_cur, const VecField& J_cur,
                        const std::vector<Real>& a, const std::deque<Field>& phist, const std::array<std::deque<Field>, 3>& Ahist,
                        const std::deque<Field>& rhist) {
            auto dt1 = [&](const Field& un, const std::deque<Field>& h) {
                Field o(g.N);
                for (int id = 0; id < g.N; ++id) o[id] = a[0] * un[id];
                for (std::size_t l = 1; l < a.size(); ++l) for (int id = 0; id < g.N; ++id) o[id] += a[l] * h[l - 1][id];
                for (Real& v : o) v /= p->dt;
                return o;
            };
            const Field phit = dt1(phi_new, phist);
            const Field divA = p2_div(A_new, g);
            Field gauge(g.N);
            for (int id = 0; id < g.N; ++id) gauge[id] = phit[id] / (p->c * p->c) + divA[id];
            const Field gpx = p2_deriv(phi_new, g, 0), gpy = p2_deriv(phi_new, g, 1);
            VecField E = zero_vecfield(g.N);
            for (int c = 0; c < 3; ++c) {
                const Field At = dt1(A_new[c], Ahist[c]);
                for (int id = 0; id < g.N; ++id) E[c][id] = -((c == 0) ? gpx[id] : (c == 1) ? gpy[id] : 0.0) - At[id];
            }
            const Field divE = p2_div(E, g);
            Field gauss(g.N);
            for (int id = 0; id < g.N; ++id) gauss[id] = divE[id] - rho_cur[id] / p->eps0;
            const Field divJ = p2_div(J_cur, g);
            const Field rhot = dt1(rho_cur, rhist);
            Field cont(g.N);
            for (int id = 0; id < g.N; ++id) cont[id] = rhot[id] + divJ[id];
            const Field dAzx = p2_deriv(A_new[2], g, 0), dAzy = p2_deriv(A_new[2], g, 1), dAyx = p2_deriv(A_new[1], g, 0),
                        dAxy = p2_deriv(A_new[0], g, 1);
            VecField B = zero_vecfield(g.N);
            for (int id = 0; id < g.N; ++id) { B[0][id] = dAzy[id]; B[1][id] = -dAzx[id]; B[2][id] = dAyx[id] - dAxy[id]; }
            auto isq = [&](const VecField& V) {
                long double ss = 0.0;
                for (int id = 0; id < g.N; ++id)
                    ss += static_cast<long double>(V[0][id]) * V[0][id] + static_cast<long double>(V[1][id]) * V[1][id] +
                          static_cast<long double>(V[2][id]) * V[2][id];
                return static_cast<Real>(ss * (g.dx[0] * g.dx[1]));
            };
            const Real fe = 0.5 * p->eps0 * isq(E) + 0.5 / p->mu0 * isq(B);
            long double kin = 0.0;
            Real maxs = 0.0;
            for (int64_t i = 0; i < np; ++i) {
                const Real beta2 = dot3(s.v[i], s.v[i]) / (p->c * p->c);
                const Real gam = 1.0 / std::sqrt(std::max(1.0e-14, 1.0 - beta2));
                kin += (gam - 1.0) * s.m[i] * p->c * p->c;
                maxs = std::max(maxs, norm3(s.v[i]));
            }
            double* r = rows + static_cast<std::size_t>(n) * 8;
            r[0] = rms_field(gauge); r[1] = rms_field(gauss); r[2] = rms_field(cont); r[3] = rms_field(B[2]);
            r[4] = fe; r[5] = static_cast<Real>(kin); r[6] = fe + static_cast<Real>(kin); r[7] = maxs;
        };
        const std::vector<Real> a1{1.0, -1.0};
        diag(0, s.phi.front(), {s.A[0].front(), s.A[1].front(), s.A[2].front()}, s.rho.front(), s.J, a1, s.phi, s.A, s.rho);
        for (int n = 1; n <= n_steps; ++n) {
            // the step mutates histories; run it on a copy of the history to
            // evaluate diagnostics against the pre-push histories as P2:1175-1182
            const auto phist = s.phi; const auto Ahist = s.A; const auto rhist = s.rho;
            p2_step(C, s);
            diag(n, s.phi.front(), {s.A[0].front(), s.A[1].front(), s.A[2].front()}, s.rho.front(), s.J,
                 bdf_a(p->bdf_order), phist, Ahist, rhist);
        }
        std::vector<Vec3> Xo(np);
        for (int64_t i = 0; i < np; ++i) Xo[i] = {s.x[i], s.y[i], 0.0};
        store3(Xo, st->x); store3(s.v, st->v); store3(s.v_prev, st->v_prev); store3(s.P, st->P);
    });
}

// libstdc++-compatible mt19937 + uniform_real_distribution<double> draws
// (P2:633-634 use them for the Weibel initial velocities).  generate_canonical
// with 53 bits takes two 32-bit outputs: (x0 + x1 * 2^32) / 2^64.
int orc_mt19937_uniform(uint32_t seed, double lo, double hi, int64_t n, double* out) {
    uint32_t mt[624];
    mt[0] = seed;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
    int idx = 624;
    auto next = [&]() -> uint32_t {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= (y >> 11); y ^= (y << 7) & 0x9d2c5680u; y ^= (y << 15) & 0xefc60000u; y ^= (y >> 18);
        return y;
    };
    for (int64_t i = 0; i < n; ++i) {
        // libstdc++ accumulates in the result type (double)
        const double r = 4294967296.0;
        double sum = 0.0, tmp = 1.0;
        for (int k = 0; k < 2; ++k) { sum += static_cast<double>(next()) * tmp; tmp *= r; }
        double c = sum / tmp;
        if (c >= 1.0) c = std::nextafter(1.0, 0.0);
        out[i] = c * (hi - lo) + lo;
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned-guess network (no reference counterpart): FP64 restatement of the
// 12 -> 64 -> 64 -> 3 SiLU MLP of the library (parameter layout of
// UPIC_MLP_NPARAM in upic.h) and of its training loss and gradient.
//   x = (f - in_mean) * in_scale; h1 = silu(W1 x + b1); h2 = silu(W2 h1 + b2);
//   y = W3 h2 + b3; output = y * out_scale.
//   loss = sum_b sum_c (y_bc - t_bc / out_scale_c)^2 / (3 n)
// ---------------------------------------------------------------------------
namespace {
constexpr int kIn = 12, kH = 64, kOut = 3;
constexpr int oMean = 0, oScale = 12, oOut = 24, oW1 = 27, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, oN = oB3 + kOut;
inline double sig(double z) { return 1.0 / (1.0 + std::exp(-z)); }
struct MlpPass {
    double x[kIn], z1[kH], h1[kH], z2[kH], h2[kH], y[kOut];
};
void mlp_fwd(const float* P, const float* f, MlpPass& s) {
    for (int k = 0; k < kIn; ++k) s.x[k] = ((double)f[k] - P[oMean + k]) * (double)P[oScale + k];
    for (int i = 0; i < kH; ++i) {
        double a = P[oB1 + i];
        for (int k = 0; k < kIn; ++k) a += (double)P[oW1 + i * kIn + k] * s.x[k];
        s.z1[i] = a;
        s.h1[i] = a * sig(a);
    }
    for (int i = 0; i < kH; ++i) {
        double a = P[oB2 + i];
        for (int k = 0; k < kH; ++k) a += (double)P[oW2 + i * kH + k] * s.h1[k];
        s.z2[i] = a;
        s.h2[i] = a * sig(a);
    }
    for (int c = 0; c < kOut; ++c) {
        double a = P[oB3 + c];
        for (int k = 0; k < kH; ++k) a += (double)P[oW3 + c * kH + k] * s.h2[k];
        s.y[c] = a;
    }
}
}  // namespace

extern "C" {
int orc_mlp_forward(const float* P, const float* X, int64_t n, double* Y) {
    MlpPass s;
    for (int64_t i = 0; i < n; ++i) {
        mlp_fwd(P, X + 12 * i, s);
        for (int c = 0; c < kOut; ++c) Y[3 * i + c] = s.y[c] * (double)P[oOut + c];
    }
    return 0;
}

int orc_mlp_loss_grad(const float* P, const float* rows, int64_t n, double* G, double* loss) {
    for (int i = 0; i < oN; ++i) G[i] = 0.0;
    double L = 0.0;
    MlpPass s;
    for (int64_t b = 0; b < n; ++b) {
        const float* r = rows + 15 * b;
        mlp_fwd(P, r, s);
        double dy[kOut];
        for (int c = 0; c < kOut; ++c) {
            const double e = s.y[c] - (double)r[12 + c] / (double)P[oOut + c];
            L += e * e;
            dy[c] = 2.0 * e / (3.0 * (double)n);
            G[oB3 + c] += dy[c];
            for (int k = 0; k < kH; ++k) G[oW3 + c * kH + k] += dy[c] * s.h2[k];
        }
        double dz2[kH];
        for (int k = 0; k < kH; ++k) {
            double dh = 0.0;
            for (int c = 0; c < kOut; ++c) dh += (double)P[oW3 + c * kH + k] * dy[c];
            const double sg = sig(s.z2[k]);
            dz2[k] = dh * sg * (1.0 + s.z2[k] * (1.0 - sg));
        }
        double dh1[kH] = {};
        for (int i = 0; i < kH; ++i) {
            G[oB2 + i] += dz2[i];
            for (int k = 0; k < kH; ++k) {
                G[oW2 + i * kH + k] += dz2[i] * s.h1[k];
                dh1[k] += (double)P[oW2 + i * kH + k] * dz2[i];
            }
        }
        for (int i = 0; i < kH; ++i) {
            const double sg = sig(s.z1[i]);
            const double d = dh1[i] * sg * (1.0 + s.z1[i] * (1.0 - sg));
            G[oB1 + i] += d;
            for (int k = 0; k < kIn; ++k) G[oW1 + i * kIn + k] += d * s.x[k];
        }
    }
    *loss = L / (3.0 * (double)n);
    return 0;
}
}  // extern "C"
