"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the
reference harnesses (oracle/_ref/libref_p*.so, present only where the
reference tree was available to build them).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from unstaggered_pic_amd.abi import HostState, Params, StepInfo, Diagnostics, PathStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
ip64 = C.POINTER(C.c_int64)
PP = C.POINTER(Params)
HSP = C.POINTER(HostState)


def ptr(a, t=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(t))


class State:
    """Numpy-backed particle/mesh state in the upic.h layout."""

    FIELDS_P = ("x", "v", "v_prev", "P", "vbar_last")

    def __init__(self, np_, N):
        self.np = np_
        self.N = N
        for f in self.FIELDS_P:
            setattr(self, f, np.zeros(3 * np_))
        self.q = np.zeros(np_)
        self.m = np.zeros(np_)
        for f in ("phi", "psi", "rho"):
            setattr(self, f, np.zeros(N))
        for f in ("A", "U", "J"):
            setattr(self, f, np.zeros(3 * N))

    def copy(self):
        s = State(self.np, self.N)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(s, k, v.copy())
        return s

    def host(self):
        h = HostState()
        h.np = self.np
        for f in ("x", "v", "v_prev", "P", "vbar_last", "q", "m", "phi", "psi", "rho", "A", "U", "J"):
            setattr(h, f, ptr(getattr(self, f)))
        return h


def _load(path):
    return C.CDLL(path) if os.path.exists(path) else None


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = self.L = C.CDLL(ORACLE_SO)
        L.orc_last_error.restype = C.c_char_p
        for name in ("orc_quadrature", "orc_bspline", "orc_breakpoints", "orc_fft", "orc_deposit_current_orbit",
                     "orc_gather_orbit", "orc_orbit_grad_A", "orc_gather_point", "orc_deposit_scalar_point",
                     "orc_spectral_gradient", "orc_spectral_divergence", "orc_spectral_curl", "orc_continuity",
                     "orc_cn_potentials", "orc_project_nyquist", "orc_poisson", "orc_p3_local_solve", "orc_step",
                     "orc_run", "orc_diagnostics", "orc_p1_step_diagnostics", "orc_p1_init", "orc_p2_run",
                     "orc_mt19937_uniform", "orc_init_fields", "orc_mlp_forward", "orc_mlp_loss_grad"):
            getattr(L, name).restype = C.c_int

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {self.L.orc_last_error().decode()}")

    def quadrature(self, n):
        s, w = np.zeros(n), np.zeros(n)
        self.check(self.L.orc_quadrature(n, ptr(s), ptr(w)))
        return s, w

    def bspline(self, code, degree, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        w, dw = np.zeros_like(r), np.zeros_like(r)
        self.check(self.L.orc_bspline(code, degree, len(r), ptr(r), ptr(w), ptr(dw)))
        return w, dw

    def breakpoints(self, p, x0, dx, max_br=16):
        n = len(x0) // 3
        nseg = np.zeros(n, np.int32)
        br = np.zeros(n * max_br)
        self.check(self.L.orc_breakpoints(C.byref(p), ptr(x0), ptr(dx), C.c_int64(n), ptr(nseg, C.c_int32), ptr(br), max_br))
        return nseg, br.reshape(n, max_br)

    def fft(self, p, z, inverse=False):
        d = np.ascontiguousarray(np.stack([z.real, z.imag], -1).ravel())
        self.check(self.L.orc_fft(C.byref(p), ptr(d), int(inverse)))
        d = d.reshape(-1, 2)
        return d[:, 0] + 1j * d[:, 1]

    def deposit_current_orbit(self, p, x0, dx, vbar, q):
        n = len(q)
        J = np.zeros(3 * p.nx * p.ny * p.nz)
        st = PathStats()
        self.check(self.L.orc_deposit_current_orbit(C.byref(p), ptr(x0), ptr(dx), ptr(vbar), ptr(q), C.c_int64(n), ptr(J), C.byref(st)))
        return J, st

    def gather_orbit(self, p, x0, dx, f3):
        n = len(x0) // 3
        out = np.zeros(3 * n)
        self.check(self.L.orc_gather_orbit(C.byref(p), ptr(x0), ptr(dx), ptr(f3), C.c_int64(n), ptr(out)))
        return out

    def orbit_grad_A(self, p, x0, dx, An, Anp1, th0=0.0, th1=1.0):
        n = len(x0) // 3
        D = np.zeros(9 * n)
        self.check(self.L.orc_orbit_grad_A(C.byref(p), ptr(x0), ptr(dx), ptr(An), ptr(Anp1), C.c_double(th0), C.c_double(th1), C.c_int64(n), ptr(D)))
        return D

    def gather_point(self, p, pos, f3):
        n = len(pos) // 3
        out = np.zeros(3 * n)
        self.check(self.L.orc_gather_point(C.byref(p), ptr(pos), ptr(f3), C.c_int64(n), ptr(out)))
        return out

    def deposit_scalar_point(self, p, pos, values, divide=True):
        mesh = np.zeros(p.nx * p.ny * p.nz)
        self.check(self.L.orc_deposit_scalar_point(C.byref(p), ptr(pos), ptr(values), C.c_int64(len(values)), int(divide), ptr(mesh)))
        return mesh

    def _N(self, p):
        return p.nx * p.ny * p.nz

    def spectral_gradient(self, p, f):
        out = np.zeros(3 * self._N(p))
        self.check(self.L.orc_spectral_gradient(C.byref(p), ptr(f), ptr(out)))
        return out

    def spectral_divergence(self, p, V):
        out = np.zeros(self._N(p))
        self.check(self.L.orc_spectral_divergence(C.byref(p), ptr(V), ptr(out)))
        return out

    def spectral_curl(self, p, A):
        out = np.zeros(3 * self._N(p))
        self.check(self.L.orc_spectral_curl(C.byref(p), ptr(A), ptr(out)))
        return out

    def continuity(self, p, rho, J, dt):
        out = np.zeros(self._N(p))
        self.check(self.L.orc_continuity(C.byref(p), ptr(rho), ptr(J), C.c_double(dt), ptr(out)))
        return out

    def cn_potentials(self, p, phi, psi, A, U, rho_mid, J, dt):
        N = self._N(p)
        o = [np.zeros(N), np.zeros(N), np.zeros(3 * N), np.zeros(3 * N)]
        self.check(self.L.orc_cn_potentials(C.byref(p), ptr(phi), ptr(psi), ptr(A), ptr(U), ptr(rho_mid), ptr(J),
                                            C.c_double(dt), *[ptr(a) for a in o]))
        return o

    def project_nyquist(self, p, f, ncomp=1):
        f = f.copy()
        self.check(self.L.orc_project_nyquist(C.byref(p), ptr(f), ncomp))
        return f

    def poisson(self, p, rho):
        out = np.zeros(self._N(p))
        self.check(self.L.orc_poisson(C.byref(p), ptr(rho), ptr(out)))
        return out

    def bdf_wave(self, p, src, hist, order):
        out = np.zeros(len(src))
        self.check(self.L.orc_bdf_wave(C.byref(p), ptr(src), ptr(hist), int(order), ptr(out)))
        return out

    def p3_local_solve(self, p, guess, x0, P0, q, m, h, th0, th1, An, Aend, gpm):
        n = len(q)
        vb, Pe, ve = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        it, rs = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.check(self.L.orc_p3_local_solve(C.byref(p), ptr(guess), ptr(x0), ptr(P0), ptr(q), ptr(m), C.c_double(h),
                                             C.c_double(th0), C.c_double(th1), ptr(An), ptr(Aend), ptr(gpm), C.c_int64(n),
                                             ptr(vb), ptr(Pe), ptr(ve), ptr(it, C.c_int32), ptr(rs, C.c_int32)))
        return vb, Pe, ve, it, rs

    def step(self, p, st):
        out = State(st.np, st.N)
        hin, hout = st.host(), out.host()
        info = StepInfo()
        self.check(self.L.orc_step(C.byref(p), C.byref(hin), C.byref(hout), C.byref(info)))
        return out, info

    def run(self, p, st, n_steps, mode=1):
        s = st.copy()
        rows = np.zeros((n_steps + 1, 19))
        h = s.host()
        self.check(self.L.orc_run(C.byref(p), C.byref(h), n_steps, mode, ptr(rows)))
        return s, rows

    def diagnostics(self, p, st, mode=1):
        d = Diagnostics()
        h = st.host()
        self.check(self.L.orc_diagnostics(C.byref(p), C.byref(h), mode, C.byref(d)))
        return d

    def mlp_forward(self, P, X):
        P = np.ascontiguousarray(P, np.float32)
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 12)
        Y = np.zeros((X.shape[0], 3))
        self.check(self.L.orc_mlp_forward(ptr(P, C.c_float), ptr(X, C.c_float), C.c_int64(X.shape[0]), ptr(Y)))
        return Y

    def mlp_loss_grad(self, P, rows):
        P = np.ascontiguousarray(P, np.float32)
        rows = np.ascontiguousarray(rows, np.float32).reshape(-1, 15)
        G = np.zeros(P.size)
        loss = np.zeros(1)
        self.check(self.L.orc_mlp_loss_grad(ptr(P, C.c_float), ptr(rows, C.c_float), C.c_int64(rows.shape[0]),
                                            ptr(G), ptr(loss)))
        return G, float(loss[0])

    def p1_step_diagnostics(self, p, s0, s1):
        out = np.zeros(17)
        h0, h1 = s0.host(), s1.host()
        self.check(self.L.orc_p1_step_diagnostics(C.byref(p), C.byref(h0), C.byref(h1), ptr(out)))
        return out

    def p1_init(self, p, kind, ppc, alpha=0.01, vth=1.0, v0=0.3, pert=0.02, mode=1, n0=1.0):
        h = HostState()
        args = (C.byref(p), kind, ppc, C.c_double(alpha), C.c_double(vth), C.c_double(v0), C.c_double(pert), mode, C.c_double(n0))
        self.check(self.L.orc_p1_init(*args, C.byref(h)))
        st = State(h.np, self._N(p))
        h = st.host()
        self.check(self.L.orc_p1_init(*args, C.byref(h)))
        return st

    def init_fields(self, p, x, v, q, m, n0=1.0):
        st = State(len(q), self._N(p))
        st.x[:] = x; st.v[:] = v; st.q[:] = q; st.m[:] = m
        h = st.host()
        self.check(self.L.orc_init_fields(C.byref(p), C.byref(h), C.c_double(n0)))
        return st

    def p2_run(self, p, st, fixed_ion_rho, n_steps):
        s = st.copy()
        rows = np.zeros((n_steps + 1, 8))
        h = s.host()
        self.check(self.L.orc_p2_run(C.byref(p), C.byref(h), ptr(fixed_ion_rho), n_steps, ptr(rows)))
        return s, rows

    def mt19937_uniform(self, seed, lo, hi, n):
        out = np.zeros(n)
        self.check(self.L.orc_mt19937_uniform(C.c_uint32(seed), C.c_double(lo), C.c_double(hi), C.c_int64(n), ptr(out)))
        return out


class Ref:
    """The reference's own functions, compiled from its sources (oracle/_ref)."""

    def __init__(self):
        self.p1 = _load(os.path.join(REF_DIR, "libref_p1.so"))
        self.p2 = _load(os.path.join(REF_DIR, "libref_p2.so"))
        self.p3 = _load(os.path.join(REF_DIR, "libref_p3.so"))
        for lib, pre in ((self.p1, "ref_p1"), (self.p2, "ref_p2"), (self.p3, "ref_p3")):
            if lib is not None:
                getattr(lib, pre + "_last_error").restype = C.c_char_p

    @property
    def available(self):
        return self.p1 is not None and self.p2 is not None and self.p3 is not None

    @staticmethod
    def check(lib, pre, rc):
        if rc != 0:
            raise RuntimeError(f"reference error: {getattr(lib, pre + '_last_error')().decode()}")
