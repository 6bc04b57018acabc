"""ctypes binding of libupic.so (the HIP library) and a thin host wrapper
around the device-resident simulation handle.

The product path runs only on the HIP library: importing this module without
the built library, or using it without a HIP device, raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import (Diagnostics, HostState, LoadSpec, P3Case, Params, PathStats, StepDiagnostics, StepInfo,
                  STATUS_NAMES)

HERE = os.path.dirname(os.path.abspath(__file__))
# UPIC_LIBRARY selects a tuning variant built with build.py -o
LIB_PATH = os.environ.get("UPIC_LIBRARY") or os.path.join(HERE, "libupic.so")

dp = C.POINTER(C.c_double)
vp = C.c_void_p
MLP_NPARAM = 5214  # UPIC_MLP_NPARAM
GUESS_PLAIN, GUESS_PREDICTOR, GUESS_LEARNED = 0, 1, 2


class UpicError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None


def lib():
    """Load libupic.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH} (run python -m unstaggered_pic_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.upic_last_error.restype = C.c_char_p
        L.upic_last_error.argtypes = [vp]
        sigs = {
            "upic_ctx_create": [C.POINTER(Params), C.POINTER(vp)],
            "upic_ctx_destroy": [vp], "upic_ctx_sync": [vp],
            "upic_quadrature": [vp, dp, dp],
            "upic_deposit_current_orbit": [vp, vp, vp, vp, vp, C.c_int64, vp, C.POINTER(PathStats)],
            "upic_gather_orbit": [vp, vp, vp, vp, C.c_int64, vp],
            "upic_orbit_grad_A": [vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int64, vp],
            "upic_gather_point": [vp, vp, vp, C.c_int64, vp],
            "upic_bdf_wave": [vp, vp, vp, C.c_int, vp],
            "upic_local_substep": [vp, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, vp, vp, vp, C.c_int64,
                                   vp, vp, vp, vp, vp],
            "upic_local_substep_noalloc": [vp, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, vp, vp, vp,
                                           C.c_int64, vp, vp, vp, vp, vp],
            "upic_deposit_scalar_point": [vp, vp, vp, C.c_int64, C.c_int, vp],
            "upic_fft": [vp, vp, C.c_int],
            "upic_spectral_gradient": [vp, vp, vp], "upic_spectral_divergence": [vp, vp, vp],
            "upic_spectral_curl": [vp, vp, vp],
            "upic_continuity": [vp, vp, vp, C.c_double, vp],
            "upic_cn_potentials": [vp, vp, vp, vp, vp, vp, vp, C.c_double, vp, vp, vp, vp],
            "upic_project_nyquist": [vp, vp, C.c_int], "upic_poisson": [vp, vp, vp],
            "upic_sim_create": [vp, C.c_int64, C.POINTER(vp)], "upic_sim_destroy": [vp],
            "upic_sim_upload": [vp, C.POINTER(HostState)], "upic_sim_download": [vp, C.POINTER(HostState)],
            "upic_step_p1": [vp, C.POINTER(StepInfo)], "upic_step_p2": [vp, C.POINTER(StepInfo)],
            "upic_step_p3": [vp, C.POINTER(StepInfo)],
            "upic_sim_diagnostics": [vp, C.c_int, C.POINTER(Diagnostics)],
            "upic_sim_device_state": [vp, C.POINTER(HostState)],
            "upic_sim_failure_reasons": [vp, C.POINTER(C.c_uint8)],
            "upic_sim_init_fields": [vp, C.c_double],
            "upic_sim_init_fields_AU": [vp, C.c_double, vp, vp],
            "upic_sim_keep_step_start": [vp, C.c_int],
            "upic_sim_step_diagnostics": [vp, C.POINTER(StepDiagnostics)],
            "upic_sim_substep_paths": [vp, C.POINTER(C.c_int64), vp, vp, vp],
            "upic_p3_initial_state": [C.POINTER(Params), C.POINTER(P3Case), C.POINTER(HostState), dp],
            "upic_ctx_set_timing": [vp, C.c_int],
            "upic_device_count": [C.POINTER(C.c_int)], "upic_device_alloc": [C.c_size_t, C.POINTER(vp)],
            "upic_device_free": [vp], "upic_memcpy_htod": [vp, vp, C.c_size_t],
            "upic_memcpy_dtoh": [vp, vp, C.c_size_t], "upic_memcpy_dtod": [vp, vp, C.c_size_t],
            "upic_ctx_timing": [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)],
            "upic_load_particles": [C.POINTER(Params), C.POINTER(LoadSpec), C.POINTER(HostState)],
            "upic_sim_collect_guess_data": [vp, C.c_int64, C.c_int64],
            "upic_sim_guess_data": [vp, vp, C.c_int64, C.POINTER(C.c_int64)],
            "upic_sim_train_guess": [vp, C.c_int32, C.c_int32, C.c_double, C.c_uint64, dp],
            "upic_sim_set_guess_mode": [vp, C.c_int32],
            "upic_sim_set_deposit_variant": [vp, C.c_int32],
            "upic_ctx_set_graphs": [vp, C.c_int],
            "upic_ctx_set_deterministic": [vp, C.c_int],
            "upic_ensemble_step": [vp, C.c_int32, C.c_int32, vp, vp],
            "upic_sim_deposit_variant": [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
            "upic_sim_guess_params": [vp, vp], "upic_sim_set_guess_params": [vp, vp],
            "upic_mlp_forward": [vp, vp, vp, C.c_int64, vp],
            "upic_mlp_loss_grad": [vp, vp, vp, C.c_int64, vp, dp],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.upic_p3_case_defaults.argtypes = [C.c_int32, C.POINTER(P3Case)]
        L.upic_p3_case_defaults.restype = None
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/upic.h (checked against the library)."""
    hdr = os.path.join(os.path.dirname(HERE), "include", "upic.h")
    import re
    names = []
    for line in open(hdr):
        m = re.match(r"^(?:upic_status|const char\*|void)\s+(upic_\w+)\(", line)
        if m:
            names.append(m.group(1))
    return names


def device_count():
    n = C.c_int()
    lib().upic_device_count(C.byref(n))
    return n.value


class DeviceArray:
    """A float64 (or float32) buffer in HBM, allocated through the library."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.p = vp()
        rc = lib().upic_device_alloc(self.dtype.itemsize * max(self.n, 1), C.byref(self.p))
        if rc != 0:
            raise UpicError(rc, "device allocation failed")

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, dtype)
        if a.size:
            lib().upic_memcpy_htod(d.p, a.ctypes.data_as(vp), a.nbytes)
        return d

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            rc = lib().upic_memcpy_dtoh(out.ctypes.data_as(vp), self.p, self.dtype.itemsize * self.n)
            if rc != 0:
                raise UpicError(rc, "device copy failed")
        return out

    def numel(self):
        return self.n

    def __del__(self):
        try:
            if self.p:
                lib().upic_device_free(self.p)
                self.p = vp()
        except Exception:
            pass


class Context:
    """Grid, quadrature, stream (upic_ctx)."""

    def __init__(self, params: Params):
        if device_count() < 1:
            raise RuntimeError("no HIP device: the unstaggered_pic_amd hot path needs an MI355X")
        self.params = params
        self.h = vp()
        self._check(lib().upic_ctx_create(C.byref(params), C.byref(self.h)), None)

    def _check(self, rc, h="self"):
        if rc != 0:
            hh = self.h if h == "self" else None
            raise UpicError(rc, lib().upic_last_error(hh).decode())

    def sync(self):
        self._check(lib().upic_ctx_sync(self.h))

    def set_timing(self, on=True):
        self._check(lib().upic_ctx_set_timing(self.h, int(on)))

    def set_graphs(self, on=True):
        """hipGraph replay: a uniform P3 step's whole outer solve as one graph
        (device-controlled, gated map evaluations), else the field solves.
        on: False/0 off, True/1 on, n >= 2 on with n unrolled evaluations."""
        self._check(lib().upic_ctx_set_graphs(self.h, int(on)))

    def set_deterministic(self, on=True):
        """Bitwise-reproducible FP64 P3 steps (fixed-point deposit, stable
        cell order, one-workgroup reductions)."""
        self._check(lib().upic_ctx_set_deterministic(self.h, int(on)))

    def timing(self, name):
        n, ms = C.c_int64(), C.c_double()
        self._check(lib().upic_ctx_timing(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def quadrature(self):
        n = self.params.orbit_quad_order
        s, w = np.zeros(n), np.zeros(n)
        self._check(lib().upic_quadrature(self.h, s.ctypes.data_as(dp), w.ctypes.data_as(dp)))
        return s, w

    def mlp_forward(self, params, X):
        """MFMA inference of the learned-guess network on host arrays:
        params (5214 floats), X (n x 12) -> Y (n x 3), FP32."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 12)
        n = X.shape[0]
        dP, dX = DeviceArray.from_host(params, np.float32), DeviceArray.from_host(X, np.float32)
        dY = DeviceArray(3 * n, np.float32)
        self._check(lib().upic_mlp_forward(self.h, dP.p, dX.p, n, dY.p))
        return dY.numpy().reshape(n, 3)

    def mlp_loss_grad(self, params, rows):
        """FP32 training loss and gradient on rows (n x 15: features, target)."""
        rows = np.ascontiguousarray(rows, np.float32).reshape(-1, 15)
        n = rows.shape[0]
        dP, dR = DeviceArray.from_host(params, np.float32), DeviceArray.from_host(rows, np.float32)
        dG = DeviceArray(MLP_NPARAM, np.float32)
        loss = C.c_double()
        self._check(lib().upic_mlp_loss_grad(self.h, dP.p, dR.p, n, dG.p, C.byref(loss)))
        return dG.numpy(), loss.value

    def close(self):
        if self.h:
            lib().upic_ctx_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- group A on DeviceArrays --------------------------------------
    @staticmethod
    def dev(a):
        return DeviceArray.from_host(a)

    @staticmethod
    def ptr(t):
        return t.p if t is not None else None

    def N(self):
        p = self.params
        return p.nx * p.ny * p.nz

    def empty(self, n):
        return DeviceArray(n)

    def deposit_current_orbit(self, x0, dx, vbar, q):
        n = q.numel()
        J = self.empty(3 * self.N())
        st = PathStats()
        self._check(lib().upic_deposit_current_orbit(self.h, self.ptr(x0), self.ptr(dx), self.ptr(vbar), self.ptr(q),
                                                     n, self.ptr(J), C.byref(st)))
        return J, st

    def gather_orbit(self, x0, dx, f3):
        n = x0.numel() // 3
        out = self.empty(3 * n)
        self._check(lib().upic_gather_orbit(self.h, self.ptr(x0), self.ptr(dx), self.ptr(f3), n, self.ptr(out)))
        self.sync()
        return out

    def orbit_grad_A(self, x0, dx, An, Anp1, th0=0.0, th1=1.0):
        n = x0.numel() // 3
        out = self.empty(9 * n)
        self._check(lib().upic_orbit_grad_A(self.h, self.ptr(x0), self.ptr(dx), self.ptr(An), self.ptr(Anp1), th0, th1,
                                            n, self.ptr(out)))
        self.sync()
        return out

    def bdf_wave(self, src, hist, order):
        out = self.empty(self.N())
        self._check(lib().upic_bdf_wave(self.h, self.ptr(src), self.ptr(hist), int(order), self.ptr(out)))
        self.sync()
        return out

    def local_substep(self, guess, x0, P0, q, m, h, th0, th1, An, Aend, gpm, noalloc=False):
        """solve_local_particle_substep on device arrays; returns host
        arrays vbar, P_end, v_end, iterations, reasons.  noalloc selects
        upic_local_substep_noalloc (static node-record scratch)."""
        n = q.numel()
        vb, Pe, ve = self.empty(3 * n), self.empty(3 * n), self.empty(3 * n)
        it, rs = DeviceArray(n, np.int32), DeviceArray(n, np.uint8)
        fn = lib().upic_local_substep_noalloc if noalloc else lib().upic_local_substep
        self._check(fn(self.h, self.ptr(guess), self.ptr(x0), self.ptr(P0), self.ptr(q),
                                             self.ptr(m), h, th0, th1, self.ptr(An), self.ptr(Aend), self.ptr(gpm), n,
                                             self.ptr(vb), self.ptr(Pe), self.ptr(ve), self.ptr(it), self.ptr(rs)))
        self.sync()
        return vb.numpy(), Pe.numpy(), ve.numpy(), it.numpy(), rs.numpy()

    def gather_point(self, pos, f3):
        n = pos.numel() // 3
        out = self.empty(3 * n)
        self._check(lib().upic_gather_point(self.h, self.ptr(pos), self.ptr(f3), n, self.ptr(out)))
        self.sync()
        return out

    def deposit_scalar_point(self, pos, values, divide=True):
        out = self.empty(self.N())
        self._check(lib().upic_deposit_scalar_point(self.h, self.ptr(pos), self.ptr(values), values.numel(), int(divide),
                                                    self.ptr(out)))
        self.sync()
        return out

    def fft(self, z, inverse=False):
        self._check(lib().upic_fft(self.h, self.ptr(z), int(inverse)))
        self.sync()
        return z

    def _unary(self, name, inp, nout):
        out = self.empty(nout)
        self._check(getattr(lib(), name)(self.h, self.ptr(inp), self.ptr(out)))
        self.sync()
        return out

    def spectral_gradient(self, f):
        return self._unary("upic_spectral_gradient", f, 3 * self.N())

    def spectral_divergence(self, V):
        return self._unary("upic_spectral_divergence", V, self.N())

    def spectral_curl(self, A):
        return self._unary("upic_spectral_curl", A, 3 * self.N())

    def poisson(self, rho):
        return self._unary("upic_poisson", rho, self.N())

    def continuity(self, rho, J, dt):
        out = self.empty(self.N())
        self._check(lib().upic_continuity(self.h, self.ptr(rho), self.ptr(J), dt, self.ptr(out)))
        self.sync()
        return out

    def cn_potentials(self, phi, psi, A, U, rho_mid, J, dt):
        N = self.N()
        o = [self.empty(N), self.empty(N), self.empty(3 * N), self.empty(3 * N)]
        self._check(lib().upic_cn_potentials(self.h, *[self.ptr(t) for t in (phi, psi, A, U, rho_mid, J)], dt,
                                             *[self.ptr(t) for t in o]))
        self.sync()
        return o

    def project_nyquist(self, f, ncomp=1):
        f = DeviceArray.from_host(f.numpy())
        self._check(lib().upic_project_nyquist(self.h, self.ptr(f), ncomp))
        self.sync()
        return f


class Simulation:
    """Device-resident particle and mesh state with the per-step drivers."""

    FIELDS_P = ("x", "v", "v_prev", "P", "vbar_last")

    def __init__(self, ctx: Context, np_):
        self.ctx = ctx
        self.np = np_
        self.h = vp()
        ctx._check(lib().upic_sim_create(ctx.h, np_, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().upic_sim_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _host(self, st):
        h = HostState()
        h.np = self.np
        for f in ("x", "v", "v_prev", "P", "vbar_last", "q", "m", "phi", "psi", "rho", "A", "U", "J",
                  "phi_hist", "A_hist", "rho_hist", "fixed_ion_rho"):
            a = getattr(st, f, None)
            setattr(h, f, a.ctypes.data_as(dp) if a is not None else None)
        return h

    def upload(self, st):
        h = self._host(st)
        self.ctx._check(lib().upic_sim_upload(self.h, C.byref(h)))

    def download(self, st):
        h = self._host(st)
        self.ctx._check(lib().upic_sim_download(self.h, C.byref(h)))
        return st

    def step(self):
        code = self.ctx.params.code
        info = StepInfo()
        f = {1: lib().upic_step_p1, 2: lib().upic_step_p2, 3: lib().upic_step_p3}[code]
        self.ctx._check(f(self.h, C.byref(info)))
        return info

    def diagnostics(self, mode=1):
        d = Diagnostics()
        self.ctx._check(lib().upic_sim_diagnostics(self.h, mode, C.byref(d)))
        return d

    def init_fields(self, n0=1.0, A=None, U=None):
        """Initial mesh state from the uploaded particles; A and U (3N
        float64 host arrays, P3 only) set the vector potential at t = 0."""
        if A is None and U is None:
            self.ctx._check(lib().upic_sim_init_fields(self.h, n0))
            return
        A = None if A is None else np.ascontiguousarray(A, np.float64)
        U = None if U is None else np.ascontiguousarray(U, np.float64)
        self.ctx._check(lib().upic_sim_init_fields_AU(self.h, n0, None if A is None else A.ctypes.data_as(vp),
                                                      None if U is None else U.ctypes.data_as(vp)))

    def keep_step_start(self, enable=True):
        """Keep each step's start state on the device (step diagnostics)."""
        self.ctx._check(lib().upic_sim_keep_step_start(self.h, 1 if enable else 0))

    def step_diagnostics(self):
        """compute_step_diagnostics of the last step (needs keep_step_start)."""
        d = StepDiagnostics()
        self.ctx._check(lib().upic_sim_step_diagnostics(self.h, C.byref(d)))
        return d

    def substep_paths(self):
        """The last step's substep paths when it was selectively refined:
        (offsets (np + 1), x0 (3 x n), vbar (3 x n)) in the original
        particle order, or None."""
        n = C.c_int64()
        self.ctx._check(lib().upic_sim_substep_paths(self.h, C.byref(n), None, None, None))
        if n.value == 0:
            return None
        off = np.zeros(self.np + 1, np.int64)
        x0, vb = np.zeros(3 * n.value), np.zeros(3 * n.value)
        self.ctx._check(lib().upic_sim_substep_paths(self.h, C.byref(n), off.ctypes.data_as(vp), x0.ctypes.data_as(vp),
                                                     vb.ctypes.data_as(vp)))
        return off, x0, vb

    def failure_reasons(self):
        r = np.zeros(self.np, np.uint8)
        self.ctx._check(lib().upic_sim_failure_reasons(self.h, r.ctypes.data_as(C.POINTER(C.c_uint8))))
        return r

    # --- learned initial guess (mixed precision) -------------------------
    def collect_guess_data(self, stride, capacity):
        self.ctx._check(lib().upic_sim_collect_guess_data(self.h, int(stride), int(capacity)))

    def guess_data(self, max_rows=None):
        n = C.c_int64()
        self.ctx._check(lib().upic_sim_guess_data(self.h, None, 0, C.byref(n)))
        k = n.value if max_rows is None else min(n.value, int(max_rows))
        rows = np.zeros((k, 15), np.float32)
        self.ctx._check(lib().upic_sim_guess_data(self.h, rows.ctypes.data_as(vp), k, C.byref(n)))
        return rows

    def train_guess(self, epochs=20, batch=4096, lr=1e-3, seed=1):
        losses = np.zeros(epochs)
        self.ctx._check(lib().upic_sim_train_guess(self.h, epochs, batch, lr, seed, losses.ctypes.data_as(dp)))
        return losses

    def set_guess_mode(self, mode):
        self.ctx._check(lib().upic_sim_set_guess_mode(self.h, int(mode)))

    def set_deposit_variant(self, variant):
        """Pin the mixed-precision deposit variant (0-7) or re-enable the
        autotuner (-1)."""
        self.ctx._check(lib().upic_sim_set_deposit_variant(self.h, int(variant)))

    def deposit_variant(self):
        """(sorted-order choice, unsorted-order choice); -1 while tuning."""
        a, b = C.c_int32(-1), C.c_int32(-1)
        self.ctx._check(lib().upic_sim_deposit_variant(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def guess_params(self):
        P = np.zeros(MLP_NPARAM, np.float32)
        self.ctx._check(lib().upic_sim_guess_params(self.h, P.ctypes.data_as(vp)))
        return P

    def set_guess_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        self.ctx._check(lib().upic_sim_set_guess_params(self.h, P.ctypes.data_as(vp)))


def load_particles(params: Params, spec, np_):
    """Seeded particle load (host C++): returns x, v, q, m numpy arrays."""
    x, v = np.zeros(3 * np_), np.zeros(3 * np_)
    q, m = np.zeros(np_), np.zeros(np_)
    h = HostState()
    h.np = np_
    h.x, h.v, h.q, h.m = (a.ctypes.data_as(dp) for a in (x, v, q, m))
    rc = lib().upic_load_particles(C.byref(params), C.byref(spec), C.byref(h))
    if rc != 0:
        raise UpicError(rc, "upic_load_particles rejected its arguments")
    return x, v, q, m


def p3_case(kind, **keys):
    """upic_p3_case with the reference's defaults, overridden by `keys`."""
    cs = P3Case()
    lib().upic_p3_case_defaults(int(kind), C.byref(cs))
    for k, v in keys.items():
        if not hasattr(cs, k):
            raise KeyError(k)
        setattr(cs, k, v)
    return cs


def p3_initial_state(params: Params, case):
    """The reference's P3 initial condition (host C++): returns a dict with
    x, v, q, m (particles), A, U (mesh, 3N) and the linear-theory growth
    rate."""
    h = HostState()
    rc = lib().upic_p3_initial_state(C.byref(params), C.byref(case), C.byref(h), None)
    if rc != 0:
        raise UpicError(rc, lib().upic_last_error(None).decode())
    n = h.np
    N = params.nx * params.ny * params.nz
    out = {"x": np.zeros(3 * n), "v": np.zeros(3 * n), "q": np.zeros(n), "m": np.zeros(n),
           "A": np.zeros(3 * N), "U": np.zeros(3 * N)}
    for k, a in out.items():
        setattr(h, k, a.ctypes.data_as(dp))
    rate = C.c_double()
    rc = lib().upic_p3_initial_state(C.byref(params), C.byref(case), C.byref(h), C.byref(rate))
    if rc != 0:
        raise UpicError(rc, lib().upic_last_error(None).decode())
    out["np"] = n
    out["growth_rate"] = rate.value
    return out


def ensemble_step(sims, threads=0):
    """One step of every simulation in `sims` (each on its own Context),
    advanced concurrently; returns the list of StepInfo.  Raises with the
    first failing member's error."""
    n = len(sims)
    arr = (C.c_void_p * n)(*[m.h for m in sims])
    infos = (StepInfo * n)()
    status = (C.c_int * n)()
    rc = lib().upic_ensemble_step(arr, n, int(threads), infos, status)
    if rc != 0:
        for m, st in zip(sims, status):
            if st != 0:
                m.ctx._check(st)
        raise RuntimeError(f"upic_ensemble_step failed ({rc})")
    return list(infos)
