"""GPU parity of the HIP path against the CPU oracle (oracle/liboracle.so).

Every comparison runs the same seeded inputs through the C ABI on the GPU and
through the oracle; tolerances are stated per test.  FP64 throughout.
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import (CODE_P1, CODE_P2, CODE_P3, LOAD_BIMAXWELL, LOAD_LANDAU, make_load, make_params)

pytestmark = pytest.mark.gpu

TWO_PI = 6.283185307179586


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def gpu():
    from unstaggered_pic_amd import lib
    return lib


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    s = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / s)


def host(t):
    return t.numpy()


def random_particles(rng, p, n, vscale):
    L = np.array([p.Lx, p.Ly, p.Lz])
    x = (rng.random((3, n)) - 0.5) * L[:, None]
    if p.nz == 1:
        x[2] = 0.0
    v = rng.normal(size=(3, n)) * vscale
    return x.ravel().copy(), v.ravel().copy()


GROUP_A_CASES = [(code, r, nz) for code in (CODE_P1, CODE_P3) for r in (1, 2, 3, 4) for nz in (1, 4)]


@pytest.mark.parametrize("code,order,nz", GROUP_A_CASES)
def test_group_a_orbit_ops(orc, code, order, nz):
    """Orbit deposit / gather / discrete gradient and point ops, random inputs
    with knot crossings (P1:858-934, P3:2689-2984).  Tolerance 1e-12 relative to
    the field maximum; path statistics exact."""
    lib = gpu()
    p = make_params(code=code, nx=16, ny=8, nz=nz, Lx=TWO_PI, Ly=4.0, Lz=2.0 if nz > 1 else 1.0, spline_order=order,
                    orbit_quad_order=order + 2, nyquist_projection=1 if code == CODE_P3 else 0)
    rng = np.random.default_rng(7 + order)
    n = 777
    x0, v = random_particles(rng, p, n, 1.0)
    dx = 0.3 * v
    q = rng.normal(size=n)
    N = p.nx * p.ny * p.nz
    F = rng.normal(size=3 * N)
    A1 = rng.normal(size=3 * N)
    ctx = lib.Context(p)
    d = ctx.dev
    J, st = ctx.deposit_current_orbit(d(x0), d(dx), d(v), d(q))
    Jo, sto = orc.deposit_current_orbit(p, x0, dx, v, q)
    assert rel(host(J), Jo) < 1e-12
    assert (st.crossing_particles, st.total_segments, st.max_segments) == \
        (sto.crossing_particles, sto.total_segments, sto.max_segments)
    g = ctx.gather_orbit(d(x0), d(dx), d(F))
    assert rel(host(g), orc.gather_orbit(p, x0, dx, F)) < 1e-12
    D = ctx.orbit_grad_A(d(x0), d(dx), d(F), d(A1))
    assert rel(host(D), orc.orbit_grad_A(p, x0, dx, F, A1)) < 1e-12
    D2 = ctx.orbit_grad_A(d(x0), d(dx), d(F), d(A1), 0.25, 0.5)
    assert rel(host(D2), orc.orbit_grad_A(p, x0, dx, F, A1, 0.25, 0.5)) < 1e-12
    gp = ctx.gather_point(d(x0), d(F))
    assert rel(host(gp), orc.gather_point(p, x0, F)) < 1e-12
    ds = ctx.deposit_scalar_point(d(x0), d(q))
    assert rel(host(ds), orc.deposit_scalar_point(p, x0, q)) < 1e-12
    ctx.close()


@pytest.mark.parametrize("code", [CODE_P1, CODE_P2, CODE_P3])
@pytest.mark.parametrize("shape", [(16, 8, 1), (8, 4, 4), (32, 1, 1)])
def test_fft_and_spectral_ops(orc, code, shape):
    """Radix-2 FFT against numpy and the oracle; spectral gradient,
    divergence, curl, continuity, CN potentials, Poisson and the Nyquist
    projector against the oracle (1e-12 relative)."""
    if code == CODE_P2 and shape[2] != 1:
        pytest.skip("P2 is 2-D")
    lib = gpu()
    nx, ny, nz = shape
    p = make_params(code=code, nx=nx, ny=ny, nz=nz, Lx=TWO_PI, Ly=3.0, Lz=2.0 if nz > 1 else 1.0,
                    nyquist_projection=1 if code == CODE_P3 else 0, dt=0.1, sigma1=1.0, sigma2=1.0)
    N = nx * ny * nz
    rng = np.random.default_rng(3)
    ctx = lib.Context(p)
    z = rng.normal(size=N) + 1j * rng.normal(size=N)
    zt = ctx.dev(np.stack([z.real, z.imag], -1).ravel())
    ctx.fft(zt)
    zz = host(zt).reshape(-1, 2)
    zg = zz[:, 0] + 1j * zz[:, 1]
    ref = np.fft.fftn(z.reshape(nx, ny, nz)).ravel()
    assert rel(zg, ref) < 1e-13
    ctx.fft(zt, inverse=True)
    zz = host(zt).reshape(-1, 2)
    assert rel(zz[:, 0] + 1j * zz[:, 1], z) < 1e-14
    f = rng.normal(size=N)
    V = rng.normal(size=3 * N)
    W = rng.normal(size=3 * N)
    d = ctx.dev
    if code != CODE_P2:
        assert rel(host(ctx.spectral_gradient(d(f))), orc.spectral_gradient(p, f)) < 1e-12
        assert rel(host(ctx.spectral_divergence(d(V))), orc.spectral_divergence(p, V)) < 1e-12
        assert rel(host(ctx.spectral_curl(d(V))), orc.spectral_curl(p, V)) < 1e-12
        assert rel(host(ctx.continuity(d(f), d(V), 0.1)), orc.continuity(p, f, V, 0.1)) < 1e-12
        assert rel(host(ctx.poisson(d(f))), orc.poisson(p, f)) < 1e-12
        g = rng.normal(size=N)
        h = rng.normal(size=N)
        out = ctx.cn_potentials(d(f), d(g), d(V), d(W), d(h), d(V * 0.3), 0.1)
        ref = orc.cn_potentials(p, f, g, V, W, h, V * 0.3, 0.1)
        for a, b in zip(out, ref):
            assert rel(host(a), b) < 1e-12
    if code == CODE_P3:
        assert rel(host(ctx.project_nyquist(d(V), 3)), orc.project_nyquist(p, V, 3)) < 1e-13
    ctx.close()


def p1_small(nx=16, ny=8, order=2, nq=4, ppc=8, steps=4):
    p = make_params(code=CODE_P1, nx=nx, ny=ny, nz=1, Lx=4 * np.pi, Ly=4 * np.pi, Lz=1.0, dt=0.1,
                    spline_order=order, orbit_quad_order=nq, picard_tol=1e-10, picard_max_iter=24)
    return p


def make_state_from_oracle(orc, p, x, v, q, m, n0=1.0):
    st = orc.init_fields(p, x, v, q, m, n0)
    return st


def run_gpu(p, st, steps, mode=1):
    lib = gpu()
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, st.np)
    s = st.copy()
    sim.upload(s)
    rows, infos = [sim.diagnostics(mode)], []
    for _ in range(steps):
        infos.append(sim.step())
        rows.append(sim.diagnostics(mode))
    out = State(st.np, st.N)
    sim.download(out)
    sim.close()
    ctx.close()
    return out, rows, infos


@pytest.mark.parametrize("order,nq,sort", [(2, 3, 1), (3, 4, 1), (1, 12, 0), (3, 16, 2)])
def test_p1_steps_vs_oracle(orc, order, nq, sort):
    """P1 global-Picard steps (P1:1509-1552) on a 2-D Landau-like load:
    total/field energy histories within 1e-12 relative, Picard iteration
    counts equal, particle state within 1e-9."""
    lib = gpu()
    p = p1_small(order=order, nq=nq)
    p.sort_interval = sort
    n = p.nx * p.ny * 8
    x, v, q, m = lib.load_particles(p, make_load(LOAD_LANDAU, 8, seed=5, alpha=0.05, vth=(0.3, 0.3, 0.3)), n)
    st = make_state_from_oracle(orc, p, x, v, q, m)
    steps = 4
    out, rows, infos = run_gpu(p, st, steps)
    so, orows = orc.run(p, st, steps)
    for k in range(steps + 1):
        assert abs(rows[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12
        assert abs(rows[k].field_energy - orows[k, 2]) / max(abs(orows[k, 2]), 1e-300) < 1e-10
    assert [i.iterations for i in infos] == [int(r) for r in orows[1:, 6]]
    for f in ("x", "v", "P"):
        assert rel(getattr(out, f), getattr(so, f)) < 1e-9


def test_p1_golden_landau_weak(orc):
    """The stored 300-step Landau-weak run of the reference
    (Project1/output/landau_weak_diagnostics.csv, deck from its config echo):
    the GPU path reproduces total and field energy to 1e-12 relative and the
    Picard iteration counts exactly."""
    import csv
    import os
    golden = os.path.join(os.path.dirname(__file__), "golden", "landau_weak_diagnostics.csv")
    g = list(csv.DictReader(open(golden)))
    p = make_params(code=CODE_P1, nx=32, ny=1, nz=1, Lx=12.566370614359172, Ly=1.0, Lz=1.0, dt=0.1, spline_order=1,
                    orbit_quad_order=12, picard_tol=1e-10, picard_max_iter=24)
    st = orc.p1_init(p, 0, 32, alpha=0.01, vth=1.0, mode=1)
    steps = 300
    _, rows, infos = run_gpu(p, st, steps)
    for k in range(steps + 1):
        te = float(g[k]["total_energy"])
        fe = float(g[k]["field_energy"])
        assert abs(rows[k].total_energy - te) / te < 1e-12, k
        assert abs(rows[k].field_energy - fe) / fe < 1e-9, k
        if k > 0:
            assert infos[k - 1].iterations == int(g[k]["picard_iterations"]), k


def p3_small(order=3, nq=4, nx=16, ny=16):
    return make_params(code=CODE_P3, nx=nx, ny=ny, nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0, dt=0.05,
                       spline_order=order, orbit_quad_order=nq, nyquist_projection=1, nonlinear_rtol=1e-11,
                       nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                       local_atol=1e-14, local_max_iter=64)


@pytest.mark.parametrize("order,nq,sort", [(3, 4, 1), (2, 8, 0)])
def test_p3_steps_vs_oracle(orc, order, nq, sort):
    """P3 nested solve (outer Anderson on the mesh current, inner per-particle
    Picard with the HC rotation; P3:5878-6215, 4843-4983) on a 2-D
    bi-Maxwellian load: energies within 1e-12 relative, outer iteration counts
    equal, particle state within 1e-9."""
    lib = gpu()
    p = p3_small(order, nq)
    p.sort_interval = sort
    ppc = 8
    n = p.nx * p.ny * ppc
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, ppc, seed=11, vth=(0.05, 0.2, 0.2)), n)
    st = make_state_from_oracle(orc, p, x, v, q, m)
    steps = 3
    out, rows, infos = run_gpu(p, st, steps)
    so, orows = orc.run(p, st, steps)
    for k in range(steps + 1):
        assert abs(rows[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12, k
    assert [i.iterations for i in infos] == [int(r) for r in orows[1:, 6]]
    for f in ("x", "v", "P"):
        assert rel(getattr(out, f), getattr(so, f)) < 1e-9


def test_init_fields_vs_oracle(orc):
    """Device initial state (upic_sim_init_fields) against the oracle's
    initialiser tails (P1:1287-1303, P3:3920-3937)."""
    lib = gpu()
    for p in (p1_small(), p3_small()):
        n = p.nx * p.ny * 4
        x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, 4, seed=2, vth=(0.1, 0.1, 0.1)), n)
        so = orc.init_fields(p, x, v, q, m)
        ctx = lib.Context(p)
        sim = lib.Simulation(ctx, n)
        s = State(n, p.nx * p.ny)
        s.x[:], s.v[:], s.q[:], s.m[:] = x, v, q, m
        sim.upload(s)
        sim.init_fields(1.0)
        sim.download(s)
        for f in ("rho", "phi", "J", "P", "v_prev"):
            assert rel(getattr(s, f), getattr(so, f)) < 1e-11, f
        sim.close()
        ctx.close()


def test_p2_weibel_validation_summary(orc):
    """P2 Weibel BDF1 conserving run (Project2/examples/weibel_bdf1_conserving.txt,
    160 steps) against Project2/docs/validation_summary.json: final total
    energy and Bz_l2 within 1e-10 relative; gauss/gauge/continuity at roundoff."""
    import json
    import os
    lib = gpu()
    summ = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "p2_validation_summary.json")))
    ref = summ["runs"]["weibel_bdf1_conserving"]
    L = TWO_PI
    p = make_params(code=CODE_P2, nx=32, ny=32, nz=1, Lx=L, Ly=L, Lz=1.0, dt=0.05, bdf_order=1, charge_conserving=1)
    st = p2_weibel_state(orc, p)
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st)
    sim.init_fields(1.0)
    rows = [sim.diagnostics()]
    for _ in range(160):
        sim.step()
        rows.append(sim.diagnostics())
    assert abs(rows[-1].total_energy - ref["final_total_energy"]) / ref["final_total_energy"] < 1e-10
    assert abs(rows[-1].Bz_l2 - ref["final_Bz_l2"]) / ref["final_Bz_l2"] < 1e-8
    assert max(r.continuity_rms for r in rows[1:]) < 1e-14
    assert max(r.gauge_rms for r in rows[1:]) < 1e-12
    assert max(r.gauss_rms for r in rows[1:]) < 1e-12
    sim.close()
    ctx.close()


def p2_weibel_state(orc, p):
    """initialize_weibel_particles (P2:624-669) with the libstdc++ mt19937
    stream restated by the oracle."""
    nsub = 2
    nx, ny = p.nx, p.ny
    dx, dy = p.Lx / nx, p.Ly / ny
    xs, ys = [], []
    for i in range(nx):
        for j in range(ny):
            for a in range(nsub):
                for b in range(nsub):
                    x = (i + (a + 0.5) / nsub) * dx
                    y = (j + (b + 0.5) / nsub) * dy
                    xs += [x, x]
                    ys += [y, y]
    n = len(xs)
    u = orc.mt19937_uniform(12345, -0.01, 0.01, n)
    xs, ys = np.array(xs), np.array(ys)
    sine = 0.002 * np.sin(2 * np.pi * 1 * xs / p.Lx)
    vx = np.tile([-0.5, 0.5], n // 2)
    vy = u + sine
    st = State(n, nx * ny)
    st.x[:n], st.x[n:2 * n] = xs, ys
    st.v[:n], st.v[n:2 * n] = vx, vy
    w = 1.0 * p.Lx * p.Ly / n
    st.q[:] = -w
    st.m[:] = w
    g = 1.0 / np.sqrt(1.0 - (vx * vx + vy * vy))
    st.P[:n], st.P[n:2 * n] = w * g * vx, w * g * vy
    st.v_prev[:] = st.v
    st.fixed_ion_rho = np.ones(nx * ny)
    st.phi_hist = np.zeros(5 * nx * ny)
    st.A_hist = np.zeros(15 * nx * ny)
    st.rho_hist = np.zeros(5 * nx * ny)
    return st


@pytest.mark.parametrize("order,nz,th", [(1, 1, (0.0, 1.0)), (2, 1, (0.25, 0.5)), (3, 1, (0.0, 1.0)),
                                         (4, 1, (0.5, 1.0)), (2, 4, (0.0, 1.0)), (3, 4, (0.25, 0.5))])
def test_local_substep_vs_oracle(orc, order, nz, th):
    """upic_local_substep (solve_local_particle_substep, P3:4930-4983) on
    random particles and smooth fields against the oracle: vbar, P_end, v_end
    to 1e-12, iteration counts and failure reasons exact (a low iteration cap
    makes some particles fail with maximum_iterations)."""
    lib = gpu()
    for max_iter in (64, 3):
        p = make_params(code=CODE_P3, nx=16, ny=8, nz=nz, Lx=TWO_PI, Ly=4.0, Lz=2.0 if nz > 1 else 1.0, dt=0.2,
                        spline_order=order, orbit_quad_order=order + 2, nyquist_projection=1, local_rtol=1e-12,
                        local_atol=1e-14, local_max_iter=max_iter)
        rng = np.random.default_rng(11 + order)
        n = 513
        x0, v = random_particles(rng, p, n, 0.2)
        vv = v.reshape(3, n)
        sp = np.sqrt((vv ** 2).sum(0))
        vv *= np.minimum(1.0, 0.8 / sp)  # below c
        v = vv.ravel().copy()
        N = p.nx * p.ny * p.nz
        # smooth low-mode fields: the inner Picard map is then a contraction
        # (node noise would make it discontinuous across knots, and two
        # roundings could settle on different iterates)
        gx = (np.arange(p.nx) / p.nx - 0.5) * p.Lx
        gy = (np.arange(p.ny) / p.ny - 0.5) * p.Ly
        gz = (np.arange(p.nz) / p.nz - 0.5) * p.Lz
        X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
        kx, ky, kz = 2 * np.pi / p.Lx, 2 * np.pi / p.Ly, 2 * np.pi / p.Lz
        def smooth(amp, ph):
            return np.concatenate([amp * np.sin(kx * X + ky * Y + kz * Z * (p.nz > 1) + ph + c).ravel()
                                   for c in (0.0, 1.3, 2.6)])
        An, Ae, gpm = smooth(0.05, 0.0), smooth(0.05, 0.4), smooth(0.2, 1.1)
        if p.nz == 1:
            gpm[2 * N:] = 0.0  # a gradient on a 2-D grid has no z part (the 2-D node records keep x, y)
        q, m = -np.ones(n), np.ones(n)
        vv = v.reshape(3, n)
        P0 = (vv / np.sqrt(1 - (vv ** 2).sum(0))).ravel()
        h = p.dt * (th[1] - th[0])
        print(f"local substep order={order} nz={nz} theta={th} max_iter={max_iter}", flush=True)
        ctx = lib.Context(p)
        d = ctx.dev
        vb, Pe, ve, it, rs = ctx.local_substep(d(v), d(x0), d(P0), d(q), d(m), h, th[0], th[1], d(An), d(Ae), d(gpm))
        vbo, Peo, veo, ito, rso = orc.p3_local_solve(p, v, x0, P0, q, m, h, th[0], th[1], An, Ae, gpm)
        assert np.array_equal(rs.astype(np.int32), rso)
        assert np.array_equal(it, ito)
        if max_iter == 3:
            assert (rso == 5).any()
        for a, b in ((vb, vbo), (Pe, Peo), (ve, veo)):
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
        ctx.close()


@pytest.mark.parametrize("order", [1, 2])
def test_bdf_wave_vs_oracle(orc, order):
    """upic_bdf_wave (solve_bdf_wave_scalar, P2:829-848) against the oracle's
    (the one inside its P2 step, pinned by the reference's P2 validation
    runs) on random sources and histories, 1e-12 of the solution's scale."""
    lib = gpu()
    p = make_params(code=CODE_P2, nx=32, ny=16, nz=1, Lx=TWO_PI, Ly=3.0, Lz=1.0, dt=0.05)
    N = p.nx * p.ny
    rng = np.random.default_rng(5 + order)
    src = rng.normal(size=N)
    hist = rng.normal(size=(2 if order == 1 else 4) * N)
    ctx = lib.Context(p)
    d = ctx.dev
    u = ctx.bdf_wave(d(src), d(hist), order).numpy()
    uo = orc.bdf_wave(p, src, hist, order)
    assert np.max(np.abs(u - uo)) <= 1e-12 * np.max(np.abs(uo))
    ctx.close()
