"""Mixed-precision P3 path (FP32 particle state as cell index + in-cell
offsets, FP32 orbit push and LDS deposit tiles, FP64 mesh, FFT, diagnostics
and convergence tests) against the FP64 oracle.

The mixed mode is not bit-compatible with the reference by design; each test
states its error bound against the FP64 result:
  * state round trip: FP32 rounding of the stored quantities (2^-24 relative,
    positions to 2^-23 of a cell);
  * P3 steps vs the oracle: total energy within 1e-6 relative, field energy
    within 1e-4 relative, particle state within 1e-5 of the box / 1e-4 of the
    velocity scale; Gauss's law, the Lorenz gauge and the energy balance of
    the mixed run hold as in FP64 (Gauss and gauge residuals at roundoff,
    total-energy drift within 1e-6).
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_BIMAXWELL, make_load, make_params, with_mixed_precision

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def gpu():
    from unstaggered_pic_amd import lib
    return lib


def p3_deck(nx=16, ny=16, order=3, nq=4):
    return make_params(code=CODE_P3, nx=nx, ny=ny, nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0, dt=0.05,
                       spline_order=order, orbit_quad_order=nq, nyquist_projection=1, nonlinear_rtol=1e-11,
                       nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                       local_atol=1e-14, local_max_iter=64)


def run(p, st, steps, mode=1):
    lib = gpu()
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st.copy())
    rows, infos = [sim.diagnostics(mode)], []
    for _ in range(steps):
        infos.append(sim.step())
        rows.append(sim.diagnostics(mode))
    out = State(st.np, st.N)
    sim.download(out)
    sim.close()
    ctx.close()
    return out, rows, infos


def periodic_dist(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d)


def test_mixed_state_roundtrip(orc):
    """Upload -> FP32 state -> download reproduces the FP64 state to FP32
    rounding; q and m come back as their FP32 values (the mixed mode uses the
    rounded charge and mass on the FP64 side too)."""
    lib = gpu()
    p = with_mixed_precision(p3_deck())
    n = p.nx * p.ny * 8
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, 8, seed=5, vth=(0.05, 0.2, 0.2)), n)
    st = orc.init_fields(p, x, v, q, m, 1.0)
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, n)
    sim.upload(st.copy())
    out = State(n, st.N)
    sim.download(out)
    h = p.Lx / p.nx
    xs, xo = st.x.reshape(3, n), out.x.reshape(3, n)
    for d, L in ((0, p.Lx), (1, p.Ly)):
        assert np.max(periodic_dist(xs[d], xo[d], L)) <= 2.0 ** -22 * h
    assert np.max(periodic_dist(xs[2], xo[2], p.Lz)) <= 2.0 ** -22 * p.Lz
    for f in ("v", "P", "v_prev", "vbar_last"):
        a, b = getattr(st, f), getattr(out, f)
        assert np.all(np.abs(a - b) <= 2.0 ** -24 * np.abs(a) + 1e-300), f
    assert np.array_equal(out.q, st.q.astype(np.float32).astype(np.float64))
    assert np.array_equal(out.m, st.m.astype(np.float32).astype(np.float64))
    sim.close()
    ctx.close()


@pytest.mark.parametrize("order,nq,sort", [(3, 4, 1), (2, 3, 0), (1, 2, 1), (4, 5, 1)])
def test_mixed_p3_steps_vs_oracle(orc, order, nq, sort):
    """Five P3 steps in the mixed mode against the FP64 oracle run on the
    same inputs with the FP64 tolerances."""
    p64 = p3_deck(order=order, nq=nq)
    p64.sort_interval = sort
    pm = with_mixed_precision(p64)
    n = p64.nx * p64.ny * 8
    lib = gpu()
    x, v, q, m = lib.load_particles(p64, make_load(LOAD_BIMAXWELL, 8, seed=11, vth=(0.05, 0.2, 0.2)), n)
    st = orc.init_fields(p64, x, v, q, m, 1.0)
    steps = 5
    out, rows, infos = run(pm, st, steps)
    so, orows = orc.run(p64, st, steps)
    te0 = rows[0].total_energy
    for k in range(steps + 1):
        te, fe = rows[k].total_energy, rows[k].field_energy
        assert abs(te - orows[k, 1]) / abs(orows[k, 1]) < 1e-6, k
        assert abs(fe - orows[k, 2]) / abs(orows[k, 2]) < 1e-4, k
        assert abs(te - te0) / abs(te0) < 1e-6, k
        assert rows[k].gauss_rms < 1e-12 * max(1.0, rows[k].rho_rms), k
        assert rows[k].gauge_rms < 1e-12 * max(1e-12, rows[k].divA_rms + rows[k].psi_over_c2_rms), k
    assert all(i.converged for i in infos)
    h = p64.Lx / p64.nx
    xa, xb = out.x.reshape(3, n), so.x.reshape(3, n)
    assert np.max(periodic_dist(xa[0], xb[0], p64.Lx)) < 1e-5 * p64.Lx
    assert np.max(periodic_dist(xa[1], xb[1], p64.Ly)) < 1e-5 * p64.Ly
    vs = np.max(np.abs(so.v))
    assert np.max(np.abs(out.v - so.v)) < 1e-4 * vs
    assert np.max(np.abs(out.P - so.P)) < 1e-4 * np.max(np.abs(so.P))
    assert h > 0


def test_mixed_vs_fp64_gpu_history():
    """64x64 grid, 32 particles per cell, 20 steps: the mixed run's energy
    histories against the FP64 GPU run (itself pinned to the oracle by the
    FP64 parity tests): total energy within 1e-6, field energy within 1e-4,
    mixed total-energy drift within 1e-6."""
    lib = gpu()
    p64 = make_params(code=CODE_P3, nx=64, ny=64, nz=1, Lx=8 * np.pi, Ly=8 * np.pi, Lz=1.0, dt=0.05,
                      spline_order=3, orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11,
                      nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                      local_atol=1e-14, local_max_iter=64)
    pm = with_mixed_precision(p64)
    ppc = 32
    n = p64.nx * p64.ny * ppc
    x, v, q, m = lib.load_particles(p64, make_load(LOAD_BIMAXWELL, ppc, seed=3, vth=(0.05, 0.2, 0.2)), n)
    st = State(n, p64.nx * p64.ny)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    hist = {}
    for tag, p in (("fp64", p64), ("mixed", pm)):
        ctx = lib.Context(p)
        sim = lib.Simulation(ctx, n)
        sim.upload(st.copy())
        sim.init_fields(1.0)
        rows = [sim.diagnostics(1)]
        for _ in range(20):
            sim.step()
            rows.append(sim.diagnostics(1))
        hist[tag] = np.array([[r.total_energy, r.field_energy] for r in rows])
        sim.close()
        ctx.close()
    a, b = hist["mixed"], hist["fp64"]
    assert np.max(np.abs(a[:, 0] - b[:, 0]) / np.abs(b[:, 0])) < 1e-6
    assert np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(b[:, 1])) < 1e-4
    assert np.max(np.abs(a[:, 0] - a[0, 0])) / abs(a[0, 0]) < 1e-6


@pytest.mark.parametrize("split", [1, 0])
def test_deposit_variants_agree_and_autotune(split):
    """64x64 grid, 16 particles per cell, sorted every step: six steps with
    each deposit variant (kernels, tile shapes and cell blocks) pinned reach
    the same fixed point (total energy 1e-6, field energy 1e-4, positions
    1e-5 of the box), and the autotuner settles on a usable variant.
    Without knot splitting the sort leaves the particles in plain cell
    order, which the cell-mapped variants (1, 2 on row segments, 6, 7 on
    8 x 8 blocks) need; with it, crossing orbits are grouped after the
    others and those fall back to variant 0."""
    lib = gpu()
    p64 = make_params(code=CODE_P3, nx=64, ny=64, nz=1, Lx=8 * np.pi, Ly=8 * np.pi, Lz=1.0, dt=0.05,
                      spline_order=3, orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11,
                      nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                      local_atol=1e-14, local_max_iter=64, split_orbit_at_knots=split)
    p64.sort_interval = 1
    pm = with_mixed_precision(p64)
    ppc = 16
    n = p64.nx * p64.ny * ppc
    x, v, q, m = lib.load_particles(p64, make_load(LOAD_BIMAXWELL, ppc, seed=8, vth=(0.05, 0.2, 0.2)), n)
    st = State(n, p64.nx * p64.ny)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    ctx = lib.Context(pm)
    res = {}
    variants = tuple(range(8))
    for variant in (-1,) + variants:
        sim = lib.Simulation(ctx, n)
        sim.upload(st.copy())
        sim.init_fields(1.0)
        sim.set_deposit_variant(variant)
        rows = [sim.diagnostics(1)]
        for _ in range(6):
            assert sim.step().converged
            rows.append(sim.diagnostics(1))
        out = State(n, st.N)
        sim.download(out)
        res[variant] = (np.array([[r.total_energy, r.field_energy] for r in rows]), out.x.reshape(3, n).copy())
        if variant == -1:
            chosen = sim.deposit_variant()
        sim.close()
    ctx.close()
    if split:
        assert chosen[1] in (0, 3, 4, 5), chosen
    else:
        assert chosen[0] in variants, chosen
    e0, x0 = res[0]
    assert np.max(np.abs(e0[:, 0] - e0[0, 0])) / e0[0, 0] < 1e-6
    for variant in (-1,) + variants[1:]:
        e, xv = res[variant]
        assert np.max(np.abs(e[:, 0] - e0[:, 0]) / e0[:, 0]) < 1e-6, variant
        assert np.max(np.abs(e[:, 1] - e0[:, 1]) / e0[:, 1]) < 1e-4, variant
        assert np.max(periodic_dist(xv[0], x0[0], p64.Lx)) < 1e-5 * p64.Lx, variant
        assert np.max(periodic_dist(xv[1], x0[1], p64.Ly)) < 1e-5 * p64.Ly, variant


def test_mixed_history_follows_particles_through_sorts(orc):
    """The cell sort moves v_prev and vbar lazily (the commit overwrites
    them): after every step, in the original particle order, v_prev is the
    previous step's v exactly and x^{n+1} - x^n = dt vbar_last to FP32
    position rounding; a step that fails without changing the state leaves
    v_prev and vbar_last where they were."""
    lib = gpu()
    p = with_mixed_precision(p3_deck())
    n = p.nx * p.ny * 8
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, 8, seed=9, vth=(0.05, 0.2, 0.2)), n)
    st = orc.init_fields(p, x, v, q, m, 1.0)
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, n)
    sim.upload(st.copy())
    prev = sim.download(State(n, st.N))
    h = p.Lx / p.nx
    for _ in range(3):
        sim.step()
        cur = sim.download(State(n, st.N))
        assert np.array_equal(cur.v_prev, prev.v)
        for d, L in ((0, p.Lx), (1, p.Ly)):
            dx = cur.x.reshape(3, n)[d] - prev.x.reshape(3, n)[d]
            dx -= L * np.round(dx / L)
            assert np.max(np.abs(dx - p.dt * cur.vbar_last.reshape(3, n)[d])) <= 1e-5 * h
        prev = cur
    # a step that cannot converge (one outer iteration allowed) fails after
    # the sort and leaves the state as it was
    sim.close()
    ctx.close()
    p_fail = with_mixed_precision(p3_deck())
    p_fail.nonlinear_max_iter = 1
    ctx = lib.Context(p_fail)
    sim = lib.Simulation(ctx, n)
    sim.upload(prev)
    before = sim.download(State(n, st.N))
    with pytest.raises(lib.UpicError):
        sim.step()
    after = sim.download(State(n, st.N))
    for f in ("x", "v", "v_prev", "vbar_last", "P", "q", "m"):
        assert np.array_equal(getattr(after, f), getattr(before, f)), f
    sim.close()
    ctx.close()


def test_weak_current_members_leave_the_vbar_start():
    """A cold Landau load (weak net current) with the predictor guess: the
    vbar-started sweeps gave its outer solve a memory (5-8 evaluations, up to
    5e-8 of the energy per step); after two such steps (the run's first step,
    which cannot start from the extrapolated current, does not count) the
    simulation starts every sweep from the predictor, and the steps after
    that converge in four evaluations with energy changes at 1e-9."""
    from unstaggered_pic_amd.ensemble_bench import _Host, build
    lib = gpu()
    spec = {"kind": "landau", "side": 64, "ppc": 16, "vth": 0.033, "amp": 0.005, "seed": 12345}
    p, n, (x, v, q, m) = build(lib, spec)
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, n)
    h = _Host(n, p.nx * p.ny)
    h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
    sim.upload(h)
    sim.init_fields(1.0)
    E = [sim.diagnostics(1).total_energy]
    maps = []
    for _ in range(7):
        maps.append(sim.step().map_evaluations)
        E.append(sim.diagnostics(1).total_energy)
    print("map evaluations per step:", maps, flush=True)
    dE = np.abs(np.diff(E)) / abs(E[0])
    assert all(k <= 4 for k in maps[3:])
    assert np.all(dE[3:] < 1e-8), dE
    sim.close()
    ctx.close()
