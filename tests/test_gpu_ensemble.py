"""Batched ensembles and graph-replayed field solves (no reference
counterpart: the reference runs one simulation per process and launches
nothing).

  * upic_ensemble_step advances members of different codes, decks, sizes and
    precisions concurrently; each member matches the same member stepped on
    its own (FP64 members: total energy 1e-12 and field energy 1e-10
    relative, which is the roundoff of the FP64 atomic deposit order; mixed
    members: 1e-6 and 1e-4, the FP32 bounds of the mixed mode);
  * the hipGraph-replayed field solve gives the eager results (same bounds),
    replayed many times with eager work and host synchronisations between
    the replays, for P1 (two alternating source buffers) and P3;
  * members sharing a context are rejected.
"""
import numpy as np
import pytest

from unstaggered_pic_amd.abi import (CODE_P1, CODE_P3, LOAD_BIMAXWELL, LOAD_LANDAU, make_load, make_params,
                                     with_mixed_precision)

pytestmark = pytest.mark.gpu


def gpu():
    from unstaggered_pic_amd import lib
    return lib


def p3_deck(nx, ny, mixed, learned=0):
    p = make_params(code=CODE_P3, nx=nx, ny=ny, nz=1, Lx=2 * np.pi * nx / 16, Ly=2 * np.pi * ny / 16, Lz=1.0,
                    dt=0.05, spline_order=3, orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11,
                    nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                    local_atol=1e-14, local_max_iter=64)
    if mixed:
        p = with_mixed_precision(p)
        p.learned_guess = learned
    return p


def p1_deck():
    return make_params(code=CODE_P1, nx=16, ny=8, nz=1, Lx=4 * np.pi, Ly=4 * np.pi, Lz=1.0, dt=0.1,
                       spline_order=2, orbit_quad_order=4, picard_tol=1e-10, picard_max_iter=24)


MEMBERS = [  # (params factory, particles per cell, load, FP64?)
    (lambda: p3_deck(64, 64, True, learned=1), 16, LOAD_BIMAXWELL, False),
    (lambda: p3_deck(32, 32, False), 16, LOAD_BIMAXWELL, True),
    (lambda: p1_deck(), 8, LOAD_LANDAU, True),
    (lambda: p3_deck(32, 64, True), 8, LOAD_BIMAXWELL, False),
]


def make_member(lib, k, graphs=False):
    pf, ppc, kind, fp64 = MEMBERS[k]
    p = pf()
    n = p.nx * p.ny * ppc
    if kind == LOAD_LANDAU:
        load = make_load(LOAD_LANDAU, ppc, seed=5 + k, alpha=0.05, vth=(0.3, 0.3, 0.3))
    else:
        load = make_load(LOAD_BIMAXWELL, ppc, seed=5 + k, vth=(0.05, 0.2, 0.2))
    x, v, q, m = lib.load_particles(p, load, n)
    ctx = lib.Context(p)
    ctx.set_graphs(graphs)
    sim = lib.Simulation(ctx, n)
    from oracle_bind import State
    st = State(n, p.nx * p.ny)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    sim.upload(st)
    sim.init_fields(1.0)
    return ctx, sim, fp64


def energies(sim):
    d = sim.diagnostics(1)
    return d.total_energy, d.field_energy


def compare(a, b, fp64):
    a, b = np.array(a), np.array(b)
    te = np.max(np.abs(a[:, 0] - b[:, 0]) / np.abs(b[:, 0]))
    fe = np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(b[:, 1]))
    if fp64:
        assert te < 1e-12 and fe < 1e-10, (te, fe)
    else:
        assert te < 1e-6 and fe < 1e-4, (te, fe)


def test_ensemble_uneven_members_match_individual_runs():
    lib = gpu()
    steps = 3
    alone = []
    for k in range(len(MEMBERS)):
        ctx, sim, fp64 = make_member(lib, k)
        hist = [energies(sim)]
        for _ in range(steps):
            assert sim.step().converged
            hist.append(energies(sim))
        alone.append(hist)
        sim.close()
        ctx.close()
    members = [make_member(lib, k, graphs=(k == 3)) for k in range(len(MEMBERS))]
    sims = [m[1] for m in members]
    hist = [[energies(s)] for s in sims]
    for _ in range(steps):
        infos = lib.ensemble_step(sims, threads=4)
        assert all(i.converged for i in infos)
        for h, s in zip(hist, sims):
            h.append(energies(s))
    for k, (ctx, sim, fp64) in enumerate(members):
        compare(hist[k], alone[k], fp64)
        sim.close()
        ctx.close()


def test_ensemble_rejects_shared_context():
    lib = gpu()
    p = p3_deck(16, 16, False)
    ctx = lib.Context(p)
    a, b = lib.Simulation(ctx, 64), lib.Simulation(ctx, 64)
    with pytest.raises(Exception):
        lib.ensemble_step([a, b])
    a.close()
    b.close()
    ctx.close()


@pytest.mark.parametrize("k", [1, 2, 3])
def test_graph_field_solve_matches_eager(k):
    """Six steps (P3: four map evaluations and a field solve each; P1: one
    per Picard iteration on alternating buffers) with and without graphs."""
    lib = gpu()
    steps = 6
    res = []
    for graphs in (False, True):
        ctx, sim, fp64 = make_member(lib, k, graphs=graphs)
        hist = [energies(sim)]
        for _ in range(steps):
            assert sim.step().converged
            hist.append(energies(sim))
        res.append(hist)
        sim.close()
        ctx.close()
    compare(res[1], res[0], fp64)


def test_landau_load_stays_below_c_for_p3():
    """The ensemble draws hot Landau members (thermal width up to 0.2 c); for
    the relativistic P3 step the Landau load draws momenta, so every speed is
    below c even in the far tail (a 1e6-particle draw at width 0.3 c)."""
    lib = gpu()
    p = p3_deck(64, 64, False)
    n = p.nx * p.ny * 256
    _, v, _, _ = lib.load_particles(p, make_load(LOAD_LANDAU, 256, seed=3, alpha=0.01, vth=(0.3, 0.3, 0.3)), n)
    speed = np.sqrt(np.sum(v.reshape(3, n) ** 2, axis=0))
    assert np.max(speed) < 1.0
    # momenta u = gamma v are Maxwellian of width 0.3
    u = v.reshape(3, n) / np.sqrt(1.0 - speed ** 2)
    assert abs(np.std(u[0]) - 0.3) < 0.01


def test_cold_landau_member_energy_drift():
    """A cold Landau member of the batched ensemble's draw (member 46 of
    member_specs(64): 64x64 grid, 256 particles per cell, v_th 0.036 c,
    quadratic splines, dt 0.1), mixed precision with the physics predictor,
    graphs on: the total energy drifts less than 1e-6 over 300 steps after
    10 start-up steps.  The outer solve's start from the linear current
    extrapolation keeps it near 1e-7; a quadratic extrapolation had let it
    grow by 6e-9 per step (2e-6 at 300 steps) while every step still met
    the solver tolerance."""
    lib = gpu()
    from unstaggered_pic_amd import ensemble_bench as eb
    spec = eb.member_specs(64)[46]
    assert spec["kind"] == "landau" and spec["side"] == 64 and spec["ppc"] == 256
    p, n, (x, v, q, m) = eb.build(lib, spec)
    ctx = lib.Context(p)
    ctx.set_graphs(True)
    sim = lib.Simulation(ctx, n)
    h = eb._Host(n, p.nx * p.ny)
    h.x[:], h.v[:], h.q[:], h.m[:] = x, v, q, m
    sim.upload(h)
    sim.init_fields(1.0)
    sim.set_guess_mode(1)
    for _ in range(10):
        sim.step()
    e0 = sim.diagnostics(1).total_energy
    worst = 0.0
    for k in range(300):
        info = sim.step()
        assert info.converged
        if k % 50 == 49:
            worst = max(worst, abs(sim.diagnostics(1).total_energy - e0) / abs(e0))
    print(f"cold Landau member: largest total-energy drift {worst:.2e} over 300 steps", flush=True)
    sim.close()
    ctx.close()
    assert worst < 1e-6
