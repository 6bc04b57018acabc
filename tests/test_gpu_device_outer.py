"""Device-driven outer solve (graphs on): a P3 step is one captured graph --
the controller lanes make the host loop's decisions (Anderson type II,
backtracking, restarts, convergence; Project3/ec_pic_nyquist_projected.cpp:
5878-6215) between gated map evaluations -- replayed with one launch.

  * FP64: energies within 1e-12 of the oracle and outer iteration counts,
    map evaluations, Anderson steps, restarts and update methods equal to the
    host-driven step's, for the Anderson and the plain Picard solver, and
    with two evaluations per graph (the steps continue slot by slot);
  * a step whose particles fail the inner solve falls back to the host loop
    and the selective refinement, as the oracle does;
  * mixed precision with the learned-guess modes: the same fixed point as
    the host-driven step (the FP32 bounds of the mixed mode).
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_BIMAXWELL, make_load, make_params, with_mixed_precision

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def deck(**kw):
    base = dict(code=CODE_P3, nx=16, ny=16, nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0, dt=0.05, spline_order=3,
                orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11, nonlinear_atol=1e-13,
                nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12, local_atol=1e-14, local_max_iter=64,
                sort_interval=1)
    base.update(kw)
    return make_params(**base)


def initial(orc, p, ppc=8, seed=11, vth=(0.05, 0.2, 0.2)):
    from unstaggered_pic_amd import lib
    n = p.nx * p.ny * ppc
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, ppc, seed=seed, vth=vth), n)
    return orc.init_fields(p, x, v, q, m, 1.0)


def run(p, st, steps, graphs, mode=None, warm=0):
    """graphs: 0 host-driven, 1 device-driven (adaptive unrolling), n >= 2
    device-driven with n unrolled evaluations per step graph."""
    from unstaggered_pic_amd import lib
    ctx = lib.Context(p)
    ctx.set_graphs(graphs)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st.copy())
    if mode is not None:
        sim.set_guess_mode(mode)
    for _ in range(warm):
        sim.step()
    rows, infos = [sim.diagnostics(1)], []
    for _ in range(steps):
        infos.append(sim.step())
        rows.append(sim.diagnostics(1))
    out = State(st.np, st.N)
    sim.download(out)
    sim.close()
    ctx.close()
    return out, rows, infos


INFO_EXACT = ("iterations", "converged", "map_evaluations", "anderson_steps", "anderson_restarts",
              "picard_fallback_steps", "backtracked_steps", "update_method")


@pytest.mark.parametrize("solver,graphs", [("anderson", 1), ("picard", 1), ("anderson", 2)])
def test_fp64_device_steps_match_oracle_and_host_loop(orc, solver, graphs):
    """graphs = 2: two evaluations per step graph, so every step continues
    slot by slot after the graph."""
    p = deck(nonlinear_solver_picard=1 if solver == "picard" else 0)
    st = initial(orc, p)
    steps = 4
    print(f"device outer: fp64 {solver} 16x16 ppc 8, {steps} steps, graphs {graphs}", flush=True)
    out, rows, infos = run(p, st, steps, graphs=graphs)
    _, _, host = run(p, st, steps, graphs=False)
    so, orows = orc.run(p, st, steps)
    for k in range(steps + 1):
        assert abs(rows[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12, k
        assert abs(rows[k].field_energy - orows[k, 2]) / abs(orows[k, 2]) < 1e-10, k
    assert [i.iterations for i in infos] == [int(r) for r in orows[1:, 6]]
    assert [i.map_evaluations for i in infos] == [int(r) for r in orows[1:, 18]]
    assert all(i.graph_launch == 1 for i in infos)
    assert all(h.graph_launch == 0 for h in host)
    for d, h in zip(infos, host):
        for f in INFO_EXACT:
            assert getattr(d, f) == getattr(h, f), f
        assert d.final_rms_residual == pytest.approx(h.final_rms_residual, rel=1e-6)
        # inner iterations: equal but for a particle at a tolerance tie
        assert d.local_iterations_total == pytest.approx(h.local_iterations_total, rel=1e-3)
        assert (d.path.crossing_particles, d.path.total_segments, d.path.max_segments) == \
            (h.path.crossing_particles, h.path.total_segments, h.path.max_segments)
    for f in ("x", "v", "P"):
        d = np.max(np.abs(getattr(out, f) - getattr(so, f))) / np.max(np.abs(getattr(so, f)))
        assert d < 1e-9, f
    if graphs == 2:
        assert min(i.map_evaluations for i in infos) > 2


def test_particle_failure_falls_back_to_refinement(orc):
    """local_max_iter = 3: some particles fail over the full dt; the device
    step hands the step to the host loop, whose selective refinement matches
    the oracle's (1e-12 energies)."""
    p = deck(local_max_iter=3, selective_particle_fallback=1, selective_particle_max_substeps=64)
    st = initial(orc, p, seed=21, vth=(0.1, 0.3, 0.3))
    steps = 2
    _, rows, infos = run(p, st, steps, graphs=True)
    _, orows = orc.run(p, st, steps)
    assert any(i.selective_fallback_used for i in infos)
    for k in range(steps + 1):
        assert abs(rows[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12, k


@pytest.mark.parametrize("mode", [0, 1])
def test_mixed_device_steps_match_host_loop(orc, mode):
    """Mixed precision, 32 x 32, 16 particles per cell: ten device-driven
    steps after the deposit autotuner has settled, against the host-driven
    steps from the same state: total energy within 1e-6, field energy within
    1e-4 (the FP32 bounds; the deposit's FP32 atomics make the two runs
    differ at rounding level), the same outer iteration count in most steps."""
    p = with_mixed_precision(deck(nx=32, ny=32, Lx=4 * np.pi, Ly=4 * np.pi))
    st = initial(orc, p, ppc=16, seed=5)
    steps = 10
    _, rd, idev = run(p, st, steps, graphs=True, mode=mode, warm=4)
    _, rh, ihost = run(p, st, steps, graphs=False, mode=mode, warm=4)
    for a, b in zip(rd, rh):
        assert abs(a.total_energy - b.total_energy) <= 1e-6 * abs(b.total_energy)
        assert abs(a.field_energy - b.field_energy) <= 1e-4 * abs(b.field_energy)
    same = sum(a.iterations == b.iterations for a, b in zip(idev, ihost))
    assert same >= steps - 2, ([a.iterations for a in idev], [b.iterations for b in ihost])
    assert all(i.converged for i in idev)
