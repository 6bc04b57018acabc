"""Selective particle refinement of the P3 step (nonlinear_step_orbit's
fallback, P3:6328-6454): particles whose inner Picard solve fails over the
full dt are re-run with 2, 4, ... substeps while the macro dt stays.

The inner iteration cap is set low (local_max_iter = 3 with the plain v^n
start, which needs about 3 iterations), so some particles fail at full dt:
  * without the fallback the step reports UPIC_ENONLINEAR;
  * FP64 with the fallback: two steps agree with the oracle (which restates
    the same fallback) to 1e-12 in the energies and 1e-9 in the particle
    state, so the refined layout is the reference's;
  * mixed precision (local_max_iter = 2, dt = 0.2, where the plain step
    fails) with the fallback (the fused sweep with per-particle substeps):
    the steps converge and conserve energy to 1e-6.
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_BIMAXWELL, make_load, make_params, with_mixed_precision

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def deck(fallback):
    return make_params(code=CODE_P3, nx=16, ny=16, nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0, dt=0.05,
                       spline_order=3, orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11,
                       nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                       local_atol=1e-14, local_max_iter=3, selective_particle_fallback=fallback,
                       selective_particle_max_substeps=64)


def initial(orc, p, seed=21):
    from unstaggered_pic_amd import lib
    n = p.nx * p.ny * 8
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, 8, seed=seed, vth=(0.1, 0.3, 0.3)), n)
    return orc.init_fields(p, x, v, q, m, 1.0)


def gpu_run(p, st, steps):
    from unstaggered_pic_amd import lib
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st.copy())
    rows = [sim.diagnostics(1)]
    for _ in range(steps):
        sim.step()
        rows.append(sim.diagnostics(1))
    out = State(st.np, st.N)
    sim.download(out)
    sim.close()
    ctx.close()
    return out, rows


def test_without_fallback_the_step_fails(orc):
    from unstaggered_pic_amd import lib
    p = deck(0)
    st = initial(orc, p)
    with pytest.raises(lib.UpicError) as e:
        gpu_run(p, st, 1)
    assert "UPIC_ENONLINEAR" in str(e.value)


def test_fp64_fallback_matches_oracle(orc):
    p = deck(1)
    st = initial(orc, p)
    steps = 2
    out, rows = gpu_run(p, st, steps)
    so, orows = orc.run(p, st, steps)
    for k in range(steps + 1):
        assert abs(rows[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12, k
        assert abs(rows[k].field_energy - orows[k, 2]) / abs(orows[k, 2]) < 1e-12, k
    L = p.Lx
    d = np.abs(out.x - so.x)
    d = np.minimum(d, L - d)
    assert np.max(d) < 1e-9 * L
    assert np.max(np.abs(out.v - so.v)) < 1e-9 * np.max(np.abs(so.v))


def test_mixed_fallback_converges_and_conserves(orc):
    from unstaggered_pic_amd import lib
    # the FP32 tolerance floors need fewer iterations: at dt = 0.2 about a
    # fifth of the particles need more than 2 over the full step
    p0 = with_mixed_precision(deck(0))
    p0.local_max_iter = 2
    p0.dt = 0.2
    st = initial(orc, p0)
    with pytest.raises(lib.UpicError):
        gpu_run(p0, st, 1)
    p = with_mixed_precision(deck(1))
    p.local_max_iter = 2
    p.dt = 0.2
    steps = 3
    _, rows = gpu_run(p, st, steps)
    te = np.array([r.total_energy for r in rows])
    assert np.max(np.abs(te - te[0])) / te[0] < 1e-6


# ---------------------------------------------------------------------------
# Step diagnostics of a refined step (compute_subcycled_step_diagnostics,
# P3:4363-4473) and the refinement record (NonlinearSolveInfo, P3:5511-5518)
# ---------------------------------------------------------------------------
def ref_lib():
    import ctypes as C
    from oracle_bind import Ref
    r = Ref()
    if r.p3 is None:
        pytest.skip("reference harness (oracle/_ref) not built")
    r.p3.ref_p3_last_error.restype = C.c_char_p
    return r


def check_step_diag(d, want, E, tag):
    """upic_step_diagnostics against the reference's 17 values: energy
    changes and works to 1e-9 of the energy-change scale, R_A norms to 1e-6
    where resolved, path statistics exact."""
    scale = max(abs(want[5]), abs(want[6]), abs(want[8]), 1e-300)
    for name, k in (("delta_K", 5), ("delta_W", 6), ("particle_work", 8), ("grid_work", 9)):
        assert abs(getattr(d, name) - want[k]) <= 1e-9 * scale + 1e-14 * E, (tag, name, getattr(d, name), want[k])
    for name, k in (("A_chain_abs_rms", 0), ("A_chain_max_norm", 2), ("A_chain_mean_norm", 3)):
        g, w = getattr(d, name), want[k]
        if w > 1e-14:
            assert abs(g - w) <= 1e-6 * w, (tag, name, g, w)
        else:
            assert g <= 1e-14, (tag, name, g)
    assert abs(d.A_chain_work_defect - want[4]) <= 1e-6 * abs(want[4]) + 1e-10 * scale, tag
    for name, k in (("deposit_gather_work_residual", 10), ("particle_energy_residual", 11),
                    ("field_energy_residual", 12)):
        assert abs(getattr(d, name) - want[k]) <= 1e-9 * scale + 1e-13 * E, (tag, name, getattr(d, name), want[k])
    assert (d.crossing_particles, d.total_crossings, d.max_path_segments) == (want[13], want[14], want[16]), tag
    assert d.avg_path_segments == pytest.approx(want[15], rel=1e-14), tag


def refined_step(p, st):
    from unstaggered_pic_amd import lib
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st.copy())
    sim.keep_step_start(True)
    s0 = sim.download(State(st.np, st.N))
    info = sim.step()
    d = sim.step_diagnostics()
    paths = sim.substep_paths()
    s1 = sim.download(State(st.np, st.N))
    E = sim.diagnostics(1).total_energy
    sim.close()
    ctx.close()
    return s0, s1, info, d, paths, E


def ref_diag_on_paths(ref, p, s0, s1, paths):
    import ctypes as C
    from oracle_bind import ptr
    off, x0, vb = paths
    want = np.zeros(17)
    rc = ref.p3.ref_p3_step_diagnostics_paths(C.byref(p), C.byref(s0.host()), C.byref(s1.host()),
                                              ptr(off, C.c_int64), C.c_int64(len(x0) // 3), ptr(x0), ptr(vb),
                                              ptr(want))
    assert rc == 0, ref.p3.ref_p3_last_error()
    return want


def ref_refined(ref, p, s0):
    import ctypes as C
    from oracle_bind import ptr
    out = State(s0.np, s0.N)
    info, diag = np.zeros(16), np.zeros(17)
    rc = ref.p3.ref_p3_step_refined(C.byref(p), C.byref(s0.host()), C.byref(out.host()), ptr(info), ptr(diag))
    assert rc == 0, ref.p3.ref_p3_last_error()
    return out, info, diag


def check_refine_info(info, want):
    """upic_step_info's refinement record against the reference's
    NonlinearSolveInfo of the same step."""
    assert info.iterations == int(want[0])
    assert info.converged == int(want[1]) == 1
    assert info.uniform_substeps_used == int(want[2])
    assert info.selective_fallback_used == int(want[3]) == 1
    assert info.direct_attempt_iterations == int(want[4])
    assert info.selective_refinement_restarts == int(want[5])
    assert info.selective_refined_particles == int(want[6]) > 0
    assert info.selective_total_particle_substeps == int(want[7])
    assert info.selective_work_ratio == pytest.approx(want[8], rel=1e-15)
    assert list(info.selective_level_counts) == [int(v) for v in want[9:16]]


def test_fp64_refined_step_diagnostics_match_reference(orc):
    """FP64 refined step (16^2 grid, inner cap 3): the step diagnostics
    equal the reference's compute_step_diagnostics on the downloaded start
    and end states with the step's substep paths, and the reference's own
    refined step from the same state has the same refinement record and
    step diagnostics."""
    ref = ref_lib()
    p = deck(1)
    st = initial(orc, p)
    s0, s1, info, d, paths, E = refined_step(p, st)
    assert info.selective_fallback_used == 1 and paths is not None
    off, x0, vb = paths
    assert off[-1] == info.selective_total_particle_substeps == len(x0) // 3
    print(f"refined: {info.selective_refined_particles} particles, substeps {info.selective_total_particle_substeps}, "
          f"R_A {d.A_chain_abs_rms:.3e}", flush=True)
    check_step_diag(d, ref_diag_on_paths(ref, p, s0, s1, paths), E, "paths")
    rout, rinfo, rdiag = ref_refined(ref, p, s0)
    check_refine_info(info, rinfo)
    check_step_diag(d, rdiag, E, "reference step")
    assert np.max(np.abs(s1.v - rout.v)) < 1e-9 * np.max(np.abs(rout.v))


def test_reference_forced_refinement_deck_3d():
    """The reference's own forced-refinement check (Project3/tests/
    test_solver.cpp:48-60: 4^3 cold Weibel, dt 1, linear splines, one inner
    iteration at rtol 2e-3) on the GPU's 3-D FP64 kernels from the
    reference's initial state: without the fallback the step fails; with it
    the step converges with the reference's refinement record (every particle
    refined to dt/16), end state to 1e-9 and step diagnostics as the
    reference's."""
    import ctypes as C
    from unstaggered_pic_amd import lib
    from unstaggered_pic_amd.abi import HostState
    ref = ref_lib()
    p = make_params(code=CODE_P3, nx=4, ny=4, nz=4, Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi, dt=1.0,
                    spline_order=1, orbit_quad_order=8, nyquist_projection=0, nonlinear_rtol=1e-9,
                    nonlinear_atol=1e-11, nonlinear_max_iter=32, anderson_damping=0.5, local_rtol=2e-3,
                    local_atol=1e-10, local_max_iter=1, selective_particle_fallback=0,
                    selective_particle_max_substeps=64)
    args = (C.byref(p), 2, 1, C.c_double(0.2), C.c_double(1.0), C.c_double(0.9), C.c_double(0.0), 1,
            C.c_double(1.0), 0)
    h = HostState()
    assert ref.p3.ref_p3_init_eigen(*args, C.byref(h)) == 0
    st = State(h.np, 64)
    assert ref.p3.ref_p3_init_eigen(*args, C.byref(st.host())) == 0
    with pytest.raises(lib.UpicError):
        gpu_run(p, st, 1)
    p.selective_particle_fallback = 1
    s0, s1, info, d, paths, E = refined_step(p, st)
    rout, rinfo, rdiag = ref_refined(ref, p, s0)
    check_refine_info(info, rinfo)
    assert info.uniform_substeps_used == 16
    for f in ("v", "P"):
        a, b = getattr(s1, f), getattr(rout, f)
        assert np.max(np.abs(a - b)) < 1e-9 * np.max(np.abs(b)), f
    check_step_diag(d, rdiag, E, "forced refinement")


def test_mixed_refined_step_diagnostics_match_reference_formula(orc):
    """Mixed precision refined step (dt 0.2, inner cap 2): the step
    diagnostics (FP64, on the FP64 view and the FP32 substep paths) equal the
    reference's compute_step_diagnostics on the downloaded states and paths;
    the refinement record is consistent with the recorded paths."""
    ref = ref_lib()
    p = with_mixed_precision(deck(1))
    p.local_max_iter = 2
    p.dt = 0.2
    st = initial(orc, p)
    s0, s1, info, d, paths, E = refined_step(p, st)
    assert info.selective_fallback_used == 1 and paths is not None
    off, x0, vb = paths
    counts = np.diff(off)
    assert int(counts.max()) == info.uniform_substeps_used
    assert int((counts > 1).sum()) == info.selective_refined_particles
    assert [int((counts == 2 ** k).sum()) for k in range(7)] == list(info.selective_level_counts)
    check_step_diag(d, ref_diag_on_paths(ref, p, s0, s1, paths), E, "mixed paths")
