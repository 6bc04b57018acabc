"""Deterministic mode (upic_ctx_set_deterministic): the same state stepped
twice gives the same bits, and the mode agrees with the oracle as the
default mode does.

The reference gets a reproducible deposit from per-thread meshes summed in a
fixed order (Project3/ec_pic_nyquist_projected.cpp:2194-2216); the GPU mode
accumulates the orbit current in exact fixed-point integer words, keeps each
cell's particles in storage order through the cell sort and runs its
reductions as single workgroups.
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_BIMAXWELL, make_load, make_params

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "v_prev", "P", "vbar_last", "phi", "psi", "rho", "A", "U", "J")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def params(nz=1, nx=16, ny=16):
    return make_params(code=CODE_P3, nx=nx, ny=ny, nz=nz, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0 if nz == 1 else 2.0,
                       dt=0.05, spline_order=3 if nz == 1 else 2, orbit_quad_order=4, nyquist_projection=1,
                       nonlinear_rtol=1e-11, nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5,
                       local_rtol=1e-12, local_atol=1e-14, local_max_iter=64, sort_interval=1)


def run(p, st, steps, det):
    from unstaggered_pic_amd import lib
    ctx = lib.Context(p)
    ctx.set_deterministic(det)
    sim = lib.Simulation(ctx, st.np)
    sim.upload(st.copy())
    infos, diags = [], [sim.diagnostics(1)]
    for _ in range(steps):
        infos.append(sim.step())
        diags.append(sim.diagnostics(1))
    out = State(st.np, st.N)
    sim.download(out)
    sim.close()
    ctx.close()
    return out, infos, diags


def load(orc, p, ppc, seed):
    from unstaggered_pic_amd import lib
    n = p.nx * p.ny * p.nz * ppc
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, ppc, seed=seed, vth=(0.05, 0.2, 0.2)), n)
    return orc.init_fields(p, x, v, q, m, 1.0)


@pytest.mark.parametrize("nz,ppc", [(1, 32), (4, 4)])
def test_same_state_twice_gives_the_same_bits(orc, nz, ppc):
    """Two runs of three FP64 steps from one state (2-D cubic with the LDS
    tile, 3-D quadratic with direct fixed-point atomics; cell sort every
    step): particle and mesh state, step records and diagnostics equal bit
    for bit."""
    p = params(nz=nz, nx=16 if nz == 1 else 8, ny=16 if nz == 1 else 8)
    st = load(orc, p, ppc, seed=21)
    print(f"determinism: nz={nz} ppc={ppc} np={st.np}", flush=True)
    a, ia, da = run(p, st, 3, True)
    b, ib, db = run(p, st, 3, True)
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for x, y in zip(ia, ib):
        assert (x.iterations, x.map_evaluations, x.local_iterations_total) == \
            (y.iterations, y.map_evaluations, y.local_iterations_total)
        assert x.final_rms_residual == y.final_rms_residual
        assert x.estimated_map_contraction == y.estimated_map_contraction or \
            (np.isinf(x.estimated_map_contraction) and np.isinf(y.estimated_map_contraction))
    for x, y in zip(da, db):
        assert (x.total_energy, x.field_energy, x.kinetic_energy, x.gauss_rms) == \
            (y.total_energy, y.field_energy, y.kinetic_energy, y.gauss_rms)


def test_deterministic_mode_matches_oracle(orc):
    """The deterministic mode changes summation order only: energies within
    1e-12 of the oracle's, outer iteration counts equal, particle state within
    1e-9, as for the default mode."""
    p = params()
    st = load(orc, p, 16, seed=5)
    out, infos, diags = run(p, st, 3, True)
    so, orows = orc.run(p, st, 3)
    for k in range(4):
        assert abs(diags[k].total_energy - orows[k, 1]) / abs(orows[k, 1]) < 1e-12, k
    assert [i.iterations for i in infos] == [int(r) for r in orows[1:, 6]]
    for f in ("x", "v", "P"):
        d = np.max(np.abs(getattr(out, f) - getattr(so, f))) / np.max(np.abs(getattr(so, f)))
        assert d < 1e-9, f
    # and the default mode to roundoff
    ref, _, rdiags = run(p, st, 3, False)
    for k in range(4):
        assert abs(diags[k].total_energy - rdiags[k].total_energy) <= 1e-13 * abs(rdiags[k].total_energy)
    assert np.max(np.abs(out.J - ref.J)) <= 1e-12 * np.max(np.abs(ref.J))
