"""upic_sim_upload against work still queued on the context stream.

The mixed-precision state keeps the FP64 particle view stale after a step;
an upload first expands it (a launch on the context's non-blocking stream)
so that arrays the caller leaves out keep their values, and then copies the
caller's arrays over the same device memory.  Those copies must wait for
the expansion: otherwise the expansion can land after them and the step
starts from the previous step's particles.  Here one simulation is
re-uploaded from the same initial state and stepped again and again; every
step must land on the oracle's step (mixed bounds), as the first one does.
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_LANDAU, make_load, make_params, with_mixed_precision

pytestmark = pytest.mark.gpu


def test_reupload_after_steps_starts_from_the_uploaded_state():
    from unstaggered_pic_amd import lib
    orc = Oracle()
    p64 = make_params(code=CODE_P3, nx=32, ny=32, nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0, dt=0.1,
                      spline_order=2, orbit_quad_order=3, nyquist_projection=1, nonlinear_rtol=1e-11,
                      nonlinear_atol=1e-13, nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12,
                      local_atol=1e-14, local_max_iter=64, sort_interval=1, selective_particle_fallback=1)
    n = 32 * 32 * 4
    x, v, q, m = lib.load_particles(p64, make_load(LOAD_LANDAU, 4, seed=6, alpha=1e-2, vth=(0.1,) * 3), n)
    st = orc.init_fields(p64, x, v, q, m, 1.0)
    so, orows = orc.run(p64, st, 1)
    pm = with_mixed_precision(p64)
    pm.learned_guess = 1
    ctx = lib.Context(pm)
    sim = lib.Simulation(ctx, n)
    reps = 24
    print(f"re-upload: landau 32x32 ppc 4 mixed, {reps} upload + step rounds", flush=True)
    try:
        for r in range(reps):
            sim.upload(st.copy())
            sim.step()
            fe = sim.diagnostics(1).field_energy
            assert abs(fe / orows[1, 2] - 1) < 1e-6, (r, fe, orows[1, 2])
            out = State(n, 32 * 32)
            sim.download(out)
            dv = np.max(np.abs(out.v - so.v)) / np.max(np.abs(so.v))
            assert dv < 1e-5, (r, dv)
    finally:
        sim.close()
        ctx.close()
