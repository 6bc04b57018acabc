"""Writes tests/golden/c1_landau_p1_history.csv: the first steps of the
64x64 Landau workload in the P1 form (409,600 electrons, 100 per cell on
[0, 4 pi)^2, density 1 + 0.01 cos(0.5 x) by inverse CDF, Maxwellian
velocities of width 0.05 c, quadratic splines, 3 Gauss points per orbit
segment, dt = 0.1, Picard tolerance 1e-10 with the final consistency sweep),
computed by the reference's own P1 step compiled from its sources
(oracle/_ref/libref_p1.so: picard_step_orbit, P1:1509-1552, and
compute_diagnostics, P1:1057-1087).

The particles come from the library's seeded loader (upic_load_particles,
seed 12345, host code); the field and history tail of the initial state is
the reference's (P1:1279-1301, ref_p1_init_from_particles).  Each row holds
the step's Picard iteration count and final errors, the energies, the Gauss
RMS and three particle-state checksums (sums of x, v and v^2 over all
particles, in the reference's FP64).

Run in a container that holds the reference tree (about 20 s per step):
    make -C oracle ref && python tests/golden/make_c1_landau_p1.py [steps]
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_bind import Ref, State, ptr  # noqa: E402
from unstaggered_pic_amd import lib  # noqa: E402
from unstaggered_pic_amd.abi import CODE_P1, LOAD_LANDAU, make_load, make_params  # noqa: E402

OUT = os.path.join(HERE, "c1_landau_p1_history.csv")
COLUMNS = ["step", "picard_iterations", "current_error", "velocity_error", "field_energy", "kinetic_energy",
           "total_energy", "gauss_rms", "sum_x", "sum_v", "sum_v2"]


def c1_params():
    return make_params(code=CODE_P1, nx=64, ny=64, nz=1, Lx=4 * np.pi, Ly=4 * np.pi, Lz=1.0, dt=0.1,
                       spline_order=2, orbit_quad_order=3, split_orbit_at_knots=1, c=1.0, eps0=1.0, mu0=1.0,
                       picard_tol=1e-10, picard_max_iter=24, final_consistency_sweep=1, sort_interval=1)


def c1_load():
    return make_load(LOAD_LANDAU, 100, seed=12345, n0=1.0, alpha=0.01, mode=1, vth=(0.05, 0.05, 0.05))


def main(steps=10):
    p = c1_params()
    n = p.nx * p.ny * 100
    N = p.nx * p.ny
    x, v, q, m = lib.load_particles(p, c1_load(), n)
    ref = Ref()
    if ref.p1 is None:
        raise SystemExit("oracle/_ref/libref_p1.so missing: run `make -C oracle ref` where the reference tree is")
    st = State(n, N)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    Ref.check(ref.p1, "ref_p1", ref.p1.ref_p1_init_from_particles(C.byref(p), C.c_double(1.0), C.byref(st.host())))

    def row(k, s, it, errs):
        d = np.zeros(8)
        Ref.check(ref.p1, "ref_p1", ref.p1.ref_p1_diagnostics(C.byref(p), C.byref(s.host()), 1, ptr(d)))
        return [k, it, errs[0], errs[1], d[3], d[4], d[5], d[0], s.x.sum(), s.v.sum(), (s.v ** 2).sum()]

    rows = [row(0, st, 0, (np.nan, np.nan))]
    for k in range(1, steps + 1):
        out = State(n, N)
        it = C.c_int32()
        errs = np.zeros(2)
        Ref.check(ref.p1, "ref_p1", ref.p1.ref_p1_step(C.byref(p), C.byref(st.host()), C.byref(out.host()),
                                                        C.byref(it), ptr(errs)))
        st = out
        rows.append(row(k, st, it.value, errs))
        print(f"step {k}: {it.value} Picard iterations, total energy {rows[-1][6]!r}", flush=True)
    with open(OUT, "w") as f:
        f.write(",".join(COLUMNS) + "\n")
        for r in rows:
            f.write(",".join([str(r[0]), str(r[1])] + [repr(float(a)) for a in r[2:]]) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 10)
