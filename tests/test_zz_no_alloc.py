"""upic_local_substep_noalloc against upic_local_substep through the C ABI.

The no-allocation variant runs the same kernels with its node records in the
library's static device scratch, so its results must equal the allocating
entry point's bit for bit.  Two calls in a row (the second on a smaller
particle set and mesh, so the scratch holds the first call's records when it
starts), a 3-D call (no scratch) and the size limit's error.
"""
import numpy as np
import pytest

from unstaggered_pic_amd.abi import CODE_P3, UPIC_EINVAL, make_params

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, n, seed, order=3):
    p = make_params(code=CODE_P3, nx=nx, ny=ny, nz=nz, Lx=6.283185307179586, Ly=4.0, Lz=2.0 if nz > 1 else 1.0,
                    dt=0.2, spline_order=order, orbit_quad_order=order + 1, nyquist_projection=1, local_rtol=1e-12,
                    local_atol=1e-14, local_max_iter=64)
    rng = np.random.default_rng(seed)
    L = np.array([p.Lx, p.Ly, p.Lz])
    x = (rng.random((3, n)) - 0.5) * L[:, None]
    if nz == 1:
        x[2] = 0.0
    v = rng.normal(size=(3, n)) * 0.2
    v *= np.minimum(1.0, 0.8 / np.sqrt((v ** 2).sum(0)))
    N = nx * ny * nz
    gx = (np.arange(nx) / nx - 0.5) * p.Lx
    gy = (np.arange(ny) / ny - 0.5) * p.Ly
    gz = (np.arange(nz) / nz - 0.5) * p.Lz
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")

    def smooth(amp, ph):
        return np.concatenate([amp * np.sin(2 * np.pi * (X / p.Lx + Y / p.Ly + (Z / p.Lz) * (nz > 1)) + ph + c).ravel()
                               for c in (0.0, 1.3, 2.6)])
    An, Ae, gpm = smooth(0.05, 0.0), smooth(0.05, 0.4), smooth(0.2, 1.1)
    if nz == 1:
        gpm[2 * N:] = 0.0
    P0 = (v / np.sqrt(1 - (v ** 2).sum(0))).ravel()
    return p, dict(v=v.ravel().copy(), x=x.ravel().copy(), P0=P0, q=-np.ones(n), m=np.ones(n), An=An, Ae=Ae, gpm=gpm)


def _run(lib, p, d, noalloc, th=(0.0, 1.0)):
    ctx = lib.Context(p)
    dv = ctx.dev
    out = ctx.local_substep(dv(d["v"]), dv(d["x"]), dv(d["P0"]), dv(d["q"]), dv(d["m"]), p.dt * (th[1] - th[0]),
                            th[0], th[1], dv(d["An"]), dv(d["Ae"]), dv(d["gpm"]), noalloc=noalloc)
    ctx.close()
    return out


def test_noalloc_local_substep_matches_allocating_entry_point():
    from unstaggered_pic_amd import lib
    cases = [(64, 32, 1, 4099, 1, (0.0, 1.0)),   # first call: 2048 nodes of records
             (16, 8, 1, 513, 2, (0.25, 0.5)),    # second call: smaller mesh and particle set
             (8, 8, 4, 300, 3, (0.0, 1.0))]      # 3-D: no scratch
    for nx, ny, nz, n, seed, th in cases:
        print(f"noalloc local substep {nx}x{ny}x{nz} np={n} theta={th}", flush=True)
        p, d = _case(nx, ny, nz, n, seed)
        ref = _run(lib, p, d, False, th)
        got = _run(lib, p, d, True, th)
        assert (ref[4] == 0).all()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_noalloc_local_substep_rejects_meshes_beyond_the_scratch():
    from unstaggered_pic_amd import lib
    p = make_params(code=CODE_P3, nx=1024, ny=512, nz=1, Lx=6.283185307179586, Ly=6.283185307179586, Lz=1.0,
                    dt=0.1, spline_order=2, orbit_quad_order=3, nyquist_projection=1)
    ctx = lib.Context(p)
    n = 4
    dv = ctx.dev
    z3, z1 = np.zeros(3 * n), np.zeros(n)
    zN = np.zeros(3 * 1024 * 512)
    with pytest.raises(lib.UpicError) as e:
        ctx.local_substep(dv(z3), dv(z3), dv(z3), dv(z1), dv(np.ones(n)), 0.1, 0.0, 1.0, dv(zN), dv(zN), dv(zN),
                          noalloc=True)
    assert e.value.status == UPIC_EINVAL
    assert "static scratch" in str(e.value)
    ctx.close()
