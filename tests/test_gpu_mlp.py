"""Learned initial guess of the P3 inner solve (no reference counterpart):
the network's device kernels against the FP64 oracle, and the guess inside
the mixed-precision step.

Bounds:
  * training loss and gradient (FP32 kernel) vs the oracle (FP64): 1e-4 of
    the largest gradient entry, loss to 1e-5 relative;
  * MFMA inference (BF16 operands, FP32 accumulation) vs the oracle run on
    BF16-rounded weights and activations: 2e-3 of the output scale (FP32
    accumulation order and the device exp), and vs the exact FP64 network
    within the BF16 quantisation error (5e-2 of the output scale);
  * the step with a learned or predicted guess converges to the same fixed
    point as the plain start (energies within 1e-6 relative); the learned
    guess takes fewer inner iterations per particle and step than the
    predictor alone, which takes fewer than the plain start.
"""
import numpy as np
import pytest

from oracle_bind import Oracle, State
from unstaggered_pic_amd.abi import CODE_P3, LOAD_BIMAXWELL, make_load, make_params, with_mixed_precision

pytestmark = pytest.mark.gpu

MLP_N = 5214
W1, B1, W2, B2, W3, B3 = 27, 795, 859, 4955, 5019, 5211


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def gpu():
    from unstaggered_pic_amd import lib
    return lib


def random_params(rng, scale=0.3):
    P = np.zeros(MLP_N, np.float32)
    P[0:12] = rng.normal(size=12) * 0.1
    P[12:24] = 1.0 + rng.random(12)
    P[24:27] = 0.5 + rng.random(3)
    P[W1:] = rng.normal(size=MLP_N - W1) * scale
    return P


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(P, X):
    """The network with BF16 weights, inputs and hidden activations and wide
    accumulation (what the MFMA path computes up to summation order)."""
    silu = lambda z: z / (1.0 + np.exp(-z))
    x = bf16(((X.astype(np.float32) - P[0:12]) * P[12:24]).astype(np.float32))
    w1 = bf16(P[W1:B1]).reshape(64, 12)
    w2 = bf16(P[W2:B2]).reshape(64, 64)
    w3 = bf16(P[W3:B3]).reshape(3, 64)
    h1 = bf16(silu(x @ w1.T + P[B1:W2]).astype(np.float32))
    h2 = bf16(silu(h1 @ w2.T + P[B2:W3]).astype(np.float32))
    return (h2 @ w3.T + P[B3:B3 + 3]) * P[24:27]


def test_mlp_loss_grad_vs_oracle(orc):
    lib = gpu()
    rng = np.random.default_rng(3)
    P = random_params(rng)
    rows = np.zeros((1000, 15), np.float32)
    rows[:, :12] = rng.normal(size=(1000, 12))
    rows[:, 12:] = rng.normal(size=(1000, 3))
    ctx = lib.Context(make_params(code=CODE_P3, nx=8, ny=8))
    G, L = ctx.mlp_loss_grad(P, rows)
    Go, Lo = orc.mlp_loss_grad(P, rows)
    assert abs(L - Lo) / Lo < 1e-5
    assert np.max(np.abs(G - Go)) < 1e-4 * np.max(np.abs(Go))
    assert np.all(G[:W1] == 0.0)
    ctx.close()


@pytest.mark.parametrize("n", [64, 1000])
def test_mlp_forward_vs_oracle(orc, n):
    lib = gpu()
    rng = np.random.default_rng(11 + n)
    P = random_params(rng, scale=0.25)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    ctx = lib.Context(make_params(code=CODE_P3, nx=8, ny=8))
    Y = ctx.mlp_forward(P, X)
    Yo = orc.mlp_forward(P, X)
    Yb = forward_bf16(P, X)
    scale = np.max(np.abs(Yo))
    assert np.max(np.abs(Y - Yb)) < 2e-3 * scale
    assert np.max(np.abs(Y - Yo)) < 5e-2 * scale
    ctx.close()


def weibel_small(nx=32, ppc=32):
    L = 20.0 * np.pi * nx / 256.0  # the cell width of the 256^2 Weibel deck
    p = make_params(code=CODE_P3, nx=nx, ny=nx, nz=1, Lx=L, Ly=L, Lz=1.0, dt=0.05, spline_order=3,
                    orbit_quad_order=4, nyquist_projection=1, nonlinear_rtol=1e-11, nonlinear_atol=1e-13,
                    nonlinear_max_iter=80, anderson_damping=0.5, local_rtol=1e-12, local_atol=1e-14,
                    local_max_iter=64)
    return with_mixed_precision(p), nx * nx * ppc, ppc


def test_learned_guess_same_fixed_point_fewer_iterations():
    """Collect rows from 12 mixed steps, train on the device, then run 4 steps
    from the same state with the plain start v^n, the physics predictor and
    predictor + network: the three runs agree to the solver tolerance and the
    inner iterations per particle and step fall from plain to predictor to
    learned."""
    lib = gpu()
    p, n, ppc = weibel_small()
    x, v, q, m = lib.load_particles(p, make_load(LOAD_BIMAXWELL, ppc, seed=12345, vth=(0.05, 0.2, 0.2)), n)
    st = State(n, p.nx * p.ny)
    st.x[:], st.v[:], st.q[:], st.m[:] = x, v, q, m
    ctx = lib.Context(p)
    sim = lib.Simulation(ctx, n)
    sim.upload(st)
    sim.init_fields(1.0)
    sim.collect_guess_data(4, 12 * n // 4)
    for _ in range(12):
        sim.step()
    rows = sim.guess_data()
    assert rows.shape[0] == 12 * n // 4 and np.all(np.isfinite(rows))
    losses = sim.train_guess(epochs=30, batch=2048, lr=2e-3, seed=5)
    assert losses[-1] < 0.5 * losses[0]
    params = sim.guess_params()
    mid = State(n, p.nx * p.ny)
    sim.download(mid)
    sim.close()
    results = {}
    for mode in (0, 1, 2):
        s2 = lib.Simulation(ctx, n)
        s2.upload(mid.copy())
        if mode == 2:
            s2.set_guess_params(params)
        s2.set_guess_mode(mode)
        its, te = [], []
        for _ in range(4):
            info = s2.step()
            # inner iterations per particle summed over the step's sweeps: the
            # guess acts in the first sweep, the later ones start from the
            # previous sweep's converged vbar in modes 1 and 2
            its.append(info.local_iterations_total / n)
            te.append(s2.diagnostics(1).total_energy)
        results[mode] = (np.array(its), np.array(te))
        s2.close()
    ctx.close()
    it0, te0 = results[0]
    for mode in (1, 2):
        itm, tem = results[mode]
        assert np.max(np.abs(tem - te0) / np.abs(te0)) < 1e-6, mode
    # the network's gain over the predictor alone, and the predictor's over v^n
    assert np.mean(results[2][0]) < np.mean(results[1][0]) < np.mean(it0)
    print("inner iterations per particle per step plain/predictor/learned:",
          [float(np.mean(results[k][0])) for k in (0, 1, 2)])
