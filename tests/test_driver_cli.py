"""The deck driver's error behaviour (CPU: every case fails before the
first GPU call).  Exit codes as the reference's mains: 2 for usage
(P1:1776-1780, P2:1196-1199, P3:8325-8329), 1 with the reference's message
for a deck it rejects (P2 read_config 281-290, P3 initialisers 3715-3717,
4004-4006, 3871-3872)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "unstaggered_pic_amd", "upic_run")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="upic_run not built")


def run(tmp_path, *args, deck=None):
    if deck is not None:
        (tmp_path / "deck.txt").write_text(deck)
        args = args + (str(tmp_path / "deck.txt"),)
    return subprocess.run([EXE, *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)


def test_usage_exit_code(tmp_path):
    assert run(tmp_path).returncode == 2
    assert run(tmp_path, "p4", "deck.txt").returncode == 2


def test_missing_deck(tmp_path):
    r = run(tmp_path, "p2", str(tmp_path / "absent.txt"))
    assert r.returncode == 1 and "could not open input file" in r.stderr


@pytest.mark.parametrize("body,message", [
    ("method = bdf3", "method must be bdf1 or bdf2"),
    ("charge_update = sometimes", "charge_update must be conserving or naive"),
    ("example = plasma", "example must be weibel or cloud"),
    ("nx = 24", "powers of two"),
])
def test_p2_deck_checks(tmp_path, body, message):
    r = run(tmp_path, "p2", deck=body + "\n")
    assert r.returncode == 1 and message in r.stderr, r.stderr


@pytest.mark.parametrize("body,message", [
    ("test_case = cold_relativistic_two_stream\ntwo_stream_density_alpha = 0.3",
     "two_stream_density_alpha is a linear-mode amplitude and should be < 0.25"),
    ("test_case = weibel\nv0 = 1.0", "Weibel initialization requires |v0| < kappa"),
    ("test_case = weibel\nv0 = 0.9\nweibel_B0 = 0.5", "Weibel eigenmode density perturbation is too large"),
    ("test_case = landau_weak\nparticles_per_cell_pair = 1",
     "Landau examples need at least two velocity samples per cell"),
    ("test_case = landau_weak\nlandau_alpha = 1.5", "Landau density amplitude must satisfy |alpha| < 1"),
    ("test_case = nonsense", "unknown test_case/initial_condition"),
])
def test_p3_initial_condition_checks(tmp_path, body, message):
    r = run(tmp_path, "p3", deck=body + "\nnx = 8\nny = 8\nn_steps = 1\n")
    assert r.returncode == 1 and message in r.stderr, r.stderr


# ---------------------------------------------------------------------------
# The P3 deck reading against the reference's read_config (P3:1178-1435):
# the config echo (write_config_echo, P3:8217-8300) of the same deck, written
# by the driver before its first GPU call, equals the reference's line for
# line (compiled from its sources in oracle/_ref; skipped where absent).
# ---------------------------------------------------------------------------
def ref_echo(tmp_path, deck):
    import ctypes as C
    from oracle_bind import Ref
    ref = Ref()
    if ref.p3 is None:
        pytest.skip("reference harness (oracle/_ref) not built here")
    ref.p3.ref_p3_last_error.restype = C.c_char_p
    path = tmp_path / "ref_deck.txt"
    path.write_text(deck)
    out = tmp_path / "ref_echo.txt"
    rc = ref.p3.ref_p3_config_echo(str(path).encode(), str(out).encode())
    assert rc == 0, ref.p3.ref_p3_last_error()
    return out.read_text().splitlines()


def driver_echo(tmp_path, deck):
    r = run(tmp_path, "p3", deck=deck)
    prefix = "orbit_smoke"
    for line in deck.splitlines():
        if line.split("=")[0].strip() == "output_prefix":
            prefix = line.split("=", 1)[1].strip()
    return r, (tmp_path / (prefix + "_config_echo.txt")).read_text().splitlines()


@pytest.mark.parametrize("deck", [
    "",  # every key at the reference's Config default (P3:837-990)
    "test_case = Cold_Relativistic_Weibel\nv0 = 0.9\nweibel_B0 = 1e-3\nnx = 8\nny = 8\nnz = 1\n",
    # the manuscript decks' output switches are echoed as set
    "test_case = cold_relativistic_two_stream\nv0 = 0.9\ntwo_stream_density_alpha = 0.005\n"
    "transverse_diagnostics = true\ntransverse_spectrum_interval = 1.0\nselective_failure_diagnostics = true\n"
    "field_diagnostics = true\nfield_diagnostics_interval = 10.0\nparticle_snapshot_times = 0, 20, 40\n"
    "finite_grid_theory = false\nweibel_finite_grid_theory = true\n",
])
def test_p3_echo_matches_reference(tmp_path, deck):
    want = ref_echo(tmp_path, deck)
    r, got = driver_echo(tmp_path, deck)
    assert got == want, "\n".join(f"{a} | {b}" for a, b in zip(got, want) if a != b)
    if "transverse_diagnostics = true" in deck:
        assert "upic_run does not write these files" in r.stderr


def test_p3_deprecated_coefficient_keys(tmp_path):
    """c, eps0 and mu0 set kappa, sigma1 = 1/eps0 and sigma2 with the
    reference's deprecation warnings (P3:1308-1315)."""
    deck = "c = 2.0\neps0 = 0.25\nmu0 = 1.0\n"
    want = ref_echo(tmp_path, deck)
    r, got = driver_echo(tmp_path, deck)
    assert got == want
    assert "kappa=2" in got and "sigma1=4" in got and "sigma2=1" in got
    for k in ("c is deprecated", "eps0 is deprecated", "mu0 is deprecated"):
        assert k in r.stderr


@pytest.mark.parametrize("first,second", [("particles_per_cell", "particles_per_cell_pair"),
                                          ("particles_per_cell_pair", "particles_per_cell")])
def test_p3_alias_set_last_wins(tmp_path, first, second):
    """Two aliases of one key: the line set last wins, as in the reference's
    line-by-line parser."""
    deck = f"{first} = 2\n{second} = 3\n"
    want = ref_echo(tmp_path, deck)
    _, got = driver_echo(tmp_path, deck)
    assert got == want
    assert "particles_per_cell_pair_or_velocity_samples=3" in got


def test_p3_manuscript_controls_off(tmp_path):
    """The reference's manuscript driver refuses adaptive dt and uniform
    subcycling (P3:8344-8345)."""
    for body in ("adaptive_dt = true", "uniform_particle_subcycling = true"):
        r = run(tmp_path, "p3", deck=body + "\n")
        assert r.returncode == 1 and "manuscript runs require adaptive_dt=false" in r.stderr, r.stderr
