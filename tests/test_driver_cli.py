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
