"""The pins of the oracle fail on plausible mistakes: each planted variant of
the oracle (built from the same source with -DORC_PLANT=n, on the CPU) must
make the named pin tests fail, and the unplanted build must pass them."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC = os.path.join(ROOT, "oracle")

# planted mistake -> the pin file and the tests in it that must catch it
PLANTS = {
    1: ("tests/test_shock_sensor_oracle.py", "test_pressure_sensor_on_a_step"),
    2: ("tests/test_shock_sensor_oracle.py", "test_interior_face_amplitude_with_an_active_sensor"),
    3: ("tests/test_shock_sensor_oracle.py", "test_ducros_sensor_with_divergence"),
    # characteristic boundaries: the textbook sign of the imax density equation,
    # 0.5 for the 0.528 of the subsonic relaxation, v for v + c in the jmax
    # outgoing wave, the a04 rows for the a40 ones at imax
    4: ("tests/test_charac_oracle.py", "test_imax_subsonic_relaxation_at_rest"),
    5: ("tests/test_charac_oracle.py", "test_imax_subsonic_relaxation_at_rest"),
    6: ("tests/test_charac_oracle.py", "test_jmax_subsonic_outflow_and_inflow"),
    7: ("tests/test_charac_oracle.py", "test_imax_supersonic_outflow_and_closures"),
    # c3 walls: the tangential term of the wall extrapolation with the sign a
    # second-order dp/dn = 0 would give, the two tangents of the adiabatic
    # gradient exchanged
    8: ("tests/test_c3_wall_oracle.py", "test_sheared_wall_state_is_the_closed_form"),
    9: ("tests/test_c3_wall_oracle.py", "test_unit_c3_walls_equal_cartesian_walls"),
    # advanced walls (is_wall2): a02 at the max face without the sign of a20,
    # the unit tangents of a c3 slip wall for their dual basis
    10: ("tests/test_wall2_oracle.py", "test_advanced_rows_take_the_exact_derivative"),
    11: ("tests/test_wall2_oracle.py", "test_c3_slip_wall_keeps_exactly_the_tangential_momentum"),
}


def _build(n):
    out = os.path.join(ORC, "_build", f"plant_{n}.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    f"-DORC_PLANT={n}", "-o", out, os.path.join(ORC, "musica_oracle.c"), "-lm"], check=True)
    return out


def _run(path, test, lib):
    env = dict(os.environ)
    if lib:
        env["MUSICA_ORACLE_LIB"] = lib
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", f"{path}", "-k", test],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.mark.parametrize("n", sorted(PLANTS))
def test_planted_mistake_fails_its_pin(n):
    path, test = PLANTS[n]
    rc, out = _run(path, test, _build(n))
    assert rc == 1 and "failed" in out, out  # pytest exit 1: tests ran and some failed


def test_unplanted_oracle_passes_the_same_pins():
    for path, test in set(PLANTS.values()):
        rc, out = _run(path, test, None)
        assert rc == 0, out
