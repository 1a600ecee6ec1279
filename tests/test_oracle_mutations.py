"""Self-check of the oracle's pins: each planted mistake must make a pin fail.

oracle/oracle.cpp carries switches (ORACLE_MUTATION = 1..8) that each plant one
plausible mistake in the projection or the tile rectangle. This test builds the
oracle with each switch into a temporary directory, runs the pinning tests that
should catch it against that build (GSPLAT_ORACLE_LIB), and requires them to
fail; the unmutated build must pass the same tests. CPU only.
"""

import os
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PINS = "tests/test_oracle_projection.py"

MUTATIONS = {
    1: ("sign of j02 flipped", "test_offaxis_axes_match_finite_difference_jacobian"),
    2: ("aspect dropped from lim_x", "test_jacobian_clamp_at_frustum_edge or test_jacobian_unclamped_inside"),
    3: ("64x anisotropy cap removed", "test_anisotropy_cap"),
    4: ("upper radius clamp removed", "test_radius_upper_clamp"),
    5: ("NDC-box cull on |axis_u.x| only", "test_ndc_box_cull_edges"),
    6: ("blur term left out", "test_offaxis_axes_match_finite_difference_jacobian"),
    7: ("tile rect from the quad extent", "test_tile_rect_two_sided"),
    8: ("lower radius clamp removed", "test_radius_lower_clamp"),
}


def build(tmp: pathlib.Path, k: int) -> pathlib.Path:
    out = tmp / f"liboracle_m{k}.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                    f"-DORACLE_MUTATION={k}", "-o", str(out), str(ROOT / "oracle/oracle.cpp")], check=True)
    return out


def run_pins(lib: pathlib.Path, select: str) -> subprocess.CompletedProcess:
    env = dict(os.environ, GSPLAT_ORACLE_LIB=str(lib))
    return subprocess.run([sys.executable, "-m", "pytest", PINS, "-q", "-p", "no:cacheprovider", "-k", select],
                          cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module")
def tmp_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutations")


def test_unmutated_build_passes_the_pins(tmp_dir):
    r = run_pins(build(tmp_dir, 0), " or ".join(sel for _, sel in MUTATIONS.values()))
    assert r.returncode == 0, r.stdout[-3000:]


@pytest.mark.parametrize("k", sorted(MUTATIONS))
def test_planted_mistake_fails_a_pin(tmp_dir, k):
    what, select = MUTATIONS[k]
    r = run_pins(build(tmp_dir, k), select)
    assert r.returncode == 1 and " failed" in r.stdout, f"mutation {k} ({what}) passed every pin:\n{r.stdout[-3000:]}"
