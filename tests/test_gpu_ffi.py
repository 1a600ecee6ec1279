"""C programs against libgsplat_ffi_c.so on the GPU: this repository's C
smoke and, when it was built here from the reference sources, the
reference's own tests/ffi/ffi_smoke.c (compiled against the reference's
gsplat.h), both run unchanged."""

import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
SCENE = ROOT / "tests/golden/minimal_ascii.ply"


def run(binary):
    print(f"[ffi] {binary.name}", flush=True)
    p = subprocess.run([str(binary), str(SCENE)], cwd=str(ROOT), capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    return p


def test_repository_c_smoke(gpu):
    binary = ROOT / "gsplat-rs_amd/_build/ffi_smoke_mi355x"
    assert binary.exists(), "build() compiles the C smoke"
    p = run(binary)
    assert p.returncode == 0
    assert "ffi smoke ok" in p.stdout and "drawn=2 visible=2" in p.stdout


def test_reference_c_smoke_runs_unchanged(gpu):
    binary = ROOT / "oracle/_ref/ffi_smoke"
    if not binary.exists():
        pytest.skip("reference C smoke not built (reference sources absent at build time)")
    p = run(binary)
    assert p.returncode == 0, p.stderr
    assert "ffi smoke ok" in p.stdout and "drawn=2 visible=2" in p.stdout
