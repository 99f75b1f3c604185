"""Flag handling of the scene caller (cli/scene_plume.hip) needs no GPU: --help
prints the usage and succeeds, an unknown flag, a flag without its value and
an empty grid fail before any HIP call."""
import subprocess

import pytest

import smoke_abi as abi

EXE = abi.LIB_PATH.parent / "visual-simulation-of-smoke-scene"


def _run(*args):
    if not EXE.exists():
        pytest.skip("scene caller not built")
    return subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=60)


def test_scene_cli_flags():
    r = _run("--help")
    assert r.returncode == 0 and r.stdout.startswith("visual-simulation-of-smoke-scene [--nx N]")
    r = _run("--bogus")
    assert r.returncode != 0 and "unknown argument: --bogus" in r.stderr
    r = _run("--steps")
    assert r.returncode != 0 and "missing value for --steps" in r.stderr
    r = _run("--nx", "0")
    assert r.returncode != 0 and "invalid grid or step count" in r.stderr
