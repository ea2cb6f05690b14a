"""The drop-in's problem-setup hooks (hpcg_amd/dropin/setup_hooks.cpp:
GenerateProblem and SetupHalo, threaded, one block per matrix array) against
the reference's GenerateProblem_ref / SetupHalo_ref and CheckProblem, host
only (oracle/_ref/setup_check, built from the reference's sources by
oracle/Makefile under AddressSanitizer).  Every array, map, count and vector
of every level must be equal; the reference's driver sequence on the drop-in
hooks must pass CheckProblem on all 4 levels; DeleteMatrix must free both
storage forms."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "setup_check")


@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 16, 8), (8, 40, 16), (64, 64, 64)])
def test_setup_hooks_match_reference(dims):
    if not os.path.exists(CHECK):
        pytest.skip("setup_check not built (needs the reference sources: make -C oracle ref)")
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1")
    p = subprocess.run([CHECK, *map(str, dims)], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "setup_check ok" in p.stdout
