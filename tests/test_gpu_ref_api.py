"""The reference's own public-API solver tests (reference
tests/test_bdg_1d_uniform.c, tests/test_bdg_3d_harmonic.c), compiled unchanged
against include/bdg/bdg.h and linked with libbdg_amd.so by tests/ref_api/Makefile,
run on the GPU: the drop-in check at the boundary.  The binaries are built in
this checkout while the reference sources are present; without them the test
is skipped."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "ref_api", "_bin")


def _run(prog):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.skip(f"{prog} not built (reference sources were not present at build time)")
    print(f"running {prog}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, cwd=HERE)
    print(r.stdout[-3000:], flush=True)
    return r


def test_reference_1d_uniform():
    """Bogoliubov dispersion of the uniform 1D gas, d and z variants
    (reference tests/test_bdg_1d_uniform.c): every assertion of the program."""
    r = _run("test_bdg_1d_uniform")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[FAIL]" not in r.stdout and "[PASS]" in r.stdout


def test_reference_3d_harmonic():
    """Non-interacting 3D oscillator (reference tests/test_bdg_3d_harmonic.c).
    The program asserts convergence of nev = 10 modes, evals[0] ~ 0 within
    0.1, then evals[j] ~ 1.0 within 5% for every j in 1..9 (lines 97-104).
    The oscillator's spectrum is 0, 1 (x3), 2 (x6), so that last loop can only
    hold for j <= 3; its first failure is expected at j = 4 with a value near
    2.  Checked here: convergence, the Goldstone value, the 1-triplet, and
    that the program stops exactly at j = 4 on the 2-shell."""
    r = _run("test_bdg_3d_harmonic")
    out = r.stdout
    m = re.search(r"converged:\s*(\d+)\s*/\s*(\d+)", out)
    assert m and int(m.group(1)) >= int(m.group(2)) == 10, out[-1500:]
    ev = {int(j): float(v) for j, v in re.findall(r"evals\[(\d+)\]\s*=\s*([-0-9.eE+]+)", out)}
    assert abs(ev[0]) < 0.1
    for j in (1, 2, 3):
        assert abs(ev[j] - 1.0) < 0.05
    assert abs(ev[4] - 2.0) < 0.05
    assert set(ev) == {0, 1, 2, 3, 4}
