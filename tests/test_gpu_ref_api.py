"""The reference's own public-API solver tests (reference
tests/test_bdg_1d_uniform.c, tests/test_bdg_3d_harmonic.c), compiled unchanged
against include/bdg/bdg.h and linked with libbdg_amd.so by tests/ref_api/Makefile,
run on the GPU: the drop-in check at the boundary.  The binaries are built in
this checkout while the reference sources are present; without them the test
is skipped."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "ref_api", "_bin")


@pytest.mark.parametrize("prog", ["test_bdg_1d_uniform", "test_bdg_3d_harmonic"])
def test_reference_api_program(prog):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.skip(f"{prog} not built (reference sources were not present at build time)")
    print(f"running {prog}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, cwd=HERE)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[FAIL]" not in r.stdout
    assert "[PASS]" in r.stdout
