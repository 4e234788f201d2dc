"""bdgx_gram_noalloc (the allocation-free Gram entry point) against
bdgx_gram, through the C ABI: bit-identical results on two calls in a row,
the second with a smaller input; the context's own workspace stays
unallocated; a graph captured around the call replays it (the callers it is
for cannot allocate, e.g. under capture); oversize calls are refused."""
import numpy as np
import pytest

from tests.common import rand_block

pytestmark = pytest.mark.gpu


def _gram_both(d, A, B, prec, swap_b=False):
    import bdg_lobpcg_amd as B_
    dt = d.dtype(prec)
    A = np.asfortranarray(A, dtype=dt)
    Bm = np.asfortranarray(B, dtype=dt)
    dA, dB = d.upload(A), d.upload(Bm)
    w = 2 if d.cplx else 1
    p, q, n = A.shape[1], Bm.shape[1], A.shape[0]
    dG = B_.DeviceBuffer(8 * w * p * q)
    rc = d.gram_noalloc_device(prec, dA, n, p, dB, n, q, n, swap_b, dG)
    assert rc == 0, B_.error_string(rc)
    got = d.download(dG, (p, q), d.dtype64)
    ref = d.gram(A, Bm, prec, swap_b)
    return got, ref


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_gram_matches_gram_twice_second_smaller(cplx):
    import bdg_lobpcg_amd as B
    N, L = (64, 64, 64), (12.0, 12.0, 12.0)
    d = B.DeviceContext(N, L, cplx)
    S = d.size
    print(f"noalloc gram cplx={cplx}: fresh context workspace {d.workspace_bytes()} B", flush=True)
    for prec in (B.FP64, B.FP32):
        for (p, q) in ((48, 32), (20, 7)):  # second call smaller
            A = rand_block(S, p, cplx, seed=p)
            Bm = rand_block(S, q, cplx, seed=q + 1)
            if d.workspace_bytes() == 0:
                # nothing allocated by the entry point itself
                dt = d.dtype(prec)
                dA = d.upload(np.asfortranarray(A, dtype=dt))
                dB = d.upload(np.asfortranarray(Bm, dtype=dt))
                dG = B.DeviceBuffer(8 * (2 if cplx else 1) * p * q)
                assert d.gram_noalloc_device(prec, dA, S, p, dB, S, q, S, 0, dG) == 0
                d.sync()
                assert d.workspace_bytes() == 0
            got, ref = _gram_both(d, A, Bm, prec)
            print(f"  prec {prec} {p}x{q}: max |diff| {np.abs(got - ref).max():.1e}", flush=True)
            np.testing.assert_array_equal(got, ref)
    # the B-Gram form (swapped halves of B)
    got, ref = _gram_both(d, rand_block(S, 16, cplx, 3), rand_block(S, 12, cplx, 4), B.FP64, True)
    np.testing.assert_array_equal(got, ref)
    d.close()


def test_noalloc_gram_under_graph_capture():
    """Captured once, replayed four times with eager work between the
    replays: every replay equals the eager call."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 32), (10.0, 10.0, 10.0)
    d = B.DeviceContext(N, L, False)
    S = d.size
    A = rand_block(S, 24, False, 5)
    Bm = rand_block(S, 16, False, 6)
    dA, dB = d.upload(A), d.upload(Bm)
    dG = B.DeviceBuffer(8 * 24 * 16)
    ref = d.gram(A, Bm)
    g = d.capture(lambda: d.gram_noalloc_device(B.FP64, dA, S, 24, dB, S, 16, S, 0, dG))
    for r in range(4):
        d.gram(A[:, :5], Bm[:, :3])  # eager work between the replays
        g.launch()
        d.sync()
        got = d.download(dG, (24, 16), np.float64)
        np.testing.assert_array_equal(got, ref)
    g.close()
    d.close()


def test_noalloc_gram_refuses_oversize():
    import bdg_lobpcg_amd as B
    N, L = (64, 64, 64), (12.0, 12.0, 12.0)
    d = B.DeviceContext(N, L, False)
    S = d.size
    p = q = 96  # 96 * 96 partials per chunk (2 chunks per CU) exceed the fixed scratch
    dA = B.DeviceBuffer(8 * S * p)
    dG = B.DeviceBuffer(8 * p * q)
    rc = d.gram_noalloc_device(B.FP64, dA, S, p, dA, S, q, S, 0, dG)
    assert rc == -5  # BDGX_ETOOBIG, before any launch
    assert "fixed scratch" in B.error_string(rc)
    d.close()
