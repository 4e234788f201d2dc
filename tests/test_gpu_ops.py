"""GPU parity: every operator of the hot path (reference src/*_impl.inc) on the
device against the FP64 CPU oracle, for real (_d) and complex (_z) variants,
in FP64 (relative max error <= 1e-12) and FP32 (<= 2e-5)."""
import numpy as np
import pytest

from tests.common import device_twin, physics, rand_block, relerr

pytestmark = pytest.mark.gpu

GRIDS = [
    ((16, 16, 16), (12.0, 12.0, 12.0)),
    ((12, 10, 8), (10.0, 9.0, 8.0)),      # radix 3, 5, 2/4
    ((32, 16, 24), (16.0, 12.0, 14.0)),
    ((64,), (2 * np.pi,)),
    ((16, 12), (8.0, 6.0)),
    ((32, 32, 12), (16.0, 16.0, 10.0)),    # whole-plane x/y kernels (real)
    ((64, 64, 6), (16.0, 16.0, 6.0)),
    ((128, 128, 4), (24.0, 24.0, 4.0)),
    ((16, 12, 32), (12.0, 10.0, 16.0)),    # fixed-size z pass: Nyquist-column tiles
    ((32, 32, 64), (16.0, 16.0, 24.0)),    # plane kernels + fixed z pass
    ((24, 8, 128), (12.0, 8.0, 30.0)),     # fixed z pass, partial kx tiles
    ((48, 48, 80), (16.0, 16.0, 20.0)),    # plane and fixed z kernels, radix 3 / 5 sizes
    ((80, 80, 48), (18.0, 18.0, 12.0)),
    ((96, 96, 8), (20.0, 20.0, 8.0)),
    ((12, 10, 192), (10.0, 9.0, 60.0)),    # fixed z pass, Nz = 192
]
OPS_HALF = ["kinetic", "K", "M", "precondK", "precondM", "dipolar_conv"]
OPS_FULL = ["Lrep", "Swap", "precondLrep"]
TOL = {0: 1e-12, 1: 2e-5}


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g[0])))
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_operator_parity(grid, cplx, prec):
    N, L = grid
    c, f = physics(N, L, cplx, dipolar=True)
    d = device_twin(N, L, cplx, f)
    S = c.size
    print(f"case N={N} cplx={cplx} prec={prec}", flush=True)
    for op in OPS_HALF + OPS_FULL:
        if op == "dipolar_conv" and not f["dipolar"]:
            continue
        rows = 2 * S if op in OPS_FULL else S
        X = rand_block(rows, 3, cplx, seed=hash(op) % 1000)
        ref = c.apply(op, X)
        got = d.apply(op, X, prec)
        e = relerr(got, ref)
        assert e < TOL[prec], f"{op} N={N} cplx={cplx} prec={prec}: rel err {e:.3e}"


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("N", [(16, 12, 10), (32, 32, 10)], ids=["pencil", "plane"])
def test_dipolar_kernel_and_meanfield(cplx, N):
    L = (12.0, 10.0, 9.0)
    c, f = physics(N, L, cplx, dipolar=True)
    d = device_twin(N, L, cplx, f)
    assert relerr(d.dipolar_kernel(), c.dipolar_kernel()) < 1e-13
    assert relerr(d.dipolar_meanfield(), c.dipolar_meanfield()) < 1e-12


def test_large_column_count_and_inplace():
    """Many columns in one launch; in-place apply equals out-of-place."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 32), (16.0, 16.0, 16.0)
    c, f = physics(N, L, False, dipolar=True)
    d = device_twin(N, L, False, f)
    S = c.size
    X = rand_block(2 * S, 12, False, seed=3)
    ref = c.apply("Lrep", X)
    dX = d.upload(X)
    d.apply_device("Lrep", B.FP64, dX, dX, 12, 2 * S)
    got = d.download(dX, X.shape, np.float64)
    assert relerr(got, ref) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_gram_combine(cplx, prec):
    import bdg_lobpcg_amd as B
    N, L = (16, 16, 8), (8.0, 8.0, 8.0)
    d = B.DeviceContext(N, L, cplx)
    S = d.size
    A = rand_block(2 * S, 70, cplx, 1)
    Bm = rand_block(2 * S, 45, cplx, 2)
    dt = d.dtype(prec)
    A32, B32 = A.astype(dt), Bm.astype(dt)
    ref = A32.astype(np.complex128 if cplx else np.float64).conj().T @ B32.astype(
        np.complex128 if cplx else np.float64)
    G = d.gram(A, Bm, prec)
    assert relerr(G, ref) < 1e-12
    # swapped halves
    Bsw = np.concatenate([B32[S:], B32[:S]], axis=0).astype(ref.dtype)
    G2 = d.gram(A, Bm, prec, swap_b=True)
    assert relerr(G2, A32.astype(ref.dtype).conj().T @ Bsw) < 1e-12
    Cm = rand_block(70, 33, cplx, 4)
    Y = d.combine(A, Cm, prec)
    refY = A32.astype(ref.dtype) @ Cm
    assert relerr(Y, refY) < (1e-13 if prec == 0 else 1e-6)


@pytest.mark.parametrize("grid", [((32, 32, 64), (16.0, 16.0, 24.0)),
                                  ((16, 12, 32), (12.0, 10.0, 16.0)),
                                  ((128, 128, 4), (24.0, 24.0, 4.0)),
                                  ((48, 48, 80), (16.0, 16.0, 20.0)),
                                  ((80, 80, 48), (18.0, 18.0, 12.0)),
                                  ((96, 96, 8), (20.0, 20.0, 8.0)),
                                  ((32, 32, 128), (16.0, 16.0, 30.0))],
                         ids=lambda g: "x".join(map(str, g[0])))
def test_launch_variants_and_autotune(grid):
    """Every launch variant of the plane and fixed z kernels gives the
    oracle's result; the autotuner picks valid variants."""
    import bdg_lobpcg_amd as B
    N, L = grid
    c, f = physics(N, L, False, dipolar=True)
    d = device_twin(N, L, False, f)
    X = rand_block(c.size, 3, False, seed=5)
    refs = {op: c.apply(op, X) for op in ("K", "M", "precondK")}
    for prec in (0, 1):
        for kind in range(3):
            # planes: 4 shapes (one-plane / persistent, 4N / 8N threads) x 2
            # radix plans; z: 2 shapes x 2 radix plans (the second plan
            # exists at N = 128 only; elsewhere those variants repeat)
            for v in (range(8) if kind <= 1 else range(4)):
                d.set_launch_variant(kind, prec, v)
                for op, ref in refs.items():
                    e = relerr(d.apply(op, X, prec), ref)
                    assert e < TOL[prec], f"kind {kind} variant {v} prec {prec} {op}: {e:.2e}"
        rng = np.random.default_rng(3)
        A = rng.standard_normal((c.size, 20))
        Cm = rng.standard_normal((20, 36))
        for v in (0, 1, 2, 3):  # block combine; 2, 3 persistent
            d.set_launch_variant(3, prec, v)
            e = relerr(d.combine(A, Cm, prec), A @ Cm)
            assert e < (1e-14 if prec == 0 else 1e-6), f"combine variant {v} prec {prec}: {e:.2e}"
        # the solver's main shape (k = 96: whole groups of in-flight steps, so
        # the persistent kernels prefetch across row tiles): identical bits
        A96 = rng.standard_normal((c.size, 96))
        C96 = rng.standard_normal((96, 64))
        outs = []
        for v in (0, 1, 2, 3):
            d.set_launch_variant(3, prec, v)
            outs.append(d.combine(A96, C96, prec))
        e = relerr(outs[0], A96 @ C96)
        assert e < (1e-14 if prec == 0 else 1e-6), f"combine 96 -> 64 prec {prec}: {e:.2e}"
        for v in (1, 2, 3):
            assert np.array_equal(outs[v], outs[0]), f"combine variant {v} differs from 0"
        Ag = rng.standard_normal((c.size, 90))
        Bg = rng.standard_normal((c.size, 20))
        gs = []
        for v in (0, 1, 2, 3):
            d.set_launch_variant(4, prec, v)  # Gram tile height 96 / 48 / 32 / 64
            G = d.gram(Ag, Bg, prec)
            ref = (Ag.astype(np.float32 if prec else np.float64).astype(np.float64).T
                   @ Bg.astype(np.float32 if prec else np.float64).astype(np.float64))
            e = relerr(G, ref)
            assert e < 1e-12, f"gram variant {v} prec {prec}: {e:.2e}"
            gs.append(G)
        for v in (1, 2, 3):  # same sums in the same order
            assert np.array_equal(gs[v], gs[0]), f"gram variant {v} differs from 0"
        chosen = d.autotune(prec, 8, force=True)
        assert chosen["plane_fwd"] in range(8) and chosen["plane_inv"] in range(8)
        assert chosen["combine"] in (0, 1, 2, 3)
        assert chosen["z"] in range(4) and chosen["gram"] in (0, 1, 2, 3)
        for op, ref in refs.items():
            assert relerr(d.apply(op, X, prec), ref) < TOL[prec]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_graph_replay_matches_eager(prec):
    """A captured chain (operator applies + Gram) replayed several times, with
    the program's own eager work and synchronisations between the replays,
    reproduces the eager result exactly on every replay."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 64), (16.0, 16.0, 24.0)
    c, f = physics(N, L, False, dipolar=True)
    d = device_twin(N, L, False, f)
    S, n = c.size, 4
    dt = np.float64 if prec == 0 else np.float32
    X = np.asfortranarray(rand_block(S, n, False, seed=11).astype(dt))
    dX = d.upload(X)
    dY, dZ = B.DeviceBuffer(X.nbytes), B.DeviceBuffer(X.nbytes)
    dG = B.DeviceBuffer(8 * n * n)

    def chain():
        d.apply_device("precondK", prec, dX, dY, n, S)
        d.apply_device("M", prec, dY, dZ, n, S)
        d.apply_device("K", prec, dZ, dY, n, S)
        B.check(B.lib().bdgx_gram(d.h, prec, dX.ptr, S, n, dY.ptr, S, n, S, 0, dG.ptr), "gram")

    chain()
    d.sync()
    Y_ref = d.download(dY, (S, n), dt)
    G_ref = d.download(dG, (n, n), np.float64)
    g = d.capture(chain)
    junk = d.upload(np.asfortranarray(rand_block(S, n, False, seed=12).astype(dt)))
    for r in range(4):
        # eager work in between: overwrite the outputs, apply another operator
        d.apply_device("precondM", prec, junk, dY, n, S)
        d.apply_device("kinetic", prec, junk, dZ, n, S)
        d.sync()
        g.launch()
        d.sync()
        Y = d.download(dY, (S, n), dt)
        G = d.download(dG, (n, n), np.float64)
        assert np.array_equal(Y, Y_ref), f"replay {r}: operator chain differs"
        assert np.array_equal(G, G_ref), f"replay {r}: Gram differs"
    g.close()


def test_column_chunking_and_reserve():
    """Operator applies split into column chunks (bdgx_set_chunk_bytes) give
    the unchunked result exactly; bdgx_reserve grows the workspaces."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 64), (16.0, 16.0, 24.0)
    c, f = physics(N, L, False, dipolar=True)
    d = device_twin(N, L, False, f)
    X = rand_block(c.size, 7, False, seed=21)
    full = {op: d.apply(op, X, 0) for op in ("K", "M", "precondM")}
    kk = c.size // 32 * 17 * 16  # spectrum bytes of one FP64 channel column
    B.check(B.lib().bdgx_set_chunk_bytes(d.h, 2 * kk), "chunk")
    for op, ref in full.items():
        assert np.array_equal(d.apply(op, X, 0), ref), op
    B.check(B.lib().bdgx_set_chunk_bytes(d.h, 0), "chunk")
    B.check(B.lib().bdgx_reserve(d.h, 64 << 20, 1 << 20), "reserve")
    assert B.lib().bdgx_workspace_bytes(d.h) >= (64 << 20)
    for op, ref in full.items():
        assert np.array_equal(d.apply(op, X, 0), ref), op


def test_residual_on_selected_columns():
    """bdgx_residual_cols gives the selected columns of bdgx_residual (packed
    output, eigenvalues in selection order) bit for bit."""
    import ctypes as C
    import bdg_lobpcg_amd as B
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, False, dipolar=True)
    d = device_twin(N, L, False, f)
    S, m = c.size, 9
    rng = np.random.default_rng(4)
    AX, X, MX = (np.asfortranarray(rng.standard_normal((S, m))) for _ in range(3))
    lam = rng.standard_normal(m)
    dAX, dX, dMX, dlam = d.upload(AX), d.upload(X), d.upload(MX), d.upload(lam)
    dR, dn = B.DeviceBuffer(8 * S * m), B.DeviceBuffer(8 * 3 * m)
    B.check(B.lib().bdgx_residual(d.h, 0, dAX.ptr, dX.ptr, dMX.ptr, S, m, S, dlam.ptr, 0, dR.ptr,
                                  S, dn.ptr), "residual")
    R = d.download(dR, (S, m), np.float64)
    n2 = d.download(dn, (3 * m,), np.float64)
    np.testing.assert_allclose(R, AX - X * lam[None, :], rtol=0, atol=1e-14)
    sel = np.array([1, 4, 5, 8], dtype=np.int32)
    dls = d.upload(lam[sel])
    dRs, dns = B.DeviceBuffer(8 * S * len(sel)), B.DeviceBuffer(8 * 3 * len(sel))
    B.lib().bdgx_residual_cols.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                           C.c_uint64, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_uint64, C.c_void_p]
    B.check(B.lib().bdgx_residual_cols(d.h, 0, dAX.ptr, dX.ptr, dMX.ptr, S,
                                       sel.ctypes.data_as(C.c_void_p), len(sel), S, dls.ptr, 0,
                                       dRs.ptr, S, dns.ptr), "residual_cols")
    Rs = d.download(dRs, (S, len(sel)), np.float64)
    ns = d.download(dns, (3 * len(sel),), np.float64)
    assert np.array_equal(Rs, R[:, sel])
    assert np.array_equal(ns.reshape(-1, 3), n2.reshape(-1, 3)[sel])


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_fill_uniform_matches_oracle_start_block(cplx):
    """bdgx_fill_uniform_xorshift (jump-ahead, one thread per 64 values)
    equals the v half of the oracle's sequential uniform start block (u = v,
    the stream advancing S values per column) bit for bit, including a
    partial last run and several columns."""
    import bdg_lobpcg_amd as B
    N, L = (24, 20, 18), (10.0, 9.0, 8.0)
    c, f = physics(N, L, cplx)
    S, m = c.size, 7
    X = c.start_block_uniform(42, m)
    d = B.DeviceContext(N, L, cplx)
    V = d.fill_uniform(S, m, S, 42)
    np.testing.assert_array_equal(V, X[S:, :])
    np.testing.assert_array_equal(X[:S, :], X[S:, :])
    W = d.fill_uniform(S, 3, S, 7)
    np.testing.assert_array_equal(W, c.start_block_uniform(7, 3)[S:, :])


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_gram_pair_equals_two_grams(prec):
    """bdgx_gram_pair (A read once for [B1 | B2]) gives the two bdgx_gram
    results bit for bit, at the solver's tail shape (36 x 2 + 2), a wider
    pair (96 x 16 + 16) and a row count that ends in a partial step."""
    import bdg_lobpcg_amd as B
    N, L = (24, 20, 18), (10.0, 9.0, 8.0)
    c, f = physics(N, L, False)
    d = B.DeviceContext(N, L, False)
    rng = np.random.default_rng(4)
    for rows, p, q in ((c.size, 36, 2), (c.size, 96, 16), (c.size - 20, 20, 5)):
        A = rng.standard_normal((rows, p))
        B1 = rng.standard_normal((rows, q))
        B2 = rng.standard_normal((rows, q))
        G = d.gram_pair(A, B1, B2, prec)
        np.testing.assert_array_equal(G[:, :q], d.gram(A, B1, prec))
        np.testing.assert_array_equal(G[:, q:], d.gram(A, B2, prec))
        dt = np.float32 if prec else np.float64
        ref = A.astype(dt).astype(np.float64).T @ np.hstack([B1, B2]).astype(dt).astype(np.float64)
        assert relerr(G, ref) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_combine_cols_writes_only_listed_columns(cplx):
    """bdgx_combine_cols (the solver's update when most columns keep their
    vectors): listed output columns equal A @ C (numpy, FP64), the other
    columns of the output block are untouched, a unit coefficient column
    copies its source column bit for bit, and the sum order is fixed (two
    calls agree bit for bit).  Limits: FP32, more than 8 columns and
    repeated column indices are rejected."""
    import ctypes as Ct
    import bdg_lobpcg_amd as B
    N, L = (16, 12, 10), (8.0, 7.0, 6.0)
    d = B.DeviceContext(N, L, cplx)
    rng = np.random.default_rng(11 + cplx)
    dt = np.complex128 if cplx else np.float64

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    for rows, k, cols in ((1920, 36, [30, 31, 32, 33]), (1917, 5, [2]),
                          (1920, 96, [0, 7, 12, 40, 41, 42, 60, 63])):
        A = rnd(rows, k).astype(dt)
        Cm = rnd(k, len(cols)).astype(dt)
        Cm[:, 0] = 0
        Cm[cols[0] % k, 0] = 1.0
        Y0 = rnd(rows, 64).astype(dt)
        Y = d.combine_cols(A, Cm, cols, Y0)
        Y2 = d.combine_cols(A, Cm, cols, Y0)
        np.testing.assert_array_equal(Y, Y2)
        ref = A @ Cm
        assert relerr(Y[:, cols], ref) < 1e-13
        np.testing.assert_array_equal(Y[:, cols[0]], A[:, cols[0] % k])
        rest = [j for j in range(64) if j not in cols]
        np.testing.assert_array_equal(Y[:, rest], Y0[:, rest])
    lib = B.lib()
    cols = (Ct.c_int * 9)(*range(9))
    dup = (Ct.c_int * 2)(3, 3)
    buf = B.DeviceBuffer(8 * 64 * 16)
    assert lib.bdgx_combine_cols(d.h, B.FP32, buf.ptr, 8, 4, buf.ptr, buf.ptr, 8, cols, 2, 8) != 0
    assert lib.bdgx_combine_cols(d.h, B.FP64, buf.ptr, 8, 4, buf.ptr, buf.ptr, 8, cols, 9, 8) != 0
    assert lib.bdgx_combine_cols(d.h, B.FP64, buf.ptr, 8, 4, buf.ptr, buf.ptr, 8, dup, 2, 8) != 0


@pytest.mark.parametrize("grid", [((32, 32, 16), (16.0, 16.0, 10.0)),
                                  ((96, 96, 8), (20.0, 20.0, 8.0))],
                         ids=lambda g: "x".join(map(str, g[0])))
def test_complex_plane_launch_variants(grid):
    """Both launch shapes (4N and 8N threads) of the complex whole-plane
    kernels, forward and inverse, give the oracle's result for every channel
    kind (real weight, conj(psi0) / psi0 weights, dipolar two-channel M)."""
    N, L = grid
    c, f = physics(N, L, True, dipolar=True)
    d = device_twin(N, L, True, f)
    X = rand_block(c.size, 3, True, seed=7)
    refs = {op: c.apply(op, X) for op in ("K", "M", "precondK", "dipolar_conv")}
    for prec in (0, 1):
        for kind in (0, 1):
            for v in (0, 1):
                d.set_launch_variant(kind, prec, v)
                for op, ref in refs.items():
                    e = relerr(d.apply(op, X, prec), ref)
                    assert e < TOL[prec], f"kind {kind} variant {v} prec {prec} {op}: {e:.2e}"


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_derived_precond_weights_match_host_weights(cplx):
    """bdgx_derive_precond (the solver's preconditioner weights 1 / sqrt(max(
    floor, lt + mu)) computed on the device from the uploaded local terms)
    against the same weights computed on the host with numpy and uploaded:
    precondK and precondM agree bit for bit in FP64 and FP32."""
    N, L = (32, 32, 12), (16.0, 16.0, 10.0)
    c, f = physics(N, L, cplx, dipolar=True)
    mu = float(f["mu"])
    floor = max(mu, 1e-8)
    pK = 1.0 / np.sqrt(np.maximum(floor, f["ltK"] + mu))
    pM = 1.0 / np.sqrt(np.maximum(floor, f["ltM"] + mu))
    host = device_twin(N, L, cplx, dict(f, pK=pK, pM=pM))
    dev = device_twin(N, L, cplx, dict(f, pK=None, pM=None))
    dev.derive_precond(floor)
    X = rand_block(c.size, 3, cplx, seed=5)
    print(f"case N={N} cplx={cplx}", flush=True)
    for prec in (0, 1):
        for op in ("precondK", "precondM"):
            a = host.apply(op, X, prec)
            b = dev.apply(op, X, prec)
            assert np.array_equal(a, b), (op, prec, float(np.abs(a - b).max()))
