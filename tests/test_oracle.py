"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Pins: numpy.fft golden vectors and brute-force DFTs (FFT machinery); an
independent numpy restatement of the operators (layout and indexing); the
reference's own closed-form tests (operator semantics, cited per test); and
dense scipy eigensolves of assembled K, M on tiny grids (eigenvalues, no
iterative solver involved).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))
TOL = 1e-10  # the reference's tolerance for its closed-form tests


def _maxrel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------------ FFT --
@pytest.mark.parametrize("i", range(7))
def test_fft_golden(i):
    sh = tuple(FIX[f"fft{i}_shape"])
    xr, xc = FIX[f"fft{i}_xr"], FIX[f"fft{i}_xc"]
    assert _maxrel(O.fft_r2c(xr), FIX[f"fft{i}_r2c"]) < 1e-13
    assert _maxrel(O.fft_c2c(xc, -1), FIX[f"fft{i}_c2c"]) < 1e-13
    assert _maxrel(O.fft_c2c(xc, +1), FIX[f"fft{i}_c2c_inv"]) < 1e-13
    # c2r(r2c(x)) = size * x (reference tests/test_setup.c:225-281 round trips)
    back = O.fft_c2r(O.fft_r2c(xr), sh[-1])
    assert _maxrel(back / xr.size, xr) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 9, 11, 13, 15, 22])
def test_fft_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
    assert _maxrel(O.fft_c2c(x, -1), F) < 1e-12


def test_fft_impulse_and_parseval():
    x = np.zeros((6, 4, 10))
    x[2, 1, 3] = 1.0
    h = O.fft_c2c(x.astype(complex), -1)
    assert np.allclose(np.abs(h), 1.0, atol=1e-14)
    rng = np.random.default_rng(3)
    y = rng.standard_normal((5, 6, 9)) + 1j * rng.standard_normal((5, 6, 9))
    Y = O.fft_c2c(y, -1)
    assert abs(np.vdot(Y, Y).real / y.size - np.vdot(y, y).real) < 1e-10 * np.vdot(y, y).real


# ------------------------------------------------------ k-space arrays --
def test_k_size_and_k2_values():
    """reference tests/test_setup.c:49-190"""
    assert O.Ctx((32,), (1.0,), False).k_size == 17
    assert O.Ctx((32,), (1.0,), True).k_size == 32
    assert O.Ctx((16, 8), (1.0, 1.0), False).k_size == 9 * 8
    assert O.Ctx((8, 6, 4), (1.0,) * 3, False).k_size == 5 * 6 * 4
    assert O.Ctx((8, 6, 4), (1.0,) * 3, True).k_size == 192
    c = O.Ctx((8, 6), (2 * np.pi, 2 * np.pi), False)
    np.testing.assert_allclose(c.kx2(0), [0, 1, 4, 9, 16, 9, 4, 1], atol=1e-12)
    np.testing.assert_allclose(c.kx2(1), [0, 1, 4, 9, 4, 1], atol=1e-12)
    c = O.Ctx((8, 8, 8), (2 * np.pi,) * 3, False)
    k2 = c.k2()
    n0k = 5
    assert c.k_size == 320
    assert abs(k2[1 * 8 * n0k + 1 * n0k + 1] - 3.0) < 1e-12
    assert abs(k2[5 * 8 * n0k + 3 * n0k + 2] - 22.0) < 1e-12


# ------------------------------------------------ operators vs golden --
@pytest.mark.parametrize("tag", ["op0", "op1"])
def test_operators_vs_numpy_restatement(tag):
    N, L = tuple(FIX[f"{tag}_N"]), tuple(FIX[f"{tag}_L"])
    cplx = bool(FIX[f"{tag}_cplx"])
    c = O.Ctx(N, L, cplx)
    psi = FIX[f"{tag}_psi"]
    S = c.size
    ltK, ltM = np.zeros(S), np.zeros(S)
    c.trap_harmonic((1.0, 0.8, 1.3), ltK, ltM)
    c.contact(psi, 7.0, 7.0, ltK, ltM)
    c.set_fields(wf=psi)
    c.set_dipolar(3.0, (0.3, 0.0, 1.0), 2.5)
    assert _maxrel(c.dipolar_kernel(), FIX[f"{tag}_U"]) < 1e-13
    gphi = c.dipolar_meanfield()
    assert _maxrel(gphi, FIX[f"{tag}_gphi"]) < 1e-12
    ltK += gphi
    ltM += gphi
    pK, pM = O.setup_mu(1.1, ltK, ltM)
    np.testing.assert_allclose(ltK, FIX[f"{tag}_ltK"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pM, FIX[f"{tag}_pM"], rtol=1e-12)
    c.set_fields(ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=1.1)
    v = FIX[f"{tag}_v"]
    for op, key in [("kinetic", "kinetic"), ("K", "K"), ("M", "M"), ("dipolar_conv", "dipconv"),
                    ("precondK", "precondK"), ("precondM", "precondM")]:
        assert _maxrel(c.apply(op, v), FIX[f"{tag}_{key}"]) < 1e-12, op


# ------------------------------------- reference closed forms (KATs) --
def _x1d(N, L):
    return np.arange(N) * L / N


def test_kinetic_closed_forms():
    """reference tests/test_kinetic.c:45-214"""
    x = _x1d(64, 2 * np.pi)
    for cplx in (False, True):
        c = O.Ctx((64,), (2 * np.pi,), cplx)
        y = c.apply("kinetic", np.cos(x).astype(c.dtype))
        assert np.abs(y - 0.5 * np.cos(x)).max() < TOL
        y = c.apply("kinetic", (np.cos(x) + 0.5 * np.cos(3 * x)).astype(c.dtype))
        assert np.abs(y - (0.5 * np.cos(x) + 2.25 * np.cos(3 * x))).max() < TOL
    c = O.Ctx((32,), (2 * np.pi,), False)
    assert np.abs(c.apply("kinetic", np.ones(32))).max() < TOL
    N = (16, 16, 16)
    c = O.Ctx(N, (2 * np.pi,) * 3, False)
    g = np.arange(16) * 2 * np.pi / 16
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    f = np.cos(X + Y + Z).ravel()
    assert np.abs(c.apply("kinetic", f) - 1.5 * f).max() < TOL


def _const_ctx(N, L, cplx, V0, mu):
    c = O.Ctx((N,), (L,), cplx)
    S = c.size
    c.set_fields(ltK=np.full(S, V0 - mu), ltM=np.full(S, V0 - mu),
                 pK=np.full(S, 1 / np.sqrt(V0)), pM=np.full(S, 1 / np.sqrt(V0)), mu=mu)
    return c


def test_matmulK_M_closed_forms():
    """reference tests/test_matmulK.c:69-120, 172-216; tests/test_matmulM.c:67-115"""
    x = _x1d(64, 2 * np.pi)
    c = _const_ctx(64, 2 * np.pi, False, 2.0, 0.0)
    assert np.abs(c.apply("K", np.cos(x)) - 2.5 * np.cos(x)).max() < TOL
    assert np.abs(c.apply("M", np.cos(x)) - 2.5 * np.cos(x)).max() < TOL
    c = _const_ctx(32, 2 * np.pi, False, 3.0, 1.0)
    assert np.abs(c.apply("K", np.ones(32)) - 2.0).max() < TOL
    cz = _const_ctx(64, 2 * np.pi, True, 2.0, 0.5)
    cd = _const_ctx(64, 2 * np.pi, False, 2.0, 0.5)
    v = np.sin(2 * x) + 0.3 * np.cos(5 * x)
    assert np.abs(cz.apply("K", v.astype(complex)) - cd.apply("K", v)).max() < TOL


def test_precond_closed_forms():
    """reference tests/test_precond.c:69-126 (planewave), 366-390 (zero mode)"""
    x = _x1d(64, 2 * np.pi)
    for cplx in (False, True):
        c = _const_ctx(64, 2 * np.pi, cplx, 2.0, 0.5)
        y = c.apply("precondK", np.cos(x).astype(c.dtype))
        assert np.abs(y - np.cos(x) / (2.0 * (0.5 + 0.5))).max() < TOL
    c = _const_ctx(64, 2 * np.pi, False, 2.0, 1.5)
    assert np.abs(c.apply("precondK", np.ones(64)) - 1.0 / (2.0 * 1.5)).max() < TOL


def test_swap_and_block_structure():
    """reference tests/test_matmulSwap.c:52-124, test_matmulLrep.c:68-104,
    test_precond.c:211-315"""
    c = _const_ctx(32, 2 * np.pi, False, 2.0, 0.5)
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal(32), rng.standard_normal(32)
    sw = c.apply("Swap", np.concatenate([u, v]))
    assert np.array_equal(sw, np.concatenate([v, u]))
    L = c.apply("Lrep", np.concatenate([u, v]))
    assert np.abs(L - np.concatenate([c.apply("K", u), c.apply("M", v)])).max() < 1e-14
    T = c.apply("precondLrep", np.concatenate([u, v]))
    assert np.abs(T - np.concatenate([c.apply("precondK", u), c.apply("precondM", v)])).max() < 1e-14


def test_dipolar_kernel_known_values():
    """reference tests/test_dipolar.c:44-144: U(0)=0; kR=pi: 2f along the
    dipole, -f transverse; r2c and c2c kernels agree on shared k-points"""
    N, L = (8, 8, 8), (2 * np.pi,) * 3
    cr = O.Ctx(N, L, False)
    cr.set_dipolar(1.0, (0, 0, 1), np.pi)
    Ur = cr.dipolar_kernel()
    f = 1 + 3 * np.cos(np.pi) / np.pi ** 2 - 3 * np.sin(np.pi) / np.pi ** 3
    assert abs(Ur[0]) < 1e-14
    assert abs(Ur[1 * 8 * 5] - 2 * f) < 1e-12
    assert abs(Ur[1] + f) < 1e-12
    cc = O.Ctx(N, L, True)
    cc.set_dipolar(1.0, (0, 0, 1), np.pi)
    Uc = cc.dipolar_kernel().reshape(8, 8, 8)
    assert np.abs(Uc[:, :, :5] - Ur.reshape(8, 8, 5)).max() < 1e-14


def test_dipolar_meanfield_uniform_and_conv_closed_form():
    """Uniform density gives Phi = 0 (reference tests/test_dipolar.c:150-173).
    For uniform psi0 = c and x = cos(k.r), D x = g c^2 U(k) x exactly."""
    N, L = (8, 8, 8), (2 * np.pi,) * 3
    c = O.Ctx(N, L, False)
    S = c.size
    c.set_fields(wf=np.full(S, 0.7))
    c.set_dipolar(2.0, (0, 0, 1), 1.3)
    assert np.abs(c.dipolar_meanfield()).max() < 1e-13
    g = np.arange(8) * 2 * np.pi / 8
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    x = np.cos(X + 2 * Z).ravel()
    kR = np.sqrt(5.0) * 1.3
    U = (3 * 4 / 5 - 1) * (1 + 3 * np.cos(kR) / kR ** 2 - 3 * np.sin(kR) / kR ** 3)
    y = c.apply("dipolar_conv", x)
    assert np.abs(y - 2.0 * 0.49 * U * x).max() < 1e-12
    # M x = (k^2/2 + ltM + 2 g c^2 U) x for constant ltM
    c.set_fields(ltK=np.full(S, 0.4), ltM=np.full(S, 0.9), pK=np.ones(S), pM=np.ones(S), mu=0.1)
    assert np.abs(c.apply("M", x) - (2.5 + 0.9 + 2 * 2.0 * 0.49 * U) * x).max() < 1e-12


def test_setup_contact_mu_trap():
    """reference tests/test_setup.c:299-350, 415-500"""
    c = O.Ctx((16,), (8.0,), False)
    ltK, ltM = np.zeros(16), np.zeros(16)
    c.trap_harmonic((1.0,), ltK, ltM)
    x = (np.arange(16) - 8.0) * 0.5
    assert np.abs(ltK - 0.5 * x ** 2).max() < 1e-14
    ltK, ltM = np.zeros(16), np.zeros(16)
    c.contact(np.full(16, 0.7), 2.5, 2.5, ltK, ltM)
    assert np.abs(ltK - 2.5 * 0.49).max() < 1e-14 and np.abs(ltM - 3 * 2.5 * 0.49).max() < 1e-14
    before = ltK.copy()
    pK, pM = O.setup_mu(3.0, ltK, ltM)
    assert np.abs(ltK - (before - 3.0)).max() < 1e-14
    assert np.abs(pK - 1 / np.sqrt(before)).max() < 1e-14
    lt = np.array([-1.0, 0.0, 4.0])
    pK, _ = O.setup_mu(0.0, lt.copy(), lt.copy())
    np.testing.assert_allclose(pK, [1e4, 1e4, 0.5])   # fmax(1e-8, lt), not fabs


def test_xrand_stream():
    """reference src/bdg_internal.h:53-65: xorshift32 seed 42"""
    s = 42
    out = []
    for _ in range(3):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        out.append((s & 0x7FFFFFFF) / 0x7FFFFFFF)
    np.testing.assert_array_equal(O.xrand_stream(42, 3), out)


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_reuse_block_recipe(cplx):
    """Oracle reuse block (reference src/bdg.c:112-207) pinned by what the
    reference's recipe fixes: the shape and a nonzero buffer
    (tests/test_io.c:332-365); noise 0 copies the modes exactly and fills
    the extra columns with |z0| |psi0| >= 0; the stream position of every
    (row, column): with unit-norm modes and noise 1 the added pair is
    (z0, z1) of xrand values 2 (j S + i) + 1, + 2 of seed 12345 (a swapped
    pair, a restarted stream per column or a dropped column shifts it); and
    the Box-Muller law: z0, z1 standard normal and uncorrelated."""
    rng = np.random.default_rng(4)
    S, nev, m = 40, 2, 4
    dt = np.complex128 if cplx else np.float64
    U = rng.standard_normal((S, nev)).astype(dt)
    V = rng.standard_normal((S, nev)).astype(dt)
    if cplx:
        U += 1j * rng.standard_normal((S, nev))
        V += 1j * rng.standard_normal((S, nev))
    wf = rng.standard_normal(S) + (1j * rng.standard_normal(S) if cplx else 0)
    X0 = O.reuse_block(U, V, m, 0.0, wf)
    assert X0.shape == (2 * S, m)
    np.testing.assert_array_equal(X0[:S, :nev], U)
    np.testing.assert_array_equal(X0[S:, :nev], V)
    assert np.all(X0[:, nev:].real >= 0) and np.all(X0[:, nev:].imag == 0)
    np.testing.assert_array_equal(X0[:S, nev:], X0[S:, nev:])
    # unit columns, noise 1: the noise is the Box-Muller pair of the stream
    nrm = np.sqrt(np.sum(np.abs(U) ** 2 + np.abs(V) ** 2, axis=0))
    Un, Vn = U / nrm, V / nrm
    X1 = O.reuse_block(Un, Vn, m, 1.0, wf)
    xr = O.xrand_stream(12345, 2 * S * m).reshape(m, S, 2)
    u1 = np.maximum(xr[..., 0], 1e-10)
    rr = np.sqrt(-2 * np.log(u1))
    z0 = (rr * np.cos(2 * np.pi * xr[..., 1])).T
    z1 = (rr * np.sin(2 * np.pi * xr[..., 1])).T
    np.testing.assert_allclose((X1[:S, :nev] - Un).real, z0[:, :nev], atol=1e-13)
    np.testing.assert_allclose((X1[S:, :nev] - Vn).real, z1[:, :nev], atol=1e-13)
    np.testing.assert_array_equal(X1[:S, :nev].imag, Un.imag)
    np.testing.assert_allclose(X1[:S, nev:].real, np.abs(z0[:, nev:]) * np.abs(wf)[:, None],
                               rtol=1e-14)
    # the law of the noise over a long column
    Sl = 100000
    Ul = np.zeros((Sl, 1), dt)
    Ul[0, 0] = 1.0
    Xl = O.reuse_block(Ul, np.zeros((Sl, 1), dt), 1, 1.0)
    a, b = Xl[1:Sl, 0].real, Xl[Sl + 1:, 0].real
    assert abs(a.mean()) < 0.02 and abs(b.mean()) < 0.02
    assert abs(a.var() - 1) < 0.03 and abs(b.var() - 1) < 0.03
    assert abs(np.mean(a * b)) < 0.02


# -------------------------------------------------- eigenvalue pins --
def _dense_case(tag):
    N, L = tuple(FIX[f"{tag}_N"]), tuple(FIX[f"{tag}_L"])
    cplx = bool(FIX[f"{tag}_cplx"])
    c = O.Ctx(N, L, cplx)
    psi = FIX[f"{tag}_psi"]
    S = c.size
    ltK, ltM = np.zeros(S), np.zeros(S)
    c.trap_harmonic(tuple(FIX[f"{tag}_omega"]), ltK, ltM)
    g, gdd = float(FIX[f"{tag}_g"]), float(FIX[f"{tag}_gdd"])
    c.contact(psi, g, g, ltK, ltM)
    c.set_fields(wf=psi)
    if gdd:
        c.set_dipolar(gdd, (0, 0, 1), float(FIX[f"{tag}_R"]))
        gphi = c.dipolar_meanfield()
        ltK += gphi
        ltM += gphi
    pK, pM = O.setup_mu(0.0, ltK, ltM)
    c.set_fields(ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=0.0)
    return c, FIX[f"{tag}_w"]


@pytest.mark.parametrize("tag", ["eig0", "eig1", "eig2", "eig3"])
def test_oracle_dense_spectrum(tag):
    """eig(MKM, M) from the oracle's operators matches the fixture from the
    independent numpy restatement."""
    import scipy.linalg as sla
    c, w_ref = _dense_case(tag)
    S = c.size
    E = np.eye(S, dtype=c.dtype)
    K = c.apply("K", E)
    M = c.apply("M", E)
    K, M = 0.5 * (K + K.conj().T), 0.5 * (M + M.conj().T)
    w2 = sla.eigh(M @ K @ M, M, eigvals_only=True)[: len(w_ref)]
    np.testing.assert_allclose(np.sqrt(w2), w_ref, rtol=1e-9)


@pytest.mark.parametrize("tag", ["eig0", "eig1", "eig2"])
def test_reference_solver_vs_dense(tag):
    """The numpy iLOBPCG (CPU reference solver) reproduces the dense spectra.
    These fixtures use mu = 0, where the reference preconditioner divides the
    k = 0 mode by safe_val(0) = 1e-8; the solver math is pinned with the
    identity preconditioner instead."""
    from oracle.ilobpcg_ref import ilobpcg
    c, w_ref = _dense_case(tag)
    S = c.size
    nev, m = 8, 10
    X0 = c.start_block_uniform(42, m)[S:]
    r = ilobpcg(lambda X: c.apply("K", X), lambda X: c.apply("M", X),
                lambda X: X, lambda X: X, X0, nev, tol=1e-11, maxiter=500)
    np.testing.assert_allclose(r["w"], w_ref[:nev], rtol=1e-9)


@pytest.mark.parametrize("cplx", [False, True])
def test_reference_solver_1d_uniform(cplx):
    """reference tests/test_bdg_1d_uniform.c:60-138: 0, E1, E1, E2."""
    from oracle.ilobpcg_ref import ilobpcg
    c = O.Ctx((64,), (2 * np.pi,), cplx)
    ltK, ltM = np.zeros(64), np.zeros(64)
    c.contact(np.ones(64, dtype=c.dtype), 1.0, 1.0, ltK, ltM)
    pK, pM = O.setup_mu(1.0, ltK, ltM)
    c.set_fields(ltK=ltK, ltM=ltM, wf=np.ones(64, dtype=c.dtype), pK=pK, pM=pM, mu=1.0)
    X0 = c.start_block_uniform(42, 10)[64:]
    r = ilobpcg(lambda X: c.apply("K", X), lambda X: c.apply("M", X),
                lambda X: c.apply("precondK", X), lambda X: c.apply("precondM", X),
                X0, 4, tol=1e-10, maxiter=200)
    E1, E2 = np.sqrt(1.25), np.sqrt(8.0)
    assert abs(r["w"][0]) < 1e-3
    np.testing.assert_allclose(r["w"][1:], [E1, E1, E2], rtol=1e-9)
    # residual of the returned pair through the original pencil A x = w B x
    u, v = r["u"][:, 1], r["v"][:, 1]
    Ku, Mv = c.apply("K", u), c.apply("M", v)
    assert np.linalg.norm(Ku - r["w"][1] * v) < 1e-8 and np.linalg.norm(Mv - r["w"][1] * u) < 1e-8


# --------------------------------------------------------------- ITP --
def test_itp_noninteracting_is_stationary():
    """g = 0 harmonic trap: the normalised Gaussian of width 1 is the ground
    state, so ITP keeps it and mu = dim/2 (to the grid's spectral accuracy)."""
    N, L = (24, 24, 24), (12.0,) * 3
    c = O.Ctx(N, L, False)
    g = (np.arange(24) - 12) * 0.5
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).ravel()
    out, mu = c.itp_groundstate(psi, (1, 1, 1), 0.0, 0.0, (0, 0, 1), 6.0, 1e-3, 50)
    psi_n = psi / np.sqrt(np.sum(psi ** 2) * 0.125)
    assert abs(mu - 1.5) < 1e-8
    assert np.abs(out - psi_n).max() < 1e-8
    assert abs(np.sum(out ** 2) * 0.125 - 1.0) < 1e-12


def _gp_numpy(N, L, omega, g, gdd, R):
    """H_GP restated in numpy (tests/golden/make_golden.py's NumpyBdG for
    the FFTs and the cutoff kernel, independent of the oracle's C): returns
    (H, energy) for real psi on the grid of reference src/setup.c:35,68-72."""
    from tests.golden.make_golden import NumpyBdG
    nb = NumpyBdG(N, L, False)
    U = nb.kernel((0, 0, 1), R)
    xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(N, L)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    V = 0.5 * ((omega[0] * X) ** 2 + (omega[1] * Y) ** 2 + (omega[2] * Z) ** 2).ravel()
    dV = float(np.prod([l / n for n, l in zip(N, L)]))

    def parts(psi, gg, gd):
        n = psi ** 2
        phi = nb.inv(nb.fwd(n) * U).real / nb.S
        T = nb.inv(nb.fwd(psi) * nb.k2 / 2).real / nb.S
        return T, n, phi

    def H(psi, gg=g, gd=gdd):
        T, n, phi = parts(psi, gg, gd)
        return T + (V + gg * n + gd * phi) * psi

    def energy(psi):
        T, n, phi = parts(psi, g, gdd)
        return float(np.sum(psi * (T + (V + 0.5 * g * n + 0.5 * gdd * phi) * psi)) * dV)

    def residual(psi, gg=g, gd=gdd):
        Hp = H(psi, gg, gd)
        mu = float(np.sum(psi * Hp) * dV)
        return float(np.linalg.norm(Hp - mu * psi) / np.linalg.norm(psi)), mu

    return residual, energy, X, Y, Z, dV


def test_itp_energy_decreases_with_interactions():
    """Interacting dipolar ITP (orc_itp_groundstate): the GP energy, computed
    by an independent numpy H_GP, decreases monotonically along the
    imaginary-time path, the returned mu equals <psi|H_GP|psi>, and the norm
    is conserved."""
    N, L, g, gdd, R = (16, 16, 16), (12.0,) * 3, 50.0, 20.0, 6.0
    residual, energy, X, Y, Z, dV = _gp_numpy(N, L, (1, 1, 1), g, gdd, R)
    c = O.Ctx(N, L, False)
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).ravel()
    E = []
    for n in (20, 100, 400, 2000):
        p, mu = c.itp_groundstate(psi, (1, 1, 1), g, gdd, (0, 0, 1), R, 1e-3, n)
        assert abs(np.sum(p ** 2) * dV - 1.0) < 1e-12
        _, mu_np = residual(p)
        assert abs(mu - mu_np) < 1e-12 * abs(mu_np)
        E.append(energy(p))
    assert all(b < a for a, b in zip(E, E[1:])), E


def test_itp_relaxes_to_the_gp_ground_state():
    """The relaxed interacting state satisfies H_GP psi = mu psi with the
    numpy H_GP up to the split-step's O(dt) floor: halving dt halves the GP
    residual.  The same residual under a Hamiltonian with a wrong sign or
    factor on g n or g_dd Phi is 30x larger, so such a mistake in the
    oracle's step would fail here."""
    N, L, g, gdd, R = (16, 16, 16), (12.0,) * 3, 50.0, 20.0, 6.0
    residual, energy, X, Y, Z, dV = _gp_numpy(N, L, (1, 1, 1), g, gdd, R)
    c = O.Ctx(N, L, False)
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).ravel()
    p1, _ = c.itp_groundstate(psi, (1, 1, 1), g, gdd, (0, 0, 1), R, 5e-3, 4000)
    p2, _ = c.itp_groundstate(psi, (1, 1, 1), g, gdd, (0, 0, 1), R, 2.5e-3, 8000)
    r1, r2 = residual(p1)[0], residual(p2)[0]
    assert r1 < 1e-2 and 0.4 < r2 / r1 < 0.6, (r1, r2)
    for gg, gd in ((-g, gdd), (g, 2 * gdd), (g, -gdd), (g, 0.0), (2 * g, gdd)):
        assert residual(p2, gg, gd)[0] > 30 * r2


def test_itp_complex_path_equals_real_path_times_phase():
    """A constant phase commutes with every term of H_GP, so the complex
    path of the oracle's ITP relaxes psi e^{i a} to (real result) e^{i a}
    with the same mu."""
    N, L, g, gdd, R = (12, 12, 16), (10.0, 10.0, 12.0), 40.0, 15.0, 5.0
    _, _, X, Y, Z, _ = _gp_numpy(N, L, (1, 1.2, 0.8), g, gdd, R)
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).ravel()
    pr, mur = O.Ctx(N, L, False).itp_groundstate(psi, (1, 1.2, 0.8), g, gdd, (0, 0, 1), R,
                                                 1e-3, 300)
    ph = np.exp(0.7j)
    pc, muc = O.Ctx(N, L, True).itp_groundstate(psi * ph, (1, 1.2, 0.8), g, gdd, (0, 0, 1), R,
                                                1e-3, 300)
    assert np.abs(pc - pr * ph).max() < 1e-12 * np.abs(pr).max()
    assert abs(muc - mur) < 1e-12 * abs(mur)


def test_reference_solver_soft_mode_meets_the_residual_test():
    """A mode with |w| below zero_tol * scale but w != 0 is reported as a
    zero mode, yet converges only through its absolute residual (it is not
    counted converged on its magnitude alone): M = I, K = diag(d) with
    w_0 = 1e-4, far below 1e-3 * max|w|."""
    from oracle.ilobpcg_ref import ilobpcg
    n, nev = 60, 4
    d = np.concatenate([[1e-8], np.linspace(0.5, 9.0, n - 1)])
    rng = np.random.default_rng(7)
    K = lambda X: d[:, None] * X
    M = lambda X: X
    T = lambda X: X / (d[:, None] + 0.5)
    X0 = rng.standard_normal((n, 6))
    X0[:, 0] = 1e-4 * X0[:, 0]
    X0[0, 0] = 1.0  # near the soft eigenvector: its Ritz value is flagged at once
    one = ilobpcg(K, M, T, lambda X: X, X0, nev, tol=1e-10, maxiter=1)
    assert one["zero"][0] and one["hist"][-1] > 1e-10  # not passed on magnitude
    out = ilobpcg(K, M, T, lambda X: X, rng.standard_normal((n, 6)), nev, tol=1e-10,
                  maxiter=300)
    assert out["zero"][0] and out["res"][0] < 1e-10
    assert abs(out["w"][0] - 1e-4) < 1e-9
    np.testing.assert_allclose(out["w"][1:], np.sqrt(d[1:nev]), rtol=1e-9)


@pytest.mark.parametrize("tag", ["eig0", "eig1", "eig2"])
def test_first_order_reference_solver_vs_dense(tag):
    """The numpy first-order solver (oracle/lrep_ref.py, the restatement of
    the device solver's iteration) reproduces the dense spectra (identity
    preconditioner, as for the product-form reference solver above)."""
    from oracle.lrep_ref import lrep_lobpcg
    c, w_ref = _dense_case(tag)
    S = c.size
    nev, m = 8, 10
    X0 = c.start_block_uniform(42, m)
    r = lrep_lobpcg(lambda X: c.apply("K", X), lambda X: c.apply("M", X),
                    lambda X: X, lambda X: X, X0[:S], X0[S:], nev, tol=1e-11, maxiter=800)
    np.testing.assert_allclose(r["w"], w_ref[:nev], rtol=1e-9)
    # the modes satisfy the first-order pencil under the oracle's operators
    Ku, Mv = c.apply("K", r["u"]), c.apply("M", r["v"])
    assert np.abs(Ku - r["v"] * r["w"]).max() < 1e-8 * max(1, r["w"].max())
    assert np.abs(Mv - r["u"] * r["w"]).max() < 1e-8 * max(1, r["w"].max())


def test_first_order_reference_solver_fp32_phase_then_fp64():
    """3D dipolar problem with the solver's preconditioner: the first-order
    iteration in FP64 gives the product-form reference solver's omega, and
    the mixed policy (FP32-rounded storage and operators to residual 1e-5,
    then FP64 from those vectors) gives the same omega to 1e-9."""
    from oracle.ilobpcg_ref import ilobpcg, solver_precond_weights
    from oracle.lrep_ref import lrep_lobpcg
    from tests.common import physics
    N, L = (12, 12, 16), (10.0, 10.0, 12.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=8.0, mu=1.0)
    S = c.size
    pKs, pMs = solver_precond_weights(f["ltK"], f["ltM"], f["mu"])
    c.set_fields(pK=pKs, pM=pMs, mu=f["mu"])
    ops = [lambda X, o=o: c.apply(o, X) for o in ("K", "M", "precondK", "precondM")]
    X0 = c.start_block_uniform(42, 8)
    nev = 6
    ref = ilobpcg(*ops, X0[S:], nev, tol=1e-11, maxiter=600)
    r64 = lrep_lobpcg(*ops, X0[:S], X0[S:], nev, tol=1e-10, maxiter=600)
    small = np.abs(ref["w"]) < 1e-2
    scale = np.where(small, 1.0, np.abs(ref["w"]))
    assert np.all(np.abs(r64["w"] - ref["w"]) / scale < 1e-9)
    f32 = lambda X: np.asarray(X, np.complex64 if np.iscomplexobj(X) else np.float32)
    o32 = [lambda X, op=op: op(np.asarray(X, np.float32).astype(np.float64)).astype(np.float32)
           for op in ops]
    r32 = lrep_lobpcg(*o32, X0[:S], X0[S:], 8, tol=1e-5, maxiter=600, rnd=f32)
    U = np.asarray(r32["u"], np.float64)
    V = np.asarray(r32["v"], np.float64)
    rfin = lrep_lobpcg(*ops, U, V, nev, tol=1e-10, maxiter=600)
    assert np.all(np.abs(rfin["w"] - ref["w"]) / scale < 1e-9)
    assert rfin["iters"] < r64["iters"]
