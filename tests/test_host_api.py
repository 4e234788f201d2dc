"""CPU tests of the product's host layer (include/bdg/bdg.h implemented in
C): exported symbols, setup arrays, misuse exits, start blocks and file I/O.
No device call happens here (the device context is created on first use)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import bdg_lobpcg_amd as B
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for hdr in ("include/bdg/bdg.h", "include/bdg/bdg_amd.h", "include/bdgx.h"):
        txt = open(os.path.join(ROOT, hdr)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(bdgx?_[a-z0-9_]+)\s*\(", txt):
            if not txt[m.end():].lstrip().startswith("*"):
                names.add(m.group(1))
    names.discard("bdg_init_fn")
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = B.lib()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(_declared_functions()) >= 50


def test_reference_api_complete():
    """The 22 functions of reference include/bdg/bdg.h:67-244."""
    ref = ["bdg_alloc", "bdg_free", "bdg_reset", "bdg_set_system", "bdg_set_trap",
           "bdg_set_wavefunction", "bdg_set_local_interactions", "bdg_set_mu", "bdg_set_dipolar",
           "bdg_set_solver_params", "bdg_set_init_mode", "bdg_reuse_modes", "bdg_solve",
           "bdg_converged", "bdg_eigenvalues", "bdg_modes_u", "bdg_modes_v",
           "bdg_load_wavefunction", "bdg_load_wavefunction_fmt", "bdg_write_eigenvalues",
           "bdg_write_mode_u", "bdg_write_mode_v"]
    assert len(ref) == 22
    lib = B.lib()
    assert all(hasattr(lib, n) for n in ref)


def test_setup_arrays_match_oracle():
    """trap + contact + mu through the C host layer vs the oracle
    (reference tests/test_setup.c:299-500)."""
    N, L = (12, 10, 8), (6.0, 5.0, 4.0)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal(int(np.prod(N)))
    p = B.Problem(N, L, False)
    p.set_system()
    p.set_trap_harmonic((1.0, 0.7, 1.4))
    p.set_wavefunction(psi)
    p.set_contact(3.0)
    p.set_mu(0.9)
    a = p.arrays()
    c = O.Ctx(N, L, False)
    ltK, ltM = np.zeros(c.size), np.zeros(c.size)
    c.trap_harmonic((1.0, 0.7, 1.4), ltK, ltM)
    c.contact(psi, 3.0, 3.0, ltK, ltM)
    pK, pM = O.setup_mu(0.9, ltK, ltM)
    np.testing.assert_allclose(a["ltK"], ltK, rtol=0, atol=1e-14)
    np.testing.assert_allclose(a["ltM"], ltM, rtol=0, atol=1e-14)
    np.testing.assert_allclose(a["pK"], pK, rtol=1e-15)
    np.testing.assert_allclose(a["pM"], pM, rtol=1e-15)
    assert a["mu"] == 0.9


def test_local_interactions_asymmetric_and_trap_additive():
    """U_K != U_M: K += U_K, M += U_K + 2 U_M; two traps add
    (reference tests/test_setup.c:415-443, 505-524)."""
    p = B.Problem((8,), (4.0,), False)
    p.set_system()
    p.set_trap_harmonic((1.0,))
    p.set_trap_harmonic((1.0,))
    p.set_wavefunction(np.full(8, 0.5))
    p.set_local_interactions(lambda n: 2.0 * n, lambda n: 5.0 * n)
    p.set_mu(0.0)
    a = p.arrays()
    x = (np.arange(8) - 4) * 0.5
    np.testing.assert_allclose(a["ltK"], x ** 2 + 0.5, atol=1e-14)
    np.testing.assert_allclose(a["ltM"], x ** 2 + 0.5 + 2 * 1.25, atol=1e-14)


MISUSE = {
    "interactions_before_wf": "p.set_contact(1.0)",
    "trap_before_system": "q = B.Problem((8,), (4.0,), False); q.set_trap_harmonic((1.0,))",
    "interactions_twice": "p.set_wavefunction(np.ones(8)); p.set_contact(1.0); p.set_contact(1.0)",
    "dipolar_after_mu": "p.set_wavefunction(np.ones(8)); p.set_mu(1.0); p.set_dipolar(1.0, (0, 0, 1), 1.0)",
    "solve_without_mu": "p.solve()",
    "wf_wrong_size": "p.set_wavefunction(np.ones(7))",
}


@pytest.mark.parametrize("case", sorted(MISUSE))
def test_misuse_exits_with_failure(case):
    """Out-of-order setup exits with EXIT_FAILURE (reference
    tests/test_setup.c:531-592, which forks for the same check)."""
    code = ("import sys, numpy as np; sys.path.insert(0, %r); import bdg_lobpcg_amd as B;"
            "p = B.Problem((8,), (4.0,), False); p.set_system(); %s; print('no-exit')"
            % (ROOT, MISUSE[case]))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "ERROR" in r.stderr and "no-exit" not in r.stdout


def _ref_planewave_block(N, L, psi, m, hint="auto"):
    """Restatement of reference src/solver.c:79-129, 196-258 in numpy."""
    dim = len(N)
    need = (m + 1) // 2
    kx2 = [((2 * np.pi / l) * np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)) ** 2
           for n, l in zip(N, L)]
    rng3 = [1, 1, 1]
    for d in range(dim):
        rng3[d] = min(N[d], 2 * need)
    ent = []
    for iz in range(rng3[2]):
        for iy in range(rng3[1]):
            for ix in range(rng3[0]):
                k2 = kx2[0][ix] + (kx2[1][iy] if dim > 1 else 0) + (kx2[2][iz] if dim > 2 else 0)
                ent.append((k2, (ix, iy, iz)))
    ent.sort(key=lambda e: e[0])    # stable, like glibc's merge-sort qsort
    nk = min(len(ent), need)
    S = int(np.prod(N))
    X = np.zeros((2 * S, m))
    seed = [42]

    def xr():
        s = seed[0]
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        seed[0] = s
        return (s & 0x7FFFFFFF) / 0x7FFFFFFF

    idx = np.arange(S)
    ids = [idx % N[0], (idx // N[0]) % (N[1] if dim > 1 else 1), idx // (N[0] * (N[1] if dim > 1 else 1))]
    for j in range(m):
        kv = (j + 1) // 2
        if kv >= nk:
            for i in range(S):
                X[i, j] = X[i + S, j] = (xr() - 0.5) * abs(psi[i])
            continue
        ki = ent[kv][1]
        kr = np.zeros(S)
        for d in range(dim):
            f = ki[d] if ki[d] <= N[d] // 2 else ki[d] - N[d]
            kr += 2 * np.pi * f / L[d] * ids[d] * L[d] / N[d]
        pw = np.sin(kr) if j % 2 == 1 else np.cos(kr)
        for i in range(S):
            X[i, j] = X[i + S, j] = (pw[i] + 1e-4 * (xr() - 0.5)) * abs(psi[i])
    return X


def test_planewave_start_block_matches_reference_recipe():
    N, L = (6, 5, 4), (3.0, 4.0, 2.0)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal(int(np.prod(N)))
    p = B.Problem(N, L, False)
    p.set_system()
    p.set_wavefunction(psi)
    p.set_solver_params(5, 7, 10, 1e-6)
    S = psi.size
    X = p.start_block((2 * S, 7))
    np.testing.assert_allclose(X, _ref_planewave_block(N, L, psi, 7), rtol=0, atol=1e-13)


def test_uniform_custom_block_matches_oracle():
    N, L = (8, 6, 4), (4.0, 3.0, 2.0)
    p = B.Problem(N, L, False)
    p.set_system()
    p.set_solver_params(3, 4, 10, 1e-6)
    p.set_init_uniform(42)
    S = int(np.prod(N))
    X = p.start_block((2 * S, 4))
    c = O.Ctx(N, L, False)
    np.testing.assert_array_equal(X, c.start_block_uniform(42, 4))


def test_io_formats_and_error_codes(tmp_path):
    """reference tests/test_io.c:45-321: loader formats and return codes."""
    lib = B.lib()
    p = B.Problem((4,), (2.0,), True)
    p.set_system()
    f = tmp_path / "wf.dat"
    f.write_text("1.0 0.5\n2.0 -1.0\n3.0 0.0\n4.0 2.5\n")
    assert lib.bdg_load_wavefunction(p.h, str(f).encode()) == 0
    assert lib.bdg_load_wavefunction(p.h, str(tmp_path / "missing.dat").encode()) == -1
    g = tmp_path / "short.dat"
    g.write_text("1.0 0.5\n2.0\n")
    assert lib.bdg_load_wavefunction(p.h, str(g).encode()) == -2
    assert lib.bdg_write_eigenvalues(p.h, str(tmp_path / "ev.dat").encode()) == -3
    assert lib.bdg_write_mode_u(p.h, 0, str(tmp_path / "m").encode()) == -3
    assert lib.bdg_reuse_modes(p.h, 0.1) == -3


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_binary_wavefunction_round_trip_and_codes(tmp_path, cplx):
    """Binary fast path beside the text loader: a saved psi0 loads back bit
    for bit into a handle of the same grid; a file of another grid or
    variant, a truncated file and a missing file give -2 / -2 / -1; nothing
    to save gives -3; modes without results give -3."""
    N, L = (6, 5, 4), (3.0, 2.5, 2.0)
    rng = np.random.default_rng(1)
    S = int(np.prod(N))
    psi = rng.standard_normal(S) + (1j * rng.standard_normal(S) if cplx else 0)
    p = B.Problem(N, L, cplx)
    p.set_system()
    f = tmp_path / "psi.bin"
    assert p.save_wavefunction_bin(f) == -3
    p.set_wavefunction(psi)
    assert p.save_wavefunction_bin(f) == 0
    assert f.stat().st_size == 64 + S * 8 * (2 if cplx else 1)
    hdr, got = B.read_bin(f)
    assert hdr == dict(kind=1, complex=cplx, dim=3, N=N, count=1)
    np.testing.assert_array_equal(got, psi)
    q = B.Problem(N, L, cplx)
    q.set_system()
    assert q.load_wavefunction_bin(f) == 0
    g = tmp_path / "psi2.bin"
    assert q.save_wavefunction_bin(g) == 0
    assert f.read_bytes() == g.read_bytes()
    other = B.Problem((6, 5, 5), L, cplx)
    other.set_system()
    assert other.load_wavefunction_bin(f) == -2
    variant = B.Problem(N, L, not cplx)
    variant.set_system()
    assert variant.load_wavefunction_bin(f) == -2
    t = tmp_path / "short.bin"
    t.write_bytes(f.read_bytes()[:-8])
    r = B.Problem(N, L, cplx)
    r.set_system()
    assert r.load_wavefunction_bin(t) == -2
    assert r.load_wavefunction_bin(tmp_path / "missing.bin") == -1
    assert r.write_modes_bin(tmp_path / "modes.bin") == -3


@pytest.mark.parametrize("k,m", [(12, 5), (40, 16), (96, 32)])
def test_host_rayleigh_ritz_matches_generalized_eigh(k, m):
    """The solver's host Rayleigh-Ritz (dense.c / rr_impl.inc: Euclidean
    orthonormalisation, M-Cholesky, Householder tridiagonalisation, implicit
    QR, blocked products) returns the m smallest eigenpairs of the pencil
    (G_A, G_M) of an independent basis: values against scipy's generalized
    eigh, vectors through G_A y = w2 G_M y and G_M-orthonormality."""
    import ctypes as C
    from scipy.linalg import eigh
    lib = B.lib()
    fn = lib.dn_rayleigh_ritz
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5 + [C.c_void_p]
    rng = np.random.default_rng(k)
    V = rng.standard_normal((3 * k, k))
    d = np.linspace(0.5, 3.0, 3 * k)
    GV = V.T @ V
    GM = V.T @ (d[:, None] * V)
    Q, _ = np.linalg.qr(rng.standard_normal((3 * k, 3 * k)))
    H = Q @ np.diag(np.linspace(-0.01, 40.0, 3 * k)) @ Q.T
    GA = V.T @ H @ V
    GA = 0.5 * (GA + GA.T)
    w2 = np.zeros(m)
    Y = np.zeros((k, m), order="F")
    zf = np.zeros(m, dtype=np.int32)
    args = [np.asfortranarray(a) for a in (GV, GM, GA)]
    got = fn(k, m, 0, *[a.ctypes.data for a in args], w2.ctypes.data, Y.ctypes.data,
             zf.ctypes.data)
    assert got == m
    ref = eigh(GA, GM, eigvals_only=True)[:m]
    np.testing.assert_allclose(w2, ref, rtol=1e-10, atol=1e-10 * abs(ref).max())
    assert np.abs(Y.T @ GM @ Y - np.eye(m)).max() < 1e-10
    res = GA @ Y - (GM @ Y) * w2
    assert np.abs(res).max() < 1e-9 * abs(ref).max()
    assert not zf.any()


@pytest.mark.parametrize("cplx", [0, 1])
def test_host_gram_projection_and_independence_test(cplx):
    """dn_qhgq (the Gram set of a combined basis V Q from the Gram of V,
    used by the solver between full Gram recomputes) against numpy's Q^H G Q;
    dn_chol_ok (the solver's dependence test of [X P W]) against numpy's
    Cholesky of the diagonally scaled Gram: independent columns pass, a
    column that is a combination of two others fails, a zero column fails."""
    import ctypes as C
    lib = B.lib()
    lib.dn_qhgq.restype = None
    lib.dn_qhgq.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.dn_chol_ok.restype = C.c_int
    lib.dn_chol_ok.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_double]
    rng = np.random.default_rng(7 + cplx)
    dt = np.complex128 if cplx else np.float64

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return a + 1j * rng.standard_normal(shape) if cplx else a

    for k, q in ((5, 3), (36, 34), (96, 64), (97, 65)):
        V = rnd(4 * k, k)
        G = np.asfortranarray((V.conj().T @ V).astype(dt))
        Q = np.asfortranarray(rnd(k, q).astype(dt))
        out = np.zeros((q, q), dtype=dt, order="F")
        lib.dn_qhgq(k, q, cplx, G.ctypes.data, Q.ctypes.data, out.ctypes.data)
        ref = Q.conj().T @ G @ Q
        assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
        assert lib.dn_chol_ok(k, cplx, G.ctypes.data, 1e-10) == 1
        d = 1.0 / np.sqrt(np.real(np.diag(G)))
        np.linalg.cholesky(d[:, None] * G * d[None, :])
        Vd = V.copy()
        Vd[:, k // 2] = 2.0 * V[:, 0] - 0.5 * V[:, k - 1]
        Gd = np.asfortranarray((Vd.conj().T @ Vd).astype(dt))
        assert lib.dn_chol_ok(k, cplx, Gd.ctypes.data, 1e-10) == 0
        Vz = V.copy()
        Vz[:, k - 1] = 0
        Gz = np.asfortranarray((Vz.conj().T @ Vz).astype(dt))
        assert lib.dn_chol_ok(k, cplx, Gz.ctypes.data, 1e-10) == 0


def _lrep_lib():
    import ctypes as C
    lib = B.lib()
    lib.dn_lrep_rr.restype = C.c_int
    lib.dn_lrep_rr.argtypes = [C.c_int] * 4 + [C.c_void_p] * 6
    lib.dn_lrep_update.restype = C.c_int
    lib.dn_lrep_update.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_double, C.c_void_p, C.c_void_p]
    lib.dn_orth_coef.restype = C.c_int
    lib.dn_orth_coef.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                 C.c_void_p]
    return lib


def _herm_pd(rng, n, cplx, lo, hi):
    a = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    Q, _ = np.linalg.qr(a)
    return (Q * np.linspace(lo, hi, n)[None, :]) @ Q.conj().T


@pytest.mark.parametrize("cplx", [0, 1])
def test_host_first_order_rayleigh_ritz(cplx):
    """dn_lrep_rr (the first-order pencil solver's host Rayleigh-Ritz): on
    the full space its w^2 are the exact eigenvalues of M K (numpy eigvals of
    the dense product, the first-order pencil K u = w v, M v = w u; one pair
    with w^2 < 0 from a K with a negative eigenvalue); on random orthonormal
    subspaces of unequal sizes its pairs satisfy the Galerkin conditions
    G_K a = s|w| G_W b, G_M b = |w| G_W^H a with s = sign(w^2), normalised."""
    lib = _lrep_lib()
    rng = np.random.default_rng(11 + cplx)
    dt = np.complex128 if cplx else np.float64
    n, m = 40, 6
    K = _herm_pd(rng, n, cplx, -0.02, 30.0)
    M = _herm_pd(rng, n, cplx, 0.5, 25.0)
    F = lambda a: np.asfortranarray(a.astype(dt))
    # full space: Z_U = Z_V = I
    w2 = np.zeros(m)
    A = np.zeros((n, m), dtype=dt, order="F")
    Bc = np.zeros((n, m), dtype=dt, order="F")
    I = np.eye(n)
    Kf, Mf, If = F(K), F(M), F(I)  # kept alive through the call
    got = lib.dn_lrep_rr(n, n, m, cplx, Kf.ctypes.data, Mf.ctypes.data, If.ctypes.data,
                         w2.ctypes.data, A.ctypes.data, Bc.ctypes.data)
    assert got == m
    ref = np.sort(np.linalg.eigvals(M @ K).real)[:m]
    assert ref[0] < 0 < ref[1]
    np.testing.assert_allclose(w2, ref, rtol=1e-10, atol=1e-12)
    # subspaces of 17 and 14 columns (uncoupled directions of Z_U: the
    # eigen-decomposition path with the Schur complement) and of 14 and 17
    # (the Cholesky path)
    for ku, kv in ((17, 14), (14, 17)):
        _galerkin_case(lib, rng, cplx, dt, F, K, M, n, m, ku, kv)


def _galerkin_case(lib, rng, cplx, dt, F, K, M, n, m, ku, kv):
    w2 = np.zeros(m)
    ZU, _ = np.linalg.qr(rng.standard_normal((n, ku)) + (1j * rng.standard_normal((n, ku)) if cplx else 0))
    ZV, _ = np.linalg.qr(rng.standard_normal((n, kv)) + (1j * rng.standard_normal((n, kv)) if cplx else 0))
    GK, GM, GW = ZU.conj().T @ K @ ZU, ZV.conj().T @ M @ ZV, ZU.conj().T @ ZV
    A = np.zeros((ku, m), dtype=dt, order="F")
    Bc = np.zeros((kv, m), dtype=dt, order="F")
    GKf, GMf, GWf = F(GK), F(GM), F(GW)
    got = lib.dn_lrep_rr(ku, kv, m, cplx, GKf.ctypes.data, GMf.ctypes.data, GWf.ctypes.data,
                         w2.ctypes.data, A.ctypes.data, Bc.ctypes.data)
    assert got == m
    wt, sg = np.sqrt(np.abs(w2)), np.sign(w2)
    r1 = GK @ A - GW @ Bc * (sg * wt)[None, :]
    r2 = GM @ Bc - GW.conj().T @ A * wt[None, :]
    assert max(np.abs(r1).max(), np.abs(r2).max()) < 1e-11 * 30
    nrm = (np.abs(A) ** 2).sum(0) + (np.abs(Bc) ** 2).sum(0)
    np.testing.assert_allclose(nrm, 1.0, rtol=1e-12)
    assert np.all(np.diff(w2) >= 0)


@pytest.mark.parametrize("cplx", [0, 1])
def test_host_first_order_update_and_orthonormalisation(cplx):
    """dn_lrep_update: [Q1 | Q2] orthonormal, A = Q1 R11 with R11 upper
    triangular, span(Q2) = the part of A[m:, act] outside span(A), a
    dependent P column dropped.  dn_orth_coef: [Q | W] C is orthonormal and
    orthogonal to Q; a W column inside span(Q) is left out."""
    lib = _lrep_lib()
    rng = np.random.default_rng(3 + cplx)
    dt = np.complex128 if cplx else np.float64

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return a + 1j * rng.standard_normal(shape) if cplx else a

    k, m = 30, 8
    A = np.asfortranarray(rnd(k, m).astype(dt))
    act = np.array([1, 3, 4, 6], dtype=np.int32)
    A[m:, 6] = 2.0 * A[m:, 3]  # its P part duplicates column 3's
    A = np.asfortranarray(A)
    CQ = np.zeros((k, k), dtype=dt, order="F")
    R11 = np.zeros((m, m), dtype=dt, order="F")
    nc = lib.dn_lrep_update(k, m, cplx, A.ctypes.data, act.ctypes.data, len(act), 1e-3,
                            CQ.ctypes.data, R11.ctypes.data)
    assert nc == m + 3
    Q = CQ[:, :nc]
    assert np.abs(Q.conj().T @ Q - np.eye(nc)).max() < 1e-13
    assert np.abs(Q[:, :m] @ R11 - A).max() < 1e-13 * np.abs(A).max()
    assert np.abs(np.tril(R11, -1)).max() == 0
    A0 = A[:, act].copy()
    A0[:m] = 0
    # every P direction lies in span(Q)
    assert np.abs(A0 - Q @ (Q.conj().T @ A0)).max() < 1e-12 * np.abs(A0).max()
    # orthonormalisation of W against Q
    n, nq, nw = 50, 7, 5
    Qb, _ = np.linalg.qr(rnd(n, nq))
    W = rnd(n, nw)
    W[:, 2] = Qb @ rnd(nq) + 1e-6 * rnd(n)  # almost inside span(Q)
    GQW = np.asfortranarray((Qb.conj().T @ W).astype(dt))
    GWW = np.asfortranarray((W.conj().T @ W).astype(dt))
    C = np.zeros((nq + nw, nw), dtype=dt, order="F")
    r = lib.dn_orth_coef(nq, nw, cplx, GQW.ctypes.data, GWW.ctypes.data, 1e-3, C.ctypes.data)
    assert r == nw - 1
    Wn = np.concatenate([Qb, W], axis=1) @ C[:, :r]
    assert np.abs(Wn.conj().T @ Wn - np.eye(r)).max() < 1e-10
    assert np.abs(Qb.conj().T @ Wn).max() < 1e-10


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("kind", ["random", "degenerate_pairs", "block_diagonal"])
def test_host_lowest_eigenpairs(cplx, kind):
    """dn_eig_low (Householder tridiagonalisation; up to n = 128 implicit QR
    with the rotations accumulated, beyond it eigenvalues by QR without
    vectors and inverse iteration for the lowest m vectors, Gram-Schmidt
    inside clusters) against numpy.linalg.eigh: eigenvalues to 1e-12
    relative to ||G||, residuals ||G x - w x|| <= 1e-12 ||G||, orthonormal
    vectors; exactly degenerate pairs (the spectra of the solver's
    projected problems have many) and a matrix whose off-diagonal blocks
    vanish (some reflectors are the identity)."""
    import ctypes as C
    import bdg_lobpcg_amd as B
    lib = B.lib()
    lib.dn_eig_low.restype = C.c_int
    lib.dn_eig_low.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(7 + 3 * cplx)
    dt = np.complex128 if cplx else np.float64
    for n, m in ((96, 32), (192, 64), (40, 40), (7, 3), (2, 1), (1, 1)):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0))
        if kind == "random":
            lam = rng.standard_normal(n) * 3.0
        elif kind == "degenerate_pairs":
            lam = np.repeat(np.sort(rng.standard_normal((n + 1) // 2)), 2)[:n]
        else:
            lam = rng.standard_normal(n)
        G = (Q * lam[None, :]) @ Q.conj().T
        if kind == "block_diagonal" and n > 4:
            G[: n // 2, n // 2:] = 0
            G[n // 2:, : n // 2] = 0
        G = np.asfortranarray(0.5 * (G + G.conj().T)).astype(dt)
        w = np.zeros(m)
        X = np.zeros((n, m), dtype=dt, order="F")
        assert lib.dn_eig_low(n, m, int(cplx), G.ctypes.data, w.ctypes.data, X.ctypes.data) == 0
        ref = np.linalg.eigvalsh(G)[:m]
        gn = max(np.abs(G).max(), 1e-300)
        np.testing.assert_allclose(w, ref, atol=1e-12 * gn * n)
        R = G @ X - X * w[None, :]
        assert np.abs(R).max() < 1e-12 * gn * n, (n, m, np.abs(R).max())
        np.testing.assert_allclose(X.conj().T @ X, np.eye(m), atol=1e-12 * n)
