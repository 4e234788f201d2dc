"""GPU solver tests through the public C API (include/bdg/bdg.h)."""
import numpy as np
import pytest

from tests.common import cluster_angles, physics, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cplx", [False, True])
def test_1d_uniform_bogoliubov(cplx):
    """reference tests/test_bdg_1d_uniform.c:60-138 and 142-228: uniform gas,
    N=64, L=2pi, g=n0=mu=1: spectrum 0, E(1) (x2), E(2)."""
    import bdg_lobpcg_amd as B
    N, L = 64, 2 * np.pi
    p = B.Problem((N,), (L,), cplx)
    p.set_system()
    p.set_solver_params(4, 10, 200, 1e-10)
    p.set_wavefunction(np.ones(N))
    p.set_contact(1.0)
    p.set_mu(1.0)
    assert p.solve() == 0
    assert p.converged() >= 4
    w = p.eigenvalues(4)
    E1, E2 = np.sqrt(0.5 * 2.5), np.sqrt(2.0 * 4.0)
    assert abs(w[0]) < 1e-3
    np.testing.assert_allclose(w[1:], [E1, E1, E2], rtol=1e-9)


_REF = {}  # CPU reference solves, per psi0 kind (shared by the parameter cases)


@pytest.mark.parametrize("form", [0, 1], ids=["first_order", "product"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_3d_dipolar_vs_oracle(prec, cplx, form):
    """Small 3D dipolar problem: device solve (first-order or product form;
    FP64, or mixed: FP32 phase + FP64 phase for the first-order form) vs the
    CPU reference solver (numpy iLOBPCG on the oracle operators), same psi0
    and start block: omega to 1e-9, and the modes' subspaces to 1e-6 rad."""
    import bdg_lobpcg_amd as B
    from oracle.ilobpcg_ref import ilobpcg, solver_precond_weights
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, cplx, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    S = c.size
    nev, m = 6, 8
    pKs, pMs = solver_precond_weights(f["ltK"], f["ltM"], f["mu"])
    c.set_fields(pK=pKs, pM=pMs, mu=f["mu"])
    X0 = c.start_block_uniform(42, m)
    if cplx not in _REF:
        _REF[cplx] = ilobpcg(lambda X: c.apply("K", X), lambda X: c.apply("M", X),
                             lambda X: c.apply("precondK", X), lambda X: c.apply("precondM", X),
                             X0[S:], nev, tol=1e-11, maxiter=400)
    ref = _REF[cplx]
    p = B.Problem(N, L, cplx)
    p.set_system()
    p.set_trap_harmonic((1.0, 1.0, 1.0))
    p.set_wavefunction(f["wf"])
    p.set_contact(20.0)
    p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
    p.set_mu(1.0)
    p.set_solver_params(nev, m, 400, 1e-10)
    p.set_precision(prec)
    p.set_formulation(form)
    p.set_init_custom_block(X0)
    print(f"3d dipolar vs oracle: prec {prec} cplx {cplx} form {form}", flush=True)
    assert p.solve() == 0
    st = p.stats()
    print(f"  iterations {st['iterations']} (fp32 {st['iterations_fp32']}), "
          f"fp32 applies {st['column_applies_fp32']}", flush=True)
    if prec == 1 and form == 0:
        assert st["iterations_fp32"] > 0 and st["column_applies_fp32"] > 0
    w = p.eigenvalues(nev)
    z = p.zero_modes(nev) | ref["zero"]
    nz = ~z
    np.testing.assert_allclose(w[nz], ref["w"][nz], rtol=1e-9)
    assert np.all(np.abs(w[z] - ref["w"][z]) < 1e-6)
    assert p.converged() >= nev
    # eigenvectors: principal angles between the device and reference modes
    # (v half), per cluster of equal omega
    ang = cluster_angles(w, p.modes(nev)[1], ref["v"], z)
    print(f"max subspace angle vs reference solver {ang:.2e}", flush=True)
    assert ang < 1e-6


def test_graph_capture_gives_identical_solve():
    """The solver's graph-replayed iteration chain gives the same iterations
    and eigenvalues as the eager chain."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 32), (14.0, 14.0, 14.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    X0 = c.start_block_uniform(42, 8)
    out = {}
    for on in (False, True):
        p = B.Problem(N, L, False)
        p.set_system()
        p.set_trap_harmonic((1.0, 1.0, 1.0))
        p.set_wavefunction(f["wf"])
        p.set_contact(20.0)
        p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
        p.set_mu(1.0)
        p.set_solver_params(6, 8, 300, 1e-9)
        p.set_init_custom_block(X0)
        p.set_formulation(B.FORM_PRODUCT)  # the product form replays graphs
        p.set_graph_capture(on)
        assert p.solve() == 0
        out[on] = (p.eigenvalues(6), p.stats())
        p.close()
    (w0, s0), (w1, s1) = out[False], out[True]
    assert s0["graph_launches"] == 0
    assert s1["graph_launches"] > 0 and s1["graph_captures"] > 0
    assert s0["iterations"] == s1["iterations"]
    np.testing.assert_array_equal(w0, w1)


def test_solve_batch_matches_sequential():
    """bdg_solve_batch (concurrent worker threads, one stream per problem)
    gives the sequential solves' eigenvalues bit for bit."""
    import bdg_lobpcg_amd as B
    setups = [((16, 16, 16), (12.0, 12.0, 12.0), 20.0, 10.0),
              ((12, 12, 16), (10.0, 10.0, 14.0), 30.0, 5.0),
              ((16, 16, 12), (12.0, 12.0, 10.0), 10.0, 8.0)]

    def build():
        out = []
        for N, L, g, gdd in setups:
            c, f = physics(N, L, False, dipolar=True, g=g, gdd=gdd, mu=1.0)
            p = B.Problem(N, L, False)
            p.set_system()
            p.set_trap_harmonic((1.0, 1.0, 1.0))
            p.set_wavefunction(f["wf"])
            p.set_contact(g)
            p.set_dipolar(gdd, (0.0, 0.0, 1.0), f["R"])
            p.set_mu(1.0)
            p.set_solver_params(4, 6, 300, 1e-9)
            p.set_init_uniform(42)
            out.append(p)
        return out

    seq = build()
    assert [p.solve() for p in seq] == [0, 0, 0]
    w_seq = [p.eigenvalues(4) for p in seq]
    st_seq = [(p.stats()["iterations"], p.converged()) for p in seq]
    bat = build()
    assert B.Problem.solve_batch(bat, 0) == [0, 0, 0]
    for p, w, st in zip(bat, w_seq, st_seq):
        np.testing.assert_array_equal(p.eigenvalues(4), w)
        assert (p.stats()["iterations"], p.converged()) == st


def test_device_start_block_gives_identical_solve():
    """The built-in uniform start block generated on the device gives the
    same iterations and eigenvalues as the same block filled on the host
    through a callback."""
    import bdg_lobpcg_amd as B
    N, L = (32, 32, 32), (14.0, 14.0, 14.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    out = {}
    for host in (False, True):
        p = B.Problem(N, L, False)
        p.set_system()
        p.set_trap_harmonic((1.0, 1.0, 1.0))
        p.set_wavefunction(f["wf"])
        p.set_contact(20.0)
        p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
        p.set_mu(1.0)
        p.set_solver_params(6, 8, 300, 1e-9)
        p.set_init_uniform(42, host=host)
        assert p.solve() == 0
        out[host] = (p.eigenvalues(6), p.stats()["iterations"], p.modes(6)[1])
        p.close()
    (w0, i0, v0), (w1, i1, v1) = out[False], out[True]
    assert i0 == i1
    np.testing.assert_array_equal(w0, w1)
    np.testing.assert_array_equal(v0, v1)


@pytest.mark.parametrize("cplx", [False, True])
def test_frozen_columns_match_soft_locking(cplx):
    """Converged columns that the Rayleigh-Ritz leaves in place keep their
    vectors and the block update skips them (bdg_set_freeze, default on):
    same eigenvalues as plain soft locking (every column updated) to 1e-9,
    every mode converged in both, and skipped column writes only with the
    freeze on."""
    import bdg_lobpcg_amd as B
    N, L = (24, 24, 24), (12.0, 12.0, 12.0)
    c, f = physics(N, L, cplx, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    nev, m = 6, 8
    X0 = c.start_block_uniform(42, m)
    out = {}
    for tol in (0.0, 1e-6):
        p = B.Problem(N, L, cplx)
        p.set_system()
        p.set_trap_harmonic((1.0, 1.0, 1.0))
        p.set_wavefunction(f["wf"])
        p.set_contact(20.0)
        p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
        p.set_mu(1.0)
        p.set_solver_params(nev, m, 400, 1e-9)
        p.set_init_custom_block(X0)
        p.set_formulation(B.FORM_PRODUCT)  # frozen columns are a product-form feature
        p.set_freeze(tol)
        assert p.solve() == 0
        out[tol] = (p.eigenvalues(nev), p.zero_modes(nev), p.stats(), p.converged(),
                    p.residuals(nev))
        p.close()
    (w0, z0, s0, c0, r0), (w1, z1, s1, c1, r1) = out[0.0], out[1e-6]
    print("iterations", s0["iterations"], s1["iterations"], "skips", s1["frozen_skips"],
          flush=True)
    assert c0 >= nev and c1 >= nev
    assert s0["frozen_skips"] == 0 and s1["frozen_skips"] > 0
    nz = ~(z0 | z1)
    np.testing.assert_allclose(w1[nz], w0[nz], rtol=1e-9)
    assert np.all(r1[nz] < 1e-9)


def test_mixed_fp64_groups():
    """Mixed first-order solve whose FP64 phase refines the modes in groups
    (the 256^3 configuration's path, forced here at 16^3 with groups of 3 of
    9 modes): each group iterated in FP64 from the FP32 phase's vectors and
    deflated against the groups before it.  Omega to 1e-9 of the FP64 solve,
    every mode converged to the tolerance."""
    import bdg_lobpcg_amd as B
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    out = {}
    for grouped in (False, True):
        p = B.Problem(N, L, False)
        p.set_system()
        p.set_trap_harmonic((1.0, 1.0, 1.0))
        p.set_wavefunction(f["wf"])
        p.set_contact(20.0)
        p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
        p.set_mu(1.0)
        p.set_solver_params(9, 12, 600, 1e-10)
        p.set_precision(1 if grouped else 0)
        p.set_init_uniform(42)
        p.set_fp64_groups(3 if grouped else 0)
        print(f"fp64 groups case: {grouped}", flush=True)
        assert p.solve() == 0
        out[grouped] = (p.eigenvalues(9), p.residuals(9), p.stats(), p.converged())
        p.close()
    w64 = out[False][0]
    w, res, st, conv = out[True]
    print(f"  iterations {st['iterations']} (fp32 {st['iterations_fp32']}), res {res.max():.2e}",
          flush=True)
    assert st["fp64_group_size"] == 3 and conv >= 9
    assert np.all(res < 1e-10)
    small = np.abs(w64) < 1e-2
    err = np.abs(w - w64) / np.where(small, 1.0, np.abs(w64))
    assert np.all(err < np.where(small, 1e-7, 1e-9)), (w, w64)


def test_mixed_fp64_sweep_only():
    """Mixed first-order solve whose FP64 phase is the refinement sweep alone
    (the path the 256^3 configuration takes because its FP64 basis does not
    fit the device; forced here at 16^3): the FP32 phase's vectors, one FP64
    Rayleigh-Ritz, FP64 residuals.  Omega within 1e-7 of the FP64 solve (the
    Ritz values of vectors with residual r are off by O(r^2)), residuals at
    the FP32 phase's level, no FP64 iterations."""
    import bdg_lobpcg_amd as B
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    out = {}
    for sweep in (False, True):
        p = B.Problem(N, L, False)
        p.set_system()
        p.set_trap_harmonic((1.0, 1.0, 1.0))
        p.set_wavefunction(f["wf"])
        p.set_contact(20.0)
        p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
        p.set_mu(1.0)
        p.set_solver_params(6, 8, 400, 1e-10)
        p.set_precision(0 if not sweep else 1)
        p.set_init_uniform(42)
        p.set_fp64_sweep_only(sweep)
        print(f"sweep-only case: {sweep}", flush=True)
        assert p.solve() == 0
        out[sweep] = (p.eigenvalues(6), p.residuals(6), p.stats())
        p.close()
    w64, _, _ = out[False]
    w, res, st = out[True]
    print(f"  iterations {st['iterations']} (fp32 {st['iterations_fp32']}), res {res.max():.2e}",
          flush=True)
    assert st["fp64_sweep_only"] == 1 and st["iterations"] == st["iterations_fp32"]
    small = np.abs(w64) < 1e-2
    err = np.abs(w - w64) / np.where(small, 1.0, np.abs(w64))
    assert np.all(err < 1e-7), (w, w64)
    assert np.all(res < 1e-4)
