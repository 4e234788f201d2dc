"""Device-resident parameter sweeps: the sequence of reference
tests/test_io.c:390-446 (solve, bdg_reuse_modes, bdg_reset, re-setup with
mu x 1.01, solve) on a 3D dipolar problem, with the modes, the reuse block
(reference src/bdg.c:112-207) and the warm start kept on the device."""
import numpy as np
import pytest

from tests.common import physics

pytestmark = pytest.mark.gpu


def _setup(p, f, mu):
    """the physics of the problem (after bdg_set_system or bdg_reset)"""
    p.set_trap_harmonic((1.0, 1.0, 1.0))
    p.set_wavefunction(f["wf"])
    p.set_contact(20.0)
    p.set_dipolar(10.0, (0.0, 0.0, 1.0), f["R"])
    p.set_mu(mu)


def _problem(N, L, cplx, f, mu, nev, m, device_sweeps=True):
    import bdg_lobpcg_amd as B
    p = B.Problem(N, L, cplx)
    p.set_system()
    _setup(p, f, mu)
    p.set_solver_params(nev, m, 400, 1e-9)
    p.set_init_uniform(42)
    p.set_device_sweeps(device_sweeps)
    return p


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_device_sweep_reuse_block_and_warm_start(cplx):
    """(1) The reuse block built on the device from the device-resident modes
    equals the block the host builds from the downloaded modes (device sweeps
    off) and the oracle's restatement of the recipe, to 1e-13 (the device
    log / cos / sin and the norm's summation order differ in the last bits).
    (2) After bdg_reset and a re-setup with mu x 1.01, the solve started from
    the device block reaches the cold solve's omega (1e-9 relative, 1e-7
    absolute for the Goldstone pair) in fewer iterations."""
    from oracle import oracle as O
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, cplx, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    S = c.size
    nev, m, noise = 6, 8, 1e-3
    print(f"sweep: cplx {cplx} N {N} nev {nev} m {m}", flush=True)
    blocks = {}
    for dev in (True, False):
        p = _problem(N, L, cplx, f, 1.0, nev, m, device_sweeps=dev)
        assert p.solve() == 0 and p.converged() >= nev
        U, V = p.modes(nev)
        assert p.reuse_modes(noise) == 0
        blocks[dev] = (p, U, V, p.start_block((2 * S, m)))
    p, U, V, Xd = blocks[True]
    _, U2, V2, Xh = blocks[False]
    np.testing.assert_array_equal(U, U2)  # the two solves are the same solve
    ref = O.reuse_block(U, V, m, noise, f["wf"])
    scale = np.abs(ref).max()
    e_dev = np.abs(Xd - ref).max() / scale
    e_host = np.abs(Xh - ref).max() / scale
    print(f"  reuse block vs oracle: device {e_dev:.2e}, host {e_host:.2e}", flush=True)
    assert e_dev < 1e-13 and e_host < 1e-13
    blocks[False][0].close()
    # warm start after reset and re-setup (reference tests/test_io.c:419-428)
    it_first = p.stats()["iterations"]
    p.reset()
    _setup(p, f, 1.01)
    assert p.solve() == 0 and p.converged() >= nev
    warm_w, warm_it = p.eigenvalues(nev), p.stats()["iterations"]
    q = _problem(N, L, cplx, f, 1.01, nev, m)
    assert q.solve() == 0 and q.converged() >= nev
    cold_w, cold_it = q.eigenvalues(nev), q.stats()["iterations"]
    print(f"  iterations: first {it_first}, warm {warm_it}, cold {cold_it}; "
          f"w warm {warm_w}", flush=True)
    small = np.abs(cold_w) < 1e-2
    err = np.abs(warm_w - cold_w) / np.where(small, 1.0, np.abs(cold_w))
    assert np.all(err < np.where(small, 1e-7, 1e-9)), (warm_w, cold_w)
    assert warm_it < cold_it
    # the block is consumed by the solve: the next one starts WF-weighted
    p.close()
    q.close()


def test_binary_modes_file(tmp_path):
    """bdg_write_modes_bin holds the solve's eigenvalues and modes bit for
    bit, and agrees with the reference's text writers (%.8e, reference
    src/io.c:72-165) to their 9 significant digits."""
    import bdg_lobpcg_amd as B
    N, L = (16, 16, 16), (12.0, 12.0, 12.0)
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    nev, m = 6, 8
    p = _problem(N, L, False, f, 1.0, nev, m)
    assert p.solve() == 0
    path = tmp_path / "modes.bin"
    assert p.write_modes_bin(path) == 0
    hdr, w, U, V = B.read_bin(path)
    assert hdr == dict(kind=2, complex=False, dim=3, N=N, count=nev)
    Ur, Vr = p.modes(nev)
    np.testing.assert_array_equal(w, p.eigenvalues(nev))
    np.testing.assert_array_equal(U, Ur)
    np.testing.assert_array_equal(V, Vr)
    base = str(tmp_path / "mode")
    assert B.lib().bdg_write_mode_u(p.h, 2, base.encode()) == 0
    txt = np.loadtxt(f"{base}_u_2.dat").reshape(-1, 2)[:, 0]
    np.testing.assert_allclose(txt, U[:, 2], rtol=1e-8, atol=1e-8 * np.abs(U[:, 2]).max())
    p.close()
