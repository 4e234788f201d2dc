"""Learned preconditioner: the numpy restatement (oracle/learned_ref.py) pinned
on the CPU, and (GPU) the device forward and parameter gradient against it."""
import numpy as np
import pytest

from oracle import learned_ref as LR
from tests.common import physics, rand_block, relerr


def _small(N=(8, 6, 4), L=(6.0, 5.0, 4.0), seed=0):
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    net = LR.LearnedRef(N, [c.kx2(d) for d in range(3)], f["pK"], f["pM"], f["wf"], f["mu"])
    return c, f, net


def test_zero_theta_is_the_analytic_preconditioner():
    """With v = 0 and alpha = beta = kappa = 0 the network is T_M T_K exactly
    (reference src/precond_impl.inc:21-111 as restated by the C oracle)."""
    c, f, net = _small()
    th = LR.init_theta(seed=1)  # random stencil weights act only through v = 0
    R = rand_block(c.size, 3, False, seed=4)
    ref = c.apply("precondM", c.apply("precondK", R))
    assert relerr(net.forward(th, R), ref) < 1e-12


def test_gradient_matches_finite_differences():
    c, f, net = _small()
    rng = np.random.default_rng(7)
    th = LR.init_theta(seed=2, wscale=0.5)
    th[:8] = 0.3 * rng.standard_normal(8)
    th[8:8 + 2 * LR.NKNOT] = 0.2 * rng.standard_normal(2 * LR.NKNOT)
    o = 8 + 2 * LR.NKNOT + 27 * LR.HID
    th[o:] = 0.3 * rng.standard_normal(2 * LR.HID)
    R = rand_block(c.size, 2, False, seed=5)
    dW = rand_block(c.size, 2, False, seed=6)
    g = net.grad(th, R, dW)
    f0 = lambda t: float(np.sum(dW * net.forward(t, R)))
    h = 1e-5
    idx = list(range(0, 8 + 2 * LR.NKNOT)) + list(rng.choice(np.arange(o - 27 * LR.HID, o), 20,
                                                              replace=False)) + list(range(o, LR.NPARAM))
    fd = np.zeros(len(idx))
    for n, i in enumerate(idx):
        e = np.zeros_like(th)
        e[i] = h
        fd[n] = (f0(th + e) - f0(th - e)) / (2 * h)
    np.testing.assert_allclose(g[idx], fd, rtol=2e-5, atol=1e-7 * np.abs(fd).max())


@pytest.mark.parametrize("layer", [0, 1], ids=["K", "M"])
def test_layer_alone_zero_theta_and_gradient(layer):
    """One spectral layer (the first-order solver's per-half preconditioner):
    at theta = 0 it is the oracle's precondK / precondM; its gradient against
    central finite differences; entries outside the layer are zero."""
    c, f, net = _small()
    R = rand_block(c.size, 2, False, seed=8)
    ref = c.apply("precondK" if layer == 0 else "precondM", R)
    assert relerr(net.layer_forward(np.zeros(LR.NPARAM), R, layer), ref) < 1e-12
    rng = np.random.default_rng(9)
    th = np.zeros(LR.NPARAM)
    th[:8] = 0.3 * rng.standard_normal(8)
    th[8:8 + 2 * LR.NKNOT] = 0.2 * rng.standard_normal(2 * LR.NKNOT)
    dW = rand_block(c.size, 2, False, seed=10)
    g = net.layer_grad(th, R, dW, layer)
    f0 = lambda t: float(np.sum(dW * net.layer_forward(t, R, layer)))
    own = (list(range(0, 4)) + list(range(8, 8 + LR.NKNOT)) if layer == 0 else
           list(range(4, 8)) + list(range(8 + LR.NKNOT, 8 + 2 * LR.NKNOT)))
    h = 1e-5
    for i in own:
        e = np.zeros_like(th)
        e[i] = h
        fd = (f0(th + e) - f0(th - e)) / (2 * h)
        assert abs(g[i] - fd) <= 2e-5 * abs(fd) + 1e-7 * np.abs(g).max(), (i, g[i], fd)
    assert np.all(np.delete(g, own) == 0)


def test_cosine_loss_gradient():
    rng = np.random.default_rng(3)
    W, E = rng.standard_normal((50, 3)), rng.standard_normal((50, 3))
    loss, dW = LR.cosine_loss(W, E)
    h = 1e-6
    for (i, j) in [(0, 0), (7, 1), (49, 2)]:
        Wp, Wm = W.copy(), W.copy()
        Wp[i, j] += h
        Wm[i, j] -= h
        fd = (LR.cosine_loss(Wp, E)[0] - LR.cosine_loss(Wm, E)[0]) / (2 * h)
        assert abs(fd - dW[i, j]) < 1e-7
    assert LR.cosine_loss(E * 3.0, E)[0] < 1e-14


def _device(N, L):
    from tests.common import device_twin
    c, f = physics(N, L, False, dipolar=True, g=20.0, gdd=10.0, mu=1.0)
    d = device_twin(N, L, False, f)
    net = LR.LearnedRef(N, [c.kx2(k) for k in range(3)], f["pK"], f["pM"], f["wf"], f["mu"])
    return c, f, d, net


def _theta(seed):
    rng = np.random.default_rng(seed)
    th = LR.init_theta(seed=seed, wscale=0.5)
    th[:8] = 0.3 * rng.standard_normal(8)
    th[8:8 + 2 * LR.NKNOT] = 0.2 * rng.standard_normal(2 * LR.NKNOT)
    o = 8 + 2 * LR.NKNOT + 27 * LR.HID
    th[o:] = 0.3 * rng.standard_normal(2 * LR.HID)
    return th


@pytest.mark.gpu
@pytest.mark.parametrize("N,L", [((16, 12, 10), (12.0, 10.0, 9.0)), ((32, 32, 32), (16.0, 16.0, 12.0))],
                         ids=["pencil", "plane"])
def test_device_layer_forward_and_gradient(N, L):
    """bdgx_lp_apply_layer / bdgx_lp_backward_layer (the first-order solver's
    per-half learned preconditioner) against the numpy restatement."""
    c, f, d, net = _device(N, L)
    th = _theta(13)
    d.lp_set_params(th.astype(np.float32))
    t64 = th.astype(np.float32).astype(np.float64)
    R = rand_block(c.size, 3, False, seed=23)
    G = rand_block(c.size, 3, False, seed=24)
    R64, G64 = R.astype(np.float32).astype(np.float64), G.astype(np.float32).astype(np.float64)
    print(f"layer case N={N}", flush=True)
    for layer in (0, 1):
        W = d.lp_apply_layer(layer, R)
        assert relerr(W, net.layer_forward(t64, R64, layer)) < 2e-5
        dR, dG = d.upload(R.astype(np.float32)), d.upload(G.astype(np.float32))
        g = d.lp_backward_layer_device(layer, dR, dG, 3)
        gr = net.layer_grad(t64, R64, G64, layer)
        assert np.abs(g - gr).max() < 1e-3 * np.abs(gr).max(), (layer, g[:8], gr[:8])


@pytest.mark.gpu
@pytest.mark.parametrize("N,L", [((16, 12, 10), (12.0, 10.0, 9.0)), ((32, 32, 32), (16.0, 16.0, 12.0))],
                         ids=["pencil", "plane"])
def test_device_forward_and_gradient(N, L):
    """Device network forward and parameter gradient (hand-written HIP forward
    and adjoint kernels) against the numpy restatement, FP32 tolerances."""
    c, f, d, net = _device(N, L)
    th = _theta(11)
    d.lp_set_params(th.astype(np.float32))
    R = rand_block(c.size, 3, False, seed=21)
    G = rand_block(c.size, 3, False, seed=22)
    print(f"case N={N}", flush=True)
    W = d.lp_apply(R)
    Wr = net.forward(th.astype(np.float32).astype(np.float64), R.astype(np.float32).astype(np.float64))
    assert relerr(W, Wr) < 2e-5
    g = d.lp_backward(R, G)
    gr = net.grad(th.astype(np.float32).astype(np.float64), R.astype(np.float32).astype(np.float64),
                  G.astype(np.float32).astype(np.float64))
    scale = np.abs(gr).max()
    # per parameter group
    K = LR.NKNOT
    o = 8 + 2 * K
    for name, sl in [("pointwise", slice(0, 8)), ("knots", slice(8, o)),
                     ("stencil", slice(o, LR.NPARAM))]:
        err = np.abs(g[sl] - gr[sl]).max() / max(np.abs(gr[sl]).max(), 1e-30)
        assert err < 1e-3, f"{name}: rel err {err:.2e}"
    assert np.abs(g - gr).max() < 1e-3 * scale


@pytest.mark.gpu
def test_device_zero_theta_is_the_analytic_preconditioner():
    c, f, d, net = _device((16, 16, 16), (12.0, 12.0, 12.0))
    th = LR.init_theta(seed=3)
    d.lp_set_params(th.astype(np.float32))
    R = rand_block(c.size, 4, False, seed=5)
    ref = c.apply("precondM", c.apply("precondK", R))
    assert relerr(d.lp_apply(R), ref) < 2e-5


@pytest.mark.gpu
def test_trained_network_gain_on_32c():
    """What the trained network gains over the method without it, on the
    32^3 dipolar configuration (box [-8,8]^3, g = 100, eps_dd = 0.5, nev =
    block = 10, tol 1e-8): the network is trained on the GPU on pairs from a
    related problem (g x 0.75, eps_dd x 0.875), then the same mixed
    product-form solve (the form whose preconditioner T_M T_K the network
    replaces) runs from the same start block with T_M T_K and with the
    network.  Both reach
    the same omega (<= 1e-9 relative, non-zero modes) with every mode
    converged; the network lowers the training loss and the iteration count."""
    from bdg_lobpcg_amd.workloads import CONFIGS, ground_state, make_problem, train_learned
    cfg = CONFIGS["cfg0_32c_d"]
    theta, secs, info = train_learned(cfg)
    print(f"training {secs:.1f} s, loss {info['loss_analytic']:.4f} -> "
          f"{info['loss_learned']:.4f}", flush=True)
    assert info["loss_learned"] < info["loss_analytic"]
    psi, mu, _ = ground_state(cfg)
    out = {}
    for name, th in (("analytic", None), ("learned", theta)):
        p = make_problem(cfg, psi, mu, precision=1)
        p.set_formulation(1)  # product form
        if th is not None:
            p.set_learned(th)
        assert p.solve() == 0
        out[name] = (p.eigenvalues(cfg["nev"]), p.zero_modes(cfg["nev"]), p.stats(),
                     p.converged())
        p.close()
    (w0, z0, s0, c0), (w1, z1, s1, c1) = out["analytic"], out["learned"]
    print(f"iterations: T_M T_K {s0['iterations']}, network {s1['iterations']} "
          f"(learned iterations {s1['learned_iterations']}, fallback at "
          f"{s1['learned_fallback_at']})", flush=True)
    assert c0 >= cfg["nev"] and c1 >= cfg["nev"]
    nz = ~(z0 | z1)
    np.testing.assert_allclose(w1[nz], w0[nz], rtol=1e-9)
    assert s1["learned_iterations"] > 0
    assert s1["iterations"] < s0["iterations"]


@pytest.mark.gpu
def test_stall_fallback_with_a_useless_network():
    """A network that only passes high-k content (radial gain exp(-40) at
    k = 0 rising to 1 at the top knot, in both spectral layers) gives search
    directions that do not reduce the residual: the stall rule hands over to
    T_M T_K (fallback recorded), and the solve still converges every mode to
    the analytic solve's omega (<= 1e-9 relative, non-zero modes)."""
    from bdg_lobpcg_amd import LP_NKNOT, LP_NPARAM
    from bdg_lobpcg_amd.workloads import CONFIGS, ground_state, make_problem
    cfg = CONFIGS["cfg0_32c_d"]
    psi, mu, _ = ground_state(cfg)
    theta = np.zeros(LP_NPARAM)
    ramp = np.linspace(-40.0, 0.0, LP_NKNOT)
    theta[8:8 + LP_NKNOT] = ramp               # kappa_K
    theta[8 + LP_NKNOT:8 + 2 * LP_NKNOT] = ramp  # kappa_M
    out = {}
    for name, th in (("analytic", None), ("useless", theta)):
        p = make_problem(cfg, psi, mu, precision=1)
        p.set_formulation(1)  # product form (the network replaces its T_M T_K)
        if th is not None:
            p.set_learned(th)
        assert p.solve() == 0
        out[name] = (p.eigenvalues(cfg["nev"]), p.zero_modes(cfg["nev"]), p.stats(),
                     p.converged())
        p.close()
    (w0, z0, s0, c0), (w1, z1, s1, c1) = out["analytic"], out["useless"]
    print(f"iterations: T_M T_K {s0['iterations']}, useless network {s1['iterations']} "
          f"(learned iterations {s1['learned_iterations']}, fallback at "
          f"{s1['learned_fallback_at']})", flush=True)
    assert s1["learned_fallback_at"] > 0
    assert c0 >= cfg["nev"] and c1 >= cfg["nev"]
    nz = ~(z0 | z1)
    np.testing.assert_allclose(w1[nz], w0[nz], rtol=1e-9)
