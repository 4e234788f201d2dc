"""Synthetic BdG workloads (the benchmark configurations) built through the
public API in the reference's setup order (examples/3d_jens.c:53-67):
system -> trap -> wavefunction -> local interactions -> dipolar -> mu.

Conventions settled here (the configuration text leaves them open):
  * psi0: normalised Gaussian of width 1 (sum |psi|^2 dV = 1), relaxed by
    imaginary-time split-step steps of `dt` on the device
    (bdgx_itp_groundstate), mu = <psi| H_GP |psi> of the final state: the
    stated `itp_steps` where the configuration gives a step count (32^3:
    2000, 128^3: 4000); where it only asks for a relaxed state (the complex
    96x96x192 tube, the batch) chunks of `itp_steps` until the imaginary-time
    residual ||psi' - psi|| / (dt ||psi||) (psi' one more step from psi, not
    kept) is at most `itp_tol` (at most `itp_max` steps);
  * contact: U_K = U_M = g * density;
  * dipolar coupling g_ddi = g * eps_dd (the reference examples' convention,
    examples/3d_dipolar_atom.c:39-48), dipoles along z, cutoff R as given
    (configs without one use R = min(L) / 2);
  * start block: BDG_INIT_CUSTOM with xorshift32 seed 42, u = v = xrand - 0.5.
"""
from __future__ import annotations

import time

import numpy as np

CONFIGS = {
    "cfg0_32c_d": dict(N=(32, 32, 32), L=(16.0, 16.0, 16.0), omega=(1.0, 1.0, 1.0), g=100.0,
                       edd=0.5, R=8.0, itp_steps=2000, dt=1e-3, nev=10, m=10, tol=1e-8,
                       cplx=False),
    "cfg1_128c_d": dict(N=(128, 128, 128), L=(24.0, 24.0, 48.0), omega=(1.0, 1.0, 0.5),
                        g=2000.0, edd=0.8, R=12.0, itp_steps=4000, dt=1e-3, nev=32, m=32,
                        tol=1e-8, cplx=False),
    "cfg2_96x96x192_z": dict(N=(96, 96, 192), L=(20.0, 20.0, 80.0), omega=(3.0, 3.0, 0.3),
                             g=3000.0, edd=0.95, R=10.0, itp_steps=2000, dt=1e-3, nev=64, m=64,
                             tol=1e-8, cplx=True, modulation=0.3, mod_period=4.0, vortex=True,
                             itp_tol=1e-8, itp_max=40000),
    "cfg4_256c_d": dict(N=(256, 256, 256), L=(32.0, 32.0, 32.0), omega=(1.0, 1.0, 2.0),
                        g=5000.0, edd=0.7, R=16.0, itp_steps=2000, dt=1e-3, nev=128, m=128,
                        tol=1e-8, cplx=False, itp_tol=1e-8, itp_max=40000),
}


def coords(N, L):
    xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(N, L)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def initial_psi(cfg):
    X, Y, Z = coords(cfg["N"], cfg["L"])
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2.0)
    if cfg.get("modulation"):
        psi = psi * (1.0 + cfg["modulation"] * np.cos(2 * np.pi * Z / cfg["mod_period"]))
    if cfg["cplx"]:
        psi = psi.astype(np.complex128)
        if cfg.get("vortex"):
            psi = psi * np.exp(1j * np.arctan2(Y, X))
    dV = np.prod([l / n for n, l in zip(cfg["N"], cfg["L"])])
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dV)


def ground_state(cfg, steps=None, info=None):
    """psi0 and mu by device imaginary-time evolution; returns (psi, mu,
    seconds).  `steps` overrides the configuration's step rule with a fixed
    count.  `info` (a dict), if given, receives the steps taken and the
    imaginary-time residual of the last step."""
    import bdg_lobpcg_amd as B

    d = B.DeviceContext(cfg["N"], cfg["L"], cfg["cplx"])
    t0 = time.perf_counter()

    def run(psi, n):
        return d.itp_groundstate(psi, cfg["omega"], cfg["g"], cfg["g"] * cfg["edd"],
                                 (0.0, 0.0, 1.0), cfg["R"], cfg["dt"], n)

    def residual(psi):  # one more step's change (the step is not kept)
        nxt, _ = run(psi, 1)
        return float(np.linalg.norm(nxt - psi) / (cfg["dt"] * np.linalg.norm(psi)))

    chunk = cfg["itp_steps"] if steps is None else steps
    relax = steps is None and cfg.get("itp_tol") is not None
    psi, mu = run(initial_psi(cfg), chunk)
    total = chunk
    while relax and total + chunk <= cfg["itp_max"] and residual(psi) > cfg["itp_tol"]:
        psi, mu = run(psi, chunk)
        total += chunk
    dt = time.perf_counter() - t0
    res = residual(psi) if info is not None else float("nan")
    d.close()
    if info is not None:
        info.update(itp_steps=total, itp_residual=res)
    return psi, mu, dt


# graph replay of the solver's iteration chain for the problems built here
# (bdg_set_graph_capture; on by default, as in the library)
GRAPHS = True


def make_problem(cfg, psi, mu, precision=0, maxiter=1000, tol=None, nev=None, m=None):
    import bdg_lobpcg_amd as B

    p = B.Problem(cfg["N"], cfg["L"], cfg["cplx"])
    p.set_system()
    p.set_trap_harmonic(cfg["omega"])
    p.set_wavefunction(psi)
    p.set_contact(cfg["g"])
    p.set_dipolar(cfg["g"] * cfg["edd"], (0.0, 0.0, 1.0), cfg["R"])
    p.set_mu(mu)
    p.set_solver_params(nev or cfg["nev"], m or cfg["m"], maxiter, tol or cfg["tol"])
    p.set_precision(precision)
    p.set_init_uniform(42)
    p.set_graph_capture(GRAPHS)
    return p


def batch_configs(n=16, seed=20260601):
    """The many-problems workload: n independent problems with grids drawn from
    {48^3, 64^3, 80^3, 96^3, 64^2 x 128}, eps_dd ~ U[0.3, 0.95], g ~ U[500, 3000]
    and trap aspect w_z / w_perp ~ U[0.25, 4] (w_perp = 1), by xorshift32 from
    `seed`; nev = 16.  Box: 16 across, along z 16 / sqrt(aspect) (elongated
    traps get longer boxes), times Nz / Nx for the 64^2 x 128 grid."""
    grids = [(48, 48, 48), (64, 64, 64), (80, 80, 80), (96, 96, 96), (64, 64, 128)]
    st = [seed & 0xFFFFFFFF or 1]

    def xr():
        x = st[0]
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        st[0] = x
        return (x & 0x7FFFFFFF) / 0x7FFFFFFF

    out = []
    for i in range(n):
        N = grids[min(int(xr() * len(grids)), len(grids) - 1)]
        edd = 0.3 + 0.65 * xr()
        g = 500.0 + 2500.0 * xr()
        asp = 0.25 * 16.0 ** xr()  # log-uniform on [0.25, 4]
        lz = 16.0 / np.sqrt(asp) * N[2] / N[0]
        out.append(dict(N=N, L=(16.0, 16.0, lz), omega=(1.0, 1.0, asp), g=g, edd=edd,
                        R=min(16.0, lz) / 2, itp_steps=2000, itp_tol=1e-8, itp_max=40000,
                        dt=1e-3, nev=16, m=16, tol=1e-8, cplx=False, name=f"batch{i}"))
    return out


def train_learned(cfg, steps=300, lr=1e-3, reg=0.3, snapshots=(4, 8, 14, 22, 32, 48, 70),
                  log=None, form="product", weighted=False):
    """Learned-preconditioner parameters for cfg's grid, trained on the GPU on
    pairs from a related problem (g x 0.75, eps_dd x 0.875: same grid and trap,
    different condensate), so that the evaluation problem is not in the
    training set.  form "product": the whole network, for the product form's
    T_M T_K; "first_order": the two spectral layers, for the first-order
    form's per-half T_K, T_M.  Returns (theta, seconds, info)."""
    from bdg_lobpcg_amd import learned as LP

    t0 = time.perf_counter()
    tcfg = dict(cfg, g=cfg["g"] * 0.75, edd=cfg["edd"] * 0.875)
    psi_t, mu_t, _ = ground_state(tcfg)
    if form == "first_order":
        def mk_fo(maxiter):
            return make_problem(tcfg, psi_t, mu_t, maxiter=maxiter)

        ref, d, snaps = LP.collect_fo(mk_fo, tcfg["nev"], snapshots=snapshots, log=log)
        th0 = np.zeros(len(LP.init_theta(0)), dtype=np.float32)
        loss0 = LP.evaluate_fo(d, th0, snaps, weighted)
        th, hist = LP.train_fo(d, snaps, th0, steps=steps, lr=lr, reg=reg, log=log,
                               weighted=weighted)
        loss1 = LP.evaluate_fo(d, th, snaps, weighted)
        d.close()
        ref.close()
        return th, time.perf_counter() - t0, dict(
            form=form, weighted=weighted, loss_analytic=loss0, loss_learned=loss1, steps=steps,
            lr=lr, reg=reg,
            snapshots=list(snapshots), train_problem=f"g={tcfg['g']}, eps_dd={tcfg['edd']}")
    def mk(maxiter):
        # the network replaces the product form's T_M T_K: its training pairs
        # come from that iteration
        p = make_problem(tcfg, psi_t, mu_t, maxiter=maxiter)
        p.set_formulation(1)
        return p

    ref, d, snaps = LP.collect(mk, tcfg["nev"], snapshots=snapshots)
    th0 = LP.init_theta(0)
    loss0 = LP.evaluate(d, th0, snaps)
    th, hist = LP.train(d, snaps, th0, steps=steps, lr=lr, reg=reg, log=log)
    loss1 = LP.evaluate(d, th, snaps)
    d.close()
    ref.close()
    return th, time.perf_counter() - t0, dict(loss_analytic=loss0, loss_learned=loss1,
                                              steps=steps, lr=lr, reg=reg,
                                              snapshots=list(snapshots),
                                              train_problem=f"g={tcfg['g']}, eps_dd={tcfg['edd']}")
