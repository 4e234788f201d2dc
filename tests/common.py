"""Shared fixtures: the same seeded physics set up on the CPU oracle and,
for GPU tests, on the device context."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def grid_coords(N, L):
    xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(N, L)]
    if len(N) == 1:
        return (xs[0],)
    if len(N) == 2:
        Y, X = np.meshgrid(xs[1], xs[0], indexing="ij")
        return X.ravel(), Y.ravel()
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def gaussian_psi(N, L, cplx=False, width=1.0, vortex=False):
    r = grid_coords(N, L)
    r2 = sum(c ** 2 for c in r)
    psi = np.exp(-r2 / (2 * width ** 2))
    if cplx:
        psi = psi.astype(np.complex128)
        if vortex and len(N) >= 2:
            psi = psi * np.exp(1j * np.arctan2(r[1], r[0]))
        else:
            psi = psi * np.exp(0.3j * r[0])
    return psi


def physics(N, L, cplx=False, dipolar=True, g=20.0, gdd=10.0, cutoff=None, mu=1.0,
            omega=None):
    """Oracle context with trap + contact (+ dipolar) + mu; returns (ctx, fields)."""
    c = O.Ctx(N, L, cplx)
    S = c.size
    psi = gaussian_psi(N, L, cplx)
    ltK, ltM = np.zeros(S), np.zeros(S)
    c.trap_harmonic(omega or (1.0,) * len(N), ltK, ltM)
    c.contact(psi, g, g, ltK, ltM)
    c.set_fields(wf=psi)
    f = dict(wf=psi, dipolar=False)
    if dipolar and len(N) == 3:
        R = cutoff if cutoff is not None else min(L) / 2
        c.set_dipolar(gdd, (0.0, 0.0, 1.0), R)
        gphi = c.dipolar_meanfield()
        ltK += gphi
        ltM += gphi
        f.update(dipolar=True, gdd=gdd, R=R)
    pK, pM = O.setup_mu(mu, ltK, ltM)
    c.set_fields(ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=mu)
    f.update(ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=mu)
    return c, f


def device_twin(N, L, cplx, f):
    import bdg_lobpcg_amd as B

    d = B.DeviceContext(N, L, cplx)
    d.set_fields(wf=f["wf"])
    if f.get("dipolar"):
        d.set_dipolar(f["gdd"], (0.0, 0.0, 1.0), f["R"])
    d.set_fields(ltK=f["ltK"], ltM=f["ltM"], pK=f["pK"], pM=f["pM"], mu=f["mu"])
    return d


def rand_block(rows, cols, cplx, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, cols))
    if cplx:
        X = X + 1j * rng.standard_normal((rows, cols))
    return np.asfortranarray(X)


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def cluster_angles(w, A, B, skip, rtol=1e-7):
    """Largest principal angle (radians) between span(A[:, J]) and
    span(B[:, J]) over the clusters J of equal w (relative rtol; degenerate
    modes are compared as subspaces), leaving out the indices in skip (zero
    modes) and the cluster holding the last index (its partners may lie
    beyond the computed modes)."""
    from scipy.linalg import subspace_angles
    n = len(w)
    worst, j0 = 0.0, 0
    while j0 < n:
        j1 = j0 + 1
        while j1 < n and abs(w[j1] - w[j0]) <= rtol * max(1.0, abs(w[j0])):
            j1 += 1
        J = [j for j in range(j0, j1) if not skip[j]]
        if J and j1 < n:
            worst = max(worst, float(np.max(subspace_angles(A[:, J], B[:, J]))))
        j0 = j1
    return worst


def oracle_problem(cfg, psi, mu):
    """Oracle context of a benchmark configuration (bdg_lobpcg_amd.workloads
    CONFIGS entry) for psi0 / mu: trap, contact U = g n, dipolar mean field
    with g_ddi = g eps_dd along z, mu (reference src/setup.c:17-177 order).
    Returns (ctx, fields)."""
    N, L, cplx = cfg["N"], cfg["L"], cfg["cplx"]
    c = O.Ctx(N, L, cplx)
    S = c.size
    ltK, ltM = np.zeros(S), np.zeros(S)
    c.trap_harmonic(cfg["omega"], ltK, ltM)
    c.contact(psi, cfg["g"], cfg["g"], ltK, ltM)
    c.set_fields(wf=psi)
    c.set_dipolar(cfg["g"] * cfg["edd"], (0.0, 0.0, 1.0), cfg["R"])
    gphi = c.dipolar_meanfield()
    ltK += gphi
    ltM += gphi
    pK, pM = O.setup_mu(mu, ltK, ltM)
    c.set_fields(ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=mu)
    return c, dict(wf=psi, ltK=ltK, ltM=ltM, pK=pK, pM=pM, mu=mu)


def oracle_mode_residuals(c, w, U, V, chunk=8):
    """Residuals of device modes under the oracle's operators (FP64 C,
    reference formulas), column by column:
      first order  the reference pencil's ||A x - w B x|| / (||x|| max(1, |w|))
                   for x = [u; v] (A = diag(K, M), B = Swap): K u = |w| v,
                   M v = w u -- for w < 0 (a pair with imaginary frequency,
                   reported as -|w| with M v = w u) K u = |w| v still holds;
      product      (||K M v - w^2 v|| / |w|) / (||[M v / w; v]|| max(1, |w|)),
                   w^2 carrying the sign of w (the product-form solver's
                   convergence measure).
    Returns (first_order[nev], product[nev])."""
    nev = len(w)
    first = np.zeros(nev)
    pen = np.zeros(nev)
    for j0 in range(0, nev, chunk):
        J = slice(j0, min(nev, j0 + chunk))
        u, v = U[:, J], V[:, J]
        Ku, Mv = c.apply("K", u), c.apply("M", v)
        KMv = c.apply("K", Mv)
        wj = w[J]
        aw = np.maximum(np.abs(wj), 1e-300)
        w2 = np.sign(wj) * wj ** 2
        rn = np.linalg.norm(KMv - v * w2[None, :], axis=0)
        xnorm = np.sqrt(np.linalg.norm(v, axis=0) ** 2 + (np.linalg.norm(Mv, axis=0) / aw) ** 2)
        pen[J] = (rn / aw) / (xnorm * np.maximum(1.0, aw))
        xn = np.sqrt(np.linalg.norm(u, axis=0) ** 2 + np.linalg.norm(v, axis=0) ** 2)
        r1 = np.sqrt(np.linalg.norm(Ku - v * aw[None, :], axis=0) ** 2 +
                     np.linalg.norm(Mv - u * wj[None, :], axis=0) ** 2)
        first[J] = r1 / (xn * np.maximum(1.0, aw))
    return first, pen
